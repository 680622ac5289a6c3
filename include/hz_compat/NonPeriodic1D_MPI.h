// NonPeriodic1D_MPI.h (drop-in) -- the interface of gpritam/MPI
// HelmholtzSolvers/DirectPseudoSpectral/include/NonPeriodic1D_MPI.h:15-42
// (GLL elements in x, Robin x faces), same names, globals and argument
// meaning, running on one GPU through the C ABI of hz.h (a plan with
// nx = ny = 1).  One process holds every element (Rank = 0, size = 1), so
// Ns = ceil(Nx/N) elements of N GLL points and x runs from 0 as in the
// reference (NonPeriodic1D_MPI.cpp:74-91).
//
// Differences a caller can observe: SolvePDE leaves rhsPDE, leftBC and
// rightBC unchanged; `determinant` is not updated while SolvePDE records the
// operator; SolverForm selects the GPU method (HZ_FORM_PLAIN_LU by default,
// HZ_FORM_DIAG for the standard shifted operator).
#pragma once

#include "../hz.h"
#include "General.h"
#include "LU_Decomposition.h"

struct FieldVariable {
    double *phi;
};

extern int Rank, size;
extern int Nx, N, Ns;
extern double xl, xr, Lx;
extern double alphal, betal, alphar, betar, Lambda;
extern double **Dx, **Dxx, **Operator, *x, **rhsPDE, leftBC, rightBC, determinant;
extern int *npivot;
extern int SolverForm;

void CreatePDEPlan(const int Nx, const int N, const double xl, const double xr);
void DestroyPDEPlan();

// The particular solution of one column in every element, from rhsPDE with
// the operator the caller factored into Operator / npivot.
void ConstructFundamentalSolutions(FieldVariable U[]);
void SolvePDE(FieldVariable U[], void (*ConstructOperator)());

void FirstDerivativeX(FieldVariable Ux[], FieldVariable U[]);
void SecondDerivativeX(FieldVariable Uxx[], FieldVariable U[]);

void ComputeError(FieldVariable U[], double (*ExactFunction)(const double));
void WriteFile(FieldVariable U[], double (*ExactFunction)(const double));
