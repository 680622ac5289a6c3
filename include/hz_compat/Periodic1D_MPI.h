// Periodic1D_MPI.h (drop-in) -- the interface of gpritam/MPI
// HelmholtzSolvers/DirectPseudoSpectral/include/Periodic1D_MPI.h:15-40
// (GLL elements in a periodic x), same names, globals and argument meaning,
// running on one GPU through the C ABI of hz.h (a plan with nx = ny = 1,
// periodic_z).  One process holds every element (Rank = 0, size = 1), so
// Ns = ceil(Nx/N) elements of N GLL points and x runs from 0 as in the
// reference (Periodic1D_MPI.cpp:70-87).
//
// Differences a caller can observe: SolvePDE leaves rhsPDE unchanged;
// `determinant` is not updated while SolvePDE records the operator;
// SolverForm selects the GPU method.  alphaPDE (the closure coefficient of
// the cyclic interface system, Periodic1D_MPI.cpp:418) is filled from the
// GPU solve, as the reference does with more than one element.
#pragma once

#include "../hz.h"
#include "General.h"
#include "LU_Decomposition.h"

struct FieldVariable {
    double *phi;
};

extern int Rank, size;
extern int Nx, N, Ns;
extern double xl, xr, Lambda, Lx;
extern double **Dx, **Dxx, **Operator, *x, **rhsPDE, determinant, alphaPDE;
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
