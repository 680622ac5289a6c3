// Periodic2D_MPI.h (drop-in) -- the interface of gpritam/MPI
// HelmholtzSolvers/DirectPseudoSpectral/include/Periodic2D_MPI.h:15-45
// (periodic x, GLL elements in a periodic y), same names, globals and
// argument meaning, running on one GPU through the C ABI of hz.h (a plan with
// nx = 1, ny = Nx, periodic_z).  One process holds every element (Rank = 0,
// size = 1), so Ns = ceil(Ny/N) elements of N GLL points and y runs from 0 as
// in the reference (Periodic2D_MPI.cpp:91-108).
//
// Differences a caller can observe: SolvePDE leaves rhsPDE unchanged;
// `determinant` is not updated while SolvePDE records operators; SolverForm
// selects the GPU method.  alphaPDE (the per-column closure coefficients of
// the cyclic interface system, Periodic2D_MPI.cpp:556) is filled from the
// GPU solve in the reference's layout, as the reference does with more
// than one element.
#pragma once

#include "../hz.h"
#include "General.h"
#include "LU_Decomposition.h"
#include "Fourier_Quadrature.h"

struct FieldVariable {
    double **phi;
};

extern int Rank, size;
extern int Nx, Ny, NX, N, Ns;
extern double xl, xr, yl, yr, Lx, Ly, dx, Lambda;
extern double **Dy, **Dyy, **Operator, *y, ***rhsPDE, determinant, *alphaPDE;
extern int *npivot;
extern int SolverForm;

void CreatePDEPlan(const int Nx, const int Ny, const int N, const double xl, const double xr, const double yl,
                   const double yr);
void DestroyPDEPlan();

// The particular solution of one column in every element, from rhsPDE with
// the operator the caller factored into Operator / npivot.
void ConstructFundamentalSolutions(FieldVariable U[], const int i);
void SolvePDE(FieldVariable U[], void (*ConstructOperator)(const int));

// The reference leaves the X versions empty (Periodic2D_MPI.cpp:696-709);
// here they are Fourier derivatives along x on the GPU.
void FirstDerivativeX(FieldVariable Ux[], FieldVariable U[]);
void SecondDerivativeX(FieldVariable Uxx[], FieldVariable U[]);
void FirstDerivativeY(FieldVariable Uy[], FieldVariable U[]);
void SecondDerivativeY(FieldVariable Uyy[], FieldVariable U[]);

void ComputeError(FieldVariable U[], double (*ExactFunction)(const double, const double));
void WriteFile(FieldVariable U[], double (*ExactFunction)(const double, const double));
