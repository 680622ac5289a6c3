// NonPeriodic3D_MPI.h (drop-in) -- the interface of gpritam/MPI
// HelmholtzSolvers/DirectPseudoSpectral/include/NonPeriodic3D_MPI.h:15-47,
// same names, globals and argument meaning, running on one GPU through the
// C ABI of hz.h.  One process holds every element (Rank = 0, size = 1), so
// Ns = ceil(Nz/N) elements of N GLL points and z runs from 0 as in the
// reference (NonPeriodic3D_MPI.cpp:111-120).
//
// Differences a caller can observe:
//  - SolvePDE leaves rhsPDE, leftBC and rightBC unchanged (the reference
//    overwrites them with their spectra; callers refill them before every
//    solve either way);
//  - `determinant` is not updated while SolvePDE records operators;
//  - SolverForm selects the GPU method (HZ_FORM_PLAIN_LU, the reference
//    algorithm, by default; HZ_FORM_DIAG for the z-diagonalised form, used
//    when the callback builds the standard shifted operator to the bit; any
//    other operator runs the reference algorithm).  A PLAIN_LU plan that the
//    reference-algorithm kernel cannot hold (N - 2 > 128, or too many
//    elements) is created in the DIAG form, and its solves then need the
//    standard operator.
#pragma once

#include "../hz.h"
#include "General.h"
#include "LU_Decomposition.h"
#include "Fourier_Quadrature.h"

struct FieldVariable {
    double ***phi;
};

extern int Rank, size;
extern int Nx, Ny, Nz, NY, N, Ns;
extern double xl, xr, yl, yr, zl, zr, Lx, Ly, Lz, dx, dy;
extern double alphal, betal, alphar, betar, Lambda;
extern double **Dz, **Dzz, **Operator, *z, ****rhsPDE, **leftBC, **rightBC, determinant;
extern int *npivot;
extern int SolverForm;

void CreatePDEPlan(const int Nx, const int Ny, const int Nz, const int N, const double xl, const double xr,
                   const double yl, const double yr, const double zl, const double zr);
void DestroyPDEPlan();

// The particular solution of one column in every element, from rhsPDE with
// the operator the caller factored into Operator / npivot.
void ConstructFundamentalSolutions(FieldVariable U[], const int i, const int j);
void SolvePDE(FieldVariable U[], void (*ConstructOperator)(const int, const int));

// The reference leaves the X/Y versions empty (NonPeriodic3D_MPI.cpp:596-627);
// here they are the Fourier derivatives of Fourier_Quadrature.cpp:1494-1800
// applied to every element, on the GPU.
void FirstDerivativeX(FieldVariable Ux[], FieldVariable U[]);
void SecondDerivativeX(FieldVariable Uxx[], FieldVariable U[]);
void FirstDerivativeY(FieldVariable Uy[], FieldVariable U[]);
void SecondDerivativeY(FieldVariable Uyy[], FieldVariable U[]);
void FirstDerivativeZ(FieldVariable Uz[], FieldVariable U[]);
void SecondDerivativeZ(FieldVariable Uzz[], FieldVariable U[]);

void ComputeError(FieldVariable U[], double (*ExactFunction)(const double, const double, const double));
void WriteFile(FieldVariable U[], double (*ExactFunction)(const double, const double, const double));
