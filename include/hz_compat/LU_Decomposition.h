// LU_Decomposition.h (drop-in) -- gpritam/MPI
// HelmholtzSolvers/DirectPseudoSpectral/include/LU_Decomposition.h:8-9.
// Inside SolvePDE's operator callback, LUDecomposition records the operator
// for the batched GPU solve; elsewhere it factors on the GPU (hz_lu_factor_batched)
// and returns the reference's storage, pivots and determinant.
#pragma once
#include "General.h"

void LUDecomposition(double **&A, int *&npivot, double &determinant, const int N);
void LUSolve(double **A, int *npivot, double *&b, const int N);
