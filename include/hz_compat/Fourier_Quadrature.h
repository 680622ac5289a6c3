// Fourier_Quadrature.h (drop-in) -- the FFT plan calls of gpritam/MPI
// HelmholtzSolvers/DirectPseudoSpectral/include/Fourier_Quadrature.h:43-44
// that callers of SolvePDE make.  The transforms run inside the GPU solve;
// the plan only checks the size contract (N >= max(Nx, Ny)).
#pragma once
#include "General.h"

void CreateFFTPlan(const int N);
void DestroyFFTPlan(const int N);
