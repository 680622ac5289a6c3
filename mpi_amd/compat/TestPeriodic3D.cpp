// TestPeriodic3D.cpp -- a caller of the periodic-z drop-in written the way
// gpritam/MPI's own test program uses the reference
// (HelmholtzSolvers/DirectPseudoSpectral/TestPeriodic3D_MPI.cpp): the
// manufactured problem u = sin(8 pi x) cos(6 pi y) cos(4 pi z) on [0,1]^3
// with z periodic, lambda = 1, grid 64 x 32 with Nz = 250 requested z points
// of N = 10 per element.  Prints the reference's "Error = ..." line.
#include "../../include/hz_compat/Periodic3D_MPI.h"

#include <iostream>

double ExactFunction(const double x, const double y, const double zz)
{
    return sin(8.0 * PI * (x - xl) / Lx) * cos(6.0 * PI * (y - yl) / Ly) * cos(4.0 * PI * (zz - zl) / Lz);
}

double RHS(const double x, const double y, const double zz)
{
    return -(64.0 * PI * PI / (Lx * Lx) + 36.0 * PI * PI / (Ly * Ly) + 16.0 * PI * PI / (Lz * Lz) + Lambda) *
           sin(8.0 * PI * (x - xl) / Lx) * cos(6.0 * PI * (y - yl) / Ly) * cos(4.0 * PI * (zz - zl) / Lz);
}

void ConstructOperator(const int iw, const int jw)
{
    const double shift = Lambda + 4.0 * PI * PI * iw * iw / (Lx * Lx) + 4.0 * PI * PI * jw * jw / (Ly * Ly);
    for (int i = 0; i < N - 2; ++i)
        for (int j = 0; j < N - 2; ++j) Operator[i][j] = Dzz[i + 1][j + 1] - (i == j ? shift : 0.0);
    LUDecomposition(Operator, npivot, determinant, N - 2);
}

int main()
{
    std::cout.flags(std::ios::dec);
    std::cout.precision(8);
    Lambda = 1.0;
    CreatePDEPlan(64, 32, 250, 10, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0);
    CreateFFTPlan(Maximum(Nx, Ny));
    FieldVariable *V = new FieldVariable[Ns];
    for (int kk = 0; kk < Ns; ++kk) Allocate(V[kk].phi, 2 * Nx, NY, N);
    for (int kk = 0; kk < Ns; ++kk)
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                for (int k = 0; k < N; ++k) rhsPDE[kk][i][j][k] = RHS(xl + i * dx, yl + j * dy, z[kk * (N - 1) + k]);
    SolvePDE(V, ConstructOperator);
    ComputeError(V, ExactFunction);
    for (int kk = 0; kk < Ns; ++kk) Deallocate(V[kk].phi, 2 * Nx, NY, N);
    delete[] V;
    DestroyFFTPlan(Maximum(Nx, Ny));
    DestroyPDEPlan();
    return 0;
}
