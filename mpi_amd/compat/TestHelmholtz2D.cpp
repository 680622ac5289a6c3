// TestHelmholtz2D.cpp -- a caller of the 2-D drop-ins written the way
// gpritam/MPI's own test programs use the reference
// (HelmholtzSolvers/DirectPseudoSpectral/TestNonPeriodic2D_MPI.cpp,
// TestPeriodic2D_MPI.cpp): u = sin(8 pi x) cos(6 pi y) on [0,1]^2, lambda = 1,
// Nx = 512 points in x, Ny = 512 requested y points of N = 10 per element;
// the non-periodic problem has Neumann data on the first y face and
// Dirichlet on the last.  Built against either drop-in (HZC_PERIODIC).
// Prints the reference's "Error = ..." line, then the SecondDerivativeX error
// of the solution against -(8 pi)^2 u.
#if HZC_PERIODIC
#include "../../include/hz_compat/Periodic2D_MPI.h"
#else
#include "../../include/hz_compat/NonPeriodic2D_MPI.h"
#endif

#include <iostream>

double ExactFunction(const double x, const double yy)
{
    return sin(8.0 * PI * (x - xl) / Lx) * cos(6.0 * PI * (yy - yl) / Ly);
}

double RHS(const double x, const double yy)
{
    return -(64.0 * PI * PI / (Lx * Lx) + 36.0 * PI * PI / (Ly * Ly) + Lambda) * sin(8.0 * PI * (x - xl) / Lx) *
           cos(6.0 * PI * (yy - yl) / Ly);
}

#if !HZC_PERIODIC
double LeftBC(const double x)
{
    const double yy = yl;
    return (alphal * 6.0 * PI * sin(6.0 * PI * (yy - yl) / Ly) / Ly + betal * cos(6.0 * PI * (yy - yl) / Ly)) *
           sin(8.0 * PI * (x - xl) / Lx);
}

double RightBC(const double x)
{
    const double yy = yr;
    return (-alphar * 6.0 * PI * sin(6.0 * PI * (yy - yl) / Ly) / Ly + betar * cos(6.0 * PI * (yy - yl) / Ly)) *
           sin(8.0 * PI * (x - xl) / Lx);
}
#endif

void ConstructOperator(const int iw)
{
    for (int i = 0; i < N - 2; ++i)
        for (int j = 0; j < N - 2; ++j)
            Operator[i][j] = Dyy[i + 1][j + 1] - (i == j ? (Lambda + 4.0 * PI * PI * iw * iw / (Lx * Lx)) : 0.0);
    LUDecomposition(Operator, npivot, determinant, N - 2);
}

int main()
{
    std::cout.flags(std::ios::dec);
    std::cout.precision(8);
#if !HZC_PERIODIC
    alphal = 1.0;
    betal = 0.0;
    alphar = 0.0;
    betar = 1.0;
#endif
    Lambda = 1.0;
    CreatePDEPlan(512, 512, 10, 0.0, 1.0, 0.0, 1.0);
    CreateFFTPlan(Nx);
    FieldVariable *V = new FieldVariable[Ns], *W = new FieldVariable[Ns];
    for (int jj = 0; jj < Ns; ++jj) {
        Allocate(V[jj].phi, NX, N);
        Allocate(W[jj].phi, NX, N);
    }
    for (int jj = 0; jj < Ns; ++jj)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < Nx; ++i) rhsPDE[jj][i][j] = RHS(xl + i * dx, y[jj * (N - 1) + j]);
#if !HZC_PERIODIC
    for (int i = 0; i < Nx; ++i) {
        leftBC[i] = LeftBC(xl + i * dx);
        rightBC[i] = RightBC(xl + i * dx);
    }
#endif
    SolvePDE(V, ConstructOperator);
    ComputeError(V, ExactFunction);
    SecondDerivativeX(W, V);
    double derr = 0.0;
    for (int jj = 0; jj < Ns; ++jj)
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < N; ++j)
                derr = Maximum(derr, Absolute(W[jj].phi[i][j] + 64.0 * PI * PI * V[jj].phi[i][j]));
    std::cout << "SecondDerivativeX error = " << derr << std::endl;
    for (int jj = 0; jj < Ns; ++jj) {
        Deallocate(V[jj].phi, NX, N);
        Deallocate(W[jj].phi, NX, N);
    }
    delete[] V;
    delete[] W;
    DestroyFFTPlan(Nx);
    DestroyPDEPlan();
    return 0;
}
