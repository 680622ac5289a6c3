// TestNonPeriodic3D.cpp -- a caller of the drop-in interface written the way
// gpritam/MPI's own test program uses the reference
// (HelmholtzSolvers/DirectPseudoSpectral/TestNonPeriodic3D_MPI.cpp): the same
// manufactured problem u = sin(8 pi x) cos(6 pi y) cos(4 pi z) on [0,1]^3,
// Neumann data on the first z face, Dirichlet on the last, lambda = 1, grid
// 64 x 32 with Nz = 250 requested z points of N = 10 per element.
//
//   TestNonPeriodic3D [callback]   callback = shift (default) | full | singular | diag | large
// "full" changes an off-diagonal entry of the operator of mode (1, 2), whose
// right-hand side is zero for this problem, so SolvePDE must record whole
// matrices while the solution is unchanged; "singular" zeroes a column of
// the (0, 0) operator, which the reference reports as singular; "diag" sets
// SolverForm = HZ_FORM_DIAG before CreatePDEPlan (the standard operator, so
// the z-diagonalised solve runs); "large" keeps the default SolverForm with
// one element of N = 258, more than the reference-algorithm kernel holds
// (N - 2 <= 128), so CreatePDEPlan makes a z-diagonalised plan and the
// standard operator runs on it.
#include "../../include/hz_compat/NonPeriodic3D_MPI.h"

#include <chrono>
#include <cstring>
#include <iostream>

static int g_mode = 0;

double ExactFunction(const double x, const double y, const double zz)
{
    return sin(8.0 * PI * (x - xl) / Lx) * cos(6.0 * PI * (y - yl) / Ly) * cos(4.0 * PI * (zz - zl) / Lz);
}

double LeftBC(const double x, const double y)
{
    const double zz = zl;
    return (alphal * 4.0 * PI * sin(4.0 * PI * (zz - zl) / Lz) / Lz + betal * cos(4.0 * PI * (zz - zl) / Lz)) *
           sin(8.0 * PI * (x - xl) / Lx) * cos(6.0 * PI * (y - yl) / Ly);
}

double RightBC(const double x, const double y)
{
    const double zz = zr;
    return (-alphar * 4.0 * PI * sin(4.0 * PI * (zz - zl) / Lz) / Lz + betar * cos(4.0 * PI * (zz - zl) / Lz)) *
           sin(8.0 * PI * (x - xl) / Lx) * cos(6.0 * PI * (y - yl) / Ly);
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
    if (g_mode == 1 && iw == 1 && jw == 2) Operator[0][1] *= 1.5;
    if (g_mode == 2 && iw == 0 && jw == 0)
        for (int i = 0; i < N - 2; ++i) Operator[i][0] = 0.0;
    LUDecomposition(Operator, npivot, determinant, N - 2);
}

int main(int argc, char **argv)
{
    if (argc > 1 && !std::strcmp(argv[1], "full")) g_mode = 1;
    if (argc > 1 && !std::strcmp(argv[1], "singular")) g_mode = 2;
    if (argc > 1 && !std::strcmp(argv[1], "diag")) SolverForm = HZ_FORM_DIAG;
    const bool large = argc > 1 && !std::strcmp(argv[1], "large");
    std::cout.flags(std::ios::dec);
    std::cout.precision(8);

    alphal = 1.0;
    betal = 0.0;
    alphar = 0.0;
    betar = 1.0;
    Lambda = 1.0;
    if (large)
        CreatePDEPlan(64, 32, 258, 258, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0);
    else
        CreatePDEPlan(64, 32, 250, 10, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0);
    CreateFFTPlan(Maximum(Nx, Ny));

    FieldVariable *V = new FieldVariable[Ns];
    for (int kk = 0; kk < Ns; ++kk) Allocate(V[kk].phi, 2 * Nx, NY, N);

    for (int rep = 0; rep < 2; ++rep) {
        const auto start = std::chrono::high_resolution_clock::now();
        for (int kk = 0; kk < Ns; ++kk)
            for (int i = 0; i < Nx; ++i)
                for (int j = 0; j < Ny; ++j)
                    for (int k = 0; k < N; ++k) rhsPDE[kk][i][j][k] = RHS(xl + i * dx, yl + j * dy, z[kk * (N - 1) + k]);
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j) {
                leftBC[i][j] = LeftBC(xl + i * dx, yl + j * dy);
                rightBC[i][j] = RightBC(xl + i * dx, yl + j * dy);
            }
        SolvePDE(V, ConstructOperator);
        const auto stop = std::chrono::high_resolution_clock::now();
        std::cout << std::endl
                  << "Total time elapsed : "
                  << std::chrono::duration_cast<std::chrono::milliseconds>(stop - start).count() << " milliseconds."
                  << std::endl
                  << std::endl;
    }
    ComputeError(V, ExactFunction);

    // d2u/dz2 of the exact solution is -(4 pi)^2 u: check SecondDerivativeZ too
    FieldVariable *W = new FieldVariable[Ns];
    for (int kk = 0; kk < Ns; ++kk) Allocate(W[kk].phi, 2 * Nx, NY, N);
    SecondDerivativeZ(W, V);
    double derr = 0.0;
    for (int kk = 0; kk < Ns; ++kk)
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                for (int k = 0; k < N; ++k)
                    derr = Maximum(derr, Absolute(W[kk].phi[i][j][k] +
                                                  16.0 * PI * PI * ExactFunction(xl + i * dx, yl + j * dy, z[kk * (N - 1) + k])));
    std::cout << "SecondDerivativeZ error = " << derr << std::endl;
    // and along x: the exact solution's x dependence is sin(8 pi x)
    SecondDerivativeX(W, V);
    derr = 0.0;
    for (int kk = 0; kk < Ns; ++kk)
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                for (int k = 0; k < N; ++k)
                    derr = Maximum(derr, Absolute(W[kk].phi[i][j][k] +
                                                  64.0 * PI * PI * ExactFunction(xl + i * dx, yl + j * dy, z[kk * (N - 1) + k])));
    std::cout << "SecondDerivativeX error = " << derr << std::endl;

    for (int kk = 0; kk < Ns; ++kk) {
        Deallocate(V[kk].phi, 2 * Nx, NY, N);
        Deallocate(W[kk].phi, 2 * Nx, NY, N);
    }
    delete[] V;
    delete[] W;
    DestroyFFTPlan(Maximum(Nx, Ny));
    DestroyPDEPlan();
    return 0;
}
