// TestHelmholtz1D.cpp -- a caller of the 1-D drop-ins written the way
// gpritam/MPI's own test programs use the reference
// (HelmholtzSolvers/DirectPseudoSpectral/TestNonPeriodic1D_MPI.cpp,
// TestPeriodic1D_MPI.cpp): Nx = 512 requested points of N = 8 per element on
// [0,1].  Non-periodic: u = sin(2 pi x), first with Neumann data on the left
// face and Dirichlet on the right at lambda = 1, then Robin (alpha = beta = 1)
// on the left at lambda = 5, as the reference's program does; periodic:
// u = sin(8 pi x) at lambda = 1.  Built against either drop-in
// (HZC_PERIODIC).  Prints the reference's "Error = ..." line per problem.
#if HZC_PERIODIC
#include "../../include/hz_compat/Periodic1D_MPI.h"
#else
#include "../../include/hz_compat/NonPeriodic1D_MPI.h"
#endif

#include <iostream>

#if HZC_PERIODIC
double ExactFunction(const double xx) { return sin(8.0 * PI * (xx - xl) / Lx); }
double RHS(const double xx) { return -(64.0 * PI * PI / (Lx * Lx) + Lambda) * sin(8.0 * PI * (xx - xl) / Lx); }
#else
double ExactFunction(const double xx) { return sin(2.0 * PI * (xx - xl) / Lx); }
double RHS(const double xx) { return -(4.0 * PI * PI / (Lx * Lx) + Lambda) * sin(2.0 * PI * (xx - xl) / Lx); }
double LeftBC()
{
    const double xx = xl;
    return -alphal * 2.0 * PI * cos(2.0 * PI * (xx - xl) / Lx) / Lx + betal * sin(2.0 * PI * (xx - xl) / Lx);
}
double RightBC()
{
    const double xx = xr;
    return alphar * 2.0 * PI * cos(2.0 * PI * (xx - xl) / Lx) / Lx + betar * sin(2.0 * PI * (xx - xl) / Lx);
}
#endif

void ConstructOperator()
{
    for (int i = 0; i < N - 2; ++i)
        for (int j = 0; j < N - 2; ++j) Operator[i][j] = Dxx[i + 1][j + 1] - (i == j ? Lambda : 0.0);
    LUDecomposition(Operator, npivot, determinant, N - 2);
}

static void solve_and_report(FieldVariable *V)
{
    for (int ii = 0; ii < Ns; ++ii)
        for (int i = 0; i < N; ++i) rhsPDE[ii][i] = RHS(x[ii * (N - 1) + i]);
#if !HZC_PERIODIC
    leftBC = LeftBC();
    rightBC = RightBC();
#endif
    SolvePDE(V, ConstructOperator);
    ComputeError(V, ExactFunction);
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
    CreatePDEPlan(512, 8, 0.0, 1.0);
    FieldVariable *V = new FieldVariable[Ns];
    for (int ii = 0; ii < Ns; ++ii) Allocate(V[ii].phi, N);
    solve_and_report(V);
#if !HZC_PERIODIC
    alphal = 1.0;
    betal = 1.0;
    alphar = 0.0;
    betar = 1.0;
    Lambda = 5.0;
    solve_and_report(V);
#endif
    for (int ii = 0; ii < Ns; ++ii) Deallocate(V[ii].phi, N);
    delete[] V;
    DestroyPDEPlan();
    return 0;
}
