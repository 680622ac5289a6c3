// Helmholtz1D.cpp -- drop-in implementation of gpritam/MPI's DirectPseudoSpectral
// 1-D solver interfaces (GLL elements in x) on top of the C ABI of hz.h,
// built twice:
//   HZC_PERIODIC = 0: include/hz_compat/NonPeriodic1D_MPI.h (Robin x faces),
//   HZC_PERIODIC = 1: include/hz_compat/Periodic1D_MPI.h (x periodic).
// Citations: HelmholtzSolvers/DirectPseudoSpectral/.
//
// The GPU plan has nx = ny = 1 (no transform; its only mode is 0) and the
// reference's x as its GLL axis.  ConstructOperator() is called once as in
// the reference (NonPeriodic1D_MPI.cpp:199, Periodic1D_MPI.cpp:195) and
// recorded by LUDecomposition (compat_core.cpp).
#if HZC_PERIODIC
#include "../../include/hz_compat/Periodic1D_MPI.h"
#else
#include "../../include/hz_compat/NonPeriodic1D_MPI.h"
#endif
#include "compat_core.h"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <vector>

int Rank = 0, size = 1;
int Nx, N, Ns;
double xl, xr, Lx;
double Lambda;
#if HZC_PERIODIC
double **Dx, **Dxx, **Operator, *x, **rhsPDE, determinant, alphaPDE;
#else
double alphal, betal, alphar, betar;
double **Dx, **Dxx, **Operator, *x, **rhsPDE, leftBC, rightBC, determinant;
#endif
int *npivot;
int SolverForm = HZ_FORM_PLAIN_LU;

namespace {

hzc::Solver g;

void derivative(FieldVariable out[], FieldVariable in[], int order)
{
    std::vector<double> host((size_t)Ns * N);
    for (int ii = 0; ii < Ns; ++ii) std::memcpy(&host[(size_t)ii * N], in[ii].phi, sizeof(double) * N);
    g.derivative_z(host, order);
    for (int ii = 0; ii < Ns; ++ii) std::memcpy(out[ii].phi, &host[(size_t)ii * N], sizeof(double) * N);
}

}  // namespace

// NonPeriodic1D_MPI.cpp:360-399, Periodic1D_MPI.cpp:523-562
void FirstDerivativeX(FieldVariable Ux[], FieldVariable U[]) { derivative(Ux, U, 1); }
void SecondDerivativeX(FieldVariable Uxx[], FieldVariable U[]) { derivative(Uxx, U, 2); }

// NonPeriodic1D_MPI.cpp:40-105, Periodic1D_MPI.cpp:36-101
void CreatePDEPlan(const int nx, const int n, const double xl_, const double xr_)
{
    if (size < 1) size = 1;
    ::Nx = nx;
    ::N = n;
    ::xl = xl_;
    ::xr = xr_;
    Lx = xr - xl;
    Ns = (int)std::ceil(((double)Nx) / (size * N));

    hz_plan_desc d{};
    d.nx = 1;
    d.ny = 1;
    d.n = N;
    d.ne = Ns;
    d.xl = d.yl = 0.0;  // unit axes: any positive length, their only mode is 0
    d.xr = d.yr = 1.0;
    d.zl = xl;
    d.zr = xr;
    d.lambda = Lambda;
#if HZC_PERIODIC
    d.periodic_z = 1;
    d.beta_l = d.beta_r = 1.0;
#else
    d.alpha_l = alphal;
    d.beta_l = betal;
    d.alpha_r = alphar;
    d.beta_r = betar;
#endif
    d.form = SolverForm;
    g.create(d);

    Allocate(x, (N - 1) * Ns + 1);
    Allocate(rhsPDE, Ns, N);
    Allocate(Dx, N, N);
    Allocate(Dxx, N, N);
    Allocate(Operator, N - 2, N - 2);
    Allocate(npivot, N - 2);
    hzc::hz_check(hz_plan_z_nodes(g.plan, x));
    hzc::hz_check(hz_plan_operators(g.plan, &Dx[0][0], &Dxx[0][0]));
}

// NonPeriodic1D_MPI.cpp:110-130, Periodic1D_MPI.cpp:106-126
void DestroyPDEPlan()
{
    Deallocate(x, (N - 1) * Ns + 1);
    Deallocate(rhsPDE, Ns, N);
    Deallocate(Dx, N, N);
    Deallocate(Dxx, N, N);
    Deallocate(Operator, N - 2, N - 2);
    Deallocate(npivot, N - 2);
    g.destroy();
}

// NonPeriodic1D_MPI.cpp:144-190, Periodic1D_MPI.cpp:140-186: the particular
// solution in every element from rhsPDE, with the operator the caller
// factored into Operator / npivot.
void ConstructFundamentalSolutions(FieldVariable U[])
{
    std::vector<double *> cols(Ns);
    std::vector<const double *> rhs(Ns);
    for (int jj = 0; jj < Ns; ++jj) {
        cols[jj] = U[jj].phi;
        rhs[jj] = rhsPDE[jj];
    }
    hzc::particular_solutions(cols.data(), rhs.data(), Ns, N, Operator, npivot);
}

// NonPeriodic1D_MPI.cpp:195-355, Periodic1D_MPI.cpp:191-518
void SolvePDE(FieldVariable U[], void (*ConstructOperator)())
{
    hzc::Recorder &r = hzc::recorder();
    r.begin(g.modes, N - 2, Operator, Dxx);
    r.mode = 0;
    ConstructOperator();
    r.on = false;
    for (int ii = 0; ii < Ns; ++ii) std::memcpy(&g.h_f[(size_t)ii * (N - 1)], rhsPDE[ii], sizeof(double) * N);
#if HZC_PERIODIC
    g.solve_recorded(Lambda, 0.0, 1.0, 0.0, 1.0);
    {  // the closure coefficient (Periodic1D_MPI.cpp), one real column
        std::vector<double> al;
        g.closure_full(al);
        alphaPDE = al[0];
    }
#else
    g.h_g[0] = leftBC;
    g.h_g[1] = rightBC;
    g.solve_recorded(Lambda, alphal, betal, alphar, betar);
#endif
    for (int ii = 0; ii < Ns; ++ii) std::memcpy(U[ii].phi, &g.h_f[(size_t)ii * (N - 1)], sizeof(double) * N);
}

// NonPeriodic1D_MPI.cpp:404-422, Periodic1D_MPI.cpp:567-585 (one process: no reduction)
void ComputeError(FieldVariable U[], double (*ExactFunction)(const double))
{
    double err = 0.0;
    for (int ii = 0; ii < Ns; ++ii)
        for (int i = 0; i < N; ++i) err = Maximum(err, Absolute(U[ii].phi[i] - ExactFunction(x[ii * (N - 1) + i])));
    if (Rank == 0) std::cout << "Error = " << err << std::endl << std::endl;
}

// NonPeriodic1D_MPI.cpp:427-449, Periodic1D_MPI.cpp:590-613: one text file per rank.
void WriteFile(FieldVariable U[], double (*ExactFunction)(const double))
{
    char name[200];
    std::snprintf(name, sizeof(name), "Output/Solution-%04d.tec", Rank);
    std::ofstream out(name, std::ios::out);
    if (!out) ErrorMessage("Output file couldnot be opened!");
    out.flags(std::ios::dec | std::ios::fixed);
    out.precision(8);
    for (int ii = 0; ii < Ns; ++ii)
        for (int i = 0; i < N; ++i)
            out << x[ii * (N - 1) + i] << "\t" << U[ii].phi[i] << "\t" << ExactFunction(x[ii * (N - 1) + i])
                << std::endl;
}
