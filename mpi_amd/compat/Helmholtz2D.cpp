// Helmholtz2D.cpp -- drop-in implementation of gpritam/MPI's DirectPseudoSpectral
// 2-D solver interfaces (periodic x, GLL elements in y) on top of the C ABI
// of hz.h, built twice:
//   HZC_PERIODIC = 0: include/hz_compat/NonPeriodic2D_MPI.h (Robin y faces),
//   HZC_PERIODIC = 1: include/hz_compat/Periodic2D_MPI.h (y periodic).
// Citations: HelmholtzSolvers/DirectPseudoSpectral/.
//
// The GPU plan has nx = 1 and ny = Nx: the reference's periodic x is the
// plan's half-spectrum axis (Nx/2 + 1 mode columns), its y the plan's GLL
// axis.  ConstructOperator(iw) is called for iw = 0..Nx/2 as in the
// reference (NonPeriodic2D_MPI.cpp:240-242, Periodic2D_MPI.cpp:233-235) and
// recorded by LUDecomposition (compat_core.cpp).
#if HZC_PERIODIC
#include "../../include/hz_compat/Periodic2D_MPI.h"
#else
#include "../../include/hz_compat/NonPeriodic2D_MPI.h"
#endif
#include "compat_core.h"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <vector>

int Rank = 0, size = 1;
int Nx, Ny, NX, N, Ns;
double xl, xr, yl, yr, Lx, Ly, dx;
double Lambda;
#if HZC_PERIODIC
double **Dy, **Dyy, **Operator, *y, ***rhsPDE, determinant, *alphaPDE;
#else
double alphal, betal, alphar, betar;
double **Dy, **Dyy, **Operator, *y, ***rhsPDE, *leftBC, *rightBC, determinant;
#endif
int *npivot;
int SolverForm = HZ_FORM_PLAIN_LU;

namespace {

hzc::Solver g;

// Element arrays [Ns][1][Nx][N] <-> the caller's FieldVariable.
void gather(std::vector<double> &host, FieldVariable in[])
{
    host.resize((size_t)Ns * Nx * N);
    for (int jj = 0; jj < Ns; ++jj)
        for (int i = 0; i < Nx; ++i) std::memcpy(&host[((size_t)jj * Nx + i) * N], in[jj].phi[i], sizeof(double) * N);
}

void scatter(FieldVariable out[], const std::vector<double> &host)
{
    for (int jj = 0; jj < Ns; ++jj)
        for (int i = 0; i < Nx; ++i) std::memcpy(out[jj].phi[i], &host[((size_t)jj * Nx + i) * N], sizeof(double) * N);
}

void derivative_y(FieldVariable out[], FieldVariable in[], int order)
{
    std::vector<double> host;
    gather(host, in);
    g.derivative_z(host, order);
    scatter(out, host);
}

// Fourier derivative along the periodic x: the stack [1][Nx][Ns*N].
void derivative_x(FieldVariable out[], FieldVariable in[], int order)
{
    const size_t lanes = (size_t)Ns * N;
    std::vector<double> host((size_t)Nx * lanes);
    for (int i = 0; i < Nx; ++i)
        for (int jj = 0; jj < Ns; ++jj)
            std::memcpy(&host[(size_t)i * lanes + (size_t)jj * N], in[jj].phi[i], sizeof(double) * N);
    g.derivative_xy(host, (int)lanes, 1, order);
    for (int i = 0; i < Nx; ++i)
        for (int jj = 0; jj < Ns; ++jj)
            std::memcpy(out[jj].phi[i], &host[(size_t)i * lanes + (size_t)jj * N], sizeof(double) * N);
}

}  // namespace

// The reference leaves the X versions empty (NonPeriodic2D_MPI.cpp:479-492,
// Periodic2D_MPI.cpp:696-709); here they are Fourier derivatives on the GPU.
void FirstDerivativeX(FieldVariable Ux[], FieldVariable U[]) { derivative_x(Ux, U, 1); }
void SecondDerivativeX(FieldVariable Uxx[], FieldVariable U[]) { derivative_x(Uxx, U, 2); }
// NonPeriodic2D_MPI.cpp:497-542, Periodic2D_MPI.cpp:714-759
void FirstDerivativeY(FieldVariable Uy[], FieldVariable U[]) { derivative_y(Uy, U, 1); }
void SecondDerivativeY(FieldVariable Uyy[], FieldVariable U[]) { derivative_y(Uyy, U, 2); }

// NonPeriodic2D_MPI.cpp:40-126, Periodic2D_MPI.cpp:36-122
void CreatePDEPlan(const int nx, const int ny, const int n, const double xl_, const double xr_, const double yl_,
                   const double yr_)
{
    if (size < 1) size = 1;
    ::Nx = nx;
    ::Ny = ny;
    ::N = n;
    ::xl = xl_;
    ::xr = xr_;
    ::yl = yl_;
    ::yr = yr_;
    NX = 2 * Nx;
    Lx = xr - xl;
    Ly = yr - yl;
    dx = Lx / Nx;
    Ns = (int)std::ceil(((double)Ny) / (size * N));

    hz_plan_desc d{};
    d.nx = 1;
    d.ny = Nx;
    d.n = N;
    d.ne = Ns;
    d.xl = 0.0;  // the unit axis: any positive length, its only mode is 0
    d.xr = 1.0;
    d.yl = xl;
    d.yr = xr;
    d.zl = yl;
    d.zr = yr;
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

    Allocate(y, Ns * (N - 1) + 1);
    Allocate(rhsPDE, Ns, NX, N);
#if HZC_PERIODIC
    Allocate(alphaPDE, NX);
#else
    Allocate(leftBC, NX);
    Allocate(rightBC, NX);
#endif
    Allocate(Dy, N, N);
    Allocate(Dyy, N, N);
    Allocate(Operator, N - 2, N - 2);
    Allocate(npivot, N - 2);
    hzc::hz_check(hz_plan_z_nodes(g.plan, y));
    hzc::hz_check(hz_plan_operators(g.plan, &Dy[0][0], &Dyy[0][0]));
}

// NonPeriodic2D_MPI.cpp:131-162, Periodic2D_MPI.cpp:127-158
void DestroyPDEPlan()
{
    Deallocate(y, Ns * (N - 1) + 1);
    Deallocate(rhsPDE, Ns, NX, N);
#if HZC_PERIODIC
    Deallocate(alphaPDE, NX);
#else
    Deallocate(leftBC, NX);
    Deallocate(rightBC, NX);
#endif
    Deallocate(Dy, N, N);
    Deallocate(Dyy, N, N);
    Deallocate(Operator, N - 2, N - 2);
    Deallocate(npivot, N - 2);
    g.destroy();
}

// NonPeriodic2D_MPI.cpp:176-223, Periodic2D_MPI.cpp:172-219: the particular
// solution of column i in every element from rhsPDE, with the operator the
// caller factored into Operator / npivot.
void ConstructFundamentalSolutions(FieldVariable U[], const int i)
{
    std::vector<double *> cols(Ns);
    std::vector<const double *> rhs(Ns);
    for (int jj = 0; jj < Ns; ++jj) {
        cols[jj] = U[jj].phi[i];
        rhs[jj] = rhsPDE[jj][i];
    }
    hzc::particular_solutions(cols.data(), rhs.data(), Ns, N, Operator, npivot);
}

// NonPeriodic2D_MPI.cpp:228-474, Periodic2D_MPI.cpp:224-691
void SolvePDE(FieldVariable U[], void (*ConstructOperator)(const int))
{
    if (hzc::fft_plan_size() == -1) ErrorMessage("Create a FFT plan first!");
    if (Nx > hzc::fft_plan_size()) ErrorMessage("Computation of DFT is not possible with the present FFT plan!");
    hzc::Recorder &r = hzc::recorder();
    r.begin(g.modes, N - 2, Operator, Dyy);
    for (int iw = 0; iw <= Nx / 2; iw++) {
        r.mode = iw;
        ConstructOperator(iw);
    }
    r.on = false;
    for (int jj = 0; jj < Ns; ++jj)
        for (int i = 0; i < Nx; ++i)
            std::memcpy(&g.h_f[(size_t)i * g.K + (size_t)jj * (N - 1)], rhsPDE[jj][i], sizeof(double) * N);
#if HZC_PERIODIC
    g.solve_recorded(Lambda, 0.0, 1.0, 0.0, 1.0);
    {  // the closure coefficients (Periodic2D_MPI.cpp), [NX]
        std::vector<double> al;
        g.closure_full(al);
        std::memcpy(alphaPDE, al.data(), sizeof(double) * NX);
    }
#else
    for (int i = 0; i < Nx; ++i) {
        g.h_g[i] = leftBC[i];
        g.h_g[Nx + i] = rightBC[i];
    }
    g.solve_recorded(Lambda, alphal, betal, alphar, betar);
#endif
    for (int jj = 0; jj < Ns; ++jj)
        for (int i = 0; i < Nx; ++i)
            std::memcpy(U[jj].phi[i], &g.h_f[(size_t)i * g.K + (size_t)jj * (N - 1)], sizeof(double) * N);
}

// NonPeriodic2D_MPI.cpp:547-566, Periodic2D_MPI.cpp:764-783 (one process: no reduction)
void ComputeError(FieldVariable U[], double (*ExactFunction)(const double, const double))
{
    double err = 0.0;
    for (int jj = 0; jj < Ns; ++jj)
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < N; ++j)
                err = Maximum(err, Absolute(U[jj].phi[i][j] - ExactFunction(xl + i * dx, y[jj * (N - 1) + j])));
    if (Rank == 0) std::cout << "Error = " << err << std::endl << std::endl;
}

// NonPeriodic2D_MPI.cpp:571-598, Periodic2D_MPI.cpp:788-815: Tecplot POINT file, one per rank.
void WriteFile(FieldVariable U[], double (*ExactFunction)(const double, const double))
{
    char name[200];
    std::snprintf(name, sizeof(name), "Output/Solution-%04d.tec", Rank);
    std::ofstream out(name, std::ios::out);
    if (!out) ErrorMessage("Output file couldnot be opened!");
    out.flags(std::ios::dec | std::ios::fixed);
    out.precision(8);
    out << "TITLE = \"Helmholtz solver\"" << std::endl;
    out << "Variables = \"X\",\"Y\",\"U\",\"Exact\"" << std::endl;
    out << "Zone I = " << Nx << ", J = " << Ns * N << ", DATAPACKING=POINT" << std::endl;
    for (int jj = 0; jj < Ns; ++jj)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < Nx; ++i)
                out << (xl + i * dx) << "\t" << y[jj * (N - 1) + j] << "\t" << U[jj].phi[i][j] << "\t"
                    << ExactFunction(xl + i * dx, y[jj * (N - 1) + j]) << std::endl;
}
