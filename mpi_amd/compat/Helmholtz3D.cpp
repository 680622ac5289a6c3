// Helmholtz3D.cpp -- drop-in implementation of gpritam/MPI's DirectPseudoSpectral
// 3-D solver interfaces on top of the C ABI of hz.h, built twice:
//   HZC_PERIODIC = 0: include/hz_compat/NonPeriodic3D_MPI.h (Robin z faces),
//   HZC_PERIODIC = 1: include/hz_compat/Periodic3D_MPI.h (z periodic).
// Citations: HelmholtzSolvers/DirectPseudoSpectral/.
//
// SolvePDE keeps the reference's callback contract
// (NonPeriodic3D_MPI.cpp:248-253, Periodic3D_MPI.cpp:246-251): for every
// (|kx|,|ky|) mode it calls the user's ConstructOperator(iw, jw), which fills
// `Operator` and calls LUDecomposition.  While SolvePDE runs,
// LUDecomposition records the operator instead of factoring it (the
// diagonal alone when the off-diagonal part equals Dzz's interior, the whole
// matrix otherwise, compat_core.cpp); the recorded operators are then
// factored and solved on the GPU in one batched launch.
#if HZC_PERIODIC
#include "../../include/hz_compat/Periodic3D_MPI.h"
#else
#include "../../include/hz_compat/NonPeriodic3D_MPI.h"
#endif
#include "compat_core.h"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <vector>

int Rank = 0, size = 1;
int Nx, Ny, Nz, NY, N, Ns;
double xl, xr, yl, yr, zl, zr, Lx, Ly, Lz, dx, dy;
double Lambda;
#if HZC_PERIODIC
double **Dz, **Dzz, **Operator, *z, ****rhsPDE, determinant, **alphaPDE;
#else
double alphal, betal, alphar, betar;
double **Dz, **Dzz, **Operator, *z, ****rhsPDE, **leftBC, **rightBC, determinant;
#endif
int *npivot;
int SolverForm = HZ_FORM_PLAIN_LU;

namespace {

hzc::Solver g;

// Element arrays [Ns][Nx][Ny][N] <-> the caller's FieldVariable.
void gather(std::vector<double> &host, FieldVariable in[])
{
    host.resize((size_t)Ns * Nx * Ny * N);
    for (int kk = 0; kk < Ns; ++kk)
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                std::memcpy(&host[(((size_t)kk * Nx + i) * Ny + j) * N], in[kk].phi[i][j], sizeof(double) * N);
}

void scatter(FieldVariable out[], const std::vector<double> &host)
{
    for (int kk = 0; kk < Ns; ++kk)
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                std::memcpy(out[kk].phi[i][j], &host[(((size_t)kk * Nx + i) * Ny + j) * N], sizeof(double) * N);
}

void derivative_z(FieldVariable out[], FieldVariable in[], int order)
{
    std::vector<double> host;
    gather(host, in);
    g.derivative_z(host, order);
    scatter(out, host);
}

// x/y Fourier derivative of every element: the stack [Nx][Ny][Ns*N] (z
// lanes innermost), one GPU call.
void derivative_xy(FieldVariable out[], FieldVariable in[], int axis, int order)
{
    const size_t lanes = (size_t)Ns * N;
    std::vector<double> host((size_t)Nx * Ny * lanes);
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int kk = 0; kk < Ns; ++kk)
                std::memcpy(&host[((size_t)i * Ny + j) * lanes + (size_t)kk * N], in[kk].phi[i][j], sizeof(double) * N);
    g.derivative_xy(host, (int)lanes, axis, order);
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int kk = 0; kk < Ns; ++kk)
                std::memcpy(out[kk].phi[i][j], &host[((size_t)i * Ny + j) * lanes + (size_t)kk * N], sizeof(double) * N);
}

}  // namespace

void FirstDerivativeX(FieldVariable Ux[], FieldVariable U[]) { derivative_xy(Ux, U, 0, 1); }
void SecondDerivativeX(FieldVariable Uxx[], FieldVariable U[]) { derivative_xy(Uxx, U, 0, 2); }
void FirstDerivativeY(FieldVariable Uy[], FieldVariable U[]) { derivative_xy(Uy, U, 1, 1); }
void SecondDerivativeY(FieldVariable Uyy[], FieldVariable U[]) { derivative_xy(Uyy, U, 1, 2); }
void FirstDerivativeZ(FieldVariable Uz[], FieldVariable U[]) { derivative_z(Uz, U, 1); }
void SecondDerivativeZ(FieldVariable Uzz[], FieldVariable U[]) { derivative_z(Uzz, U, 2); }

// NonPeriodic3D_MPI.cpp:40-134, Periodic3D_MPI.cpp:36-130
void CreatePDEPlan(const int nx, const int ny, const int nz, const int n, const double xl_, const double xr_,
                   const double yl_, const double yr_, const double zl_, const double zr_)
{
    if (size < 1) size = 1;
    ::Nx = nx;
    ::Ny = ny;
    ::Nz = nz;
    ::N = n;
    ::xl = xl_;
    ::xr = xr_;
    ::yl = yl_;
    ::yr = yr_;
    ::zl = zl_;
    ::zr = zr_;
    NY = 2 * Ny;
    Lx = xr - xl;
    Ly = yr - yl;
    Lz = zr - zl;
    dx = Lx / Nx;
    dy = Ly / Ny;
    Ns = (int)std::ceil(((double)Nz) / (size * N));

    hz_plan_desc d{};
    d.nx = Nx;
    d.ny = Ny;
    d.n = N;
    d.ne = Ns;
    d.xl = xl;
    d.xr = xr;
    d.yl = yl;
    d.yr = yr;
    d.zl = zl;
    d.zr = zr;
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

    Allocate(z, Ns * (N - 1) + 1);
    Allocate(rhsPDE, Ns, 2 * Nx, NY, N);
#if HZC_PERIODIC
    Allocate(alphaPDE, Nx, NY);
#else
    Allocate(leftBC, 2 * Nx, NY);
    Allocate(rightBC, 2 * Nx, NY);
#endif
    Allocate(Dz, N, N);
    Allocate(Dzz, N, N);
    Allocate(Operator, N - 2, N - 2);
    Allocate(npivot, N - 2);
    hzc::hz_check(hz_plan_z_nodes(g.plan, z));
    hzc::hz_check(hz_plan_operators(g.plan, &Dz[0][0], &Dzz[0][0]));
}

// NonPeriodic3D_MPI.cpp:139-170, Periodic3D_MPI.cpp:135-166
void DestroyPDEPlan()
{
    Deallocate(z, Ns * (N - 1) + 1);
    Deallocate(rhsPDE, Ns, 2 * Nx, NY, N);
#if HZC_PERIODIC
    Deallocate(alphaPDE, Nx, NY);
#else
    Deallocate(leftBC, 2 * Nx, NY);
    Deallocate(rightBC, 2 * Nx, NY);
#endif
    Deallocate(Dz, N, N);
    Deallocate(Dzz, N, N);
    Deallocate(Operator, N - 2, N - 2);
    Deallocate(npivot, N - 2);
    g.destroy();
}

// NonPeriodic3D_MPI.cpp:184-232, Periodic3D_MPI.cpp:180-228: the particular
// solution of column (i, j) in every element from rhsPDE, with the operator
// the caller factored into Operator / npivot.
void ConstructFundamentalSolutions(FieldVariable U[], const int i, const int j)
{
    std::vector<double *> cols(Ns);
    std::vector<const double *> rhs(Ns);
    for (int kk = 0; kk < Ns; ++kk) {
        cols[kk] = U[kk].phi[i][j];
        rhs[kk] = rhsPDE[kk][i][j];
    }
    hzc::particular_solutions(cols.data(), rhs.data(), Ns, N, Operator, npivot);
}

// NonPeriodic3D_MPI.cpp:237-591, Periodic3D_MPI.cpp:233-864
void SolvePDE(FieldVariable U[], void (*ConstructOperator)(const int, const int))
{
    if (hzc::fft_plan_size() == -1) ErrorMessage("Create a FFT plan first!");  // Fourier_Quadrature.cpp:1103-1104
    if (Maximum(Nx, Ny) > hzc::fft_plan_size())
        ErrorMessage("Computation of DFT is not possible with the present FFT plan!");
    const int n = N - 2, M = Ny / 2 + 1;
    // ---- Step 2's callbacks, recorded ----
    hzc::Recorder &r = hzc::recorder();
    r.begin(g.modes, n, Operator, Dzz);
    for (int iw = 0; iw <= Nx / 2; iw++)
        for (int jw = 0; jw <= Ny / 2; jw++) {
            r.mode = iw * M + jw;
            ConstructOperator(iw, jw);
        }
    r.on = false;
    // ---- inputs in the plan's layout ----
    for (int kk = 0; kk < Ns; ++kk)
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                std::memcpy(&g.h_f[((size_t)i * Ny + j) * g.K + (size_t)kk * (N - 1)], rhsPDE[kk][i][j],
                            sizeof(double) * N);
#if HZC_PERIODIC
    g.solve_recorded(Lambda, 0.0, 1.0, 0.0, 1.0);
    {  // the closure coefficients (Periodic3D_MPI.cpp:698), [Nx][NY]
        std::vector<double> al;
        g.closure_full(al);
        for (int i = 0; i < Nx; ++i) std::memcpy(alphaPDE[i], &al[(size_t)i * NY], sizeof(double) * NY);
    }
#else
    const size_t ng = (size_t)Nx * Ny;
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j) {
            g.h_g[(size_t)i * Ny + j] = leftBC[i][j];
            g.h_g[ng + (size_t)i * Ny + j] = rightBC[i][j];
        }
    g.solve_recorded(Lambda, alphal, betal, alphar, betar);
#endif
    // ---- output in the caller's element arrays ----
    for (int kk = 0; kk < Ns; ++kk)
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                std::memcpy(U[kk].phi[i][j], &g.h_f[((size_t)i * Ny + j) * g.K + (size_t)kk * (N - 1)],
                            sizeof(double) * N);
}

// NonPeriodic3D_MPI.cpp:688-708, Periodic3D_MPI.cpp:961-981 (one process: no reduction)
void ComputeError(FieldVariable U[], double (*ExactFunction)(const double, const double, const double))
{
    double err = 0.0;
    for (int kk = 0; kk < Ns; ++kk)
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                for (int k = 0; k < N; ++k)
                    err = Maximum(err, Absolute(U[kk].phi[i][j][k] -
                                                ExactFunction(xl + i * dx, yl + j * dy, z[kk * (N - 1) + k])));
    if (Rank == 0) std::cout << "Error = " << err << std::endl << std::endl;
}

// NonPeriodic3D_MPI.cpp:713-741, Periodic3D_MPI.cpp:986-1014: Tecplot POINT file, one per rank.
void WriteFile(FieldVariable U[], double (*ExactFunction)(const double, const double, const double))
{
    char name[200];
    std::snprintf(name, sizeof(name), "Output/Solution-%04d.tec", Rank);
    std::ofstream out(name, std::ios::out);
    if (!out) ErrorMessage("Output file couldnot be opened!");
    out.flags(std::ios::dec | std::ios::fixed);
    out.precision(8);
    out << "TITLE = \"Helmholtz solver\"" << std::endl;
    out << "Variables = \"X\",\"Y\",\"Z\",\"U\",\"Exact\"" << std::endl;
    out << "Zone I = " << Nx << ", J = " << Ny << ", K = " << Ns * N << ", DATAPACKING=POINT" << std::endl;
    for (int kk = 0; kk < Ns; ++kk)
        for (int k = 0; k < N; ++k)
            for (int j = 0; j < Ny; ++j)
                for (int i = 0; i < Nx; ++i)
                    out << (xl + i * dx) << "\t" << (yl + j * dy) << "\t" << z[kk * (N - 1) + k] << "\t"
                        << U[kk].phi[i][j][k] << "\t" << ExactFunction(xl + i * dx, yl + j * dy, z[kk * (N - 1) + k])
                        << std::endl;
}
