// NonPeriodic3D.cpp -- drop-in implementation of gpritam/MPI's
// DirectPseudoSpectral 3-D solver interface (include/hz_compat/*.h) on top of
// the C ABI of hz.h.  Citations: HelmholtzSolvers/DirectPseudoSpectral/.
//
// SolvePDE keeps the reference's callback contract
// (NonPeriodic3D_MPI.cpp:248-253): for every (|kx|,|ky|) mode it calls the
// user's ConstructOperator(iw, jw), which fills `Operator` and calls
// LUDecomposition.  While SolvePDE runs, LUDecomposition records the
// operator instead of factoring it: the diagonal alone when the
// off-diagonal part equals Dzz's interior (the usual shifted operator), the
// whole matrix otherwise.  The recorded operators are then factored and
// solved on the GPU in one batched launch.
#include "../../include/hz_compat/NonPeriodic3D_MPI.h"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <vector>

int Rank = 0, size = 1;
int Nx, Ny, Nz, NY, N, Ns;
double xl, xr, yl, yr, zl, zr, Lx, Ly, Lz, dx, dy;
double alphal, betal, alphar, betar, Lambda;
double **Dz, **Dzz, **Operator, *z, ****rhsPDE, **leftBC, **rightBC, determinant;
int *npivot;
int SolverForm = HZ_FORM_PLAIN_LU;

namespace {

hz_plan *g_plan = nullptr;
int g_fft_n = -1;
int g_K = 0, g_modes = 0, g_M = 0;
// recording state of the operator callback
bool g_recording = false;
int g_cur_mode = -1;
bool g_full = false;
std::vector<double> g_diag, g_full_ops;
// device and pinned staging buffers
double *d_f = nullptr, *d_gl = nullptr, *d_gr = nullptr, *d_u = nullptr, *d_ops = nullptr;
double *h_f = nullptr, *h_g = nullptr;

void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        std::cout << what << ": " << hipGetErrorString(e) << std::endl;
        std::exit(1);
    }
}

void hz_check(int code)
{
    if (code != HZ_OK) ErrorMessage(hz_last_error_string());
}

void free_device()
{
    for (double *p : {d_f, d_gl, d_gr, d_u, d_ops}) (void)hipFree(p);
    (void)hipHostFree(h_f);
    (void)hipHostFree(h_g);
    d_f = d_gl = d_gr = d_u = d_ops = h_f = h_g = nullptr;
}

// Element-layout staging for FirstDerivativeZ / SecondDerivativeZ.
void derivative(FieldVariable out[], FieldVariable in[], int order)
{
    const size_t col = (size_t)N, cols = (size_t)Ns * Nx * Ny;
    std::vector<double> host(cols * col);
    for (int kk = 0; kk < Ns; ++kk)
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                std::memcpy(&host[(((size_t)kk * Nx + i) * Ny + j) * col], in[kk].phi[i][j], sizeof(double) * N);
    double *d_in = nullptr, *d_out = nullptr;
    hip_check(hipMalloc(&d_in, sizeof(double) * host.size()), "hipMalloc");
    hip_check(hipMalloc(&d_out, sizeof(double) * host.size()), "hipMalloc");
    hip_check(hipMemcpy(d_in, host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice), "hipMemcpy");
    hz_check(hz_derivative_z(g_plan, order, d_in, d_out, nullptr));
    hip_check(hipMemcpy(host.data(), d_out, sizeof(double) * host.size(), hipMemcpyDeviceToHost), "hipMemcpy");
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    for (int kk = 0; kk < Ns; ++kk)
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                std::memcpy(out[kk].phi[i][j], &host[(((size_t)kk * Nx + i) * Ny + j) * col], sizeof(double) * N);
}

// x/y Fourier derivative of every element: gather to [Nx][Ny][Ns*N] (z lanes
// innermost), one hz_derivative_xy call, scatter back.
void derivative_xy(FieldVariable out[], FieldVariable in[], int axis, int order)
{
    const size_t lanes = (size_t)Ns * N;
    std::vector<double> host((size_t)Nx * Ny * lanes);
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int kk = 0; kk < Ns; ++kk)
                std::memcpy(&host[((size_t)i * Ny + j) * lanes + (size_t)kk * N], in[kk].phi[i][j], sizeof(double) * N);
    double *d_in = nullptr, *d_out = nullptr;
    hip_check(hipMalloc(&d_in, sizeof(double) * host.size()), "hipMalloc");
    hip_check(hipMalloc(&d_out, sizeof(double) * host.size()), "hipMalloc");
    hip_check(hipMemcpy(d_in, host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice), "hipMemcpy");
    hz_check(hz_derivative_xy(g_plan, axis, order, d_in, d_out, (int)lanes, nullptr));
    hip_check(hipMemcpy(host.data(), d_out, sizeof(double) * host.size(), hipMemcpyDeviceToHost), "hipMemcpy");
    (void)hipFree(d_in);
    (void)hipFree(d_out);
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

// General.cpp:6-11
void ErrorMessage(const char *message)
{
    std::cout << message << std::endl;
    std::exit(1);
}

// Fourier_Quadrature.cpp:16-69: the plan is grow-only; solves check the size.
void CreateFFTPlan(const int n)
{
    if (g_fft_n < 0 || n > g_fft_n) g_fft_n = n;
}

void DestroyFFTPlan(const int)
{
    if (g_fft_n < 0) ErrorMessage("At first, create a FFT plan!");
    g_fft_n = -1;
}

// NonPeriodic3D_MPI.cpp:40-134
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
    d.alpha_l = alphal;
    d.beta_l = betal;
    d.alpha_r = alphar;
    d.beta_r = betar;
    d.form = SolverForm;
    hz_check(hz_plan_create(&d, &g_plan));
    hz_check(hz_plan_nz_total(g_plan, &g_K));
    hz_check(hz_plan_num_modes(g_plan, &g_modes));
    g_M = Ny / 2 + 1;

    Allocate(z, Ns * (N - 1) + 1);
    Allocate(rhsPDE, Ns, 2 * Nx, NY, N);
    Allocate(leftBC, 2 * Nx, NY);
    Allocate(rightBC, 2 * Nx, NY);
    Allocate(Dz, N, N);
    Allocate(Dzz, N, N);
    Allocate(Operator, N - 2, N - 2);
    Allocate(npivot, N - 2);
    hz_check(hz_plan_z_nodes(g_plan, z));
    hz_check(hz_plan_operators(g_plan, &Dz[0][0], &Dzz[0][0]));

    const size_t nf = (size_t)Nx * Ny * g_K, ng = (size_t)Nx * Ny;
    hip_check(hipMalloc(&d_f, sizeof(double) * nf), "hipMalloc");
    hip_check(hipMalloc(&d_u, sizeof(double) * nf), "hipMalloc");
    hip_check(hipMalloc(&d_gl, sizeof(double) * ng), "hipMalloc");
    hip_check(hipMalloc(&d_gr, sizeof(double) * ng), "hipMalloc");
    hip_check(hipHostMalloc(&h_f, sizeof(double) * nf), "hipHostMalloc");
    hip_check(hipHostMalloc(&h_g, sizeof(double) * 2 * ng), "hipHostMalloc");
}

// NonPeriodic3D_MPI.cpp:139-170
void DestroyPDEPlan()
{
    Deallocate(z, Ns * (N - 1) + 1);
    Deallocate(rhsPDE, Ns, 2 * Nx, NY, N);
    Deallocate(leftBC, 2 * Nx, NY);
    Deallocate(rightBC, 2 * Nx, NY);
    Deallocate(Dz, N, N);
    Deallocate(Dzz, N, N);
    Deallocate(Operator, N - 2, N - 2);
    Deallocate(npivot, N - 2);
    free_device();
    hz_plan_destroy(g_plan);
    g_plan = nullptr;
}

// LU_Decomposition.cpp:7-72 (GPU), or recording inside SolvePDE.
void LUDecomposition(double **&A, int *&piv, double &det, const int n)
{
    if (g_recording && g_plan && n == N - 2 && A == Operator) {
        const size_t m = (size_t)g_cur_mode;
        for (int i = 0; i < n; ++i) g_diag[m * n + i] = A[i][i];
        bool standard = true;
        for (int i = 0; i < n && standard; ++i)
            for (int j = 0; j < n; ++j)
                if (i != j && A[i][j] != Dzz[i + 1][j + 1]) {
                    standard = false;
                    break;
                }
        if (!standard && !g_full) {
            // switch to whole matrices: rebuild the modes recorded so far
            g_full = true;
            g_full_ops.assign((size_t)g_modes * n * n, 0.0);
            for (size_t q = 0; q < m; ++q)
                for (int i = 0; i < n; ++i)
                    for (int j = 0; j < n; ++j)
                        g_full_ops[(q * n + i) * n + j] = (i == j) ? g_diag[q * n + i] : Dzz[i + 1][j + 1];
        }
        if (g_full)
            for (int i = 0; i < n; ++i) std::memcpy(&g_full_ops[(m * n + i) * n], A[i], sizeof(double) * n);
        return;
    }
    double *dA = nullptr, *dDet = nullptr;
    int *dPiv = nullptr, *dSt = nullptr;
    hip_check(hipMalloc(&dA, sizeof(double) * n * n), "hipMalloc");
    hip_check(hipMalloc(&dDet, sizeof(double)), "hipMalloc");
    hip_check(hipMalloc(&dPiv, sizeof(int) * (n > 1 ? n - 1 : 1)), "hipMalloc");
    hip_check(hipMalloc(&dSt, sizeof(int)), "hipMalloc");
    for (int i = 0; i < n; ++i) hip_check(hipMemcpy(dA + (size_t)i * n, A[i], sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
    hz_check(hz_lu_factor_batched(dA, n, 1, dPiv, dDet, dSt, nullptr));
    int st = 0;
    hip_check(hipMemcpy(&st, dSt, sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy");
    for (int i = 0; i < n; ++i) hip_check(hipMemcpy(A[i], dA + (size_t)i * n, sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy");
    if (n > 1) hip_check(hipMemcpy(piv, dPiv, sizeof(int) * (n - 1), hipMemcpyDeviceToHost), "hipMemcpy");
    hip_check(hipMemcpy(&det, dDet, sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
    for (void *p : {(void *)dA, (void *)dDet, (void *)dPiv, (void *)dSt}) (void)hipFree(p);
    if (st) ErrorMessage("ERROR: Matrix is singular!");
}

// LU_Decomposition.cpp:78-108 (GPU)
void LUSolve(double **A, int *piv, double *&b, const int n)
{
    double *dA = nullptr, *db = nullptr;
    int *dPiv = nullptr;
    hip_check(hipMalloc(&dA, sizeof(double) * n * n), "hipMalloc");
    hip_check(hipMalloc(&db, sizeof(double) * n), "hipMalloc");
    hip_check(hipMalloc(&dPiv, sizeof(int) * (n > 1 ? n - 1 : 1)), "hipMalloc");
    for (int i = 0; i < n; ++i) hip_check(hipMemcpy(dA + (size_t)i * n, A[i], sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
    if (n > 1) hip_check(hipMemcpy(dPiv, piv, sizeof(int) * (n - 1), hipMemcpyHostToDevice), "hipMemcpy");
    hip_check(hipMemcpy(db, b, sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
    hz_check(hz_lu_solve_batched(dA, dPiv, db, n, 1, nullptr));
    hip_check(hipMemcpy(b, db, sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy");
    for (void *p : {(void *)dA, (void *)db, (void *)dPiv}) (void)hipFree(p);
}

// NonPeriodic3D_MPI.cpp:237-591
void SolvePDE(FieldVariable U[], void (*ConstructOperator)(const int, const int))
{
    if (g_fft_n == -1) ErrorMessage("Create a FFT plan first!");  // Fourier_Quadrature.cpp:1103-1104
    if (Maximum(Nx, Ny) > g_fft_n) ErrorMessage("Computation of DFT is not possible with the present FFT plan!");
    hz_check(hz_plan_set_lambda(g_plan, Lambda));
    hz_check(hz_plan_set_robin(g_plan, alphal, betal, alphar, betar));
    const int n = N - 2;
    // ---- Step 2's callbacks, recorded ----
    g_diag.assign((size_t)g_modes * n, 0.0);
    g_full = false;
    g_recording = true;
    for (int iw = 0; iw <= Nx / 2; iw++)
        for (int jw = 0; jw <= Ny / 2; jw++) {
            g_cur_mode = iw * g_M + jw;
            ConstructOperator(iw, jw);
        }
    g_recording = false;
    // ---- inputs in the plan's layout ----
    const size_t nf = (size_t)Nx * Ny * g_K, ng = (size_t)Nx * Ny;
    for (int kk = 0; kk < Ns; ++kk)
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                std::memcpy(&h_f[((size_t)i * Ny + j) * g_K + (size_t)kk * (N - 1)], rhsPDE[kk][i][j],
                            sizeof(double) * N);
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j) {
            h_g[(size_t)i * Ny + j] = leftBC[i][j];
            h_g[ng + (size_t)i * Ny + j] = rightBC[i][j];
        }
    hip_check(hipMemcpy(d_f, h_f, sizeof(double) * nf, hipMemcpyHostToDevice), "hipMemcpy");
    hip_check(hipMemcpy(d_gl, h_g, sizeof(double) * ng, hipMemcpyHostToDevice), "hipMemcpy");
    hip_check(hipMemcpy(d_gr, h_g + ng, sizeof(double) * ng, hipMemcpyHostToDevice), "hipMemcpy");
    // ---- operators to the device ----
    const std::vector<double> &ops = g_full ? g_full_ops : g_diag;
    (void)hipFree(d_ops);
    hip_check(hipMalloc(&d_ops, sizeof(double) * ops.size()), "hipMalloc");
    hip_check(hipMemcpy(d_ops, ops.data(), sizeof(double) * ops.size(), hipMemcpyHostToDevice), "hipMemcpy");
    hz_check(hz_solve_operators(g_plan, d_f, d_gl, d_gr, g_full ? nullptr : d_ops, g_full ? d_ops : nullptr,
                                d_u, nullptr));
    if (hz_solve_status(g_plan, nullptr) == HZ_ERR_SINGULAR) ErrorMessage("ERROR: Matrix is singular!");
    // ---- output in the caller's element arrays ----
    hip_check(hipMemcpy(h_f, d_u, sizeof(double) * nf, hipMemcpyDeviceToHost), "hipMemcpy");
    for (int kk = 0; kk < Ns; ++kk)
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                std::memcpy(U[kk].phi[i][j], &h_f[((size_t)i * Ny + j) * g_K + (size_t)kk * (N - 1)],
                            sizeof(double) * N);
}

void FirstDerivativeZ(FieldVariable Uz[], FieldVariable U[]) { derivative(Uz, U, 1); }
void SecondDerivativeZ(FieldVariable Uzz[], FieldVariable U[]) { derivative(Uzz, U, 2); }

// NonPeriodic3D_MPI.cpp:688-708 (one process: no reduction)
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

// NonPeriodic3D_MPI.cpp:713-741: Tecplot POINT file, one per rank.
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
