// compat_core.cpp -- shared engine of the drop-in solvers (compat_core.h)
// and the reference-named helpers every drop-in library exports:
// LUDecomposition / LUSolve (LU_Decomposition.cpp:7-108), ErrorMessage
// (General.cpp:6-11), CreateFFTPlan / DestroyFFTPlan
// (Fourier_Quadrature.cpp:16-69).  Citations: HelmholtzSolvers/DirectPseudoSpectral/.
#include "compat_core.h"

#include "../../include/hz_compat/Fourier_Quadrature.h"
#include "../../include/hz_compat/General.h"
#include "../../include/hz_compat/LU_Decomposition.h"

#include <cstring>
#include <iostream>

namespace hzc {

namespace {
Recorder g_rec;
int g_fft_n = -1;
}  // namespace

Recorder &recorder() { return g_rec; }
int fft_plan_size() { return g_fft_n; }

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

void Recorder::begin(int modes_, int n_, double **op_, double **dxx_)
{
    on = true;
    full = false;
    mode = -1;
    modes = modes_;
    n = n_;
    op = op_;
    dxx = dxx_;
    diag.assign((size_t)modes * n, 0.0);
    ops.clear();
}

bool Recorder::take(double **A, int n_)
{
    if (!on || n_ != n || A != op || mode < 0 || mode >= modes) return false;
    const size_t m = (size_t)mode;
    for (int i = 0; i < n; ++i) diag[m * n + i] = A[i][i];
    bool standard = true;
    for (int i = 0; i < n && standard; ++i)
        for (int j = 0; j < n; ++j)
            if (i != j && A[i][j] != dxx[i + 1][j + 1]) {
                standard = false;
                break;
            }
    if (!standard && !full) {
        // switch to whole matrices: rebuild the modes recorded so far
        full = true;
        ops.assign((size_t)modes * n * n, 0.0);
        for (size_t q = 0; q < m; ++q)
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j)
                    ops[(q * n + i) * n + j] = (i == j) ? diag[q * n + i] : dxx[i + 1][j + 1];
    }
    if (full)
        for (int i = 0; i < n; ++i) std::memcpy(&ops[(m * n + i) * n], A[i], sizeof(double) * n);
    return true;
}

void Solver::create(const hz_plan_desc &d)
{
    desc = d;
    int st = hz_plan_create(&desc, &plan);
    if (st == HZ_ERR_UNSUPPORTED && desc.form == HZ_FORM_PLAIN_LU) {
        desc.form = HZ_FORM_DIAG;
        st = hz_plan_create(&desc, &plan);
    }
    hz_check(st);
    hz_check(hz_plan_nz_total(plan, &K));
    hz_check(hz_plan_num_modes(plan, &modes));
    nx = d.nx;
    ny = d.ny;
    N = d.n;
    ne = d.ne;
    periodic = d.periodic_z != 0;
    const size_t nf = (size_t)nx * ny * K, ng = (size_t)nx * ny;
    hip_check(hipMalloc(&d_f, sizeof(double) * nf), "hipMalloc");
    hip_check(hipMalloc(&d_u, sizeof(double) * nf), "hipMalloc");
    hip_check(hipHostMalloc(&h_f, sizeof(double) * nf), "hipHostMalloc");
    if (!periodic) {
        hip_check(hipMalloc(&d_gl, sizeof(double) * ng), "hipMalloc");
        hip_check(hipMalloc(&d_gr, sizeof(double) * ng), "hipMalloc");
        hip_check(hipHostMalloc(&h_g, sizeof(double) * 2 * ng), "hipHostMalloc");
    }
}

void Solver::destroy()
{
    for (double *p : {d_f, d_gl, d_gr, d_u, d_ops}) (void)hipFree(p);
    (void)hipHostFree(h_f);
    (void)hipHostFree(h_g);
    d_f = d_gl = d_gr = d_u = d_ops = h_f = h_g = nullptr;
    if (plan) hz_plan_destroy(plan);
    plan = nullptr;
}

void Solver::solve_recorded(double lambda, double al, double bl, double ar, double br)
{
    hz_check(hz_plan_set_lambda(plan, lambda));
    if (!periodic) hz_check(hz_plan_set_robin(plan, al, bl, ar, br));
    const size_t nf = (size_t)nx * ny * K, ng = (size_t)nx * ny;
    hip_check(hipMemcpy(d_f, h_f, sizeof(double) * nf, hipMemcpyHostToDevice), "hipMemcpy");
    if (!periodic) {
        hip_check(hipMemcpy(d_gl, h_g, sizeof(double) * ng, hipMemcpyHostToDevice), "hipMemcpy");
        hip_check(hipMemcpy(d_gr, h_g + ng, sizeof(double) * ng, hipMemcpyHostToDevice), "hipMemcpy");
    }
    const Recorder &r = g_rec;
    if (desc.form != HZ_FORM_PLAIN_LU && recorded_standard(lambda)) {
        hz_check(hz_solve(plan, d_f, d_gl, d_gr, d_u, nullptr));
    } else {
        const std::vector<double> &ops = r.full ? r.ops : r.diag;
        (void)hipFree(d_ops);
        d_ops = nullptr;
        hip_check(hipMalloc(&d_ops, sizeof(double) * ops.size()), "hipMalloc");
        hip_check(hipMemcpy(d_ops, ops.data(), sizeof(double) * ops.size(), hipMemcpyHostToDevice), "hipMemcpy");
        hz_check(hz_solve_operators(plan, d_f, d_gl, d_gr, r.full ? nullptr : d_ops, r.full ? d_ops : nullptr, d_u,
                                    nullptr));
    }
    if (hz_solve_status(plan, nullptr) == HZ_ERR_SINGULAR) ErrorMessage("ERROR: Matrix is singular!");
    hip_check(hipMemcpy(h_f, d_u, sizeof(double) * nf, hipMemcpyDeviceToHost), "hipMemcpy");
}

bool Solver::recorded_standard(double lambda) const
{
    const Recorder &r = g_rec;
    if (r.full || r.n != N - 2 || (int)r.diag.size() != modes * r.n) return false;
    const double PI = 3.1415926535897932;  // include/General.h:14
    const double Lx = desc.xr - desc.xl, Ly = desc.yr - desc.yl;
    const int M = ny / 2 + 1;
    for (int iw = 0; iw <= nx / 2; ++iw)
        for (int jw = 0; jw < M; ++jw) {
            const double sigma = lambda + 4.0 * PI * PI * iw * iw / (Lx * Lx) + 4.0 * PI * PI * jw * jw / (Ly * Ly);
            const double *dg = &r.diag[((size_t)iw * M + jw) * r.n];
            for (int i = 0; i < r.n; ++i)
                if (dg[i] != r.dxx[i + 1][i + 1] - sigma) return false;
        }
    return true;
}

void Solver::derivative_z(std::vector<double> &elements, int order)
{
    double *d_in = nullptr, *d_out = nullptr;
    const size_t bytes = sizeof(double) * elements.size();
    hip_check(hipMalloc(&d_in, bytes), "hipMalloc");
    hip_check(hipMalloc(&d_out, bytes), "hipMalloc");
    hip_check(hipMemcpy(d_in, elements.data(), bytes, hipMemcpyHostToDevice), "hipMemcpy");
    hz_check(hz_derivative_z(plan, order, d_in, d_out, nullptr));
    hip_check(hipMemcpy(elements.data(), d_out, bytes, hipMemcpyDeviceToHost), "hipMemcpy");
    (void)hipFree(d_in);
    (void)hipFree(d_out);
}

void Solver::derivative_xy(std::vector<double> &stack, int lanes, int axis, int order)
{
    double *d_in = nullptr, *d_out = nullptr;
    const size_t bytes = sizeof(double) * stack.size();
    hip_check(hipMalloc(&d_in, bytes), "hipMalloc");
    hip_check(hipMalloc(&d_out, bytes), "hipMalloc");
    hip_check(hipMemcpy(d_in, stack.data(), bytes, hipMemcpyHostToDevice), "hipMemcpy");
    hz_check(hz_derivative_xy(plan, axis, order, d_in, d_out, lanes, nullptr));
    hip_check(hipMemcpy(stack.data(), d_out, bytes, hipMemcpyDeviceToHost), "hipMemcpy");
    (void)hipFree(d_in);
    (void)hipFree(d_out);
}

void Solver::closure_full(std::vector<double> &alpha)
{
    const int M = ny / 2 + 1;
    std::vector<double> half((size_t)2 * nx * M);
    hz_check(hz_plan_periodic_closure(plan, half.data()));
    alpha.assign((size_t)nx * 2 * ny, 0.0);
    for (int i = 0; i < nx; ++i)
        for (int m = 0; m < ny; ++m) {
            const bool low = m < M;
            const int ii = low ? i : (nx - i) % nx, mm = low ? m : ny - m;
            const size_t h = (size_t)ii * M + mm;
            alpha[(size_t)i * 2 * ny + 2 * m] = half[h];
            alpha[(size_t)i * 2 * ny + 2 * m + 1] = low ? half[(size_t)nx * M + h] : -half[(size_t)nx * M + h];
        }
}

void particular_solutions(double *const *cols, const double *const *rhs, int count, int N, double **A, int *piv)
{
    const int n = N - 2;
    if (count <= 0 || n < 1) return;
    const size_t nn = (size_t)n * n, np = n > 1 ? (size_t)(n - 1) : 1;
    std::vector<double> hA(nn * count), hb((size_t)n * count);
    std::vector<int> hp(np * count, 0);
    for (int c = 0; c < count; ++c) {
        for (int i = 0; i < n; ++i) std::memcpy(&hA[c * nn + (size_t)i * n], A[i], sizeof(double) * n);
        if (n > 1) std::memcpy(&hp[c * np], piv, sizeof(int) * (n - 1));
        std::memcpy(&hb[(size_t)c * n], rhs[c] + 1, sizeof(double) * n);
    }
    double *dA = nullptr, *db = nullptr;
    int *dp = nullptr;
    hip_check(hipMalloc(&dA, sizeof(double) * hA.size()), "hipMalloc");
    hip_check(hipMalloc(&db, sizeof(double) * hb.size()), "hipMalloc");
    hip_check(hipMalloc(&dp, sizeof(int) * hp.size()), "hipMalloc");
    hip_check(hipMemcpy(dA, hA.data(), sizeof(double) * hA.size(), hipMemcpyHostToDevice), "hipMemcpy");
    hip_check(hipMemcpy(db, hb.data(), sizeof(double) * hb.size(), hipMemcpyHostToDevice), "hipMemcpy");
    hip_check(hipMemcpy(dp, hp.data(), sizeof(int) * hp.size(), hipMemcpyHostToDevice), "hipMemcpy");
    hz_check(hz_lu_solve_batched(dA, dp, db, n, count, nullptr));
    hip_check(hipMemcpy(hb.data(), db, sizeof(double) * hb.size(), hipMemcpyDeviceToHost), "hipMemcpy");
    for (void *p : {(void *)dA, (void *)db, (void *)dp}) (void)hipFree(p);
    for (int c = 0; c < count; ++c) {
        cols[c][0] = 0.0;
        cols[c][N - 1] = 0.0;
        std::memcpy(cols[c] + 1, &hb[(size_t)c * n], sizeof(double) * n);
    }
}

}  // namespace hzc

using namespace hzc;

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

// LU_Decomposition.cpp:7-72 (GPU), or recording inside SolvePDE.
void LUDecomposition(double **&A, int *&piv, double &det, const int n)
{
    if (g_rec.take(A, n)) return;
    double *dA = nullptr, *dDet = nullptr;
    int *dPiv = nullptr, *dSt = nullptr;
    hip_check(hipMalloc(&dA, sizeof(double) * n * n), "hipMalloc");
    hip_check(hipMalloc(&dDet, sizeof(double)), "hipMalloc");
    hip_check(hipMalloc(&dPiv, sizeof(int) * (n > 1 ? n - 1 : 1)), "hipMalloc");
    hip_check(hipMalloc(&dSt, sizeof(int)), "hipMalloc");
    for (int i = 0; i < n; ++i)
        hip_check(hipMemcpy(dA + (size_t)i * n, A[i], sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
    hz_check(hz_lu_factor_batched(dA, n, 1, dPiv, dDet, dSt, nullptr));
    int st = 0;
    hip_check(hipMemcpy(&st, dSt, sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy");
    for (int i = 0; i < n; ++i)
        hip_check(hipMemcpy(A[i], dA + (size_t)i * n, sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy");
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
    for (int i = 0; i < n; ++i)
        hip_check(hipMemcpy(dA + (size_t)i * n, A[i], sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
    if (n > 1) hip_check(hipMemcpy(dPiv, piv, sizeof(int) * (n - 1), hipMemcpyHostToDevice), "hipMemcpy");
    hip_check(hipMemcpy(db, b, sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
    hz_check(hz_lu_solve_batched(dA, dPiv, db, n, 1, nullptr));
    hip_check(hipMemcpy(b, db, sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy");
    for (void *p : {(void *)dA, (void *)db, (void *)dPiv}) (void)hipFree(p);
}
