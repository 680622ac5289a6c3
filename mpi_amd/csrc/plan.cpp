// plan.cpp -- the C ABI (include/hz.h): plan construction, device workspace
// and the solve pipelines.
//
// Solve pipeline (all forms), z-fastest fields, half spectrum in x-y:
//   f  --R2C along y-->  spec  --C2C along x-->  spec         (fft.hip)
//   g_l, g_r: the same two passes into the boundary spectra
//   spec --mode solve (form-specific)--> spec                  (lu.hip / diag.hip / mixed.hip)
//   spec --C2C^-1 along x--> spec --C2R along y--> u
#include "hz_internal.h"

#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>

namespace hz {

static thread_local std::string g_last_error;

int set_error(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

}  // namespace hz

using namespace hz;

namespace {

bool is_pow2(int v) { return v >= 1 && (v & (v - 1)) == 0; }

template <class T>
int dmalloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) return HZ_OK;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) return set_error(HZ_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
    return HZ_OK;
}

template <class T>
int upload(T **p, const std::vector<T> &v)
{
    int st = dmalloc(p, v.size());
    if (st) return st;
    HZ_HIP(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return HZ_OK;
}

std::vector<double2> twiddles(int L)
{
    std::vector<double2> t(L);
    for (int i = 0; i < L; ++i) {
        const long double ang = 2.0L * 3.14159265358979323846264338327950288L * i / L;
        t[i].x = (double)cosl(ang);
        t[i].y = (double)sinl(ang);
    }
    return t;
}

void free_plan(hz_plan *p)
{
    if (!p) return;
    for (auto &v : p->ev_solves)
        for (auto e : v) (void)hipEventDestroy(e);
    (void)hipFree(p->d_Dz);
    (void)hipFree(p->d_Dzz);
    (void)hipFree(p->d_DzzT);
    (void)hipFree(p->d_shift);
    (void)hipFree(p->d_twx);
    (void)hipFree(p->d_twy);
    (void)hipFree(p->spec.re);  // one allocation holds both planes
    (void)hipFree(p->bc.re);
    for (double *q : {p->d_mu, p->d_w0, p->d_wN, p->d_r0, p->d_rN, p->d_VinvT, p->d_VT, p->d_hom, p->d_Vrow,
                      p->d_Vinvrow, p->d_eo_vec, p->d_eo_B1, p->d_eo_B2, p->d_eo_B1a, p->d_eo_B2a, p->d_eo_hom,
                      p->d_partial, p->d_DzzTP, p->d_UL, p->d_UR, p->d_hist, p->d_closure})
        (void)hipFree(q);
    (void)hipFree(p->d_eo_M);
    (void)hipFree(p->d_eo_Ma);
    (void)hipFree(p->d_VinvT32);
    (void)hipFree(p->d_VT32);
    (void)hipFree(p->d_xspec);
    (void)hipFree(p->d_status);
    if (p->h_lambda_bar) (void)hipHostFree(p->h_lambda_bar);
    delete p;
}

// Per-mode end-point derivatives of the homogeneous solutions for the mode's
// shift: Dz[0|N-1].ul and Dz[0|N-1].ur, in the full eigenbasis and in the
// even/odd one (the diag kernel's one-element Dirichlet solves).
static void homogeneous(const hz_plan *p, int len, const double *mu, const double *w0, const double *wN,
                        const double *r0, const double *rN, std::vector<double> &hom)
{
    const int N = p->N;
    hom.assign((size_t)4 * p->modes, 0.0);
    for (int m = 0; m < p->modes; ++m) {
        long double a0 = p->Dz[0], aN = p->Dz[(size_t)(N - 1) * N];
        long double b0 = p->Dz[N - 1], bN = p->Dz[(size_t)(N - 1) * N + N - 1];
        const long double sig = p->shift[m];
        for (int j = 0; j < len; ++j) {
            if (w0[j] == 0 && wN[j] == 0) continue;  // padding (and loads with no component)
            const long double iv = 1.0L / ((long double)mu[j] - sig);
            a0 += (long double)r0[j] * w0[j] * iv;
            aN += (long double)rN[j] * w0[j] * iv;
            b0 += (long double)r0[j] * wN[j] * iv;
            bN += (long double)rN[j] * wN[j] * iv;
        }
        hom[4 * m + 0] = (double)a0;
        hom[4 * m + 1] = (double)aN;
        hom[4 * m + 2] = (double)b0;
        hom[4 * m + 3] = (double)bN;
    }
}

int compute_homogeneous(hz_plan *p)
{
    if (!p->d_hom) return HZ_OK;
    std::vector<double> hom;
    homogeneous(p, p->n, p->mu.data(), p->w0.data(), p->wN.data(), p->r0.data(), p->rN.data(), hom);
    HZ_HIP(hipMemcpy(p->d_hom, hom.data(), sizeof(double) * hom.size(), hipMemcpyHostToDevice));
    const double *v = p->eo_vec.data();
    const int NP = p->NP;
    homogeneous(p, NP, v, v + NP, v + 2 * NP, v + 3 * NP, v + 4 * NP, hom);
    HZ_HIP(hipMemcpy(p->d_eo_hom, hom.data(), sizeof(double) * hom.size(), hipMemcpyHostToDevice));
    return HZ_OK;
}

// MIXED form: the homogeneous solutions ul, ur of every mode in FP64 (once
// per lambda; the assembly of several elements or Robin faces needs them).
int mixed_operators(hz_plan *p)
{
    MixedSetupArgs a{};
    a.n = p->n;
    a.NP = p->NP;
    a.modes = p->modes;
    a.shift = p->d_shift;
    a.mu = p->d_mu;
    a.w0 = p->d_w0;
    a.wN = p->d_wN;
    a.VT = p->d_VT;
    a.UL = p->d_UL;
    a.UR = p->d_UR;
    int st = launch_mixed_setup(a, nullptr);
    if (st) return st;
    HZ_HIP(hipDeviceSynchronize());
    return HZ_OK;
}

// MIXED form: FP32 copies of the even/odd eigenbasis (Winv^T, (D W)^T, as
// the diag kernel's forward operands) and the padded FP64 Dzz_int^T of the
// residual, all NP x NP; the history and homogeneous solution buffers.
int setup_mixed(hz_plan *p)
{
    p->sweeps = p->d.refine_sweeps > 0 ? p->d.refine_sweeps : (p->d.refine_sweeps < 0 ? 0 : 2);
    if (p->sweeps > 16) return set_error(HZ_ERR_ARG, "refine_sweeps must be <= 16");
    if (2 * p->ne > mixed_tile_rows(p->NP))
        return set_error(HZ_ERR_UNSUPPORTED, "MIXED form: too many elements for one tile (at most 8, or 16 "
                                             "when N-2 rounds up to 32 or 96)");
    const int n = p->n, N = p->N, NP = p->NP;
    std::vector<float> VinvT32((size_t)NP * NP, 0.0f), VT32((size_t)NP * NP, 0.0f);
    std::vector<double> DzzTP((size_t)NP * NP, 0.0);
    for (int r = 0; r < NP; ++r) {
        const double dr = (n % 2 && r == n / 2) ? 1.0 : 0.5;
        for (int c = 0; c < NP; ++c) {
            VinvT32[(size_t)c * NP + r] = (float)p->eo_Wi[(size_t)r * NP + c];
            VT32[(size_t)c * NP + r] = (float)(dr * p->eo_W[(size_t)r * NP + c]);
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) DzzTP[(size_t)j * NP + i] = p->Dzz[(size_t)(i + 1) * N + (j + 1)];
    int st;
    if ((st = upload(&p->d_VinvT32, VinvT32)) || (st = upload(&p->d_VT32, VT32)) ||
        (st = upload(&p->d_DzzTP, DzzTP)) || (st = dmalloc(&p->d_UL, (size_t)p->modes * n)) ||
        (st = dmalloc(&p->d_UR, (size_t)p->modes * n)) ||
        (st = dmalloc(&p->d_hist, (size_t)std::max(p->sweeps, 1) * p->nx * p->M * 2 * p->ne)))
        return st;
    return HZ_OK;
}

int compute_shifts(hz_plan *p)
{
    p->shift.resize(p->modes);
    for (int iw = 0; iw <= p->nx / 2; ++iw)
        for (int jw = 0; jw < p->M; ++jw)
            p->shift[(size_t)iw * p->M + jw] = mode_shift(p->d.lambda, iw, jw, p->Lx, p->Ly);
    HZ_HIP(hipMemcpy(p->d_shift, p->shift.data(), sizeof(double) * p->modes, hipMemcpyHostToDevice));
    int st = compute_homogeneous(p);
    if (st) return st;
    if (p->d.form == HZ_FORM_MIXED && p->d_UL) return mixed_operators(p);
    return HZ_OK;
}

// The diag kernel's operands in the even/odd eigenbasis (eigen_decompose_split)
// for one element with Dirichlet faces.  With F the fold, F^-1 its inverse
// and W, Winv, mu the decomposition of As = (A + JAJ)/2 (padded to NP),
//   z = (Winv F f + zl w0 + zr wN) / (mu - sigma),  w0|wN = Winv F (-Dzz[1..N-2][0|N-1]),
// and u = F^-1 W z solves with As.  The reference's A = As + Ea, Ea =
// (A - JAJ)/2 of the size of the GLL nodes' rounding, so to first order
//   (A - sigma)^-1 = F^-1 W (I - Lambda^-1 M) Lambda^-1 Winv F,  M = Winv F Ea F^-1 W,
// Lambda = diag(mu - sigma): the kernel solves with A by z -= (M z) / (mu -
// sigma) before the second contraction (M z is about 1e-12 |z|, so bf16
// operands with FP32 sums suffice; without the correction the solution moves by up to 2e-12
// relative at N = 257).  M maps even to odd coordinates and back (Ea
// anticommutes with J).  The kernel applies F and F^-1 as sums and
// differences of mirrored entries (the halves of F^-1 are folded into D W,
// D = 1/2 except on the centre row of an odd n).  Operands, NP x NP
// row-major: B1 = Winv^T, B2 = (D W)^T, adjoint B1 = D W, B2 = Winv (FP64),
// M and M^T in bf16 (correction_operator); r0|rN =
// Dz[0|N-1][1..N-2] F^-1 W (the homogeneous coefficients' basis).
std::array<std::vector<short>, 2> correction_operator(const std::vector<double> &A, int n, int NP,
                                                      const std::vector<double> &W, const std::vector<double> &Wi);

int setup_even_odd(hz_plan *p, const std::vector<double> &A)
{
    const int n = p->n, N = p->N, NP = p->NP, H = NP / 2, m = n / 2;
    static std::mutex cache_mu;
    static std::map<std::pair<std::vector<double>, int>, std::array<std::vector<double>, 3>> cache;
    static std::map<std::pair<std::vector<double>, int>, std::array<std::vector<short>, 2>> cache_m;
    std::vector<double> mu, W, Wi;
    std::array<std::vector<short>, 2> Mq;
    const auto key = std::make_pair(A, NP);
    {
        std::lock_guard<std::mutex> lk(cache_mu);
        auto it = cache.find(key);
        if (it != cache.end()) {
            mu = it->second[0];
            W = it->second[1];
            Wi = it->second[2];
            Mq = cache_m[key];
        } else {
            const int st = eigen_decompose_split(A, n, NP, mu, W, Wi, nullptr);
            if (st) return set_error(st, "even/odd eigendecomposition of the GLL operator failed");
        }
    }
    // padded index of the fold's coordinates: even i -> i, odd i -> H + i
    auto fold = [&](const std::vector<long double> &x) {
        std::vector<long double> y(NP, 0);
        for (int i = 0; i < m; ++i) {
            y[i] = x[i] + x[n - 1 - i];
            y[H + i] = x[i] - x[n - 1 - i];
        }
        if (n % 2) y[m] = x[m];
        return y;
    };
    auto unfold_row = [&](const std::vector<long double> &d) {  // d F^-1
        std::vector<long double> y(NP, 0);
        for (int i = 0; i < m; ++i) {
            y[i] = (d[i] + d[n - 1 - i]) / 2;
            y[H + i] = (d[i] - d[n - 1 - i]) / 2;
        }
        if (n % 2) y[m] = d[m];
        return y;
    };
    std::vector<long double> g0(n), gN(n), d0(n), dN(n);
    for (int k = 0; k < n; ++k) {
        g0[k] = -p->Dzz[(size_t)(k + 1) * N];
        gN[k] = -p->Dzz[(size_t)(k + 1) * N + N - 1];
        d0[k] = p->Dz[k + 1];
        dN[k] = p->Dz[(size_t)(N - 1) * N + k + 1];
    }
    const std::vector<long double> fg0 = fold(g0), fgN = fold(gN), ud0 = unfold_row(d0), udN = unfold_row(dN);
    p->eo_vec.assign((size_t)5 * NP, 0.0);
    double *v = p->eo_vec.data();
    for (int j = 0; j < NP; ++j) {
        long double s0 = 0, sN = 0, t0 = 0, tN = 0;
        for (int c = 0; c < NP; ++c) {
            s0 += (long double)Wi[(size_t)j * NP + c] * fg0[c];
            sN += (long double)Wi[(size_t)j * NP + c] * fgN[c];
            t0 += ud0[c] * W[(size_t)c * NP + j];
            tN += udN[c] * W[(size_t)c * NP + j];
        }
        v[j] = mu[j];
        v[NP + j] = (double)s0;
        v[2 * NP + j] = (double)sN;
        v[3 * NP + j] = (double)t0;
        v[4 * NP + j] = (double)tN;
    }
    p->eo_W = W;
    p->eo_Wi = Wi;
    std::vector<double> B1((size_t)NP * NP), B2((size_t)NP * NP), B1a((size_t)NP * NP), B2a((size_t)NP * NP);
    for (int r = 0; r < NP; ++r) {
        const double dr = (n % 2 && r == m) ? 1.0 : 0.5;
        for (int c = 0; c < NP; ++c) {
            const double dw = dr * W[(size_t)r * NP + c];  // (D W)[r][c]
            B1[(size_t)c * NP + r] = Wi[(size_t)r * NP + c];
            B2[(size_t)c * NP + r] = dw;
            B1a[(size_t)r * NP + c] = dw;
            B2a[(size_t)r * NP + c] = Wi[(size_t)r * NP + c];
        }
    }
    if (Mq[0].empty()) {
        Mq = correction_operator(A, n, NP, W, Wi);
        std::lock_guard<std::mutex> lk(cache_mu);
        if (cache.size() < 256) {
            cache[key] = {mu, W, Wi};
            cache_m[key] = Mq;
        }
    }
    int st;
    if ((st = upload(&p->d_eo_vec, p->eo_vec)) || (st = upload(&p->d_eo_B1, B1)) || (st = upload(&p->d_eo_B2, B2)) ||
        (st = upload(&p->d_eo_B1a, B1a)) || (st = upload(&p->d_eo_B2a, B2a)) || (st = upload(&p->d_eo_M, Mq[0])) ||
        (st = upload(&p->d_eo_Ma, Mq[1])))
        return st;
    return HZ_OK;
}

// The even/odd solve's first-order correction M = Winv (F Ea F^-1) W, Ea =
// (A - JAJ)/2 (the middle factor column by column: F^-1 of a unit vector has
// one or two entries; then the two products, in long double), as the bf16
// B operands of c = M z (forward) and c = M^T q (adjoint):
// [0][(k / 4) * NP + j][k % 4] = bf16(M[j][k]), [1][...] = bf16(M[k][j]),
// rounded to nearest even.
std::array<std::vector<short>, 2> correction_operator(const std::vector<double> &A, int n, int NP,
                                                      const std::vector<double> &W, const std::vector<double> &Wi)
{
    const int H = NP / 2, m = n / 2;
    typedef long double ld;
    std::vector<ld> Et((size_t)NP * NP, 0), T((size_t)NP * NP, 0);
    std::vector<ld> y(n);
    auto ea = [&](int r, int c) {
        return ((ld)A[(size_t)r * n + c] - (ld)A[(size_t)(n - 1 - r) * n + (n - 1 - c)]) / 2;
    };
    for (int c = 0; c < NP; ++c) {
        const bool even = c < H;
        const int j = even ? c : c - H;
        if (j >= (even ? n - m : m)) continue;
        for (int r = 0; r < n; ++r)
            y[r] = even && 2 * j + 1 == n ? ea(r, j) : (ea(r, j) + (even ? 1 : -1) * ea(r, n - 1 - j)) / 2;
        for (int i = 0; i < m; ++i) {
            Et[(size_t)i * NP + c] = y[i] + y[n - 1 - i];
            Et[(size_t)(H + i) * NP + c] = y[i] - y[n - 1 - i];
        }
        if (n % 2) Et[(size_t)m * NP + c] = y[m];
    }
    for (int r = 0; r < NP; ++r)      // T = Et W
        for (int k = 0; k < NP; ++k) {
            const ld e = Et[(size_t)r * NP + k];
            if (e == 0) continue;
            for (int c = 0; c < NP; ++c) T[(size_t)r * NP + c] += e * (ld)W[(size_t)k * NP + c];
        }
    std::array<std::vector<short>, 2> Mq;
    Mq[0].assign((size_t)NP * NP, 0);
    Mq[1].assign((size_t)NP * NP, 0);
    for (int r = 0; r < NP; ++r)      // M = Winv T
        for (int c = 0; c < NP; ++c) {
            ld acc = 0;
            for (int k = 0; k < NP; ++k) acc += (ld)Wi[(size_t)r * NP + k] * T[(size_t)k * NP + c];
            const float f = (float)acc;
            uint32_t u;
            std::memcpy(&u, &f, 4);
            u += 0x7FFFu + ((u >> 16) & 1u);
            Mq[0][((size_t)(c / 4) * NP + r) * 4 + c % 4] = (short)(u >> 16);  // B[k = c][j = r]
            Mq[1][((size_t)(r / 4) * NP + c) * 4 + r % 4] = (short)(u >> 16);  // B[k = r][j = c]
        }
    return Mq;
}

// Eigen-decomposition of the interior operator and the derived vectors of
// the diagonalised form, uploaded padded to NP x NP.
int setup_diag(hz_plan *p)
{
    const int n = p->n, N = p->N;
    p->NP = diag_padded(n);
    if (!p->NP) return set_error(HZ_ERR_UNSUPPORTED, "DIAG form: N-2 must be <= 256");
    std::vector<double> A((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) A[(size_t)i * n + j] = p->Dzz[(size_t)(i + 1) * N + (j + 1)];
    // The decomposition depends only on the interior operator: plans with
    // the same (N, ne, Lz) share it (many-problem batches repeat sizes).
    static std::mutex cache_mu;
    static std::map<std::vector<double>, std::array<std::vector<double>, 3>> cache;
    int st = HZ_OK;
    {
        std::lock_guard<std::mutex> lk(cache_mu);
        auto it = cache.find(A);
        if (it != cache.end()) {
            p->mu = it->second[0];
            p->V = it->second[1];
            p->Vinv = it->second[2];
        } else {
            st = eigen_decompose(A, n, p->mu, p->V, p->Vinv);
            if (!st && cache.size() < 256) cache[A] = {p->mu, p->V, p->Vinv};
        }
    }
    if (st) return set_error(st, "eigendecomposition of the GLL operator did not converge");
    p->w0.assign(n, 0.0);
    p->wN.assign(n, 0.0);
    p->r0.assign(n, 0.0);
    p->rN.assign(n, 0.0);
    for (int i = 0; i < n; ++i) {
        long double s0 = 0, sN = 0, t0 = 0, tN = 0;
        for (int k = 0; k < n; ++k) {
            s0 += (long double)p->Vinv[(size_t)i * n + k] * -p->Dzz[(size_t)(k + 1) * N];
            sN += (long double)p->Vinv[(size_t)i * n + k] * -p->Dzz[(size_t)(k + 1) * N + N - 1];
            t0 += (long double)p->Dz[k + 1] * p->V[(size_t)k * n + i];
            tN += (long double)p->Dz[(size_t)(N - 1) * N + k + 1] * p->V[(size_t)k * n + i];
        }
        p->w0[i] = (double)s0;
        p->wN[i] = (double)sN;
        p->r0[i] = (double)t0;
        p->rN[i] = (double)tN;
    }
    const int NP = p->NP;
    std::vector<double> VinvT((size_t)NP * NP, 0.0), VT((size_t)NP * NP, 0.0);
    std::vector<double> Vrow((size_t)NP * NP, 0.0), Vinvrow((size_t)NP * NP, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            VinvT[(size_t)j * NP + i] = p->Vinv[(size_t)i * n + j];
            VT[(size_t)j * NP + i] = p->V[(size_t)i * n + j];
            Vrow[(size_t)i * NP + j] = p->V[(size_t)i * n + j];
            Vinvrow[(size_t)i * NP + j] = p->Vinv[(size_t)i * n + j];
        }
    if ((st = setup_even_odd(p, A))) return st;
    if ((st = upload(&p->d_mu, p->mu)) || (st = upload(&p->d_w0, p->w0)) || (st = upload(&p->d_wN, p->wN)) ||
        (st = upload(&p->d_r0, p->r0)) || (st = upload(&p->d_rN, p->rN)) || (st = upload(&p->d_VinvT, VinvT)) ||
        (st = upload(&p->d_VT, VT)) || (st = upload(&p->d_Vrow, Vrow)) || (st = upload(&p->d_Vinvrow, Vinvrow)) ||
        (st = dmalloc(&p->d_hom, (size_t)4 * p->modes)) || (st = dmalloc(&p->d_eo_hom, (size_t)4 * p->modes)))
        return st;
    return HZ_OK;
}

}  // namespace

// Arguments of the MIXED mode solve on `p`'s workspace.
MixedArgs hz::mixed_args(const hz_plan *p)
{
    MixedArgs a{};
    a.nx = p->nx;
    a.M = p->M;
    a.K = p->K;
    a.N = p->N;
    a.n = p->n;
    a.ne = p->ne;
    a.sweeps = p->sweeps;
    a.spec = p->spec.re;
    a.plane_off = p->spec.im - p->spec.re;
    a.xs = p->xs;
    a.bc_re = p->bc.re;
    a.bc_im = p->bc.im;
    a.shift = p->d_shift;
    a.mu = p->d_eo_vec;  // the even/odd eigenvalues, padded to NP
    a.VinvT32 = p->d_VinvT32;
    a.VT32 = p->d_VT32;
    a.DzzTP = p->d_DzzTP;
    a.Dzz = p->d_Dzz;
    a.Dz = p->d_Dz;
    a.UL = p->d_UL;
    a.UR = p->d_UR;
    a.hom = p->d_hom;
    a.alpha_l = p->d.alpha_l;
    a.beta_l = p->d.beta_l;
    a.alpha_r = p->d.alpha_r;
    a.beta_r = p->d.beta_r;
    a.hist = p->sweeps > 0 ? p->d_hist : nullptr;
    a.hist_rows = (long long)p->nx * p->M * 2 * p->ne;
    a.periodic = p->d.periodic_z;
    a.closure = p->d_closure;
    return a;
}

// Arguments of the four line-transform passes of a solve (see hz_internal.h).
FftArgs hz::xy_pass_args(const hz_plan *p, int pass, const double *in, double *re, double *im, double *out,
                         int lanes, const double *gl, const double *gr, double *gl_out, double *gr_out,
                         long long xs)
{
    const long long side = (long long)p->nx * p->M;
    const long long xrow = xs ? xs : (long long)p->M * lanes;  // spectrum x-row stride
    FftArgs a{};
    a.lanes = lanes;
    a.w = p->fft_w[pass];
    if (pass == 0) {  // R2C along y: [nx][ny][lanes] -> [nx][M][lanes]
        a.in_re = in;
        a.out_re = re;
        a.out_im = im;
        a.in_batch_stride = (long long)p->ny * lanes;
        a.in_point_stride = lanes;
        a.out_batch_stride = xrow;
        a.out_point_stride = lanes;
        a.nbatch = p->nx;
        a.tw = p->d_twy;
        if (gl && gr) {
            a.xlanes = 2;
            a.xin_re[0] = gl;
            a.xin_re[1] = gr;
            a.xout_re[0] = p->bc.re;
            a.xout_re[1] = p->bc.re + side;
            a.xout_im[0] = p->bc.im;
            a.xout_im[1] = p->bc.im + side;
            a.x_in_batch_stride = p->ny;
            a.x_in_point_stride = 1;
            a.x_out_batch_stride = p->M;
            a.x_out_point_stride = 1;
        }
    } else if (pass == 1 || pass == 2) {  // C2C along x, in place
        a.in_re = re;
        a.in_im = im;
        a.out_re = re;
        a.out_im = im;
        a.in_batch_stride = lanes;
        a.in_point_stride = xrow;
        a.out_batch_stride = lanes;
        a.out_point_stride = xrow;
        a.nbatch = p->M;
        a.tw = p->d_twx;
        if ((pass == 1 && gl && gr) || (pass == 2 && gl_out && gr_out)) {
            a.xlanes = 2;
            for (int sx = 0; sx < 2; ++sx) {
                a.xin_re[sx] = a.xout_re[sx] = p->bc.re + sx * side;
                a.xin_im[sx] = a.xout_im[sx] = p->bc.im + sx * side;
            }
            a.x_in_batch_stride = a.x_out_batch_stride = 1;
            a.x_in_point_stride = a.x_out_point_stride = p->M;
        }
    } else {  // C2R along y: [nx][M][lanes] -> [nx][ny][lanes]
        a.in_re = re;
        a.in_im = im;
        a.out_re = out;
        a.in_batch_stride = xrow;
        a.in_point_stride = lanes;
        a.out_batch_stride = (long long)p->ny * lanes;
        a.out_point_stride = lanes;
        a.nbatch = p->nx;
        a.tw = p->d_twy;
        if (gl_out && gr_out) {
            a.xlanes = 2;
            for (int sx = 0; sx < 2; ++sx) {
                a.xin_re[sx] = p->bc.re + sx * side;
                a.xin_im[sx] = p->bc.im + sx * side;
            }
            a.xout_re[0] = gl_out;
            a.xout_re[1] = gr_out;
            a.x_in_batch_stride = p->M;
            a.x_in_point_stride = 1;
            a.x_out_batch_stride = p->ny;
            a.x_out_point_stride = 1;
        }
    }
    return a;
}

namespace {

// Forward x-y transform of a [nx][ny][lanes] field into split half spectra;
// with gl/gr the two face planes ride along as extra lanes of the same
// launches and land in the boundary spectra.
// Forward transform; with `cubic` the R2C pass reads in * (c[0] in^2 + c[1])
// (hz_solve_cubic's right-hand side).
int forward_xy(hz_plan *p, const double *in, double *re, double *im, int lanes, hipStream_t s,
               const double *gl = nullptr, const double *gr = nullptr, long long xs = 0,
               const double *cubic = nullptr)
{
    FftArgs r2c = xy_pass_args(p, 0, in, re, im, nullptr, lanes, gl, gr, nullptr, nullptr, xs);
    if (cubic) {
        r2c.pre_cubic = 1;
        r2c.pre_c3 = cubic[0];
        r2c.pre_c1 = cubic[1];
    }
    int st = launch_fft(p->ny, FFT_R2C, r2c, s);
    if (st) return st;
    return launch_fft(p->nx, FFT_C2C_FWD,
                      xy_pass_args(p, 1, nullptr, re, im, nullptr, lanes, gl, gr, nullptr, nullptr, xs), s);
}

// Inverse transform; with gl_out/gr_out the boundary spectra are inverted
// into them in the same launches.
int inverse_xy(hz_plan *p, double *re, double *im, double *out, int lanes, hipStream_t s,
               double *gl_out = nullptr, double *gr_out = nullptr, long long xs = 0)
{
    int st = launch_fft(p->nx, FFT_C2C_INV,
                        xy_pass_args(p, 2, nullptr, re, im, nullptr, lanes, nullptr, nullptr, gl_out, gr_out, xs), s);
    if (st) return st;
    return launch_fft(p->ny, FFT_C2R,
                      xy_pass_args(p, 3, nullptr, re, im, out, lanes, nullptr, nullptr, gl_out, gr_out, xs), s);
}

}  // namespace

// The diag kernel's operands: the even/odd eigenbasis for one element with
// Dirichlet faces (the kernels' EO instances), the full eigenbasis otherwise.
void hz::diag_basis(const hz_plan *p, DiagArgs &a, bool adjoint)
{
    const int NP = p->NP;
    a.eo = !diag_general(p);
    a.eoM = adjoint ? p->d_eo_Ma : p->d_eo_M;
    if (a.eo) {
        a.mu = p->d_eo_vec;
        a.w0 = p->d_eo_vec + NP;
        a.wN = p->d_eo_vec + 2 * NP;
        a.r0 = p->d_eo_vec + 3 * NP;
        a.rN = p->d_eo_vec + 4 * NP;
        a.VinvT = adjoint ? p->d_eo_B1a : p->d_eo_B1;
        a.VT = adjoint ? p->d_eo_B2a : p->d_eo_B2;
        a.hom = p->d_eo_hom;
    } else {
        a.mu = p->d_mu;
        a.w0 = p->d_w0;
        a.wN = p->d_wN;
        a.r0 = p->d_r0;
        a.rN = p->d_rN;
        a.VinvT = adjoint ? p->d_Vrow : p->d_VinvT;
        a.VT = adjoint ? p->d_Vinvrow : p->d_VT;
        a.hom = p->d_hom;
    }
}

bool hz::diag_general(const hz_plan *p)
{
    return p->ne > 1 || p->d.alpha_l != 0.0 || p->d.alpha_r != 0.0 || p->d.periodic_z;
}

// Kernel arguments of the z-diagonalised mode solve (forward map).
DiagArgs hz::diag_args(const hz_plan *p)
{
    DiagArgs a{};
    a.nx = p->nx;
    a.M = p->M;
    a.K = p->K;
    a.N = p->N;
    a.n = p->n;
    a.ne = p->ne;
    a.spec = p->spec.re;
    a.plane_off = p->spec.im - p->spec.re;
    a.xs = p->xs;
    a.bc_re = p->bc.re;
    a.bc_im = p->bc.im;
    a.shift = p->d_shift;
    diag_basis(p, a, false);
    a.alpha_l = p->d.alpha_l;
    a.beta_l = p->d.beta_l;
    a.alpha_r = p->d.alpha_r;
    a.beta_r = p->d.beta_r;
    a.rt = p->diag_rt;
    a.periodic = p->d.periodic_z;
    a.closure = p->d_closure;
    return a;
}

namespace {

// Kernel arguments of the reference-algorithm (PLAIN_LU) mode solve.
LuModeArgs lu_args(const hz_plan *p, const double *op_diag = nullptr, const double *op_full = nullptr)
{
    LuModeArgs a{};
    a.nx = p->nx;
    a.ny = p->ny;
    a.M = p->M;
    a.K = p->K;
    a.N = p->N;
    a.n = p->n;
    a.ne = p->ne;
    a.DzzT = p->d_DzzT;
    a.Dzz = p->d_Dzz;
    a.Dz = p->d_Dz;
    a.shift = p->d_shift;
    a.op_diag = op_diag;
    a.op_full = op_full;
    a.spec_re = p->spec.re;
    a.spec_im = p->spec.im;
    a.xs = p->xs;
    a.bc_re = p->bc.re;
    a.bc_im = p->bc.im;
    a.alpha_l = p->d.alpha_l;
    a.beta_l = p->d.beta_l;
    a.alpha_r = p->d.alpha_r;
    a.beta_r = p->d.beta_r;
    a.status = p->d_status;
    a.periodic = p->d.periodic_z;
    a.closure = p->d_closure;
    a.cols = p->lu_cols;
    return a;
}

// Plan-time autotuning of the reference-algorithm mode kernel: every
// instance that holds the plan (column capacity 16, 32, 64 or 128) is timed
// on noise (rounds over the candidates in turn, each one's best round of 3
// back-to-back launches kept) and the fastest kept.  One candidate: nothing
// to time.  The spectrum and face buffers are cleared afterwards.
int tune_lu(hz_plan *p)
{
    int cols[4];
    const int k = std::min(lu_modes_shapes(p->nx, p->n, p->ne, cols, 4), 4);
    p->lu_cols = k ? cols[0] : 0;
    p->lu_cols_ms.assign(k, 0.0);
    if (k < 2) return HZ_OK;
    const size_t nspec = 2 * (size_t)p->nx * p->xs, nbc = 4 * (size_t)p->nx * p->M;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int st = HZ_OK;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess)
        st = set_error(HZ_ERR_HIP, "tuning stream failed");
    if (!st) st = launch_fill_noise(p->spec.re, nspec, 6, s);
    if (!st) st = launch_fill_noise(p->bc.re, nbc, 7, s);
    std::vector<double> ms_of(k, 1e300);
    for (int round = 0; round < 3 && !st; ++round) {
        for (int i = 0; i < k && !st; ++i) {
            LuModeArgs a = lu_args(p);
            a.cols = cols[i];
            if (round == 0 && (st = launch_lu_modes(a, p->modes, s))) break;  // warm-up
            (void)hipEventRecord(e0, s);
            for (int r = 0; r < 3 && !st; ++r) st = launch_lu_modes(a, p->modes, s);
            (void)hipEventRecord(e1, s);
            if (hipEventSynchronize(e1) != hipSuccess) st = set_error(HZ_ERR_HIP, "tuning run failed");
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            ms_of[i] = std::min(ms_of[i], ms / 3.0);
        }
    }
    if (s && hipStreamSynchronize(s) != hipSuccess && !st) st = set_error(HZ_ERR_HIP, "tuning run failed");
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (!st) {
        double best = 1e300;
        for (int i = 0; i < k; ++i) {
            p->lu_cols_ms[i] = ms_of[i];
            if (ms_of[i] < best) {
                best = ms_of[i];
                p->lu_cols = cols[i];
            }
        }
    }
    if (!st && (hipMemset(p->spec.re, 0, sizeof(double) * nspec) != hipSuccess ||
                hipMemset(p->bc.re, 0, sizeof(double) * nbc) != hipSuccess))
        st = set_error(HZ_ERR_HIP, "tuning buffer clear failed");
    return st;
}

// Plan-time autotuning of the diagonalised mode kernel's tile height: each
// choice is timed (one warm-up, three timed launches) on the plan's own
// (zeroed) spectrum and the fastest kept.
int tune_diag(hz_plan *p)
{
    int choices[8];
    const int k = diag_tile_choices(p->NP, choices, 8);
    const size_t nspec = 2 * (size_t)p->nx * p->xs, nbc = 4 * (size_t)p->nx * p->M;
    hipStream_t s;
    hipEvent_t e0, e1;
    HZ_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    // timed on noise (zeros run faster); the buffers are cleared afterwards
    int fst = launch_fill_noise(p->spec.re, nspec, 1, s);
    if (!fst) fst = launch_fill_noise(p->bc.re, nbc, 2, s);
    if (fst) {
        (void)hipStreamDestroy(s);
        return fst;
    }
    HZ_HIP(hipEventCreate(&e0));
    HZ_HIP(hipEventCreate(&e1));
    double best = 1e300;
    int st = HZ_OK;
    std::vector<double> ms_of(k, 1e300);
    // rounds over the choices in turn (so that clock drift hits all of them
    // alike), each choice's best round of 3 back-to-back launches kept
    for (int round = 0; round < 3 && !st; ++round) {
        for (int i = 0; i < k && !st; ++i) {
            if (choices[i] < 2 * p->ne) continue;  // a tile must hold one column's elements
            DiagArgs a = diag_args(p);
            a.rt = choices[i];
            if (round == 0 && (st = launch_diag_modes(a, s))) break;  // warm-up
            (void)hipEventRecord(e0, s);
            for (int r = 0; r < 3 && !st; ++r) st = launch_diag_modes(a, s);
            (void)hipEventRecord(e1, s);
            if (hipEventSynchronize(e1) != hipSuccess) st = set_error(HZ_ERR_HIP, "tuning run failed");
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            ms_of[i] = std::min(ms_of[i], ms / 3.0);
        }
    }
    p->diag_rt_ms.assign(k, 0.0);
    for (int i = 0; i < k && !st; ++i) {
        if (ms_of[i] == 1e300) continue;
        p->diag_rt_ms[i] = ms_of[i];
        if (ms_of[i] < best) {
            best = ms_of[i];
            p->diag_rt = choices[i];
        }
    }
    if (hipStreamSynchronize(s) != hipSuccess && !st) st = set_error(HZ_ERR_HIP, "tuning run failed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    if (!st && (hipMemset(p->spec.re, 0, sizeof(double) * nspec) != hipSuccess ||
                hipMemset(p->bc.re, 0, sizeof(double) * nbc) != hipSuccess))
        st = set_error(HZ_ERR_HIP, "tuning buffer clear failed");
    if (!st && best == 1e300)
        st = set_error(HZ_ERR_UNSUPPORTED, "DIAG form: 2*ne must fit one tile (ne <= 64 when N-2 <= 64, "
                                           "32 when N-2 <= 128, 16 when N-2 <= 256)");
    return st;
}

// Plan-time autotuning of the transform passes' lane width: each choice of
// each pass is timed (a warm-up, then four rounds over the choices of three
// timed launches each, the best round kept) on scratch buffers
// of the plan's shape and the fastest kept.
int tune_fft(hz_plan *p)
{
    const size_t phys = (size_t)p->nx * p->ny * p->K;
    double *in = nullptr;
    if (hipMalloc((void **)&in, sizeof(double) * phys) != hipSuccess)
        return set_error(HZ_ERR_ALLOC, "hipMalloc: tuning buffer");
    int st = HZ_OK;
    const size_t nspec = 2 * (size_t)p->nx * p->xs, nbc = 4 * (size_t)p->nx * p->M;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess)
        st = set_error(HZ_ERR_HIP, "tuning stream failed");
    // timed on noise (zeros run faster); the plan's buffers are cleared afterwards
    if (!st) st = launch_fill_noise(in, phys, 3, s);
    if (!st) st = launch_fill_noise(p->spec.re, nspec, 4, s);
    if (!st) st = launch_fill_noise(p->bc.re, nbc, 5, s);
    static const FftKind kinds[4] = {FFT_R2C, FFT_C2C_FWD, FFT_C2C_INV, FFT_C2R};
    for (int pass = 0; pass < 4 && !st; ++pass) {
        const int L = (pass == 0 || pass == 3) ? p->ny : p->nx;
        int w[4];
        const int k = fft_width_choices(L, w, 4);
        // rounds over the choices in turn (so that clock drift hits all of
        // them alike), each choice's best round kept
        double ms_of[4] = {1e300, 1e300, 1e300, 1e300};
        for (int round = 0; round < 4 && !st; ++round) {
            for (int i = 0; i < k && !st; ++i) {
                p->fft_w[pass] = w[i];
                // the faces ride along as in a solve (passes 0 and 1)
                FftArgs a = xy_pass_args(p, pass, in, p->spec.re, p->spec.im, in, p->K, pass < 2 ? in : nullptr,
                                         pass < 2 ? in : nullptr, nullptr, nullptr, p->xs);
                if (round == 0 && (st = launch_fft(L, kinds[pass], a, s))) break;  // warm-up
                (void)hipEventRecord(e0, s);
                for (int r = 0; r < 3 && !st; ++r) st = launch_fft(L, kinds[pass], a, s);
                (void)hipEventRecord(e1, s);
                if (hipEventSynchronize(e1) != hipSuccess) st = set_error(HZ_ERR_HIP, "tuning run failed");
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                ms_of[i] = std::min(ms_of[i], (double)ms);
            }
        }
        int bestw = 0;
        double best = 1e300;
        for (int i = 0; i < k; ++i)
            if (ms_of[i] < best) {
                best = ms_of[i];
                bestw = w[i];
            }
        p->fft_w[pass] = bestw;
        p->fft_w_ms[pass] = k ? best / 3.0 : 0.0;
    }
    if (s && hipStreamSynchronize(s) != hipSuccess && !st) st = set_error(HZ_ERR_HIP, "tuning run failed");
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    (void)hipFree(in);
    if (!st && (hipMemset(p->spec.re, 0, sizeof(double) * nspec) != hipSuccess ||
                hipMemset(p->bc.re, 0, sizeof(double) * nbc) != hipSuccess))
        st = set_error(HZ_ERR_HIP, "tuning buffer clear failed");
    return st;
}

// Records stage boundaries of one solve when profiling is on.
struct StageRecorder {
    hz_plan *p;
    hipStream_t s;
    std::vector<hipEvent_t> ev;
    std::vector<std::string> names;
    StageRecorder(hz_plan *plan, hipStream_t stream) : p(plan), s(stream)
    {
        if (p->profiling) mark("");
    }
    void mark(const char *stage_done)
    {
        if (!p->profiling) return;
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        (void)hipEventRecord(e, s);
        ev.push_back(e);
        if (stage_done[0]) names.push_back(stage_done);
    }
    ~StageRecorder()
    {
        if (!p->profiling || ev.size() < 2) {
            for (auto e : ev) (void)hipEventDestroy(e);
            return;
        }
        if (p->stage_names.empty()) p->stage_names = names;
        if (names == p->stage_names && p->ev_solves.size() < 4096) {
            p->ev_solves.push_back(ev);
        } else {
            for (auto e : ev) (void)hipEventDestroy(e);
        }
    }
};

}  // namespace

extern "C" {

const char *hz_last_error_string(void) { return g_last_error.c_str(); }
const char *hz_version(void) { return "hz 0.1 (gfx950)"; }

int hz_plan_create(const hz_plan_desc *desc, hz_plan **plan)
{
    if (!desc || !plan) return set_error(HZ_ERR_ARG, "null argument");
    *plan = nullptr;
    const hz_plan_desc &d = *desc;
    if (!is_pow2(d.nx) || !is_pow2(d.ny))
        return set_error(HZ_ERR_NOT_POW2, "N has to be 2^n for some integer n.");
    if (d.nx > 1024 || d.ny > 1024) return set_error(HZ_ERR_UNSUPPORTED, "nx, ny must be <= 1024");
    if (d.n < 3) return set_error(HZ_ERR_N_TOO_SMALL, "Number of points has to be more than or equal to 3.");
    if (d.ne < 1) return set_error(HZ_ERR_ARG, "ne must be >= 1");
    if (!(d.xr > d.xl) || !(d.yr > d.yl) || !(d.zr > d.zl)) return set_error(HZ_ERR_ARG, "empty domain");
    if (d.form < HZ_FORM_PLAIN_LU || d.form > HZ_FORM_MIXED) return set_error(HZ_ERR_ARG, "unknown form");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(HZ_ERR_NO_DEVICE, "no HIP device");

    hz_plan *p = new hz_plan();
    p->d = d;
    HZ_HIP(hipGetDevice(&p->device));
    p->nx = d.nx;
    p->ny = d.ny;
    p->N = d.n;
    p->n = d.n - 2;
    p->ne = d.ne;
    p->K = d.ne * (d.n - 1) + 1;
    p->M = d.ny / 2 + 1;
    p->modes = (d.nx / 2 + 1) * p->M;
    p->Lx = d.xr - d.xl;
    p->Ly = d.yr - d.yl;
    p->Lz = d.zr - d.zl;
    int st = plan_operators(p->N, p->ne, p->Lz, p->z, p->Dz, p->Dzz);
    if (st) {
        free_plan(p);
        return set_error(st, "GLL setup failed");
    }
    std::vector<double> DzzT((size_t)p->n * p->n);
    for (int i = 0; i < p->n; ++i)
        for (int j = 0; j < p->n; ++j) DzzT[(size_t)j * p->n + i] = p->Dzz[(size_t)(i + 1) * p->N + (j + 1)];
    p->xs = spectrum_xs(p->M, p->K);
    const size_t specN = (size_t)p->nx * p->xs, bcN = (size_t)2 * p->nx * p->M;
    if ((st = upload(&p->d_Dz, p->Dz)) || (st = upload(&p->d_Dzz, p->Dzz)) ||
        (st = upload(&p->d_DzzT, DzzT)) || (st = dmalloc(&p->d_shift, p->modes)) ||
        (st = upload(&p->d_twx, twiddles(p->nx))) || (st = upload(&p->d_twy, twiddles(p->ny))) ||
        (st = dmalloc(&p->spec.re, 2 * specN)) || (st = dmalloc(&p->bc.re, 2 * bcN)) ||
        (st = dmalloc(&p->d_status, p->modes)) ||
        (d.periodic_z && (st = dmalloc(&p->d_closure, 2 * (size_t)p->nx * p->M))) ||
        ((d.form == HZ_FORM_DIAG || d.form == HZ_FORM_MIXED) && (st = setup_diag(p))) ||
        (d.form == HZ_FORM_MIXED && (st = setup_mixed(p))) || (st = compute_shifts(p))) {
        free_plan(p);
        return st;
    }
    p->spec.im = p->spec.re + specN;
    p->bc.im = p->bc.re + bcN;
    HZ_HIP(hipMemset(p->d_status, 0, sizeof(int) * p->modes));
    if ((d.form == HZ_FORM_PLAIN_LU && ((st = lu_modes_check(p->nx, p->n, p->ne)) || (st = tune_lu(p)))) ||
        (st = tune_fft(p)) ||
        (d.form == HZ_FORM_DIAG && (st = tune_diag(p)))) {
        free_plan(p);
        return st;
    }
    // after tuning: the tuned mode-kernel runs write statuses, and closures
    // for a periodic plan
    HZ_HIP(hipMemset(p->d_status, 0, sizeof(int) * p->modes));
    if (p->d_closure) HZ_HIP(hipMemset(p->d_closure, 0, sizeof(double) * 2 * (size_t)p->nx * p->M));
    *plan = p;
    return HZ_OK;
}

int hz_plan_destroy(hz_plan *plan)
{
    free_plan(plan);
    return HZ_OK;
}

int hz_plan_nz_total(const hz_plan *p, int *nz)
{
    if (!p || !nz) return set_error(HZ_ERR_ARG, "null argument");
    *nz = p->K;
    return HZ_OK;
}

int hz_plan_z_nodes(const hz_plan *p, double *z)
{
    if (!p || !z) return set_error(HZ_ERR_ARG, "null argument");
    std::memcpy(z, p->z.data(), sizeof(double) * p->z.size());
    return HZ_OK;
}

int hz_plan_operators(const hz_plan *p, double *Dz, double *Dzz)
{
    if (!p) return set_error(HZ_ERR_ARG, "null argument");
    if (Dz) std::memcpy(Dz, p->Dz.data(), sizeof(double) * p->Dz.size());
    if (Dzz) std::memcpy(Dzz, p->Dzz.data(), sizeof(double) * p->Dzz.size());
    return HZ_OK;
}

int hz_plan_num_modes(const hz_plan *p, int *m)
{
    if (!p || !m) return set_error(HZ_ERR_ARG, "null argument");
    *m = p->modes;
    return HZ_OK;
}

int hz_plan_set_lambda(hz_plan *p, double lambda)
{
    if (!p) return set_error(HZ_ERR_ARG, "null argument");
    p->d.lambda = lambda;
    ++p->generation;
    return compute_shifts(p);
}

int hz_plan_set_robin(hz_plan *p, double alpha_l, double beta_l, double alpha_r, double beta_r)
{
    if (!p) return set_error(HZ_ERR_ARG, "null argument");
    p->d.alpha_l = alpha_l;
    p->d.beta_l = beta_l;
    p->d.alpha_r = alpha_r;
    p->d.beta_r = beta_r;
    ++p->generation;
    return HZ_OK;
}

static int solve_impl(hz_plan *p, const double *f, const double *gl, const double *gr,
                      const double *op_diag, const double *op_full, double *u, hipStream_t s,
                      const double *cubic = nullptr)
{
    if (!p || !f || !u) return set_error(HZ_ERR_ARG, "null argument");
    if (p->d.periodic_z) {
        gl = gr = nullptr;  // no face data in the periodic problem
    } else if (!gl || !gr) {
        return set_error(HZ_ERR_ARG, "null face data");
    }
    StageRecorder rec(p, s);
    int st = forward_xy(p, f, p->spec.re, p->spec.im, p->K, s, gl, gr, p->xs, cubic);
    rec.mark("fft_forward");
    if (st) return st;
    if (p->d.form == HZ_FORM_PLAIN_LU || op_diag || op_full) {
        st = launch_lu_modes(lu_args(p, op_diag, op_full), p->modes, s);
        rec.mark("lu_modes");
    } else if (p->d.form == HZ_FORM_DIAG) {
        DiagArgs a = diag_args(p);
        st = launch_diag_modes(a, s);
        rec.mark("diag_modes");
    } else if (p->d.form == HZ_FORM_MIXED) {
        MixedArgs a = mixed_args(p);
        st = launch_mixed_modes(a, s);
        rec.mark("mixed_modes");
    } else {
        st = set_error(HZ_ERR_UNSUPPORTED, "form not built yet");
    }
    if (st) return st;
    st = inverse_xy(p, p->spec.re, p->spec.im, u, p->K, s, nullptr, nullptr, p->xs);
    rec.mark("fft_inverse");
    return st;
}

int hz_plan_profile(hz_plan *p, int enable)
{
    if (!p) return set_error(HZ_ERR_ARG, "null argument");
    p->profiling = enable != 0;
    return HZ_OK;
}

int hz_plan_profile_read(hz_plan *p, double *stage_ms, int max_stages, int *nstages, int *nsolves,
                         char *names, int names_len)
{
    if (!p) return set_error(HZ_ERR_ARG, "null argument");
    const int ns = (int)p->stage_names.size();
    std::vector<double> acc(ns, 0.0);
    for (auto &v : p->ev_solves) {
        for (int i = 0; i < ns && i + 1 < (int)v.size(); ++i) {
            HZ_HIP(hipEventSynchronize(v[i + 1]));
            float ms = 0;
            HZ_HIP(hipEventElapsedTime(&ms, v[i], v[i + 1]));
            acc[i] += ms;
        }
    }
    const int count = (int)p->ev_solves.size();
    for (int i = 0; i < ns && i < max_stages; ++i) stage_ms[i] = count ? acc[i] / count : 0.0;
    if (nstages) *nstages = ns;
    if (nsolves) *nsolves = count;
    if (names && names_len > 0) {
        std::string j;
        for (int i = 0; i < ns; ++i) j += (i ? ";" : "") + p->stage_names[i];
        std::snprintf(names, names_len, "%s", j.c_str());
    }
    for (auto &v : p->ev_solves)
        for (auto e : v) (void)hipEventDestroy(e);
    p->ev_solves.clear();
    p->stage_names.clear();
    return HZ_OK;
}

int hz_solve(hz_plan *p, const double *f, const double *gl, const double *gr, double *u,
             hz_stream_t stream)
{
    return solve_impl(p, f, gl, gr, nullptr, nullptr, u, (hipStream_t)stream);
}

int hz_solve_cubic(hz_plan *p, const double *u, double a3, double a1, const double *gl, const double *gr,
                   double *out, hz_stream_t stream)
{
    const double c[2] = {a3, a1};
    return solve_impl(p, u, gl, gr, nullptr, nullptr, out, (hipStream_t)stream, c);
}

int hz_cubic_adjoint(hz_plan *p, const double *fbar, const double *u_in, const double *u_out, double a3, double a1,
                     double b1, double b3, double *ubar, double *acc, hz_stream_t stream)
{
    if (!p || !fbar || !u_in || !u_out || !ubar || !acc) return set_error(HZ_ERR_ARG, "null argument");
    int st;
    if (!p->d_partial && (st = dmalloc(&p->d_partial, kDotPartials))) return st;
    return launch_cubic_adjoint(fbar, u_in, u_out, (long long)p->nx * p->ny * p->K, a3, a1, b1, b3, ubar,
                                p->d_partial, acc, (hipStream_t)stream);
}

int hz_solve_operators(hz_plan *p, const double *f, const double *gl, const double *gr,
                       const double *op_diag, const double *op_full, double *u, hz_stream_t stream)
{
    if (!op_diag && !op_full) return set_error(HZ_ERR_ARG, "no operators given");
    return solve_impl(p, f, gl, gr, op_diag, op_full, u, (hipStream_t)stream);
}

int hz_solve_status(hz_plan *p, hz_stream_t stream)
{
    if (!p) return set_error(HZ_ERR_ARG, "null argument");
    HZ_HIP(hipStreamSynchronize((hipStream_t)stream));
    std::vector<int> st(p->modes);
    HZ_HIP(hipMemcpy(st.data(), p->d_status, sizeof(int) * p->modes, hipMemcpyDeviceToHost));
    for (int m = 0; m < p->modes; ++m)
        if (st[m] == HZ_ERR_HIP) return set_error(HZ_ERR_HIP, "PLAIN_LU: a mode's elimination stalled (wait bound)");
    for (int m = 0; m < p->modes; ++m)
        if (st[m]) return set_error(HZ_ERR_SINGULAR, "ERROR: Matrix is singular!");
    return HZ_OK;
}

// Adjoint of the solve map.  The x-y transform pair and the per-mode
// operators commute with translations in x and y, so the transpose of
// F^-1 M F is F^-1 M^T F: the same pipeline with the transposed column map
// (diag.hip, ADJ) in the middle.  A PLAIN_LU plan builds the diagonalised
// operators on its first adjoint call; both forms solve the same discrete
// system, so they share one adjoint.
static int adjoint_pipeline(hz_plan *p, const double *ubar, double *fbar, double *glbar, double *grbar,
                            hipStream_t s, StageRecorder *rec)
{
    int st;
    if ((st = forward_xy(p, ubar, p->spec.re, p->spec.im, p->K, s, nullptr, nullptr, p->xs))) return st;
    if (rec) rec->mark("fft_forward");
    DiagArgs a = diag_args(p);
    a.adjoint = 1;
    diag_basis(p, a, true);
    a.gbar_re = p->bc.re;
    a.gbar_im = p->bc.im;
    if ((st = launch_diag_modes(a, s))) return st;
    if (rec) rec->mark("diag_adjoint");
    if ((st = inverse_xy(p, p->spec.re, p->spec.im, fbar, p->K, s, glbar, grbar, p->xs))) return st;
    if (rec) rec->mark("fft_inverse");
    return HZ_OK;
}

static int adjoint_args(const hz_plan *p, const double *ubar, const double *u, const double *fbar,
                        const double *glbar, const double *grbar, const void *lambda_bar)
{
    if (!p || !ubar || !fbar) return set_error(HZ_ERR_ARG, "null argument");
    if (p->d.periodic_z ? (glbar || grbar) : (!glbar || !grbar))
        return set_error(HZ_ERR_ARG, "face gradients: required without periodic z, NULL with it");
    if (lambda_bar && !u) return set_error(HZ_ERR_ARG, "lambda_bar needs the forward solution u");
    return HZ_OK;
}

}  // extern "C"

namespace hz {
int adjoint_prebuilt(hz_plan *p, const double *ubar, const double *u, double *fbar, double *glbar, double *grbar,
                     double *lambda_bar_dev, double *partial, hipStream_t s)
{
    int st;
    if ((st = adjoint_args(p, ubar, u, fbar, glbar, grbar, lambda_bar_dev))) return st;
    if (!p->d_Vrow)
        return set_error(HZ_ERR_UNSUPPORTED, "the plan's diagonalised operators are not built yet (a PLAIN_LU plan "
                                             "builds them on its first hz_adjoint call)");
    if ((st = adjoint_pipeline(p, ubar, fbar, glbar, grbar, s, nullptr))) return st;
    if (lambda_bar_dev) return launch_dot(fbar, u, (long long)p->nx * p->ny * p->K, partial, lambda_bar_dev, s);
    return HZ_OK;
}
}  // namespace hz

extern "C" {

int hz_adjoint(hz_plan *p, const double *ubar, const double *u, double *fbar, double *glbar,
               double *grbar, double *lambda_bar, hz_stream_t stream)
{
    int st;
    if ((st = adjoint_args(p, ubar, u, fbar, glbar, grbar, lambda_bar))) return st;
    const hipStream_t s = (hipStream_t)stream;
    if (!p->d_Vrow) {
        if ((st = setup_diag(p)) || (st = compute_homogeneous(p))) return st;
    }
    StageRecorder rec(p, s);
    if ((st = adjoint_pipeline(p, ubar, fbar, glbar, grbar, s, &rec))) return st;
    if (lambda_bar) {
        if (!p->d_partial && (st = dmalloc(&p->d_partial, kDotPartials))) return st;
        // the final reduction writes straight into pinned host memory (no
        // separate device-to-host copy before the wait)
        if (!p->h_lambda_bar && hipHostMalloc((void **)&p->h_lambda_bar, sizeof(double), hipHostMallocCoherent) != hipSuccess) {
            p->h_lambda_bar = nullptr;
            return set_error(HZ_ERR_ALLOC, "hipHostMalloc: dL/dlambda word");
        }
        if ((st = launch_dot(fbar, u, (long long)p->nx * p->ny * p->K, p->d_partial, p->h_lambda_bar, s)))
            return st;
        HZ_HIP(hipStreamSynchronize(s));
        *lambda_bar = *(volatile double *)p->h_lambda_bar;
        rec.mark("lambda_dot");
    }
    return HZ_OK;
}

int hz_mixed_history(const hz_plan *p, double *rel_residual, int max_sweeps, int *sweeps)
{
    if (!p || !rel_residual || !sweeps) return set_error(HZ_ERR_ARG, "null argument");
    if (p->d.form != HZ_FORM_MIXED) return set_error(HZ_ERR_UNSUPPORTED, "not a MIXED plan");
    const int S = p->sweeps;
    const size_t rows = (size_t)p->nx * p->M * 2 * p->ne;
    std::vector<double> h((size_t)S * rows);
    HZ_HIP(hipDeviceSynchronize());
    if (S) HZ_HIP(hipMemcpy(h.data(), p->d_hist, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    for (int s = 0; s < S && s < max_sweeps; ++s) {
        double v = 0.0;
        for (size_t r = 0; r < rows; ++r) v = std::max(v, h[(size_t)s * rows + r]);
        rel_residual[s] = v;
    }
    *sweeps = std::min(S, max_sweeps);
    return HZ_OK;
}

int hz_plan_tuning(const hz_plan *p, int *diag_tile_rows, double *diag_ms, int max_diag, int *fft_lanes,
                   double *fft_ms)
{
    if (!p) return set_error(HZ_ERR_ARG, "null argument");
    const bool lu = p->d.form == HZ_FORM_PLAIN_LU;
    const std::vector<double> &ms = lu ? p->lu_cols_ms : p->diag_rt_ms;
    if (diag_tile_rows) *diag_tile_rows = lu ? p->lu_cols : p->diag_rt;
    if (diag_ms)
        for (int i = 0; i < (int)ms.size() && i < max_diag; ++i) diag_ms[i] = ms[i];
    for (int pass = 0; pass < 4; ++pass) {
        if (fft_lanes) fft_lanes[pass] = p->fft_w[pass];
        if (fft_ms) fft_ms[pass] = p->fft_w_ms[pass];
    }
    return HZ_OK;
}

int hz_plan_diag_candidates(const hz_plan *p, int *tile_rows, int max, int *count)
{
    if (!p || !count) return set_error(HZ_ERR_ARG, "null argument");
    int c[8];
    const int k = p->d.form == HZ_FORM_DIAG        ? diag_tile_choices(p->NP, c, 8)
                  : p->d.form == HZ_FORM_PLAIN_LU ? std::min(lu_modes_shapes(p->nx, p->n, p->ne, c, 8), 8)
                                                   : 0;
    for (int i = 0; i < k && i < max && tile_rows; ++i) tile_rows[i] = c[i];
    *count = k;
    return HZ_OK;
}

int hz_plan_set_diag_tile(hz_plan *p, int tile_rows)
{
    if (!p) return set_error(HZ_ERR_ARG, "null argument");
    int c[8];
    if (p->d.form == HZ_FORM_PLAIN_LU) {  // the LU instance, by its column capacity
        const int k = std::min(lu_modes_shapes(p->nx, p->n, p->ne, c, 8), 8);
        for (int i = 0; i < k; ++i)
            if (c[i] == tile_rows) {
                p->lu_cols = tile_rows;
                return HZ_OK;
            }
        return set_error(HZ_ERR_ARG, "not a candidate mode-kernel shape for this plan");
    }
    if (p->d.form != HZ_FORM_DIAG) return set_error(HZ_ERR_UNSUPPORTED, "not a DIAG or PLAIN_LU plan");
    const int k = diag_tile_choices(p->NP, c, 8);
    bool known = false;
    for (int i = 0; i < k; ++i) known = known || c[i] == tile_rows;
    if (!known) return set_error(HZ_ERR_ARG, "not a candidate tile shape for this plan");
    if (tile_rows < 2 * p->ne) return set_error(HZ_ERR_ARG, "the tile does not hold one column's elements");
    p->diag_rt = tile_rows;
    return HZ_OK;
}

int hz_plan_retune(hz_plan *p)
{
    if (!p) return set_error(HZ_ERR_ARG, "null argument");
    // the tuners time on the plan's own spectrum and face buffers
    HZ_HIP(hipDeviceSynchronize());
    int st = tune_fft(p);
    if (!st && p->d.form == HZ_FORM_DIAG) st = tune_diag(p);
    if (!st && p->d.form == HZ_FORM_PLAIN_LU) st = tune_lu(p);
    // the tuned runs of the mode kernels leave statuses and closures of noise
    HZ_HIP(hipMemset(p->d_status, 0, sizeof(int) * p->modes));
    if (p->d_closure) HZ_HIP(hipMemset(p->d_closure, 0, sizeof(double) * 2 * (size_t)p->nx * p->M));
    return st;
}

int hz_plan_periodic_closure(const hz_plan *p, double *alpha)
{
    if (!p || !alpha) return set_error(HZ_ERR_ARG, "null argument");
    if (!p->d_closure) return set_error(HZ_ERR_UNSUPPORTED, "not a periodic-z plan");
    HZ_HIP(hipDeviceSynchronize());
    HZ_HIP(hipMemcpy(alpha, p->d_closure, sizeof(double) * 2 * (size_t)p->nx * p->M, hipMemcpyDeviceToHost));
    return HZ_OK;
}

int hz_mixed_fp64_modes(const hz_plan *p, int *fp64_modes)
{
    if (!p || !fp64_modes) return set_error(HZ_ERR_ARG, "null argument");
    *fp64_modes = 0;  // every mode runs the FP32 eigenbasis solve (its refinement converges for all shifts)
    return HZ_OK;
}

// Fourier derivative along x or y: one R2C pass along that axis into a half
// spectrum, the wavenumber multiply, one C2R pass back (two passes over the
// field, not the four of the 2-D transform).
int hz_derivative_xy(hz_plan *p, int axis, int order, const double *u, double *out, int lanes,
                     hz_stream_t stream)
{
    if (!p || !u || !out || lanes < 1 || (axis != 0 && axis != 1) || (order != 1 && order != 2))
        return set_error(HZ_ERR_ARG, "bad argument");
    const hipStream_t s = (hipStream_t)stream;
    const long long W = lanes;
    const long long A = axis == 0 ? p->nx / 2 + 1 : p->nx;
    const long long B = axis == 0 ? p->ny : p->M;
    const size_t half = (size_t)A * B * W;
    int st;
    if (2 * half > p->xspec_n) {
        (void)hipFree(p->d_xspec);
        p->d_xspec = nullptr;
        p->xspec_n = 0;
        if ((st = dmalloc(&p->d_xspec, 2 * half))) return st;
        p->xspec_n = 2 * half;
    }
    double *re = p->d_xspec, *im = p->d_xspec + half;
    FftArgs a{};
    int L;
    if (axis == 0) {  // lines along x, batch = y index; same strides on both sides
        L = p->nx;
        a.in_batch_stride = a.out_batch_stride = W;
        a.in_point_stride = a.out_point_stride = (long long)p->ny * W;
        a.nbatch = p->ny;
        a.tw = p->d_twx;
    } else {          // lines along y, batch = x index
        L = p->ny;
        a.in_batch_stride = (long long)p->ny * W;
        a.in_point_stride = W;
        a.out_batch_stride = (long long)p->M * W;
        a.out_point_stride = W;
        a.nbatch = p->nx;
        a.tw = p->d_twy;
    }
    a.lanes = lanes;
    FftArgs b = a;
    a.in_re = u;
    a.out_re = re;
    a.out_im = im;
    if ((st = launch_fft(L, FFT_R2C, a, s))) return st;
    const double len = axis == 0 ? p->Lx : p->Ly;
    if ((st = launch_spectral_derivative(re, im, A, B, W, axis == 0, L, len, order, s))) return st;
    b.in_re = re;  // C2R: the half spectrum in, the physical layout out
    b.in_im = im;
    b.out_re = out;
    std::swap(b.in_batch_stride, b.out_batch_stride);
    std::swap(b.in_point_stride, b.out_point_stride);
    return launch_fft(L, FFT_C2R, b, s);
}

int hz_host_plan_operators(int n, int ne, double lz, double *x, double *z, double *Dz, double *Dzz)
{
    if (n < 2) return set_error(HZ_ERR_N_TOO_SMALL, "Number of points has to be more than or equal to 2.");
    if (ne < 1 || !(lz > 0)) return set_error(HZ_ERR_ARG, "bad element count or length");
    std::vector<double> xi, zz, a, b;
    int st = gll_nodes(n, xi);
    if (st) return set_error(st, "GLL setup failed");
    st = plan_operators(n, ne, lz, zz, a, b);
    if (st) return set_error(st, "GLL setup failed");
    if (x) std::memcpy(x, xi.data(), sizeof(double) * xi.size());
    if (z) std::memcpy(z, zz.data(), sizeof(double) * zz.size());
    if (Dz) std::memcpy(Dz, a.data(), sizeof(double) * a.size());
    if (Dzz) std::memcpy(Dzz, b.data(), sizeof(double) * b.size());
    return HZ_OK;
}

int hz_host_eigen(int n, int ne, double lz, double *mu, double *V, double *Vinv)
{
    if (n < 3) return set_error(HZ_ERR_N_TOO_SMALL, "need n >= 3");
    std::vector<double> zz, a, b;
    int st = plan_operators(n, ne, lz, zz, a, b);
    if (st) return set_error(st, "GLL setup failed");
    const int m = n - 2;
    std::vector<double> A((size_t)m * m), ev, vv, vi;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) A[(size_t)i * m + j] = b[(size_t)(i + 1) * n + (j + 1)];
    st = eigen_decompose(A, m, ev, vv, vi);
    if (st) return set_error(st, "eigendecomposition did not converge");
    if (mu) std::memcpy(mu, ev.data(), sizeof(double) * m);
    if (V) std::memcpy(V, vv.data(), sizeof(double) * vv.size());
    if (Vinv) std::memcpy(Vinv, vi.data(), sizeof(double) * vi.size());
    return HZ_OK;
}

int hz_host_eigen_split(int n, int ne, double lz, int np, double *mu, double *W, double *Winv, double *asym)
{
    if (n < 3) return set_error(HZ_ERR_N_TOO_SMALL, "need n >= 3");
    const int m = n - 2;
    if (np % 2 || np / 2 < m - m / 2) return set_error(HZ_ERR_ARG, "np must be even with np/2 >= ceil((n-2)/2)");
    std::vector<double> zz, a, b;
    int st = plan_operators(n, ne, lz, zz, a, b);
    if (st) return set_error(st, "GLL setup failed");
    std::vector<double> A((size_t)m * m), ev, w, wi;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) A[(size_t)i * m + j] = b[(size_t)(i + 1) * n + (j + 1)];
    double as = 0;
    st = eigen_decompose_split(A, m, np, ev, w, wi, &as);
    if (st) return set_error(st, "even/odd eigendecomposition failed");
    if (mu) std::memcpy(mu, ev.data(), sizeof(double) * ev.size());
    if (W) std::memcpy(W, w.data(), sizeof(double) * w.size());
    if (Winv) std::memcpy(Winv, wi.data(), sizeof(double) * wi.size());
    if (asym) *asym = as;
    return HZ_OK;
}

int hz_fourier_xy(hz_plan *p, int isign, const double *in, double *re, double *im, double *out,
                  int lanes, hz_stream_t stream)
{
    if (!p || !re || !im || lanes < 1) return set_error(HZ_ERR_ARG, "bad argument");
    if (isign == 1) {
        if (!in) return set_error(HZ_ERR_ARG, "null input");
        return forward_xy(p, in, re, im, lanes, (hipStream_t)stream);
    }
    if (isign == -1) {
        if (!out) return set_error(HZ_ERR_ARG, "null output");
        return inverse_xy(p, re, im, out, lanes, (hipStream_t)stream);
    }
    return set_error(HZ_ERR_ARG, "isign must be +1 or -1");
}

int hz_lu_factor_batched(double *A, int n, int count, int *piv, double *det, int *status,
                         hz_stream_t stream)
{
    if (!A || n < 1 || count < 0) return set_error(HZ_ERR_ARG, "bad argument");
    LuFactorArgs a{A, n, count, piv, det, status};
    return launch_lu_factor(a, (hipStream_t)stream);
}

int hz_lu_solve_batched(const double *LU, const int *piv, double *b, int n, int count,
                        hz_stream_t stream)
{
    if (!LU || !piv || !b || n < 1) return set_error(HZ_ERR_ARG, "bad argument");
    return launch_lu_solve(LU, piv, b, n, count, (hipStream_t)stream);
}

}  // extern "C"
