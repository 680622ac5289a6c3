// mixed.hip -- the MIXED form: FP32 solution operators with FP64 residual
// refinement, FP64 operators for the ill-conditioned low-|k| modes.
//
// Per (|kx|,|ky|) mode the interior system is A x = b with
// A = Dzz_int - sigma I (the reference's operator, TestNonPeriodic3D_MPI.cpp:71-79).
// Plan time (once per lambda): X = A^{-1} for every mode, formed in FP64 from
// the plan's eigen-decomposition Dzz_int = V diag(mu) V^{-1},
//   X = V diag(1/(mu - sigma)) V^{-1},
// and stored in FP32, except for modes whose condition number makes FP32
// refinement slow (cond * 2^-24 above a threshold), which keep FP64.  The
// homogeneous solutions ul = V (w0/(mu-sigma)), ur = V (wN/(mu-sigma)) of
// NonPeriodic3D_MPI.cpp:203-222 are stored in FP64.
//
// Per solve, one work-group per mode, for every real column of the mode
// (x index +-iw, re/im, element):
//   x = X fl32(b)                          FP32 products, FP32 sums
//   repeat S times:  r = b - A x           FP64 (Dzz_int from L2)
//                    x = x + X fl32(r)     FP32 products, FP64 update
// then the end-point derivatives, the element-interface sweep with the Robin
// rows (NonPeriodic3D_MPI.cpp:327-580) and the assembly u = x + zl ul + zr ur
// (:582-587) straight into the spectrum.  The relative max-norm residual of
// every mode after the solve and after each sweep goes to `hist`.
//
// The FP32 operator is read from HBM once per solve and stays in registers
// for all sweeps: thread (i, h) holds X[i][64h .. 64h+63].  The FP64
// residual streams Dzz_int from L2 (holding it in registers as well costs
// more in occupancy than it saves, measured).  FP64 operators are streamed
// from memory at each application (few modes).
#include "hz_internal.h"
#include "sweep.h"

namespace hz {
namespace {

constexpr int WAVE = 64;
constexpr int JC = 64;  // operator columns per thread

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// ---------------------------------------------------------------------------
// Plan time: X^T (i fastest, row stride ldx) and ul, ur of every mode.
__global__ __launch_bounds__(256) void mixed_operator_kernel(const MixedSetupArgs a)
{
    __shared__ double inv[256];
    const int m = blockIdx.x, tid = threadIdx.x;
    const int n = a.n, NP = a.NP, ldx = a.ldx;
    const double sig = a.shift[m];
    for (int k = tid; k < NP; k += 256) inv[k] = k < n ? 1.0 / (a.mu[k] - sig) : 0.0;
    __syncthreads();
    for (int i = tid; i < n; i += 256) {
        double ul = 0.0, ur = 0.0;
        for (int k = 0; k < n; ++k) {
            const double v = a.VT[(long long)k * NP + i] * inv[k];
            ul = fma(v, a.w0[k], ul);
            ur = fma(v, a.wN[k], ur);
        }
        a.UL[(long long)m * n + i] = ul;
        a.UR[(long long)m * n + i] = ur;
    }
    const int slot = a.slot[m];
    float *x32 = slot >= 0 ? a.X32 + (long long)slot * n * ldx : nullptr;
    double *x64 = slot < 0 ? a.X64 + (long long)(-slot - 1) * n * ldx : nullptr;
    // X[i][j] = sum_k V[i][k] inv[k] Vinv[k][j]; VT[k][i] = V[i][k] and
    // Vinv rows are padded with zeros to NP, so 4x4 tiles read in bounds.
    const int ti = (n + 3) / 4;
    for (int t = tid; t < ti * ti; t += 256) {
        const int i0 = 4 * (t % ti), j0 = 4 * (t / ti);
        double acc[4][4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[r][c] = 0.0;
        for (int k = 0; k < n; ++k) {
            double vi[4], vj[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) vi[r] = a.VT[(long long)k * NP + i0 + r] * inv[k];
#pragma unroll
            for (int c = 0; c < 4; ++c) vj[c] = a.Vinv[(long long)k * NP + j0 + c];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[r][c] = fma(vi[r], vj[c], acc[r][c]);
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int j = j0 + c;
            if (j >= n) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + r;
                if (i >= ldx) continue;
                const double v = i < n ? acc[r][c] : 0.0;
                if (x32) x32[(long long)j * ldx + i] = (float)v;
                else x64[(long long)j * ldx + i] = v;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Per solve.
template <int NPAD, int CMAX>
struct MixedSmem {
    static constexpr int H = NPAD / JC;
    double b[CMAX][NPAD];
    double x[CMAX][NPAD];
    double r[CMAX][NPAD];
    float r32[CMAX][NPAD];
    double part[CMAX][NPAD];
    double red[2][NPAD * H / WAVE][CMAX > 4 ? CMAX : 4];
    double dots[2][CMAX];
    double zl[CMAX], zr[CMAX];
    double bmax, rmax;
};

// part[c][i] = sum over the H column chunks of acc (deterministic order)
template <int NPAD, int CMAX, typename T>
__device__ __forceinline__ void combine(MixedSmem<NPAD, CMAX> &s, const T (&acc)[CMAX], int i, int h, int C)
{
    constexpr int H = NPAD / JC;
    for (int hh = 0; hh < H; ++hh) {
        if (h == hh) {
#pragma unroll
            for (int c = 0; c < CMAX; ++c)
                if (c < C) s.part[c][i] = (hh == 0 ? 0.0 : s.part[c][i]) + (double)acc[c];
        }
        __syncthreads();
    }
}

template <int NPAD, int CMAX>
__global__ __launch_bounds__(NPAD *(NPAD / JC)) void mixed_modes_kernel(const MixedArgs a)
{
    constexpr int H = NPAD / JC;
    constexpr int NT = NPAD * H;
    __shared__ MixedSmem<NPAD, CMAX> s;
    const int tid = threadIdx.x, i = tid % NPAD, h = tid / NPAD;
    const int lane = tid % WAVE, w = tid / WAVE;
    const int n = a.n, N = a.N, E = a.ne, K = a.K, M = a.M, ldx = a.ldx;
    const int mode = blockIdx.x;
    const int iw = mode / M, jw = mode % M;
    const int xi0 = iw, xi1 = a.nx - iw, nxi = (iw != 0 && iw != a.nx / 2) ? 2 : 1;
    const int nc = 2 * nxi, C = nc * E;
    const double sig = a.shift[mode];
    const int slot = uniform(a.slot[mode]);
    const int j0 = h * JC;

    // column c = e * nc + cc: element e, x index (cc >> 1 ? xi1 : xi0), plane cc & 1
    auto col_off = [&](int c) -> long long {
        const int e = c / nc, cc = c % nc;
        const int xi = (cc >> 1) ? xi1 : xi0;
        return (long long)xi * a.xs + (long long)jw * K + (long long)e * (N - 1) + ((cc & 1) ? a.plane_off : 0);
    };

    // ---- load b; FP32 operator into registers ----
    for (int idx = tid; idx < CMAX * NPAD; idx += NT) {
        const int c = idx / NPAD, k = idx % NPAD;
        double v = 0.0;
        if (c < C && k < n) v = a.spec[col_off(c) + 1 + k];
        s.b[c][k] = v;
        s.x[c][k] = 0.0;
        s.r[c][k] = v;
        s.r32[c][k] = (float)v;
    }
    float xr[JC];
    if (slot >= 0) {
        const float *X = a.X32 + (long long)slot * n * ldx;
#pragma unroll
        for (int jj = 0; jj < JC; ++jj) {
            const int j = j0 + jj;
            xr[jj] = (i < n && j < n) ? X[(long long)j * ldx + i] : 0.0f;
        }
    }
    __syncthreads();
    // |b| max for the relative residual
    {
        double bm = 0.0;
        for (int idx = tid; idx < CMAX * NPAD; idx += NT) bm = fmax(bm, fabs(s.b[idx / NPAD][idx % NPAD]));
        for (int o = 1; o < WAVE; o <<= 1) bm = fmax(bm, __shfl_xor(bm, o));
        if (lane == 0) s.red[0][w][0] = bm;
        __syncthreads();
        if (tid == 0) {
            double v = 0.0;
            for (int q = 0; q < NT / WAVE; ++q) v = fmax(v, s.red[0][q][0]);
            s.bmax = v;
        }
        __syncthreads();
    }

    const int S = a.sweeps;
    for (int sw = 0; sw <= S; ++sw) {
        // ---- x += X r ----
        if (slot >= 0) {
            float acc[CMAX];
#pragma unroll
            for (int c = 0; c < CMAX; ++c) acc[c] = 0.0f;
#pragma unroll
            for (int jj = 0; jj < JC; ++jj) {
#pragma unroll
                for (int c = 0; c < CMAX; ++c)
                    if (c < C) acc[c] = fmaf(xr[jj], s.r32[c][j0 + jj], acc[c]);
            }
            combine<NPAD, CMAX>(s, acc, i, h, C);
        } else {
            const double *X = a.X64 + (long long)(-slot - 1) * n * ldx;
            double acc[CMAX];
#pragma unroll
            for (int c = 0; c < CMAX; ++c) acc[c] = 0.0;
            if (i < n) {
                const int jend = min(n, j0 + JC);
                for (int j = j0; j < jend; ++j) {
                    const double xv = X[(long long)j * ldx + i];
#pragma unroll
                    for (int c = 0; c < CMAX; ++c)
                        if (c < C) acc[c] = fma(xv, s.r[c][j], acc[c]);
                }
            }
            combine<NPAD, CMAX>(s, acc, i, h, C);
        }
        if (h == 0 && i < n) {
#pragma unroll
            for (int c = 0; c < CMAX; ++c)
                if (c < C) s.x[c][i] += s.part[c][i];
        }
        __syncthreads();
        // ---- r = b - (Dzz_int x - sigma x), FP64 ----
        {
            double acc[CMAX];
#pragma unroll
            for (int c = 0; c < CMAX; ++c) acc[c] = 0.0;
            if (i < n) {
                const int jend = min(n, j0 + JC);
                for (int j = j0; j < jend; ++j) {
                    const double dv = a.DzzT[(long long)j * n + i];
#pragma unroll
                    for (int c = 0; c < CMAX; ++c)
                        if (c < C) acc[c] = fma(dv, s.x[c][j], acc[c]);
                }
            }
            combine<NPAD, CMAX>(s, acc, i, h, C);
        }
        double rm = 0.0;
        if (h == 0 && i < n) {
#pragma unroll
            for (int c = 0; c < CMAX; ++c) {
                if (c < C) {
                    const double rv = s.b[c][i] - (s.part[c][i] - sig * s.x[c][i]);
                    s.r[c][i] = rv;
                    s.r32[c][i] = (float)rv;
                    rm = fmax(rm, fabs(rv));
                }
            }
        }
        for (int o = 1; o < WAVE; o <<= 1) rm = fmax(rm, __shfl_xor(rm, o));
        if (lane == 0) s.red[1][w][0] = rm;
        __syncthreads();
        if (tid == 0) {
            double v = 0.0;
            for (int q = 0; q < NT / WAVE; ++q) v = fmax(v, s.red[1][q][0]);
            a.hist[(long long)sw * a.modes + mode] = s.bmax > 0.0 ? v / s.bmax : 0.0;
        }
    }
    __syncthreads();

    // ---- end-point derivatives of the particular solutions ----
    for (int item = w; item < 2 * C; item += NT / WAVE) {
        const int c = item >> 1, which = item & 1;
        const double *drow = a.Dz + (which ? (long long)(N - 1) * N : 0) + 1;
        double d = 0.0;
        for (int k = lane; k < n; k += WAVE) d = fma(drow[k], s.x[c][k], d);
        for (int o = 1; o < WAVE; o <<= 1) d += __shfl_xor(d, o);
        if (lane == 0) s.dots[which][c] = d;
    }
    __syncthreads();

    // ---- element-interface sweep, one thread per real column chain ----
    if (tid < nc && a.periodic) {  // cyclic system, no face data (sweep.h)
        const int cc = tid;
        const long long fc = (long long)((cc >> 1) ? xi1 : xi0) * M + jw;
        periodic_sweep(E, a.hom[4 * mode + 0], a.hom[4 * mode + 1], a.hom[4 * mode + 2], a.hom[4 * mode + 3],
                       &s.dots[0][cc], &s.dots[1][cc], &s.zl[cc], &s.zr[cc], nc,
                       a.closure ? a.closure + ((cc & 1) ? (long long)a.nx * M : 0) + fc : nullptr);
    } else if (tid < nc) {
        const int cc = tid;
        const long long bcol = (long long)((cc >> 1) ? xi1 : xi0) * M + jw;
        const double *bsrc = (cc & 1) ? a.bc_im : a.bc_re;
        const double gl = bsrc[bcol], gr = bsrc[(long long)a.nx * M + bcol];
        const double d0ul = a.hom[4 * mode + 0], dNul = a.hom[4 * mode + 1];
        const double d0ur = a.hom[4 * mode + 2], dNur = a.hom[4 * mode + 3];
        double bprev = -a.alpha_l * d0ul + a.beta_l;
        double cprev = -a.alpha_l * d0ur;
        double rprev = gl + a.alpha_l * s.dots[0][cc];
        s.zl[cc] = bprev;  // forward elimination's b and rhs kept in zl / zr
        s.zr[cc] = rprev;
        for (int e = 1; e < E; ++e) {
            const double lb = dNul;
            const double be = dNur - d0ul - (cprev / bprev) * lb;
            const double re = s.dots[0][e * nc + cc] - s.dots[1][(e - 1) * nc + cc] - (rprev / bprev) * lb;
            bprev = be;
            cprev = -d0ur;
            rprev = re;
            s.zl[e * nc + cc] = be;
            s.zr[e * nc + cc] = re;
        }
        const double lb = a.alpha_r * dNul;
        const double br = a.alpha_r * dNur + a.beta_r - (cprev / bprev) * lb;
        const double rhsr = gr - a.alpha_r * s.dots[1][(E - 1) * nc + cc] - (rprev / bprev) * lb;
        double zr = rhsr / br;
        for (int e = E - 1; e >= 0; --e) {
            const double be = s.zl[e * nc + cc], re = s.zr[e * nc + cc];
            const double ce = (e == 0) ? -a.alpha_l * d0ur : -d0ur;
            const double zl = (re - ce * zr) / be;
            s.zl[e * nc + cc] = zl;
            s.zr[e * nc + cc] = zr;
            zr = zl;
        }
    }
    __syncthreads();

    // ---- assembly into the spectrum ----
    const double *ul = a.UL + (long long)mode * n, *ur = a.UR + (long long)mode * n;
    for (int idx = tid; idx < C * NPAD; idx += NT) {
        const int c = idx / NPAD, k = idx % NPAD;
        const long long off = col_off(c);
        if (k < n) a.spec[off + 1 + k] = s.x[c][k] + (s.zl[c] * ul[k] + s.zr[c] * ur[k]);
        if (k == 0) {
            a.spec[off] = s.zl[c];
            a.spec[off + N - 1] = s.zr[c];
        }
    }
}

template <int NPAD, int CMAX>
int launch_mixed_t(const MixedArgs &a, hipStream_t s)
{
    constexpr int NT = NPAD * (NPAD / JC);
    hipLaunchKernelGGL((mixed_modes_kernel<NPAD, CMAX>), dim3(a.modes), dim3(NT), 0, s, a);
    HZ_HIP(hipGetLastError());
    return HZ_OK;
}

template <int NPAD>
int launch_mixed_c(const MixedArgs &a, hipStream_t s)
{
    const int cmax = 4 * a.ne;
    if (cmax <= 4) return launch_mixed_t<NPAD, 4>(a, s);
    if (cmax <= 8) return launch_mixed_t<NPAD, 8>(a, s);
    if (cmax <= 16) return launch_mixed_t<NPAD, 16>(a, s);
    return set_error(HZ_ERR_UNSUPPORTED, "MIXED form: at most 4 elements");
}

}  // namespace

int launch_mixed_setup(const MixedSetupArgs &a, hipStream_t s)
{
    if (a.modes <= 0) return HZ_OK;
    hipLaunchKernelGGL(mixed_operator_kernel, dim3(a.modes), dim3(256), 0, s, a);
    HZ_HIP(hipGetLastError());
    return HZ_OK;
}

int launch_mixed_modes(const MixedArgs &a, hipStream_t s)
{
    if (a.n <= 64) return launch_mixed_c<64>(a, s);
    if (a.n <= 128) return launch_mixed_c<128>(a, s);
    if (a.n <= 256) return launch_mixed_c<256>(a, s);
    return set_error(HZ_ERR_UNSUPPORTED, "MIXED form: N-2 must be <= 256");
}

}  // namespace hz
