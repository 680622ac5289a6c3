// fft_body.h -- device bodies of the batched line FFTs (fft.hip), shared
// with the fused batch kernel (batch.hip).  See fft.hip for the scheme.
#pragma once

#include "hz_internal.h"

namespace hz {
namespace {

constexpr int bitrev(int i, int bits)
{
    int r = 0;
    for (int b = 0; b < bits; ++b) r |= ((i >> b) & 1) << (bits - 1 - b);
    return r;
}
constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v >> 1); }

constexpr int kXcds = 8;  // MI355X: 8 accelerator complex dies, one L2 each

// XCD-aware tile order: work-groups are dealt round-robin to the 8 XCDs, so
// XCD x is given the contiguous logical range [x*T/8, (x+1)*T/8) of tiles
// (lane block fastest).  Neighbouring lane blocks of one line batch share
// the cache lines their 8*W-byte row segments straddle (K is odd), and they
// then share one L2.
__device__ __forceinline__ int xcd_logical(int id, int total)
{
    const int xcd = id % kXcds, full = total / kXcds, rem = total % kXcds;
    return xcd * full + min(xcd, rem) + id / kXcds;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Global-memory views of the field pointers (every FFT operand is device
// memory; the explicit address space keeps global loads when a pointer is
// chosen per thread).
typedef const double __attribute__((address_space(1))) gcdouble;
typedef double __attribute__((address_space(1))) gdouble;

// A wave-uniform argument value (kernel argument, or a field of a problem
// record read at a uniform index) moved to scalar registers: this keeps the
// compiler from turning a per-thread choice between two arguments into a
// per-thread load from the argument block, which every data load would then
// wait behind.
template <class T>
__device__ __forceinline__ T sval(T v)
{
    static_assert(sizeof(T) == 8, "64-bit values");
    const unsigned long long u = (unsigned long long)v;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
    return (T)(((unsigned long long)hi << 32) | lo);
}

// In-register R-point DFT, radix-2 decimation in time.  tw holds the
// L-point table; W_R^t = tw[t * (L/R)].  SIGN = -1 forward, +1 inverse.
template <int R, int L, int SIGN>
__device__ __forceinline__ void dft_regs(double (&re)[R], double (&im)[R], const double2 *__restrict__ tw)
{
    constexpr int bits = ilog2(R);
#pragma unroll
    for (int i = 0; i < R; ++i) {
        const int j = bitrev(i, bits);
        if (j > i) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
#pragma unroll
    for (int len = 2; len <= R; len <<= 1) {
#pragma unroll
        for (int i = 0; i < R; i += len) {
#pragma unroll
            for (int j = 0; j < len / 2; ++j) {
                const int a = i + j, b = i + j + len / 2;
                double xr = re[b], xi = im[b];
                if (j != 0) {
                    const double2 w = tw[j * (L / len)];
                    const double wr = w.x, wi = SIGN * w.y;
                    const double tr = xr * wr - xi * wi;
                    xi = xr * wi + xi * wr;
                    xr = tr;
                }
                re[b] = re[a] - xr;
                im[b] = im[a] - xi;
                re[a] = re[a] + xr;
                im[a] = im[a] + xi;
            }
        }
    }
}

// Addressing of a tile's operands.  Point p of the thread's line is at
// element index off(p) = base_offset + p * point_stride of each operand.
// Addr64T: per-thread pointers and 64-bit indices (any lane mix, any size).
// A lane with `cub` set reads x * (c3 x^2 + c1) for the real input x (the
// cubic right-hand side hz_solve_cubic forms inside the R2C pass; the same
// operations as hz_cubic_rhs).  Only the CUB instances of the tile functions
// carry the map, so the plain transforms keep their loads as they are.
__device__ __forceinline__ double cubic_map(double x, bool cub, double c3, double c1)
{
    return cub ? x * fma(c3 * x, x, c1) : x;
}

template <bool CUB>
struct Addr64T {
    gcdouble *ire, *iim;
    gdouble *ore, *oim;
    long long ib, ips, ob, ops;
    bool cub;
    double c3, c1;
    __device__ __forceinline__ double ldr(int p) const
    {
        const double v = ire[ib + (long long)p * ips];
        if constexpr (CUB) return cubic_map(v, cub, c3, c1);
        return v;
    }
    __device__ __forceinline__ double ldi(int p) const { return iim[ib + (long long)p * ips]; }
    __device__ __forceinline__ void str(int p, double v) const { ore[ob + (long long)p * ops] = v; }
    __device__ __forceinline__ void sti(int p, double v) const { oim[ob + (long long)p * ops] = v; }
};
// Addr32T: wave-uniform base pointers and 32-bit per-thread byte offsets (the
// host guarantees every offset of the launch is below 2^32 bytes), so each
// access is a scalar base plus one offset register.
template <bool CUB>
struct Addr32T {
    gcdouble *ire, *iim;
    gdouble *ore, *oim;
    unsigned ib, ips, ob, ops;
    bool cub;
    double c3, c1;
    template <class P>
    __device__ __forceinline__ static P at(P base, unsigned idx)
    {
        typedef __attribute__((address_space(1))) char gchar;
        return (P)((gchar *)base + 8u * idx);
    }
    __device__ __forceinline__ double ldr(int p) const
    {
        const double v = *at(ire, ib + (unsigned)p * ips);
        if constexpr (CUB) return cubic_map(v, cub, c3, c1);
        return v;
    }
    __device__ __forceinline__ double ldi(int p) const { return *at(iim, ib + (unsigned)p * ips); }
    __device__ __forceinline__ void str(int p, double v) const { *at(ore, ob + (unsigned)p * ops) = v; }
    __device__ __forceinline__ void sti(int p, double v) const { *at(oim, ob + (unsigned)p * ops) = v; }
};

// The transform of one tile: W lines (thread lane w), L points each, W *
// max(R1, R2) threads; `tile` holds (L + 2) * W doubles of LDS (the exchange,
// real parts, then imaginary parts).  C2R reads the half spectrum, point p
// above L/2 as the conjugate of L - p, straight from memory (the second
// read of a point comes from the caches; staging the half spectrum in LDS
// instead cost a barrier and an LDS round trip, ~1 us of a 256^2 x 129
// pass).  Starts with a barrier so that back-to-back tiles can reuse the LDS.
template <int L, int R1, int R2, int W, int KIND, class A>
__device__ __forceinline__ void fft_core(const A &ad, bool live, const double2 *__restrict__ tw, double *tile)
{
    static_assert(R1 * R2 == L, "radix split");
    constexpr int SIGN = (KIND == FFT_R2C || KIND == FFT_C2C_FWD) ? -1 : 1;
    __syncthreads();
    const int w = threadIdx.x % W, q = threadIdx.x / W;

    double sre[R1], sim[R1];
    if constexpr (KIND == FFT_C2R) {
        // mirrored points loaded again (from the caches) instead of staged
        if (q < R2) {
#pragma unroll
            for (int n1 = 0; n1 < R1; ++n1) {
                const int p = q + R2 * n1;
                const int pp = p <= L / 2 ? p : L - p;
                double vr = 0.0, vi = 0.0;
                if (live) {
                    vr = ad.ldr(pp);
                    vi = ad.ldi(pp);
                }
                sre[n1] = vr;
                sim[n1] = p > L / 2 ? -vi : vi;
            }
        }
    } else if (q < R2) {
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) {
            const int p = q + R2 * n1;
            double vr = 0.0, vi = 0.0;
            if (live) {
                vr = ad.ldr(p);
                if (KIND != FFT_R2C) vi = ad.ldi(p);
            }
            sre[n1] = vr;
            sim[n1] = vi;
        }
    }
    if (q < R2) {
        const int n2 = q;
        // inter-stage twiddles W_L^(n2 k1) = prod over the base-4 digits d_m
        // of k1 of (W_L^(n2 4^m))^(d_m): one table load per digit position
        // (loaded before any LDS store, which would otherwise serialise them)
        // and a few complex products, instead of R1 - 1 loads held in
        // registers.  At most four roundings per twiddle.
        constexpr int ND = (ilog2(R1) + 1) / 2;
        double2 pw[ND][4];
#pragma unroll
        for (int m = 0; m < ND; ++m) {
            const double2 b = tw[(n2 << (2 * m)) % L];
            pw[m][1] = b;
            pw[m][2] = cmul(b, b);
            pw[m][3] = cmul(pw[m][2], b);
        }
        dft_regs<R1, L, SIGN>(sre, sim, tw);
#pragma unroll
        for (int k1 = 1; k1 < R1; ++k1) {
            double2 t = make_double2(1.0, 0.0);
            bool first = true;
#pragma unroll
            for (int m = 0; m < ND; ++m) {
                const int d = (k1 >> (2 * m)) & 3;
                if (d == 0) continue;
                t = first ? pw[m][d] : cmul(t, pw[m][d]);
                first = false;
            }
            const double vr = sre[k1], vi = sim[k1];
            const double wr = t.x, wi = SIGN * t.y;
            sre[k1] = vr * wr - vi * wi;
            sim[k1] = vr * wi + vi * wr;
        }
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) tile[(k1 * R2 + n2) * W + w] = sre[k1];
    }
    __syncthreads();
    double re[R2], im[R2];
    if (q < R1) {
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) re[n2] = tile[(q * R2 + n2) * W + w];
    }
    __syncthreads();
    if (q < R2) {
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) tile[(k1 * R2 + q) * W + w] = sim[k1];
    }
    __syncthreads();
    if (q < R1) {
        const int k1 = q;
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) im[n2] = tile[(k1 * R2 + n2) * W + w];
        dft_regs<R2, L, SIGN>(re, im, tw);
        if (live) {
            constexpr double scale = 1.0 / L;
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) {
                const int kk = k1 + R1 * k2;
                if (KIND == FFT_R2C) {
                    if (kk <= L / 2) {
                        ad.str(kk, re[k2]);
                        ad.sti(kk, im[k2]);
                    }
                } else if (KIND == FFT_C2R) {
                    ad.str(kk, re[k2] * scale);
                } else if (KIND == FFT_C2C_INV) {
                    ad.str(kk, re[k2] * scale);
                    ad.sti(kk, im[k2] * scale);
                } else {
                    ad.str(kk, re[k2]);
                    ad.sti(kk, im[k2]);
                }
            }
        }
    }
}

// W_L^(e * k) for k < R from per-digit powers: pw[m][d] = (W_L^(e 4^m))^d,
// two table loads per digit position and at most four roundings.
template <int R, int L>
struct TwiddlePowers {
    static constexpr int ND = (ilog2(R) + 1) / 2;
    double2 pw[ND][4];
    __device__ __forceinline__ void load(const double2 *__restrict__ tw, int e)
    {
#pragma unroll
        for (int m = 0; m < ND; ++m) {
            const double2 b = tw[(e << (2 * m)) % L];
            pw[m][1] = b;
            pw[m][2] = cmul(b, b);
            pw[m][3] = cmul(pw[m][2], b);
        }
    }
    // v[k] *= W^(e k) (SIGN applied to the imaginary part)
    template <int SIGN>
    __device__ __forceinline__ void apply(double (&re)[R], double (&im)[R]) const
    {
#pragma unroll
        for (int k = 1; k < R; ++k) {
            double2 t = make_double2(1.0, 0.0);
            bool first = true;
#pragma unroll
            for (int m = 0; m < ND; ++m) {
                const int d = (k >> (2 * m)) & 3;
                if (d == 0) continue;
                t = first ? pw[m][d] : cmul(t, pw[m][d]);
                first = false;
            }
            const double vr = re[k], vi = im[k];
            const double wr = t.x, wi = SIGN * t.y;
            re[k] = vr * wr - vi * wi;
            im[k] = vr * wi + vi * wr;
        }
    }
};

// Three-stage form of fft_core for long lines (L = R1 * RA * RB, M = RA RB):
//   stage 1: thread q < M: R1-point DFT of x[q + M n1], twiddle W_L^(q k1);
//   stage 2: thread (k1, t), t < RB: RA-point DFT over a of the stage-1
//            outputs y_k1[t + RB a], twiddle W_M^(t ka);
//   stage 3: thread (k1, ka): RB-point DFT over t of z_k1[ka RB + t],
//            X[k1 + R1 (ka + RA kb)].
// (W_M^(q kM) with q = t + RB a, kM = ka + RA kb factors into
// W_RA^(a ka) W_M^(t ka) W_RB^(t kb).)  Each thread holds at most
// max(R1, RA, RB) points, so the registers stay far below the two-stage
// split's for L = 512, 1024; the exchanges go through the same tile, real
// parts then imaginary parts.  Same loads, stores and scaling as fft_core.
template <int L, int R1, int RA, int RB, int W, int KIND, class A>
__device__ __forceinline__ void fft_core3(const A &ad, bool live, const double2 *__restrict__ tw, double *tile)
{
    constexpr int M = RA * RB;
    static_assert(R1 * M == L, "radix split");
    constexpr int SIGN = (KIND == FFT_R2C || KIND == FFT_C2C_FWD) ? -1 : 1;
    __syncthreads();
    const int w = threadIdx.x % W, q = threadIdx.x / W;

    // ---- stage 1 ----
    double sre[R1], sim[R1];
    if constexpr (KIND == FFT_C2R) {
        constexpr int H = L / 2 + 1, Q = M > R1 * RB ? (M > R1 * RA ? M : R1 * RA) : (R1 * RB > R1 * RA ? R1 * RB : R1 * RA);
        constexpr int NJ = (H + Q - 1) / Q;
        double hr[NJ], hi[NJ];
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int pt = q + Q * j;
            hr[j] = hi[j] = 0.0;
            if (live && pt < H) {
                hr[j] = ad.ldr(pt);
                hi[j] = ad.ldi(pt);
            }
        }
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int pt = q + Q * j;
            if (pt < H) {
                tile[pt * W + w] = hr[j];
                tile[(H + pt) * W + w] = hi[j];
            }
        }
        __syncthreads();
        if (q < M) {
#pragma unroll
            for (int n1 = 0; n1 < R1; ++n1) {
                const int p = q + M * n1;
                const int pp = p <= L / 2 ? p : L - p;
                sre[n1] = tile[pp * W + w];
                const double vi = tile[(H + pp) * W + w];
                sim[n1] = p > L / 2 ? -vi : vi;
            }
        }
        __syncthreads();
    } else if (q < M) {
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) {
            const int p = q + M * n1;
            double vr = 0.0, vi = 0.0;
            if (live) {
                vr = ad.ldr(p);
                if (KIND != FFT_R2C) vi = ad.ldi(p);
            }
            sre[n1] = vr;
            sim[n1] = vi;
        }
    }
    if (q < M) {
        TwiddlePowers<R1, L> t1;
        t1.load(tw, q);
        dft_regs<R1, L, SIGN>(sre, sim, tw);
        t1.template apply<SIGN>(sre, sim);
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) tile[(k1 * M + q) * W + w] = sre[k1];
    }
    __syncthreads();
    // ---- stage 2: thread (k1, t) ----
    const int k1b = q / RB, tb = q % RB;
    const bool s2 = q < R1 * RB;
    double ar[RA], ai[RA];
    if (s2) {
#pragma unroll
        for (int a = 0; a < RA; ++a) ar[a] = tile[(k1b * M + tb + RB * a) * W + w];
    }
    __syncthreads();
    if (q < M) {
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) tile[(k1 * M + q) * W + w] = sim[k1];
    }
    __syncthreads();
    if (s2) {
#pragma unroll
        for (int a = 0; a < RA; ++a) ai[a] = tile[(k1b * M + tb + RB * a) * W + w];
        TwiddlePowers<RA, L> t2;
        t2.load(tw, R1 * tb);  // W_M^(t ka) = W_L^(R1 t ka)
        dft_regs<RA, L, SIGN>(ar, ai, tw);
        t2.template apply<SIGN>(ar, ai);
    }
    __syncthreads();
    if (s2) {
#pragma unroll
        for (int ka = 0; ka < RA; ++ka) tile[(k1b * M + ka * RB + tb) * W + w] = ar[ka];
    }
    __syncthreads();
    // ---- stage 3: thread (k1, ka) ----
    const int k1c = q / RA, kac = q % RA;
    const bool s3 = q < R1 * RA;
    double br[RB], bi[RB];
    if (s3) {
#pragma unroll
        for (int t = 0; t < RB; ++t) br[t] = tile[(k1c * M + kac * RB + t) * W + w];
    }
    __syncthreads();
    if (s2) {
#pragma unroll
        for (int ka = 0; ka < RA; ++ka) tile[(k1b * M + ka * RB + tb) * W + w] = ai[ka];
    }
    __syncthreads();
    if (s3) {
#pragma unroll
        for (int t = 0; t < RB; ++t) bi[t] = tile[(k1c * M + kac * RB + t) * W + w];
        dft_regs<RB, L, SIGN>(br, bi, tw);
        if (live) {
            constexpr double scale = 1.0 / L;
#pragma unroll
            for (int kb = 0; kb < RB; ++kb) {
                const int kk = k1c + R1 * (kac + RA * kb);
                if (KIND == FFT_R2C) {
                    if (kk <= L / 2) {
                        ad.str(kk, br[kb]);
                        ad.sti(kk, bi[kb]);
                    }
                } else if (KIND == FFT_C2R) {
                    ad.str(kk, br[kb] * scale);
                } else if (KIND == FFT_C2C_INV) {
                    ad.str(kk, br[kb] * scale);
                    ad.sti(kk, bi[kb] * scale);
                } else {
                    ad.str(kk, br[kb]);
                    ad.sti(kk, bi[kb]);
                }
            }
        }
    }
}

// Threads per line of a transform: max(R1, R2) for the two-stage split
// (R2 points per stage-2 thread), the largest stage for the three-stage one
// (R2 = RA * RB split into RB > 1).
template <int R1, int R2, int RB>
constexpr int fft_threads_per_line()
{
    if constexpr (RB == 1) {
        return R1 > R2 ? R1 : R2;
    } else {
        constexpr int RA = R2 / RB, a = R2, b = R1 * RB, c = R1 * RA;
        return a > b ? (a > c ? a : c) : (b > c ? b : c);
    }
}

// The two- or three-stage core (RB > 1 splits stage 2 into RA = R2 / RB and RB).
template <int L, int R1, int R2, int RB, int W, int KIND, class A>
__device__ __forceinline__ void fft_core_any(const A &ad, bool live, const double2 *__restrict__ tw, double *tile)
{
    if constexpr (RB == 1)
        fft_core<L, R1, R2, W, KIND>(ad, live, tw, tile);
    else
        fft_core3<L, R1, R2 / RB, RB, W, KIND>(ad, live, tw, tile);
}

// One (lane block, line batch) tile, lanes = the main field's z lanes then
// the extra single-lane fields (faces), per-thread addressing (the batch
// kernels' tile).
template <int L, int R1, int R2, int W, int KIND, int RB = 1, bool CUB = false>
__device__ __forceinline__ void fft_tile(const FftArgs &a, int lane_block, int batch, double *tile)
{
    const int w = threadIdx.x % W;
    const int lane = lane_block * W + w;
    // Chosen by selects between wave-uniform argument values (never by
    // indexing them with a per-thread value, which turns into vector loads
    // of the arguments that every data load then waits behind).
    const bool xl = lane >= a.lanes && lane < a.lanes + a.xlanes;
    const bool x1 = lane - a.lanes == 1;
    Addr64T<CUB> ad;
    ad.ire = (gcdouble *)(xl ? (x1 ? sval(a.xin_re[1]) : sval(a.xin_re[0])) : sval(a.in_re));
    ad.iim = (gcdouble *)(xl ? (x1 ? sval(a.xin_im[1]) : sval(a.xin_im[0])) : sval(a.in_im));
    ad.ore = (gdouble *)(xl ? (x1 ? sval(a.xout_re[1]) : sval(a.xout_re[0])) : sval(a.out_re));
    ad.oim = (gdouble *)(xl ? (x1 ? sval(a.xout_im[1]) : sval(a.xout_im[0])) : sval(a.out_im));
    ad.ib = xl ? (long long)batch * sval(a.x_in_batch_stride) : (long long)batch * sval(a.in_batch_stride) + lane;
    ad.ips = xl ? sval(a.x_in_point_stride) : sval(a.in_point_stride);
    ad.ob = xl ? (long long)batch * sval(a.x_out_batch_stride) : (long long)batch * sval(a.out_batch_stride) + lane;
    ad.ops = xl ? sval(a.x_out_point_stride) : sval(a.out_point_stride);
    ad.cub = !xl;  // the faces take no map
    ad.c3 = a.pre_c3;
    ad.c1 = a.pre_c1;
    fft_core_any<L, R1, R2, RB, W, KIND>(ad, lane < a.lanes || xl, a.tw, tile);
}

// Line batches whose lanes follow each other in memory on both sides (batch
// stride = lanes: the x passes of the half spectrum, [x][m][z]) are tiled as
// one batch of nbatch * lanes lanes, so that only the pass's last tile is
// partial instead of the last tile of every batch.
__host__ __device__ inline bool lanes_merge(int nbatch, int lanes, long long in_bs, long long out_bs)
{
    return nbatch > 1 && in_bs == lanes && out_bs == lanes;
}

// Main-field tiles of a uniform-base launch of W lanes per tile (extra-field
// tiles come on top: xlanes * ceil(nbatch / W)).
__host__ __device__ inline long long uniform_main_tiles(int nbatch, int lanes, long long in_bs, long long out_bs,
                                                        int W)
{
    const bool m = lanes_merge(nbatch, lanes, in_bs, out_bs);
    const long long ln = m ? (long long)nbatch * lanes : lanes;
    return (m ? 1 : nbatch) * ((ln + W - 1) / W);
}

// Uniform-base tiles of one launch (fft.hip): tile t < nbatch * nlb is lane
// block t % nlb of line batch t / nlb of the main field (one merged batch
// when lanes_merge holds); the rest are the extra fields,
// xlanes * ceil(nbatch / W) tiles whose W "lanes" are W consecutive line
// batches of one extra field.
template <int L, int R1, int R2, int W, int KIND, int RB = 1, bool CUB = false>
__device__ __forceinline__ void fft_tile_uniform(const FftArgs &a, int t, double *tile)
{
    const int w = threadIdx.x % W;
    const int nb = __builtin_amdgcn_readfirstlane(a.nbatch), ln = __builtin_amdgcn_readfirstlane(a.lanes);
    const long long in_bs = sval(a.in_batch_stride), out_bs = sval(a.out_batch_stride);
    const bool merged = lanes_merge(nb, ln, in_bs, out_bs);
    const int lanes = merged ? nb * ln : ln;
    const int nlb = (lanes + W - 1) / W;
    const int nmain = (merged ? 1 : nb) * nlb;
    Addr32T<CUB> ad;
    bool live;
    if (t < nmain) {
        const int batch = t / nlb, l0 = (t % nlb) * W;
        const long long ib = (long long)batch * in_bs + l0;
        const long long ob = (long long)batch * out_bs + l0;
        ad.ire = (gcdouble *)sval(a.in_re) + ib;
        ad.iim = (gcdouble *)sval(a.in_im) + ib;
        ad.ore = (gdouble *)sval(a.out_re) + ob;
        ad.oim = (gdouble *)sval(a.out_im) + ob;
        ad.ib = w;
        ad.ob = w;
        ad.ips = (unsigned)a.in_point_stride;
        ad.ops = (unsigned)a.out_point_stride;
        ad.cub = true;
        live = l0 + w < lanes;
    } else {
        const int nbb = (nb + W - 1) / W;
        const int u = t - nmain, sx = u / nbb, b0 = (u % nbb) * W;
        const bool x1 = sx == 1;
        const long long ib = (long long)b0 * a.x_in_batch_stride;
        const long long ob = (long long)b0 * a.x_out_batch_stride;
        ad.ire = (gcdouble *)(x1 ? sval(a.xin_re[1]) : sval(a.xin_re[0])) + ib;
        ad.iim = (gcdouble *)(x1 ? sval(a.xin_im[1]) : sval(a.xin_im[0])) + ib;
        ad.ore = (gdouble *)(x1 ? sval(a.xout_re[1]) : sval(a.xout_re[0])) + ob;
        ad.oim = (gdouble *)(x1 ? sval(a.xout_im[1]) : sval(a.xout_im[0])) + ob;
        ad.ib = (unsigned)(w * a.x_in_batch_stride);
        ad.ob = (unsigned)(w * a.x_out_batch_stride);
        ad.ips = (unsigned)a.x_in_point_stride;
        ad.ops = (unsigned)a.x_out_point_stride;
        ad.cub = false;
        live = b0 + w < nb;
    }
    ad.c3 = a.pre_c3;
    ad.c1 = a.pre_c1;
    fft_core_any<L, R1, R2, RB, W, KIND>(ad, live, a.tw, tile);
}

// Element t of a length-1 line pass (an axis of one point: the 2-D and 1-D
// problems): the transform of one point is the point itself (1/L = 1 on the
// inverse), so each lane is copied; R2C writes a zero imaginary part, C2R
// keeps the real part.  t runs over (batch, lane), lanes (main field, then
// the extra fields) fastest; t past the end does nothing.
template <int KIND>
__device__ __forceinline__ void unit_line_element(const FftArgs &a, long long t)
{
    const int nl = a.lanes + a.xlanes;
    if (t >= (long long)nl * a.nbatch) return;
    const int b = (int)(t / nl), l = (int)(t % nl);
    const double *ire, *iim = nullptr;
    double *ore, *oim = nullptr;
    if (l < a.lanes) {
        const long long io = (long long)b * a.in_batch_stride + l, oo = (long long)b * a.out_batch_stride + l;
        ire = a.in_re + io;
        if (KIND != FFT_R2C) iim = a.in_im + io;
        ore = a.out_re + oo;
        if (KIND != FFT_C2R) oim = a.out_im + oo;
    } else {
        const int sx = l - a.lanes;
        const long long io = (long long)b * a.x_in_batch_stride, oo = (long long)b * a.x_out_batch_stride;
        ire = a.xin_re[sx] + io;
        if (KIND != FFT_R2C) iim = a.xin_im[sx] + io;
        ore = a.xout_re[sx] + oo;
        if (KIND != FFT_C2R) oim = a.xout_im[sx] + oo;
    }
    const double re = cubic_map(*ire, KIND == FFT_R2C && a.pre_cubic && l < a.lanes, a.pre_c3, a.pre_c1);
    const double im = (KIND == FFT_C2C_FWD || KIND == FFT_C2C_INV) ? *iim : 0.0;
    *ore = re;
    if (KIND != FFT_C2R) *oim = im;
}

}  // namespace
}  // namespace hz
