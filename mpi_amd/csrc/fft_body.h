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

// One (lane block, line batch) tile of the line transform; W * max(R1, R2)
// threads; `tile` holds L * W doubles of LDS (used twice: real parts, then
// imaginary parts).  Starts with a barrier so that back-to-back tiles can
// reuse the LDS.
template <int L, int R1, int R2, int W, int KIND>
__device__ __forceinline__ void fft_tile(const FftArgs &a, int lane_block, int batch, double *tile)
{
    static_assert(R1 * R2 == L, "radix split");
    constexpr int SIGN = (KIND == FFT_R2C || KIND == FFT_C2C_FWD) ? -1 : 1;
    __syncthreads();
    const int w = threadIdx.x % W, q = threadIdx.x / W;
    const int lane = lane_block * W + w;
    const double2 *__restrict__ tw = a.tw;
    // lane source: the main field, or one of the extra single-lane fields
    const double *in_re = a.in_re, *in_im = a.in_im;
    double *out_re = a.out_re, *out_im = a.out_im;
    long long ibase = (long long)batch * a.in_batch_stride + lane, ips = a.in_point_stride;
    long long obase = (long long)batch * a.out_batch_stride + lane, ops = a.out_point_stride;
    bool live = lane < a.lanes;
    if (!live && lane < a.lanes + a.xlanes) {
        const int sx = lane - a.lanes;
        in_re = a.xin_re[sx];
        in_im = a.xin_im[sx];
        out_re = a.xout_re[sx];
        out_im = a.xout_im[sx];
        ibase = (long long)batch * a.x_in_batch_stride;
        ips = a.x_in_point_stride;
        obase = (long long)batch * a.x_out_batch_stride;
        ops = a.x_out_point_stride;
        live = true;
    }

    double sre[R1], sim[R1];
    if (q < R2) {
        const int n2 = q;
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) {
            const int p = n2 + R2 * n1;
            double vr = 0.0, vi = 0.0;
            if (live) {
                if (KIND == FFT_R2C) {
                    vr = in_re[ibase + (long long)p * ips];
                } else if (KIND == FFT_C2R) {
                    const int pp = p <= L / 2 ? p : L - p;
                    const long long off = ibase + (long long)pp * ips;
                    vr = in_re[off];
                    vi = in_im[off];
                    if (p > L / 2) vi = -vi;
                } else {
                    const long long off = ibase + (long long)p * ips;
                    vr = in_re[off];
                    vi = in_im[off];
                }
            }
            sre[n1] = vr;
            sim[n1] = vi;
        }
        dft_regs<R1, L, SIGN>(sre, sim, tw);
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) {
            double vr = sre[k1], vi = sim[k1];
            const int t = (n2 * k1) % L;
            if (t != 0) {
                const double2 tt = tw[t];
                const double wr = tt.x, wi = SIGN * tt.y;
                const double xr = vr * wr - vi * wi;
                vi = vr * wi + vi * wr;
                vr = xr;
            }
            sre[k1] = vr;
            sim[k1] = vi;
            tile[(k1 * R2 + n2) * W + w] = vr;
        }
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
            const long long base = obase;
            constexpr double scale = 1.0 / L;
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) {
                const int kk = k1 + R1 * k2;
                const long long off = base + (long long)kk * ops;
                if (KIND == FFT_R2C) {
                    if (kk <= L / 2) {
                        out_re[off] = re[k2];
                        out_im[off] = im[k2];
                    }
                } else if (KIND == FFT_C2R) {
                    out_re[off] = re[k2] * scale;
                } else if (KIND == FFT_C2C_INV) {
                    out_re[off] = re[k2] * scale;
                    out_im[off] = im[k2] * scale;
                } else {
                    out_re[off] = re[k2];
                    out_im[off] = im[k2];
                }
            }
        }
    }
}

}  // namespace
}  // namespace hz
