// fft.hip -- batched line FFTs over the x or y axis of z-fastest fields.
//
// The fields are [a][b][K] with K (z) contiguous, so a transform along a or
// b is a batch of K independent lines whose same-index points sit in one
// contiguous run of K doubles.  A work-group takes W consecutive z lanes of
// one line-batch (or, where consecutive line batches are contiguous, W
// consecutive lanes of their concatenation) and the whole transform length L:
//   stage 1: each thread loads R1 points x[n2 + R2*n1] of its lane with
//            coalesced W-wide loads, does an R1-point DFT in registers and
//            multiplies by the twiddle W_L^(n2*k1);
//   exchange through LDS (an L x W tile, lane-fastest, conflict free);
//   stage 2: R2-point DFTs in registers, stores X[k1 + R1*k2] coalesced.
// Lengths 512 and 1024 split stage 2 once more (three register stages, two
// exchanges through the same tile), which keeps each thread to 8 or 16
// points.  One HBM read and one HBM write per element; the LDS tile is the
// only transpose.  Forward transforms are unnormalised with exp(-2 pi i jk/L);
// inverse ones carry 1/L, as the reference's FourierTransformX/Y
// (Fourier_Quadrature.cpp:778, :809-817).  R2C keeps points 0..L/2; C2R
// rebuilds the upper half by conjugate symmetry and keeps the real part.
#include "fft_body.h"

namespace hz {
namespace {

// Uniform-base tiles (fft_tile_uniform): one 1-D grid of main-field tiles
// then extra-field tiles.
template <int L, int R1, int R2, int W, int KIND, int RB, bool CUB = false>
__global__ __launch_bounds__((W * fft_threads_per_line<R1, R2, RB>())) void fft_lines(const FftArgs a)
{
    __shared__ double tile[(L + 2) * W];
    fft_tile_uniform<L, R1, R2, W, KIND, RB, CUB>(a, xcd_logical(blockIdx.x, gridDim.x), tile);
}

// Per-thread addressing, for launches whose offsets do not fit 32 bits.
template <int L, int R1, int R2, int W, int KIND, int RB, bool CUB = false>
__global__ __launch_bounds__((W * fft_threads_per_line<R1, R2, RB>())) void fft_lines_wide(const FftArgs a)
{
    __shared__ double tile[(L + 2) * W];
    const int nbx = gridDim.x, logical = xcd_logical(blockIdx.x + nbx * blockIdx.y, nbx * gridDim.y);
    fft_tile<L, R1, R2, W, KIND, RB, CUB>(a, logical % nbx, logical / nbx, tile);
}

// Largest element index a uniform-base tile of this launch forms, on either
// side (inputs and outputs); below 2^29 the byte offsets fit 32 bits.
bool offsets_fit_32(int L, int W, const FftArgs &a)
{
    const long long lim = 1LL << 29;
    const long long m_in = (W - 1) + (long long)(L - 1) * a.in_point_stride;
    const long long m_out = (W - 1) + (long long)(L - 1) * a.out_point_stride;
    long long x_in = 0, x_out = 0;
    if (a.xlanes > 0) {
        x_in = (long long)(W - 1) * a.x_in_batch_stride + (long long)(L - 1) * a.x_in_point_stride;
        x_out = (long long)(W - 1) * a.x_out_batch_stride + (long long)(L - 1) * a.x_out_point_stride;
    }
    return m_in < lim && m_out < lim && x_in < lim && x_out < lim && m_in >= 0 && m_out >= 0;
}

// RB > 1: the three-stage core (stage 2 split into R2 / RB and RB).
template <int L, int R1, int R2, int W, int RB = 1>
int launch_kind(FftKind kind, const FftArgs &a, hipStream_t s)
{
    constexpr int Q = fft_threads_per_line<R1, R2, RB>();
    const dim3 block(W * Q);
    if (offsets_fit_32(L, W, a)) {
        const long long tiles = uniform_main_tiles(a.nbatch, a.lanes, a.in_batch_stride, a.out_batch_stride, W) +
                                (long long)a.xlanes * ((a.nbatch + W - 1) / W);
        const dim3 grid((unsigned)tiles);
        switch (kind) {
        case FFT_R2C:
            if (a.pre_cubic)  // the cubic right-hand side formed in the loads (hz_solve_cubic)
                hipLaunchKernelGGL((fft_lines<L, R1, R2, W, FFT_R2C, RB, true>), grid, block, 0, s, a);
            else
                hipLaunchKernelGGL((fft_lines<L, R1, R2, W, FFT_R2C, RB>), grid, block, 0, s, a);
            break;
        case FFT_C2R: hipLaunchKernelGGL((fft_lines<L, R1, R2, W, FFT_C2R, RB>), grid, block, 0, s, a); break;
        case FFT_C2C_FWD: hipLaunchKernelGGL((fft_lines<L, R1, R2, W, FFT_C2C_FWD, RB>), grid, block, 0, s, a); break;
        case FFT_C2C_INV: hipLaunchKernelGGL((fft_lines<L, R1, R2, W, FFT_C2C_INV, RB>), grid, block, 0, s, a); break;
        }
    } else {
        const dim3 grid((a.lanes + a.xlanes + W - 1) / W, a.nbatch);
        switch (kind) {
        case FFT_R2C:
            if (a.pre_cubic)
                hipLaunchKernelGGL((fft_lines_wide<L, R1, R2, W, FFT_R2C, RB, true>), grid, block, 0, s, a);
            else
                hipLaunchKernelGGL((fft_lines_wide<L, R1, R2, W, FFT_R2C, RB>), grid, block, 0, s, a);
            break;
        case FFT_C2R: hipLaunchKernelGGL((fft_lines_wide<L, R1, R2, W, FFT_C2R, RB>), grid, block, 0, s, a); break;
        case FFT_C2C_FWD:
            hipLaunchKernelGGL((fft_lines_wide<L, R1, R2, W, FFT_C2C_FWD, RB>), grid, block, 0, s, a);
            break;
        case FFT_C2C_INV:
            hipLaunchKernelGGL((fft_lines_wide<L, R1, R2, W, FFT_C2C_INV, RB>), grid, block, 0, s, a);
            break;
        }
    }
    HZ_HIP(hipGetLastError());
    return HZ_OK;
}

// Length-1 lines (the 2-D and 1-D problems): one thread per (batch, lane),
// unit_line_element in fft_body.h.
template <int KIND>
__global__ __launch_bounds__(256) void fft_unit_lines(const FftArgs a)
{
    unit_line_element<KIND>(a, (long long)blockIdx.x * 256 + threadIdx.x);
}

int launch_unit(FftKind kind, const FftArgs &a, hipStream_t s)
{
    // an in-place complex pass over one point changes nothing
    const bool inplace = a.in_re == a.out_re && a.in_im == a.out_im &&
                         (a.xlanes == 0 || (a.xin_re[0] == a.xout_re[0] && a.xin_im[0] == a.xout_im[0] &&
                                            (a.xlanes < 2 || (a.xin_re[1] == a.xout_re[1] &&
                                                              a.xin_im[1] == a.xout_im[1]))));
    if ((kind == FFT_C2C_FWD || kind == FFT_C2C_INV) && inplace &&
        a.in_batch_stride == a.out_batch_stride && a.x_in_batch_stride == a.x_out_batch_stride)
        return HZ_OK;
    const long long total = (long long)(a.lanes + a.xlanes) * a.nbatch;
    const dim3 grid((unsigned)((total + 255) / 256)), block(256);
    switch (kind) {
    case FFT_R2C: hipLaunchKernelGGL((fft_unit_lines<FFT_R2C>), grid, block, 0, s, a); break;
    case FFT_C2R: hipLaunchKernelGGL((fft_unit_lines<FFT_C2R>), grid, block, 0, s, a); break;
    case FFT_C2C_FWD: hipLaunchKernelGGL((fft_unit_lines<FFT_C2C_FWD>), grid, block, 0, s, a); break;
    case FFT_C2C_INV: hipLaunchKernelGGL((fft_unit_lines<FFT_C2C_INV>), grid, block, 0, s, a); break;
    }
    HZ_HIP(hipGetLastError());
    return HZ_OK;
}

}  // namespace

int fft_width_choices(int L, int *w, int max)
{
    if (L == 512) {  // 64 threads per line: 16 lanes fill a 1024-thread work-group
        static const int c[] = {8, 16};
        for (int i = 0; i < 2 && i < max; ++i) w[i] = c[i];
        return 2;
    }
    if (L < 64 || L > 256) return 0;
    static const int c[] = {8, 16, 32};
    for (int i = 0; i < 3 && i < max; ++i) w[i] = c[i];
    return 3;
}

int launch_fft(int L, FftKind kind, const FftArgs &a, hipStream_t s)
{
    if (a.nbatch <= 0 || a.lanes + a.xlanes <= 0) return HZ_OK;
    switch (L) {
    case 1: return launch_unit(kind, a, s);
    case 2: return launch_kind<2, 2, 1, 16>(kind, a, s);
    case 4: return launch_kind<4, 4, 1, 16>(kind, a, s);
    case 8: return launch_kind<8, 8, 1, 16>(kind, a, s);
    case 16: return launch_kind<16, 4, 4, 16>(kind, a, s);
    case 32: return launch_kind<32, 8, 4, 16>(kind, a, s);
    case 64:
        return a.w == 8 ? launch_kind<64, 8, 8, 8>(kind, a, s)
               : a.w == 32 ? launch_kind<64, 8, 8, 32>(kind, a, s) : launch_kind<64, 8, 8, 16>(kind, a, s);
    case 128:
        return a.w == 8 ? launch_kind<128, 16, 8, 8>(kind, a, s)
               : a.w == 32 ? launch_kind<128, 16, 8, 32>(kind, a, s) : launch_kind<128, 16, 8, 16>(kind, a, s);
    case 256:
        return a.w == 8 ? launch_kind<256, 16, 16, 8>(kind, a, s)
               : a.w == 32 ? launch_kind<256, 16, 16, 32>(kind, a, s) : launch_kind<256, 16, 16, 16>(kind, a, s);
    // three stages (fft_core3): 8 x 8 x 8 and 16 x 8 x 8, 59-66 and 110-114
    // VGPRs (the two-stage 32 x 16 and 32 x 32 splits held 206-256)
    case 512: return a.w == 8 ? launch_kind<512, 8, 64, 8, 8>(kind, a, s) : launch_kind<512, 8, 64, 16, 8>(kind, a, s);
    case 1024: return launch_kind<1024, 16, 64, 8, 8>(kind, a, s);
    default: return set_error(HZ_ERR_UNSUPPORTED, "FFT length must be a power of two in [1, 1024]");
    }
}

}  // namespace hz
