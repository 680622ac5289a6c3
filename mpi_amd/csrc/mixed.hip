// mixed.hip -- the MIXED form's mode solve: FP32 matrix-core solves in the
// plan's eigenbasis refined with FP64 matrix-core residuals.  Launch
// wrappers around the tile body in mixed_body.h (scheme described there),
// and the plan-time kernel of the FP64 homogeneous solutions.
#include "mixed_body.h"

namespace hz {
namespace {

// Plan time (once per lambda): ul = V (w0 / (mu - sigma)), ur = V (wN / (mu -
// sigma)) of every mode (NonPeriodic3D_MPI.cpp:203-222 through the
// eigendecomposition), interior values, FP64.
__global__ __launch_bounds__(256) void mixed_homogeneous_kernel(const MixedSetupArgs a)
{
    __shared__ double inv[256];
    const int m = blockIdx.x, tid = threadIdx.x;
    const int n = a.n, NP = a.NP;
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
}

template <int NJ, int RT, bool GENERAL>
__global__ __launch_bounds__(256, 2) void mixed_modes_kernel(const MixedArgs a)
{
    __shared__ MixedShared<NJ, RT, GENERAL> sh;
    mixed_tile<NJ, RT, GENERAL>(a, blockIdx.x, sh);
}

template <int NJ, int RT>
int launch_t(const MixedArgs &a, hipStream_t s)
{
    const int fc_per_tile = RT / (2 * a.ne);
    if (fc_per_tile < 1) return set_error(HZ_ERR_UNSUPPORTED, "MIXED form: too many elements for one tile");
    const int nfc = a.nx * a.M;
    const int blocks = (nfc + fc_per_tile - 1) / fc_per_tile;
    const bool general = a.ne > 1 || a.alpha_l != 0.0 || a.alpha_r != 0.0 || a.periodic;
    if (general)
        hipLaunchKernelGGL((mixed_modes_kernel<NJ, RT, true>), dim3(blocks), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((mixed_modes_kernel<NJ, RT, false>), dim3(blocks), dim3(256), 0, s, a);
    HZ_HIP(hipGetLastError());
    return HZ_OK;
}

}  // namespace

int mixed_tile_rows(int NP)
{
    // NP = 32, 96 split the row blocks between wave pairs: an even number of
    // 16-row blocks
    return (NP == 32 || NP == 96) ? 32 : 16;
}

int launch_mixed_setup(const MixedSetupArgs &a, hipStream_t s)
{
    if (a.modes <= 0) return HZ_OK;
    hipLaunchKernelGGL(mixed_homogeneous_kernel, dim3(a.modes), dim3(256), 0, s, a);
    HZ_HIP(hipGetLastError());
    return HZ_OK;
}

int launch_mixed_modes(const MixedArgs &a, hipStream_t s)
{
    switch (diag_padded(a.n)) {
    case 32: return launch_t<2, 32>(a, s);
    case 64: return launch_t<4, 16>(a, s);
    case 96: return launch_t<6, 32>(a, s);
    case 128: return launch_t<8, 16>(a, s);
    case 256: return launch_t<16, 16>(a, s);
    default: return set_error(HZ_ERR_UNSUPPORTED, "MIXED form: N-2 must be <= 256");
    }
}

}  // namespace hz
