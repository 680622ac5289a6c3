// diag.hip -- the z-diagonalised mode solve on FP64 matrix cores: launch
// wrappers around the tile body in diag_body.h (scheme described there).
#include "diag_body.h"

namespace hz {
namespace {

template <int NJ, int RT, bool GENERAL, bool ADJ, bool EO>
__global__ __launch_bounds__(256, RT >= 128 ? 1 : 2) void diag_modes_kernel(const DiagArgs a)
{
    __shared__ DiagShared<NJ, RT, GENERAL, ADJ> sh;
    diag_tile<NJ, RT, GENERAL, ADJ, EO>(a, blockIdx.x, sh);
}

template <int NJ, int RT>
int launch_t(const DiagArgs &a, hipStream_t s)
{
    const int rows_per_fc = 2 * a.ne;
    const int fc_per_tile = RT / rows_per_fc;
    if (fc_per_tile < 1) return set_error(HZ_ERR_UNSUPPORTED, "DIAG form: too many elements for one tile");
    const int nfc = a.nx * a.M;
    const int blocks = (nfc + fc_per_tile - 1) / fc_per_tile;
    const bool general = a.ne > 1 || a.alpha_l != 0.0 || a.alpha_r != 0.0 || a.periodic;
    if (a.eo == general) return set_error(HZ_ERR_ARG, "DIAG form: the even/odd basis is for one Dirichlet element");
    if (a.adjoint && a.eo) {
        hipLaunchKernelGGL((diag_modes_kernel<NJ, RT, false, true, true>), dim3(blocks), dim3(256), 0, s, a);
    } else if (a.adjoint) {
        hipLaunchKernelGGL((diag_modes_kernel<NJ, RT, false, true, false>), dim3(blocks), dim3(256), 0, s, a);
    } else if (general) {
        hipLaunchKernelGGL((diag_modes_kernel<NJ, RT, true, false, false>), dim3(blocks), dim3(256), 0, s, a);
    } else {
        hipLaunchKernelGGL((diag_modes_kernel<NJ, RT, false, false, true>), dim3(blocks), dim3(256), 0, s, a);
    }
    HZ_HIP(hipGetLastError());
    return HZ_OK;
}

}  // namespace

int diag_padded(int n)
{
    if (n <= 0) return 0;
    if (n <= 128) return 32 * ((n + 31) / 32);
    if (n <= 256) return 256;
    return 0;
}

int diag_tile_choices(int NP, int *rt, int max)
{
    // 16 rows: more work-groups in flight to hide the staging latency;
    // 128 rows (N-2 <= 64 only) holds a column of up to 64 elements (82 KB
    // of LDS: one work-group per CU).  NP = 32, 96 split the row blocks
    // between wave pairs, so their tiles have an even number of 16-row blocks.
    static const int c32[] = {32, 64, 128}, c64[] = {16, 32, 64, 128}, c96[] = {32, 64},
                     c128[] = {16, 32, 48, 64}, c256[] = {16, 32};
    const int *c = nullptr;
    int k = 0;
    switch (NP) {
    case 32: c = c32, k = 3; break;
    case 64: c = c64, k = 4; break;
    case 96: c = c96, k = 2; break;
    case 128: c = c128, k = 4; break;
    case 256: c = c256, k = 2; break;
    default: return 0;
    }
    for (int i = 0; i < k && i < max; ++i) rt[i] = c[i];
    return k;
}

int launch_diag_modes(const DiagArgs &a, hipStream_t s)
{
    switch (diag_padded(a.n)) {
    case 32:
        return a.rt == 128 ? launch_t<2, 128>(a, s) : a.rt == 64 ? launch_t<2, 64>(a, s) : launch_t<2, 32>(a, s);
    case 64:
        return a.rt == 16   ? launch_t<4, 16>(a, s)
               : a.rt == 32  ? launch_t<4, 32>(a, s)
               : a.rt == 128 ? launch_t<4, 128>(a, s)
                             : launch_t<4, 64>(a, s);
    case 96: return a.rt == 64 ? launch_t<6, 64>(a, s) : launch_t<6, 32>(a, s);
    case 128:
        return a.rt == 16  ? launch_t<8, 16>(a, s)
               : a.rt == 32 ? launch_t<8, 32>(a, s)
               : a.rt == 48 ? launch_t<8, 48>(a, s)
                            : launch_t<8, 64>(a, s);
    case 256:
        return a.rt == 16 ? launch_t<16, 16>(a, s) : launch_t<16, 32>(a, s);
    default: return set_error(HZ_ERR_UNSUPPORTED, "DIAG form: N-2 must be <= 256");
    }
}

}  // namespace hz
