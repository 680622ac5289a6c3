// probe.hip -- attainable ceilings of the device, measured live: the FP64
// MFMA rate of v_mfma_f64_16x16x4f64 (the instruction of the mode solve)
// and the bandwidth of a plain device-to-device copy.  The published peaks
// (78.6 TFLOP/s dense FP64 matrix, 8 TB/s HBM3E) are upper bounds that no
// kernel of this shape reaches; the roofline fractions of the solver's
// kernels are reported against both.
#include "hz_internal.h"

namespace hz {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

// Four independent accumulation chains per wave, eight waves per SIMD: the
// best shape measured for this instruction (1, 2, 4 and 8 chains gave 45.0,
// 47.8, 49.0 and 49.0 TFLOP/s in one run).  The operands are
// lane-dependent so nothing folds.
constexpr int kProbeChains = 4;
__global__ __launch_bounds__(256) void mfma_f64_probe(double *out, int iters)
{
    d4 acc[kProbeChains];
#pragma unroll
    for (int c = 0; c < kProbeChains; ++c) acc[c] = d4{0.0, 0.0, 0.0, 0.0};
    const double a = 1.0 + threadIdx.x * 1e-9, b = 0.5 - threadIdx.x * 1e-9;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int c = 0; c < kProbeChains; ++c)
            acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(c & 1 ? b : a, c & 1 ? a : b, acc[c], 0, 0, 0);
    }
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < kProbeChains; ++c) s += acc[c][0] + acc[c][1] + acc[c][2] + acc[c][3];
    out[(long long)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// One 16-byte element per thread, the grid covering the buffer: the
// fastest of the copy shapes measured (grid-stride loops and several
// elements per thread ran 5-25% slower).
__global__ __launch_bounds__(256) void copy_probe(const double2 *__restrict__ in, double2 *__restrict__ out,
                                                  long long n)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = in[i];
}

// Pseudo-random doubles in [-0.5, 0.5): the SplitMix64 finaliser
// (Steele, Lea and Flood, OOPSLA 2014) of the element index.  The plan-time
// tuners time kernels on such data: on all-zero buffers the same kernels run
// 7-9% faster (the clock depends on the operands' switching), which would
// bias the comparison.
__global__ __launch_bounds__(256) void fill_noise_kernel(double *__restrict__ p, long long n,
                                                         unsigned long long seed)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    unsigned long long z = seed + (unsigned long long)(i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    p[i] = (double)(z >> 11) * 0x1.0p-53 - 0.5;
}

// Mean milliseconds of `reps` back-to-back launches of `launch` on stream s,
// after two untimed ones.
template <class F>
int time_ms(hipStream_t s, F launch, float *ms, int reps)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return set_error(HZ_ERR_HIP, "event creation failed");
    }
    launch();
    launch();
    (void)hipEventRecord(e0, s);
    for (int r = 0; r < reps; ++r) launch();
    (void)hipEventRecord(e1, s);
    int st = HZ_OK;
    if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(ms, e0, e1) != hipSuccess)
        st = set_error(HZ_ERR_HIP, "probe run failed");
    *ms /= reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return st;
}

}  // namespace
}  // namespace hz

namespace hz {

int launch_fill_noise(double *p, size_t n, unsigned long long seed, hipStream_t s)
{
    if (n == 0) return HZ_OK;
    const size_t grid = (n + 255) / 256;
    if (!p || grid > 0x7fffffffu) return set_error(HZ_ERR_ARG, "noise fill: bad buffer");
    hipLaunchKernelGGL(fill_noise_kernel, dim3((unsigned)grid), dim3(256), 0, s, p, (long long)n, seed);
    HZ_HIP(hipGetLastError());
    return HZ_OK;
}

}  // namespace hz

using namespace hz;

int hz_probe_ceilings(double *fp64_mfma_tflops, double *copy_gbps, size_t copy_bytes)
{
    if (!fp64_mfma_tflops && !copy_gbps) return set_error(HZ_ERR_ARG, "null argument");
    int dev = 0, cus = 0;
    HZ_HIP(hipGetDevice(&dev));
    HZ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    hipStream_t s = nullptr;
    HZ_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    int st = HZ_OK;
    if (fp64_mfma_tflops) {
        const int wgs = 8 * cus, iters = 4096;  // eight waves per SIMD
        double *out = nullptr;
        if (hipMalloc((void **)&out, sizeof(double) * 256 * wgs) != hipSuccess) {
            st = set_error(HZ_ERR_ALLOC, "probe allocation failed");
        } else {
            float ms = 0.0f;
            st = time_ms(
                s, [&] { hipLaunchKernelGGL(mfma_f64_probe, dim3(wgs), dim3(256), 0, s, out, iters); }, &ms, 1);
            // 4 waves x kProbeChains chains x iters MFMAs of 16 x 16 x 4 multiply-adds per work-group
            if (!st) *fp64_mfma_tflops = (double)wgs * 4 * kProbeChains * iters * 2048.0 / (ms * 1e-3) / 1e12;
            (void)hipFree(out);
        }
    }
    if (!st && copy_gbps) {
        const long long n = (long long)(copy_bytes / sizeof(double2));  // elements of each buffer
        double2 *a = nullptr, *b = nullptr;
        if (n < 1) {
            st = set_error(HZ_ERR_ARG, "copy size too small");
        } else if (hipMalloc((void **)&a, sizeof(double2) * n) != hipSuccess ||
                   hipMalloc((void **)&b, sizeof(double2) * n) != hipSuccess) {
            st = set_error(HZ_ERR_ALLOC, "probe allocation failed");
        } else if (hipMemsetAsync(a, 0, sizeof(double2) * n, s) != hipSuccess) {
            st = set_error(HZ_ERR_HIP, "probe buffer clear failed");
        } else {
            float ms = 0.0f;
            const long long grid = (n + 255) / 256;
            if (grid > 0x7fffffffLL) {
                st = set_error(HZ_ERR_ARG, "copy size too large");
            } else {
                st = time_ms(
                    s, [&] { hipLaunchKernelGGL(copy_probe, dim3((unsigned)grid), dim3(256), 0, s, a, b, n); }, &ms,
                    10);
            }
            if (!st) *copy_gbps = 2.0 * sizeof(double2) * n / (ms * 1e-3) / 1e9;  // read + write
        }
        if (a) (void)hipFree(a);
        if (b) (void)hipFree(b);
    }
    (void)hipStreamDestroy(s);
    return st;
}
