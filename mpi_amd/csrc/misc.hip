// misc.hip -- element-wise GLL derivatives (FirstDerivativeZ /
// SecondDerivativeZ, NonPeriodic3D_MPI.cpp:632-683).
//
// out[e][i][j][k] = sum_p D[k][p] u[e][i][j][p] for every element column:
// a batch of small dense products, each thread producing one output node of
// one column (columns are contiguous, so a wave reads and writes consecutive
// z values; the N x N matrix stays in the caches).  The sum runs over
// p = 0..N-1 in order with every product rounded before it is added, as in
// the reference's loop, so the result equals the reference's bit for bit.
#include "hz_internal.h"

namespace hz {
namespace {

__device__ __forceinline__ double rounded_dot(const double *__restrict__ a, const double *__restrict__ b, int n)
{
#pragma clang fp contract(off)
    double acc = 0.0;
    for (int p = 0; p < n; ++p) acc += a[p] * b[p];
    return acc;
}

__global__ __launch_bounds__(256) void derivative_kernel(const double *__restrict__ D, int N, long long ncols,
                                                         const double *__restrict__ u, double *__restrict__ out)
{
    const long long total = ncols * N;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long col = idx / N;
        const int k = (int)(idx % N);
        const double *uc = u + col * N;
        out[idx] = rounded_dot(D + (long long)k * N, uc, N);
    }
}

__global__ __launch_bounds__(256) void dot_partial_kernel(const double *__restrict__ a, const double *__restrict__ b,
                                                          long long n, double *__restrict__ partial)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        acc = fma(a[i], b[i], acc);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// f = a3 u^3 + a1 u, one element per thread (one read, one write).
__global__ __launch_bounds__(256) void cubic_kernel(const double *__restrict__ u, double *__restrict__ f, long long n,
                                                    double a3, double a1)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double x = u[i];
    f[i] = x * fma(a3 * x, x, a1);
}

// One backward step of the semi-implicit cubic right-hand side f(u) =
// a3 u^3 + a1 u: ubar = fbar f'(u_in) = fbar (3 a3 u_in^2 + a1) (ubar may be
// fbar), and the work-group's partial sum of fbar (b1 (u_out - u_in) +
// b3 (u_in^3 - u_in)) into partial[blockIdx.x] (a fixed grid and order, so
// the sum is the same on every run).
__global__ __launch_bounds__(256) void cubic_adjoint_kernel(const double *fbar, const double *__restrict__ u_in,
                                                            const double *__restrict__ u_out, long long n, double a3,
                                                            double a1, double b1, double b3, double *ubar,
                                                            double *__restrict__ partial)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const double fb = fbar[i], x = u_in[i], x2 = x * x;
        acc = fma(fb, b1 * (u_out[i] - x) + b3 * (x2 * x - x), acc);
        ubar[i] = fb * fma(3.0 * a3, x2, a1);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// *acc += the sum of partial[0 .. np) (one work-group, fixed order).
__global__ __launch_bounds__(256) void accumulate_final_kernel(const double *__restrict__ partial, int np, double *acc)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += 256) s += partial[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *acc += red[0];
}

__global__ __launch_bounds__(256) void dot_final_kernel(const double *__restrict__ partial, int np, double *out)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < np; i += 256) acc += partial[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// Spectral derivative along one axis of a half spectrum re/im[A][B][lanes]
// (the transformed axis is A when axis_a, else B, of length L): multiplies by
// (2 pi i k / len)^order with k = j for j < L/2 and j - L otherwise, the
// wavenumber convention of Fourier_Quadrature.cpp:1512-1519 / :1739-1745
// (the Nyquist point counts as -L/2).
__global__ __launch_bounds__(256) void spectral_derivative_kernel(double *__restrict__ re, double *__restrict__ im,
                                                                  long long A, long long B, long long lanes,
                                                                  int axis_a, int L, double len, int order)
{
    const long long total = A * B * lanes;
    const double two_pi = 6.283185307179586476925286766559;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const long long ab = t / lanes;
        const int j = (int)(axis_a ? ab / B : ab % B);
        const double k = j < L / 2 ? (double)j : (double)(j - L);
        const double c = two_pi * k / len;
        const double r = re[t], q = im[t];
        if (order == 1) {
            re[t] = -c * q;
            im[t] = c * r;
        } else {
            re[t] = -(c * c) * r;
            im[t] = -(c * c) * q;
        }
    }
}

}  // namespace

int launch_spectral_derivative(double *re, double *im, long long A, long long B, long long lanes, int axis_a, int L,
                               double len, int order, hipStream_t s)
{
    const long long total = A * B * lanes;
    const int blocks = (int)std::min<long long>(4096, (total + 255) / 256);
    if (total > 0)
        hipLaunchKernelGGL(spectral_derivative_kernel, dim3(blocks), dim3(256), 0, s, re, im, A, B, lanes, axis_a, L,
                           len, order);
    HZ_HIP(hipGetLastError());
    return HZ_OK;
}

int launch_dot(const double *a, const double *b, long long n, double *partial, double *out, hipStream_t s)
{
    const int blocks = (int)std::min<long long>(kDotPartials, std::max<long long>(1, (n + 255) / 256));
    hipLaunchKernelGGL(dot_partial_kernel, dim3(blocks), dim3(256), 0, s, a, b, n, partial);
    hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(256), 0, s, partial, blocks, out);
    HZ_HIP(hipGetLastError());
    return HZ_OK;
}

int launch_cubic_adjoint(const double *fbar, const double *u_in, const double *u_out, long long n, double a3,
                         double a1, double b1, double b3, double *ubar, double *partial, double *acc, hipStream_t s)
{
    const int blocks = (int)std::min<long long>(kDotPartials, std::max<long long>(1, (n + 255) / 256));
    hipLaunchKernelGGL(cubic_adjoint_kernel, dim3(blocks), dim3(256), 0, s, fbar, u_in, u_out, n, a3, a1, b1, b3, ubar,
                       partial);
    hipLaunchKernelGGL(accumulate_final_kernel, dim3(1), dim3(256), 0, s, partial, blocks, acc);
    HZ_HIP(hipGetLastError());
    return HZ_OK;
}

int launch_derivative(const double *D, int N, long long ncols, const double *u, double *out, hipStream_t s)
{
    if (ncols <= 0) return HZ_OK;
    const long long total = ncols * N;
    const int blocks = (int)std::min<long long>((total + 255) / 256, 4096);
    hipLaunchKernelGGL(derivative_kernel, dim3(blocks), dim3(256), 0, s, D, N, ncols, u, out);
    HZ_HIP(hipGetLastError());
    return HZ_OK;
}

}  // namespace hz

using namespace hz;

extern "C" int hz_cubic_rhs(const double *u, double *f, long long n, double a3, double a1, hz_stream_t stream)
{
    if (n < 0 || (n > 0 && (!u || !f))) return set_error(HZ_ERR_ARG, "null argument or negative size");
    if (n == 0) return HZ_OK;
    const long long grid = (n + 255) / 256;
    if (grid > 0x7fffffffLL) return set_error(HZ_ERR_ARG, "size too large");
    hipLaunchKernelGGL(cubic_kernel, dim3((unsigned)grid), dim3(256), 0, (hipStream_t)stream, u, f, n, a3, a1);
    HZ_HIP(hipGetLastError());
    return HZ_OK;
}

extern "C" int hz_derivative_z(hz_plan *p, int order, const double *u, double *out, hz_stream_t stream)
{
    if (!p || !u || !out || (order != 1 && order != 2)) return set_error(HZ_ERR_ARG, "bad argument");
    const long long ncols = (long long)p->ne * p->nx * p->ny;
    return launch_derivative(order == 1 ? p->d_Dz : p->d_Dzz, p->N, ncols, u, out, (hipStream_t)stream);
}
