// lu.hip -- the reference algorithm per (|kx|,|ky|) mode, one work-group per
// mode, the whole augmented system resident in registers.
//
// For mode (iw, jw) the work-group
//   1. assembles A = Dzz_int - sigma I (or the caller's operator) and appends
//      the right-hand sides as extra columns: the two homogeneous ones
//      (-Dzz[1..N-2][0], -Dzz[1..N-2][N-1], NonPeriodic3D_MPI.cpp:211-219)
//      and the interior spectrum of every element for each real column of
//      the mode (:188-201, :248-325);
//   2. eliminates with the reference's scaled partial pivoting
//      (LU_Decomposition.cpp:7-72: scale of the ORIGINAL row at each position,
//      strictly-greater test so the first maximum wins, singular test
//      |scaled pivot| <= 1e-12, swaps of columns >= k done with the reference's
//      add/subtract exchange).  Carrying the right-hand sides through the
//      elimination is exactly LUSolve's forward phase (:86-95);
//   3. back-substitutes the right-hand sides (:97-107);
//   4. forms the end-point derivatives Dz[0|N-1].col (GLLQuadrature.cpp:234-245),
//      runs the element-interface sweep with the Robin end rows
//      (NonPeriodic3D_MPI.cpp:327-580) and assembles u = p + zl ul + zr ur
//      (:582-587) straight into the spectrum.
//
// Layout: TR x TC threads, thread (r, c) = threadIdx r + TR*c owns rows
// r + TR*a (a < RB) and columns c + TC*b (b < CB) of the augmented matrix.
// The TR threads of one column group are consecutive lanes of one wave, so
// pivot searches and back-substitution broadcasts are wave-local.
#include "hz_internal.h"

#include <climits>

namespace hz {
namespace {

constexpr double kEps = 1.000000000000E-12;  // include/General.h:15

template <int WIDTH>
__device__ __forceinline__ double shfl_d(double v, int src)
{
    return __shfl(v, src, WIDTH);
}

template <int WIDTH>
__device__ __forceinline__ double shfl_xor_d(double v, int m)
{
    return __shfl_xor(v, m, WIDTH);
}

enum { MODE_SOLVE = 0, MODE_FACTOR = 1 };

// Elimination core shared by the mode solver and the batched factorisation.
// ncols_mat = n (columns that belong to the matrix); columns >= n are
// right-hand sides.  Returns the singular flag; det accumulated in *det.
template <int TR, int TC, int RB, int CB>
struct Core {
    static constexpr int NT = TR * TC;
    static constexpr int ROWS = TR * RB;
    static constexpr int COLS = TC * CB;

    struct Smem {
        double lbuf[2][ROWS];
        double rowk[COLS], rowp[COLS];
        double srow[ROWS];
        double part[TC][ROWS];
        int p[2];
        double det;
        int singular;
    };

    // A[a][b] register tile.  Returns nothing; s.singular set on failure.
    static __device__ __forceinline__ void eliminate(double (&A)[RB][CB], int n, Smem &s, int r, int c,
                                                     int *piv_out)
    {
        // Row scales from the original matrix columns (< n), LU_Decomposition.cpp:20-29.
#pragma unroll
        for (int a = 0; a < RB; ++a) {
            double m = 0.0;
#pragma unroll
            for (int b = 0; b < CB; ++b) {
                const int j = c + TC * b;
                if (j < n) m = fmax(m, fabs(A[a][b]));
            }
            s.part[c][r + TR * a] = m;
        }
        if (threadIdx.x == 0) {
            s.det = 1.0;
            s.singular = 0;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < ROWS; i += NT) {
            double m = s.part[0][i];
            for (int cc = 1; cc < TC; ++cc) m = fmax(m, s.part[cc][i]);
            s.srow[i] = m;
        }
        __syncthreads();
        double srow[RB];
#pragma unroll
        for (int a = 0; a < RB; ++a) srow[a] = s.srow[r + TR * a];

        for (int k = 0; k < n - 1; ++k) {
            const int ck = k % TC, bk = k / TC, buf = k & 1;
            // ---- phase A: pivot search, multipliers (column-k group) ----
            if (c == ck) {
                double col[RB];
#pragma unroll
                for (int a = 0; a < RB; ++a) {
                    double v = 0.0;
#pragma unroll
                    for (int b = 0; b < CB; ++b)
                        if (b == bk) v = A[a][b];
                    col[a] = v;
                }
                double best = -1.0;
                int bi = INT_MAX;
#pragma unroll
                for (int a = 0; a < RB; ++a) {
                    const int i = r + TR * a;
                    if (i >= k && i < n) {
                        const double ratio = fabs(col[a]) / srow[a];
                        if (ratio > best || (ratio == best && i < bi)) {
                            best = ratio;
                            bi = i;
                        }
                    }
                }
#pragma unroll
                for (int m = 1; m < TR; m <<= 1) {
                    const double ob = shfl_xor_d<TR>(best, m);
                    const int oi = __shfl_xor(bi, m, TR);
                    if (ob > best || (ob == best && oi < bi)) {
                        best = ob;
                        bi = oi;
                    }
                }
                if (bi == INT_MAX) {  // no comparable candidate (NaN scales): report singular
                    bi = k;
                    best = 0.0;
                }
                const int p = bi;
                const int rk = k % TR, ak = k / TR, rp = p % TR, ap = p / TR;
                double ck_own = 0.0, cp_own = 0.0;
#pragma unroll
                for (int a = 0; a < RB; ++a) {
                    if (a == ak) ck_own = col[a];
                    if (a == ap) cp_own = col[a];
                }
                const double akk = shfl_d<TR>(ck_own, rk);
                const double apk = shfl_d<TR>(cp_own, rp);
                double piv = akk, newpk = apk;
                if (p != k) {
                    const double sum = akk + apk;
                    newpk = sum - apk;
                    piv = sum - newpk;
                }
#pragma unroll
                for (int a = 0; a < RB; ++a) {
                    const int i = r + TR * a;
                    double l = 0.0;
                    if (i > k && i < n) {
                        const double v = (i == p) ? newpk : col[a];
                        l = v / piv;
                    }
                    s.lbuf[buf][i] = l;
                    // column k of the factor: multipliers below, pivot on the diagonal
#pragma unroll
                    for (int b = 0; b < CB; ++b)
                        if (b == bk) {
                            if (i > k && i < n) A[a][b] = l;
                            if (i == k) A[a][b] = piv;
                        }
                }
                if (r == 0) {
                    s.p[buf] = p;
                    if (piv_out) piv_out[k] = p;
                    double det = s.det;
                    if (p != k) det = -det;
                    s.det = det * piv;
                    if (fabs(best) <= kEps) s.singular = 1;
                }
            }
            __syncthreads();
            // ---- phase B: publish rows k and p (columns > k) ----
            const int p = s.p[buf];
            const int rk = k % TR, ak = k / TR, rp = p % TR, ap = p / TR;
            if (r == rk) {
#pragma unroll
                for (int b = 0; b < CB; ++b) {
                    const int j = c + TC * b;
                    double v = 0.0;
#pragma unroll
                    for (int a = 0; a < RB; ++a)
                        if (a == ak) v = A[a][b];
                    if (j > k) s.rowk[j] = v;
                }
            }
            if (p != k && r == rp) {
#pragma unroll
                for (int b = 0; b < CB; ++b) {
                    const int j = c + TC * b;
                    double v = 0.0;
#pragma unroll
                    for (int a = 0; a < RB; ++a)
                        if (a == ap) v = A[a][b];
                    if (j > k) s.rowp[j] = v;
                }
            }
            __syncthreads();
            // ---- phase C: swap + rank-1 update ----
            double P[CB], NP[CB];
#pragma unroll
            for (int b = 0; b < CB; ++b) {
                const int j = c + TC * b;
                P[b] = 0.0;
                NP[b] = 0.0;
                if (TC * b + TC - 1 > k && j > k) {
                    const double vk = s.rowk[j];
                    if (p != k) {
                        const double vp = s.rowp[j];
                        const double sum = vk + vp;
                        if (j < n) {  // LUDecomposition: Swap(A[k][j], A[p][j])
                            NP[b] = sum - vp;
                            P[b] = sum - NP[b];
                        } else {      // LUSolve: Swap(b[p], b[k])
                            P[b] = sum - vk;
                            NP[b] = sum - P[b];
                        }
                    } else {
                        P[b] = vk;
                    }
                }
            }
            double Lr[RB];
#pragma unroll
            for (int a = 0; a < RB; ++a) Lr[a] = s.lbuf[buf][r + TR * a];
#pragma unroll
            for (int a = 0; a < RB; ++a) {
                if (TR * a + TR - 1 < k) continue;  // every row of this slot is final
#pragma unroll
                for (int b = 0; b < CB; ++b) {
                    if (TC * b + TC - 1 <= k) continue;  // every column of this slot is final
                    const int i = r + TR * a, j = c + TC * b;
                    double v = A[a][b];
                    if (a == ap && p != k && i == p && j > k) v = NP[b];
                    v -= Lr[a] * P[b];
                    if (a == ak && i == k && j > k) v = P[b];
                    A[a][b] = v;
                }
            }
        }
        // final pivot (not tested by the reference), determinant
        __syncthreads();
        if (n >= 1) {
            const int k = n - 1, ck = k % TC, bk = k / TC, rk = k % TR, ak = k / TR;
            if (c == ck && r == rk) {
                double v = 0.0;
#pragma unroll
                for (int a = 0; a < RB; ++a)
#pragma unroll
                    for (int b = 0; b < CB; ++b)
                        if (a == ak && b == bk) v = A[a][b];
                s.det = s.det * v;
            }
        }
        __syncthreads();
    }
};

// Mode solver: see the file comment.  NRHS_MAX = CB*TC - n columns available.
template <int TR, int TC, int RB, int CB>
__global__ __launch_bounds__(TR *TC) void lu_modes_kernel(const LuModeArgs a)
{
    using C = Core<TR, TC, RB, CB>;
    constexpr int ROWS = C::ROWS;
    __shared__ typename C::Smem sm;
    __shared__ double ubuf[2][TC][ROWS];
    __shared__ double sh_ul[ROWS], sh_ur[ROWS];
    __shared__ double dots[2][128];  // [0|N-1][rhs index]
    __shared__ double zeta[2][128];  // [left|right][element*4 + real column]

    const int tid = threadIdx.x, r = tid % TR, c = tid / TR;
    const int n = a.n, N = a.N, E = a.ne, K = a.K, M = a.M;
    const int mode = blockIdx.x;
    const int iw = mode / M, jw = mode % M;
    // real columns of this mode: (x index, re/im)
    int xi[2];
    int nxi = 1;
    xi[0] = iw;
    if (iw != 0 && iw != a.nx / 2) {
        xi[1] = a.nx - iw;
        nxi = 2;
    }
    const int nc = 2 * nxi;
    const int nrhs = 2 + 4 * E;

    // ---- assemble the augmented matrix ----
    double A[RB][CB];
    const double sig = a.shift ? a.shift[mode] : 0.0;
#pragma unroll
    for (int aa = 0; aa < RB; ++aa) {
        const int i = r + TR * aa;
#pragma unroll
        for (int b = 0; b < CB; ++b) {
            const int j = c + TC * b;
            double v = 0.0;
            if (i < n) {
                if (j < n) {
                    if (a.op_full) {
                        v = a.op_full[((long long)mode * n + i) * n + j];
                    } else {
                        v = a.DzzT[(long long)j * n + i];
                        if (i == j) v = a.op_diag ? a.op_diag[(long long)mode * n + i] : v - sig;
                    }
                } else if (j < n + nrhs) {
                    const int t = j - n;
                    if (t == 0) {
                        v = -a.Dzz[(long long)(i + 1) * N];
                    } else if (t == 1) {
                        v = -a.Dzz[(long long)(i + 1) * N + (N - 1)];
                    } else {
                        const int e = (t - 2) >> 2, cc = (t - 2) & 3;
                        if (cc < nc) {
                            const long long col = ((long long)xi[cc >> 1] * M + jw) * K;
                            const double *src = (cc & 1) ? a.spec_im : a.spec_re;
                            v = src[col + (long long)e * (N - 1) + 1 + i];
                        }
                    }
                }
            }
            A[aa][b] = v;
        }
    }

    C::eliminate(A, n, sm, r, c, nullptr);

    // ---- back substitution of the right-hand-side columns ----
    const int nblk = (n + TC - 1) / TC;
    int buf = 0;
    for (int blk = nblk - 1; blk >= 0; --blk) {
#pragma unroll
        for (int b = 0; b < CB; ++b)
            if (b == blk) {
#pragma unroll
                for (int aa = 0; aa < RB; ++aa) ubuf[buf][c][r + TR * aa] = A[aa][b];
            }
        __syncthreads();
        const int q_hi = min(n - 1, TC * blk + TC - 1), q_lo = TC * blk;
#pragma unroll
        for (int b = 0; b < CB; ++b) {
            const int j = c + TC * b;
            if (TC * b + TC - 1 < n) continue;          // slot holds no right-hand side
            if (!(j >= n && j < n + nrhs)) continue;    // this lane's column is not one
            for (int q = q_hi; q >= q_lo; --q) {
                const int rq = q % TR, aq = q / TR;
                double own = 0.0;
#pragma unroll
                for (int aa = 0; aa < RB; ++aa)
                    if (aa == aq) own = A[aa][b];
                const double yq = shfl_d<TR>(own, rq);
                const double xq = yq / ubuf[buf][q - q_lo][q];
#pragma unroll
                for (int aa = 0; aa < RB; ++aa) {
                    const int row = r + TR * aa;
                    if (row < q) A[aa][b] -= ubuf[buf][q - q_lo][row] * xq;
                    if (aa == aq && r == rq) A[aa][b] = xq;
                }
            }
        }
        buf ^= 1;
    }
    __syncthreads();

    // ---- end-point derivatives of every right-hand-side column ----
#pragma unroll
    for (int b = 0; b < CB; ++b) {
        const int j = c + TC * b;
        if (TC * b + TC - 1 < n) continue;
        const bool mine = (j >= n && j < n + nrhs);
        double d0 = 0.0, dN = 0.0;
        if (mine) {
#pragma unroll
            for (int aa = 0; aa < RB; ++aa) {
                const int row = r + TR * aa;
                if (row < n) {
                    d0 += a.Dz[row + 1] * A[aa][b];
                    dN += a.Dz[(long long)(N - 1) * N + row + 1] * A[aa][b];
                }
            }
        }
#pragma unroll
        for (int m = 1; m < TR; m <<= 1) {
            d0 += shfl_xor_d<TR>(d0, m);
            dN += shfl_xor_d<TR>(dN, m);
        }
        if (mine) {
            const int t = j - n;
            if (t == 0) {  // ul has u(0) = 1
                d0 += a.Dz[0];
                dN += a.Dz[(long long)(N - 1) * N];
            } else if (t == 1) {  // ur has u(N-1) = 1
                d0 += a.Dz[N - 1];
                dN += a.Dz[(long long)(N - 1) * N + N - 1];
            }
            if (r == 0 && t < 128) {
                dots[0][t] = d0;
                dots[1][t] = dN;
            }
#pragma unroll
            for (int aa = 0; aa < RB; ++aa) {
                const int row = r + TR * aa;
                if (t == 0) sh_ul[row] = A[aa][b];
                if (t == 1) sh_ur[row] = A[aa][b];
            }
        }
    }
    __syncthreads();

    // ---- element-interface sweep, one thread per real column ----
    if (tid < nc) {
        const int cc = tid;
        const long long bcol = (long long)xi[cc >> 1] * M + jw;
        const double *bsrc = (cc & 1) ? a.bc_im : a.bc_re;
        const double gl = bsrc[bcol], gr = bsrc[(long long)a.nx * M + bcol];
        const double d0ul = dots[0][0], dNul = dots[1][0], d0ur = dots[0][1], dNur = dots[1][1];
        double bb[32], cv[32], rh[32];
        bb[0] = -a.alpha_l * d0ul + a.beta_l;
        cv[0] = -a.alpha_l * d0ur;
        rh[0] = gl + a.alpha_l * dots[0][2 + cc];
        for (int e = 1; e < E; ++e) {
            const double lb = dNul;
            bb[e] = dNur - d0ul - (cv[e - 1] / bb[e - 1]) * lb;
            cv[e] = -d0ur;
            rh[e] = dots[0][2 + 4 * e + cc] - dots[1][2 + 4 * (e - 1) + cc] - (rh[e - 1] / bb[e - 1]) * lb;
        }
        const double lb = a.alpha_r * dNul;
        const double br = a.alpha_r * dNur + a.beta_r - (cv[E - 1] / bb[E - 1]) * lb;
        const double rhsr = gr - a.alpha_r * dots[1][2 + 4 * (E - 1) + cc] - (rh[E - 1] / bb[E - 1]) * lb;
        double zr = rhsr / br;
        for (int e = E - 1; e >= 0; --e) {
            const double zl = (rh[e] - cv[e] * zr) / bb[e];
            zeta[0][4 * e + cc] = zl;
            zeta[1][4 * e + cc] = zr;
            zr = zl;
        }
    }
    __syncthreads();

    // ---- assembly straight into the spectrum ----
#pragma unroll
    for (int b = 0; b < CB; ++b) {
        const int j = c + TC * b;
        if (TC * b + TC - 1 < n) continue;
        if (!(j >= n + 2 && j < n + nrhs)) continue;
        const int t = j - n, e = (t - 2) >> 2, cc = (t - 2) & 3;
        if (cc >= nc) continue;
        const double zl = zeta[0][4 * e + cc], zr = zeta[1][4 * e + cc];
        const long long col = ((long long)xi[cc >> 1] * M + jw) * K + (long long)e * (N - 1);
        double *dst = (cc & 1) ? a.spec_im : a.spec_re;
#pragma unroll
        for (int aa = 0; aa < RB; ++aa) {
            const int row = r + TR * aa;
            if (row < n) dst[col + 1 + row] = A[aa][b] + (zl * sh_ul[row] + zr * sh_ur[row]);
        }
        if (r == 0) {
            dst[col] = zl;
            dst[col + N - 1] = zr;
        }
    }
    if (tid == 0) a.status[mode] = sm.singular ? HZ_ERR_SINGULAR : 0;
}

// Batched factorisation in the reference's storage (no right-hand sides).
template <int TR, int TC, int RB, int CB>
__global__ __launch_bounds__(TR *TC) void lu_factor_kernel(const LuFactorArgs a)
{
    using C = Core<TR, TC, RB, CB>;
    __shared__ typename C::Smem sm;
    const int tid = threadIdx.x, r = tid % TR, c = tid / TR;
    const int n = a.n;
    double *Am = a.A + (long long)blockIdx.x * n * n;
    double A[RB][CB];
#pragma unroll
    for (int aa = 0; aa < RB; ++aa)
#pragma unroll
        for (int b = 0; b < CB; ++b) {
            const int i = r + TR * aa, j = c + TC * b;
            A[aa][b] = (i < n && j < n) ? Am[(long long)i * n + j] : 0.0;
        }
    C::eliminate(A, n, sm, r, c, a.piv ? a.piv + (long long)blockIdx.x * (n - 1) : nullptr);
#pragma unroll
    for (int aa = 0; aa < RB; ++aa)
#pragma unroll
        for (int b = 0; b < CB; ++b) {
            const int i = r + TR * aa, j = c + TC * b;
            if (i < n && j < n) Am[(long long)i * n + j] = A[aa][b];
        }
    if (tid == 0) {
        if (a.det) a.det[blockIdx.x] = sm.det;
        if (a.status) a.status[blockIdx.x] = sm.singular ? HZ_ERR_SINGULAR : 0;
    }
}

// LU_Decomposition.cpp:78-108 for many systems, one thread per system (a
// utility for callers of the reference's LUSolve; not on the solve path).
__global__ void lu_solve_kernel(const double *LU, const int *piv, double *b, int n, int count)
{
    const int sys = blockIdx.x * blockDim.x + threadIdx.x;
    if (sys >= count) return;
    const double *A = LU + (long long)sys * n * n;
    const int *pv = piv + (long long)sys * (n - 1);
    double *x = b + (long long)sys * n;
    for (int k = 0; k < n - 1; ++k) {
        const int idx = pv[k];
        if (idx != k) {  // Swap(b[index], b[k])
            double u = x[idx], v = x[k];
            u = u + v;
            v = u - v;
            u = u - v;
            x[idx] = u;
            x[k] = v;
        }
        for (int i = k + 1; i < n; ++i) x[i] -= A[(long long)i * n + k] * x[k];
    }
    x[n - 1] /= A[(long long)(n - 1) * n + (n - 1)];
    for (int p = n - 2; p >= 0; p--) {
        double temp = 0;
        for (int q = p + 1; q < n; ++q) temp += A[(long long)p * n + q] * x[q];
        x[p] = (x[p] - temp) / A[(long long)p * n + p];
    }
}

template <int TR, int TC, int RB, int CB>
int launch_modes_t(const LuModeArgs &a, int modes, hipStream_t s)
{
    hipLaunchKernelGGL((lu_modes_kernel<TR, TC, RB, CB>), dim3(modes), dim3(TR * TC), 0, s, a);
    HZ_HIP(hipGetLastError());
    return HZ_OK;
}

template <int TR, int TC, int RB, int CB>
int launch_factor_t(const LuFactorArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL((lu_factor_kernel<TR, TC, RB, CB>), dim3(a.count), dim3(TR * TC), 0, s, a);
    HZ_HIP(hipGetLastError());
    return HZ_OK;
}

}  // namespace

int launch_lu_modes(const LuModeArgs &a, int modes, hipStream_t s)
{
    const int n = a.n, cols = a.n + 2 + 4 * a.ne;
    if (a.ne > 31) return set_error(HZ_ERR_UNSUPPORTED, "PLAIN_LU form supports at most 31 elements");
    if (n <= 8 && cols <= 32) return launch_modes_t<8, 8, 1, 4>(a, modes, s);
    if (n <= 8 && cols <= 128) return launch_modes_t<8, 8, 1, 16>(a, modes, s);
    if (n <= 16 && cols <= 32) return launch_modes_t<8, 8, 2, 4>(a, modes, s);
    if (n <= 16 && cols <= 128) return launch_modes_t<8, 8, 2, 16>(a, modes, s);
    if (n <= 32 && cols <= 48) return launch_modes_t<8, 8, 4, 6>(a, modes, s);
    if (n <= 32 && cols <= 128) return launch_modes_t<8, 8, 4, 16>(a, modes, s);
    if (n <= 64 && cols <= 80) return launch_modes_t<16, 16, 4, 5>(a, modes, s);
    if (n <= 64 && cols <= 128) return launch_modes_t<16, 16, 4, 8>(a, modes, s);
    if (n <= 128 && cols <= 144) return launch_modes_t<16, 16, 8, 9>(a, modes, s);
    if (n <= 128 && cols <= 192) return launch_modes_t<16, 16, 8, 12>(a, modes, s);
    return set_error(HZ_ERR_UNSUPPORTED, "PLAIN_LU form: N-2 + 2 + 4*elements must fit 192 columns (N-2 <= 128)");
}

int launch_lu_factor(const LuFactorArgs &a, hipStream_t s)
{
    if (a.count <= 0) return HZ_OK;
    const int n = a.n;
    if (n <= 8) return launch_factor_t<8, 8, 1, 1>(a, s);
    if (n <= 16) return launch_factor_t<8, 8, 2, 2>(a, s);
    if (n <= 32) return launch_factor_t<8, 8, 4, 4>(a, s);
    if (n <= 64) return launch_factor_t<16, 16, 4, 4>(a, s);
    if (n <= 128) return launch_factor_t<16, 16, 8, 8>(a, s);
    return set_error(HZ_ERR_UNSUPPORTED, "batched FP64 LU supports n <= 128");
}

int launch_lu_solve(const double *LU, const int *piv, double *b, int n, int count, hipStream_t s)
{
    if (count <= 0) return HZ_OK;
    hipLaunchKernelGGL(lu_solve_kernel, dim3((count + 63) / 64), dim3(64), 0, s, LU, piv, b, n, count);
    HZ_HIP(hipGetLastError());
    return HZ_OK;
}

}  // namespace hz
