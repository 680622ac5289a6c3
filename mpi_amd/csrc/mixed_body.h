// mixed_body.h -- device body of the MIXED form's mode solve (mixed.hip),
// FP32 matrix-core solves refined with FP64 matrix-core residuals.
//
// Per real spectral column (x index, y index, re/im) and element e the
// interior system is A x = b with A = Dzz_int - sigma I (the reference's
// operator, TestNonPeriodic3D_MPI.cpp:71-79) and b the interior of f_e
// (NonPeriodic3D_MPI.cpp:188-201); with Dirichlet data on one element the
// end values are folded in, b = f_int - Dzz[1..N-2][0] g_l/beta_l
// - Dzz[1..N-2][N-1] g_r/beta_r, so that x is the whole interior solution.
// With the plan's even/odd eigendecomposition of the symmetrised operator,
// (Dzz_int + J Dzz_int J)/2 = F^{-1} W diag(mu) W^{-1} F (F the fold of
// mirrored entries, W block diagonal; eigen_decompose_split), held in FP32,
// one FP32 solve is
//   x = F^{-1} W32 ((W32^{-1} F fl32(b)) / (mu - sigma))   (two half-size FP32
//                                                        contractions)
// and S refinement sweeps follow, each
//   r = b - Dzz_int x + sigma x                          (FP64 contraction)
//   x = x + F^{-1} W32 ((W32^{-1} F fl32(r)) / (mu - sigma))
// The residual uses the reference's own operator, so the symmetrisation
// (a relative change of about 1e-12, the GLL nodes' rounding) is removed by
// the refinement with the FP32 solve's own error.  The FP32 solve's backward
// error does not depend on sigma -- the residual of a solve through the
// eigenbasis is V d with |d| ~ n eps_32 |V^{-1}| |r|
// -- so every mode contracts by about cond(V) n 2^-24 per sweep (cond(V) is
// 2.7 at N = 129), the ill-conditioned low-|k| modes included, and the
// iterate reaches the FP64 solution's accuracy, set by the FP64 residual.
// With several elements, Robin faces or periodic z, x is the particular
// solution (zero end values); its end-point derivatives Dz[0|N-1].x
// (GLLQuadrature.cpp:234-245), the element-interface sweep
// (NonPeriodic3D_MPI.cpp:327-580) and the assembly u = x + zl ul + zr ur with
// the FP64 homogeneous solutions of the mode (:203-222, :582-587) follow in
// FP64.
//
// A work-group owns RT tile rows (spectral column x element) and all NP
// (padded n) indices, like the diagonalised solve (diag_body.h): the
// iterate X^T in LDS in FP64, the FP32 operand of each FP32 contraction in
// LDS, the B operands (Vinv32^T, V32^T, Dzz_int^T) from L2.  The relative
// max-norm residual |r| / |b| of every row before each sweep goes to
// hist[sweep][row].
//
// Operand layouts (measured on gfx950 with in-bounds probes): for both
// v_mfma_f32_16x16x4f32 and v_mfma_f64_16x16x4f64, A[i][k] is in lane
// i + 16k and B[k][j] in lane j + 16k; the FP32 result D[i][j] is element
// i % 4 of lane j + 16 (i / 4), the FP64 result element i / 4 of lane
// j + 16 (i % 4).
#pragma once

#include "diag_body.h"

namespace hz {
namespace {

// acc[rb][jt] += A B in FP32: A = the FP32 LDS tile (rows 16 (r0 + rb) + i,
// column k, row stride LDF), B[k][j] = B[k * NP + j] from L2, loaded two
// k-steps ahead (as tile_gemm); the k loop stays rolled (the solve runs it
// in a loop of sweeps, and unrolled it spills the hoisted B loads).
template <int NP, int LDF, int RBK, int JT, int KN = NP>
__device__ __forceinline__ void tile_gemm32(f4 (&acc)[RBK][JT], const float *tile, const float *__restrict__ B,
                                            int c0, int r0, int li, int lk)
{
    constexpr int NK = KN / 4;
    static_assert(NK % 2 == 0, "two k-steps per iteration");
    const float *bcol = B + 16 * c0 + li + (long long)lk * NP;
    tile += 16 * r0 * LDF;
    float b0[JT], b1[JT];
#pragma unroll
    for (int jt = 0; jt < JT; ++jt) {
        b0[jt] = bcol[16 * jt];
        b1[jt] = bcol[(long long)4 * NP + 16 * jt];
    }
#pragma nounroll
    for (int ks = 0; ks < NK; ks += 2) {
        float n0[JT], n1[JT];
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) n0[jt] = n1[jt] = 0.0f;
        if (ks + 2 < NK) {
#pragma unroll
            for (int jt = 0; jt < JT; ++jt) n0[jt] = bcol[(long long)4 * (ks + 2) * NP + 16 * jt];
        }
#pragma unroll
        for (int rb = 0; rb < RBK; ++rb) {
            const float av = tile[(16 * rb + li) * LDF + 4 * ks + lk];
#pragma unroll
            for (int jt = 0; jt < JT; ++jt)
                acc[rb][jt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, b0[jt], acc[rb][jt], 0, 0, 0);
        }
        if (ks + 3 < NK) {
#pragma unroll
            for (int jt = 0; jt < JT; ++jt) n1[jt] = bcol[(long long)4 * (ks + 3) * NP + 16 * jt];
        }
#pragma unroll
        for (int rb = 0; rb < RBK; ++rb) {
            const float av = tile[(16 * rb + li) * LDF + 4 * (ks + 1) + lk];
#pragma unroll
            for (int jt = 0; jt < JT; ++jt)
                acc[rb][jt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, b1[jt], acc[rb][jt], 0, 0, 0);
        }
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) {
            b0[jt] = n0[jt];
            b1[jt] = n1[jt];
        }
    }
}

// LDS of one MIXED tile.
template <int NJ, int RT, bool GENERAL>
struct MixedShared {
    static constexpr int LD = 16 * NJ + 2;   // FP64 row stride (doubles)
    static constexpr int LDF = 16 * NJ + 4;  // FP32 row stride (floats)
    double x[RT * LD];                       // the iterate X^T
    float t[RT * LDF];                       // FP32 operand: fl32(residual), then the scaled eigen coordinates
    double row_sig[RT], row_zl[RT], row_zr[RT], row_d0[RT], row_dN[RT];
    long long row_col[RT];
    int row_g[RT];                           // global row index (history), -1 for an empty row
    unsigned row_bmax[RT], row_rmax[RT];     // max |b|, max |r| per row (float bits; >= 0 orders as unsigned)
};

// The FP32 solve of the rows' right-hand sides held in sh.t (fl32 of b or of
// r), added to the iterate (first: stored).  It runs in the even/odd basis
// of the symmetrised operator (eigen_decompose_split): sh.t is folded in
// place, each wave contracts over its half's k range (half the matrix work
// of the full basis), and the result is unfolded into the iterate.  The
// symmetrised operator differs from the reference's by the rounding of the
// GLL nodes (about 1e-12 relative), far below the FP32 solve's own error,
// which the FP64 residuals of the refinement remove.  Ends with a barrier.
template <int NJ, int RT, bool GENERAL>
__device__ __forceinline__ void mixed_fp32_solve(const MixedArgs &a, MixedShared<NJ, RT, GENERAL> &sh, bool first,
                                                 int c0, int r0, int li, int lk)
{
    constexpr int NP = 16 * NJ, H = NP / 2;
    constexpr int LD = MixedShared<NJ, RT, GENERAL>::LD, LDF = MixedShared<NJ, RT, GENERAL>::LDF;
    constexpr int WC = NJ % 4 == 0 ? 4 : 2, WR = 4 / WC, JT = NJ / WC, RBW = RT / 16 / WR;
    constexpr int NIT = RT * NP / 256;
    static_assert(RT * NP % 256 == 0, "whole passes");
    const int n = a.n, ho = n / 2, he = n - ho;
    const int kb = 16 * c0 >= H ? H : 0;  // the wave's half
    const int tid = threadIdx.x;
    // ---- fold: even coordinate i at i (t_i + t_{n-1-i}, the centre of an odd
    // n alone), odd coordinate i at H + i (t_i - t_{n-1-i}); zero padding ----
    {
        float v[NIT];
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int idx = tid + 256 * it, row = idx / NP, k = idx % NP;
            const bool even = k < H;
            const int i = even ? k : k - H;
            float x0 = 0.0f, x1 = 0.0f;
            if (i < (even ? he : ho)) {
                x0 = sh.t[row * LDF + i];
                if (i < ho) x1 = sh.t[row * LDF + n - 1 - i];
            }
            v[it] = even ? x0 + x1 : x0 - x1;
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int idx = tid + 256 * it;
            sh.t[(idx / NP) * LDF + idx % NP] = v[it];
        }
        __syncthreads();
    }
    f4 acc[RBW][JT];
#pragma unroll
    for (int rb = 0; rb < RBW; ++rb)
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) acc[rb][jt] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    tile_gemm32<NP, LDF, RBW, JT, H>(acc, sh.t + kb, a.VinvT32 + (long long)kb * NP, c0, r0, li, lk);
    __syncthreads();  // every wave has read the right-hand sides
#pragma unroll
    for (int rb = 0; rb < RBW; ++rb)
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) {
            const int j = 16 * (c0 + jt) + li;
            const bool jv = j < H ? j < he : j - H < ho;
            const double mu = jv ? a.mu[j] : 0.0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * (r0 + rb) + 4 * lk + r;
                const float iv = jv ? (float)(1.0 / (mu - sh.row_sig[row])) : 0.0f;
                sh.t[row * LDF + j] = acc[rb][jt][r] * iv;
            }
        }
    __syncthreads();
#pragma unroll
    for (int rb = 0; rb < RBW; ++rb)
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) acc[rb][jt] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    tile_gemm32<NP, LDF, RBW, JT, H>(acc, sh.t + kb, a.VT32 + (long long)kb * NP, c0, r0, li, lk);
    __syncthreads();  // every wave has read the scaled coordinates
#pragma unroll
    for (int rb = 0; rb < RBW; ++rb)
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) {
            const int j = 16 * (c0 + jt) + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) sh.t[(16 * (r0 + rb) + 4 * lk + r) * LDF + j] = acc[rb][jt][r];
        }
    __syncthreads();
    // ---- unfold into the iterate: x_i += e_i + o_i, x_{n-1-i} += e_i - o_i,
    // the centre of an odd n e_i; the padding stays zero ----
    for (int idx = tid; idx < RT * NP; idx += 256) {
        const int row = idx / NP, k = idx % NP;
        double &xv = sh.x[row * LD + k];
        if (k >= n) {
            if (first) xv = 0.0;
            continue;
        }
        const int i = k < n - 1 - k ? k : n - 1 - k;
        const float e = sh.t[row * LDF + i];
        const float o = 2 * k + 1 == n ? 0.0f : sh.t[row * LDF + H + i];
        const double d = (double)(k <= n - 1 - k ? e + o : e - o);
        xv = first ? d : xv + d;
    }
    __syncthreads();
}

// b of tile row `row` at interior index k (0 past the end or for an empty row).
template <int NJ, int RT, bool GENERAL>
__device__ __forceinline__ double mixed_rhs(const MixedArgs &a, const MixedShared<NJ, RT, GENERAL> &sh, int row,
                                            int k)
{
    const long long col = sh.row_col[row];
    if (col < 0 || k >= a.n) return 0.0;
    double v = a.spec[col + 1 + k];
    if constexpr (!GENERAL) {  // Dirichlet ends folded into the right-hand side
        const double *dzz = a.Dzz + (long long)(k + 1) * a.N;
        v = v - dzz[0] * sh.row_zl[row] - dzz[a.N - 1] * sh.row_zr[row];
    }
    return v;
}

// Tile `bx` (RT rows) of the MIXED mode solve; 256 threads.
template <int NJ, int RT, bool GENERAL>
__device__ __forceinline__ void mixed_tile(const MixedArgs &a, int bx, MixedShared<NJ, RT, GENERAL> &sh)
{
    constexpr int NP = 16 * NJ;
    constexpr int LD = MixedShared<NJ, RT, GENERAL>::LD, LDF = MixedShared<NJ, RT, GENERAL>::LDF;
    constexpr int WC = NJ % 4 == 0 ? 4 : 2, WR = 4 / WC, JT = NJ / WC, RBK = RT / 16, RBW = RBK / WR;
    static_assert(NJ % WC == 0 && RBK % WR == 0, "the waves split the tile evenly");
    __syncthreads();

    const int tid = threadIdx.x, lane = tid % WAVE, w = uniform(tid / WAVE);
    const int c0 = (w % WC) * JT, r0 = (w / WC) * RBW;
    const int li = lane % 16, lk = lane / 16;
    const int n = a.n, N = a.N, E = a.ne, K = a.K, M = a.M;
    const int rows_per_fc = 2 * E;
    const int fc_per_tile = RT / rows_per_fc;
    const int fc0 = bx * fc_per_tile;
    const int nfc = a.nx * M;

    // ---- row bookkeeping: row = (fc_local*2 + plane)*E + e ----
    for (int row = tid; row < RT; row += 256) {
        const int fl = row / rows_per_fc, rem = row % rows_per_fc;
        const int plane = rem / E, e = rem % E;
        const int fc = fc0 + fl;
        double sig = 0.0, zl = 0.0, zr = 0.0;
        long long col = -1;
        int g = -1;
        if (fl < fc_per_tile && fc < nfc) {
            const int i = fc / M, m = fc % M;
            const int iw = i <= a.nx / 2 ? i : a.nx - i;
            sig = a.shift[iw * M + m];
            col = (long long)i * a.xs + (long long)m * K + (long long)e * (N - 1) + (plane ? a.plane_off : 0);
            g = (fc * 2 + plane) * E + e;
            if (!GENERAL) {  // one element, Dirichlet data on both faces
                const double *bsrc = plane ? a.bc_im : a.bc_re;
                zl = bsrc[fc] / a.beta_l;
                zr = bsrc[(long long)nfc + fc] / a.beta_r;
            }
        }
        sh.row_sig[row] = sig;
        sh.row_col[row] = col;
        sh.row_g[row] = g;
        sh.row_zl[row] = zl;
        sh.row_zr[row] = zr;
        sh.row_bmax[row] = 0u;
    }
    __syncthreads();
    // ---- fl32(b) into the FP32 tile (loads in chunks, all issued before the stores) ----
    {
        constexpr int NIT = RT * NP / 256;
        constexpr int CH = NIT % 16 == 0 ? 16 : NIT % 8 == 0 ? 8 : NIT;
        static_assert(NIT % CH == 0 && RT * NP % 256 == 0, "whole chunks");
#pragma unroll
        for (int cc = 0; cc < NIT; cc += CH) {
            double v[CH];
#pragma unroll
            for (int it = 0; it < CH; ++it) {
                const int idx = tid + 256 * (cc + it);
                v[it] = mixed_rhs<NJ, RT, GENERAL>(a, sh, idx / NP, idx % NP);
            }
#pragma unroll
            for (int it = 0; it < CH; ++it) {
                const int idx = tid + 256 * (cc + it), row = idx / NP;
                const float f = (float)v[it];
                sh.t[row * LDF + idx % NP] = f;
                if (a.hist && f != 0.0f) atomicMax(&sh.row_bmax[row], __float_as_uint(fabsf(f)));
            }
        }
    }
    __syncthreads();
    mixed_fp32_solve<NJ, RT, GENERAL>(a, sh, true, c0, r0, li, lk);

    // ---- refinement sweeps: r = b - Dzz_int x + sigma x in FP64, then an FP32 solve of r ----
    constexpr int RBW64 = RBW;
    for (int s = 0; s < a.sweeps; ++s) {
        for (int row = tid; row < RT; row += 256) sh.row_rmax[row] = 0u;
        d4 acc[RBW64][JT];
#pragma unroll
        for (int rb = 0; rb < RBW64; ++rb)
#pragma unroll
            for (int jt = 0; jt < JT; ++jt) acc[rb][jt] = d4{0.0, 0.0, 0.0, 0.0};
        tile_gemm<NP, LD, RBW64, JT, true>(acc, sh.x, a.DzzTP, c0, r0, li, lk);
        __syncthreads();  // row_rmax cleared for every wave
#pragma unroll
        for (int rb = 0; rb < RBW64; ++rb)
#pragma unroll
            for (int jt = 0; jt < JT; ++jt) {
                const int j = 16 * (c0 + jt) + li;
                float m = 0.0f;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * (r0 + rb) + 4 * r + lk;  // FP64 result layout
                    const double xv = sh.x[row * LD + j];
                    const double rv = mixed_rhs<NJ, RT, GENERAL>(a, sh, row, j) - acc[rb][jt][r] +
                                      sh.row_sig[row] * xv;
                    const float rf = j < n ? (float)rv : 0.0f;
                    sh.t[row * LDF + j] = rf;
                    if (a.hist) {
                        // max over the 16 lanes that hold this row (lanes li = 0..15)
                        m = fabsf(rf);
#pragma unroll
                        for (int o = 1; o < 16; o <<= 1) m = fmaxf(m, __shfl_xor(m, o));
                        if (li == 0 && m != 0.0f) atomicMax(&sh.row_rmax[row], __float_as_uint(m));
                    }
                }
            }
        __syncthreads();
        if (a.hist)
            for (int row = tid; row < RT; row += 256) {
                const int g = sh.row_g[row];
                if (g < 0) continue;
                const float bm = __uint_as_float(sh.row_bmax[row]), rm = __uint_as_float(sh.row_rmax[row]);
                a.hist[(long long)s * a.hist_rows + g] = bm > 0.0f ? (double)rm / (double)bm : 0.0;
            }
        mixed_fp32_solve<NJ, RT, GENERAL>(a, sh, false, c0, r0, li, lk);
    }

    // ---- end values: Dirichlet (folded above) or the element-interface sweep ----
    if constexpr (GENERAL) {
        // end-point derivatives of the particular solutions (zero end values)
        for (int row = tid; row < RT; row += 256) {
            double d0 = 0.0, dN = 0.0;
            if (sh.row_col[row] >= 0) {
                const double *xr = sh.x + row * LD;
                const double *dz0 = a.Dz + 1, *dzN = a.Dz + (long long)(N - 1) * N + 1;
                for (int k = 0; k < n; ++k) {
                    d0 = fma(dz0[k], xr[k], d0);
                    dN = fma(dzN[k], xr[k], dN);
                }
            }
            sh.row_d0[row] = d0;
            sh.row_dN[row] = dN;
        }
        __syncthreads();
        for (int sc = tid; sc < 2 * fc_per_tile; sc += 256) {
            const int fl = sc >> 1, plane = sc & 1;
            const int fc = fc0 + fl;
            if (fc >= nfc) continue;
            const int i = fc / M, m = fc % M;
            const int iw = i <= a.nx / 2 ? i : a.nx - i;
            const int mode = iw * M + m;
            const double d0ul = a.hom[4 * mode + 0], dNul = a.hom[4 * mode + 1];
            const double d0ur = a.hom[4 * mode + 2], dNur = a.hom[4 * mode + 3];
            const int r0w = sc * E;
            if (a.periodic) {
                periodic_sweep(E, d0ul, dNul, d0ur, dNur, sh.row_d0 + r0w, sh.row_dN + r0w, sh.row_zl + r0w,
                               sh.row_zr + r0w, 1, a.closure ? a.closure + (plane ? nfc : 0) + fc : nullptr);
                continue;
            }
            const double *bsrc = plane ? a.bc_im : a.bc_re;
            robin_sweep(E, d0ul, dNul, d0ur, dNur, bsrc[fc], bsrc[(long long)nfc + fc], a.alpha_l, a.beta_l,
                        a.alpha_r, a.beta_r, sh.row_d0 + r0w, sh.row_dN + r0w, sh.row_zl + r0w, sh.row_zr + r0w);
        }
        __syncthreads();
    }

    // ---- store u along z: x (+ zl ul + zr ur with several elements or Robin faces) ----
    for (int idx = tid; idx < RT * NP; idx += 256) {
        const int row = idx / NP, k = idx % NP;
        const long long col = sh.row_col[row];
        if (col < 0 || k >= n) continue;
        double v = sh.x[row * LD + k];
        if constexpr (GENERAL) {
            const int fc = fc0 + row / rows_per_fc;
            const int i = fc / M, m = fc % M;
            const int mode = (i <= a.nx / 2 ? i : a.nx - i) * M + m;
            v = v + (sh.row_zl[row] * a.UL[(long long)mode * n + k] + sh.row_zr[row] * a.UR[(long long)mode * n + k]);
        }
        a.spec[col + 1 + k] = v;
    }
    for (int row = tid; row < RT; row += 256) {
        const long long col = sh.row_col[row];
        if (col >= 0) {
            a.spec[col] = sh.row_zl[row];
            a.spec[col + N - 1] = sh.row_zr[row];
        }
    }
}

}  // namespace
}  // namespace hz
