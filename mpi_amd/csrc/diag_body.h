// diag_body.h -- device body of the z-diagonalised mode solve (diag.hip),
// shared with the fused batch kernel (batch.hip).
//
// The interior GLL operator of every element is the same matrix, so with
// Dzz_int = V diag(mu) V^{-1} (computed once per plan, eig.cpp) the
// reference's per-mode solve A(sigma)^{-1} = V (diag(mu) - sigma)^{-1} V^{-1}
// needs no factorisation.  For each real spectral column (x index, y index,
// re/im) and element e, with sigma = lambda + kx^2 + ky^2:
//   y   = V^{-1} f_e                                  (MFMA contraction)
//   inv = 1 / (mu - sigma),   q = y * inv
//   p_e = V q  is the particular solution, ul = V (w0 * inv), ur = V (wN * inv)
//   with w0 = V^{-1}(-Dzz[1..N-2][0]), wN = V^{-1}(-Dzz[1..N-2][N-1])
//   (NonPeriodic3D_MPI.cpp:188-222); the end-point derivatives are
//   Dz[0|N-1].p_e = r0|rN . q with r0|rN = Dz[0|N-1][1..N-2] V (:175-179);
//   the element-interface sweep (:327-580) gives zl_e, zr_e and
//   u_e = V ((y + zl_e w0 + zr_e wN) * inv)          (MFMA contraction)
//   with u at the element ends equal to zl_e, zr_e (:582-587).
//
// Adjoint (ADJ = true, one element): the transposed column map.  With
// ubar the column of dL/du, y = V^T ubar_int, inv as above,
//   zbl = ubar_0 + (w0*inv).y,  zbr = ubar_{N-1} + (wN*inv).y
// (the adjoints of zl, zr through u = p + zl ul + zr ur), the scalar sweep
// run backwards gives dL/dg_l, dL/dg_r and the adjoints P0b, PNb of the
// end-point derivatives, and
//   fbar_int = V^{-T} ((y + P0b r0 + PNb rN) * inv),  fbar = 0 at the ends.
// The kernel is the forward one with B1 = V, B2 = V^{-1}.
//
// Even/odd instance (EO: one element with Dirichlet faces).  The GLL nodes mirror about the element centre, so Dzz_int
// commutes with the reversal J up to the nodes' rounding; with the fold
// F f = (f_i + f_{n-1-i} | f_i - f_{n-1-i}) the symmetrised operator
// (A + JAJ)/2 is block diagonal and is diagonalised block by block
// (eigen_decompose_split): V -> F^{-1} W, V^{-1} -> Winv F with W, Winv
// block diagonal (even coordinates at 0.., odd ones at NP/2..).  Each wave's
// output columns lie in one half, so its contractions run over that half's
// k range: half the matrix work.  The tile is folded while it is staged and
// unfolded through LDS before the store; between the contractions
// z -= (M z) / (mu - sigma) (M = Winv F (A - JAJ)/2 F^{-1} W, bf16 operands,
// FP32 sums) corrects to first order from the symmetrised operator to A;
// the adjoint applies the transposed correction q -= (M^T q) / (mu - sigma)
// to q = y / (mu - sigma) before the end-value dots.
//
// A work-group owns RT tile rows (spectral column x element) and all NP
// (padded n) eigen/grid indices.  Contractions are transposed so that MFMA
// output lanes run along z, which is contiguous in memory:
//   Y^T = F^T Vinv^T,   U^T = Z^T V^T
// v_mfma_f64_16x16x4f64 operand layout (wave64, measured on gfx950):
// A[i][k] in lane i + 16k, B[k][j] in lane j + 16k, D[i][j] (4 per lane) in
// lane j + 16*(i%4), element i/4.  NP is padded to a multiple of 32; the
// four waves split the output tile's column tiles four ways (NJ = NP/16 a
// multiple of 4) or two ways and its row blocks two ways (NJ = 2, 6), each
// reading A from the LDS tile and B (Vinv^T, V^T) from L2.
#pragma once

#include "hz_internal.h"
#include "sweep.h"

namespace hz {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int WAVE = 64;

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// acc[rb][jt] += A B over the k index, A = the LDS tile (rows
// 16 (r0 + rb) + i, column k), B[k][j] = B[k * NP + j] from L2 (j in the
// column tiles c0 .. c0 + JT - 1).  B is loaded two k-steps ahead so its L2
// latency overlaps the MFMAs of the steps in between.  ROLLED keeps the k
// loop rolled (kernels that run several contractions in a loop would
// otherwise spill the hoisted B loads).
template <int NP, int LD, int RBK, int JT, int KN = NP>
__device__ __forceinline__ void tile_gemm_steps(d4 (&acc)[RBK][JT], const double *tile, const double *bcol, int ks,
                                                double (&b0)[JT], double (&b1)[JT], int li, int lk)
{
    constexpr int NK = KN / 4;
    double n0[JT], n1[JT];
#pragma unroll
    for (int jt = 0; jt < JT; ++jt) n0[jt] = n1[jt] = 0.0;
    if (ks + 2 < NK) {
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) n0[jt] = bcol[(long long)4 * (ks + 2) * NP + 16 * jt];
    }
#pragma unroll
    for (int rb = 0; rb < RBK; ++rb) {
        const double av = tile[(16 * rb + li) * LD + 4 * ks + lk];
#pragma unroll
        for (int jt = 0; jt < JT; ++jt)
            acc[rb][jt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b0[jt], acc[rb][jt], 0, 0, 0);
    }
    if (ks + 3 < NK) {
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) n1[jt] = bcol[(long long)4 * (ks + 3) * NP + 16 * jt];
    }
#pragma unroll
    for (int rb = 0; rb < RBK; ++rb) {
        const double av = tile[(16 * rb + li) * LD + 4 * (ks + 1) + lk];
#pragma unroll
        for (int jt = 0; jt < JT; ++jt)
            acc[rb][jt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b1[jt], acc[rb][jt], 0, 0, 0);
    }
#pragma unroll
    for (int jt = 0; jt < JT; ++jt) {
        b0[jt] = n0[jt];
        b1[jt] = n1[jt];
    }
}

// Small tiles (at most two accumulators per wave, 8 / JT k-steps ahead):
// B loaded further ahead than the two k-steps of tile_gemm, so that the L2
// latency of a k-step's operands is covered when a wave has few MFMAs per
// step (1-3% on the 16-row tiles at 256^2 x 129).
template <int NP, int LD, int RBK, int JT, int KN, int D>
__device__ __forceinline__ void tile_gemm_deep(d4 (&acc)[RBK][JT], const double *tile, const double *__restrict__ B,
                                               int c0, int r0, int li, int lk)
{
    constexpr int NK = KN / 4;
    const double *bcol = B + 16 * c0 + li + (long long)lk * NP;
    tile += 16 * r0 * LD;
    double b[D][JT];
#pragma unroll
    for (int d = 0; d < D; ++d)
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) b[d][jt] = d < NK ? bcol[(long long)4 * d * NP + 16 * jt] : 0.0;
#pragma unroll
    for (int ks = 0; ks < NK; ++ks) {
        double cur[JT];
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) {
            cur[jt] = b[ks % D][jt];
            if (ks + D < NK) b[ks % D][jt] = bcol[(long long)4 * (ks + D) * NP + 16 * jt];
        }
#pragma unroll
        for (int rb = 0; rb < RBK; ++rb) {
            const double av = tile[(16 * rb + li) * LD + 4 * ks + lk];
#pragma unroll
            for (int jt = 0; jt < JT; ++jt)
                acc[rb][jt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, cur[jt], acc[rb][jt], 0, 0, 0);
        }
    }
}

// KN: the contraction length (the k range of A's columns and B's rows that
// the caller's pointers start at); B's row stride stays NP.
template <int NP, int LD, int RBK, int JT, bool ROLLED = false, int KN = NP>
__device__ __forceinline__ void tile_gemm(d4 (&acc)[RBK][JT], const double *tile, const double *__restrict__ B,
                                          int c0, int r0, int li, int lk)
{
    if constexpr (!ROLLED && JT * RBK <= 2) {
        tile_gemm_deep<NP, LD, RBK, JT, KN, (8 / JT < KN / 4 ? 8 / JT : KN / 4)>(acc, tile, B, c0, r0, li, lk);
        return;
    }
    constexpr int NK = KN / 4;
    static_assert(NK % 2 == 0, "two k-steps per iteration");
    const double *bcol = B + 16 * c0 + li + (long long)lk * NP;
    tile += 16 * r0 * LD;
    double b0[JT], b1[JT];
#pragma unroll
    for (int jt = 0; jt < JT; ++jt) {
        b0[jt] = bcol[16 * jt];
        b1[jt] = bcol[(long long)4 * NP + 16 * jt];
    }
    if constexpr (ROLLED) {
#pragma nounroll
        for (int ks = 0; ks < NK; ks += 2) tile_gemm_steps<NP, LD, RBK, JT, KN>(acc, tile, bcol, ks, b0, b1, li, lk);
    } else {
        for (int ks = 0; ks < NK; ks += 2) tile_gemm_steps<NP, LD, RBK, JT, KN>(acc, tile, bcol, ks, b0, b1, li, lk);
    }
}

typedef float f4 __attribute__((ext_vector_type(4)));
typedef short s4 __attribute__((ext_vector_type(4)));

// bf16 bits of x, rounded to nearest even through FP32.
__device__ __forceinline__ short bf16_bits(double x)
{
    unsigned u = __float_as_uint((float)x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

// acc[rb][jt] += A B with v_mfma_f32_16x16x16bf16_1k: A = the FP64 LDS tile
// rounded to bf16, B in bf16 from L2 laid out as Bq[k / 4][j][k % 4] (each
// lane's four k values in 8 contiguous bytes), k < KN.  Operand layout (a
// wave, 16 x 16 x 16): A[i][k] is element k % 4 of lane i + 16 (k / 4),
// B[k][j] element k % 4 of lane j + 16 (k / 4); the result D[i][j] is
// element i % 4 of lane j + 16 (i / 4), while the FP64 accumulators hold
// D[i][j] in element i / 4 of lane j + 16 (i % 4): A's rows are fed
// permuted, lane row i taking tile row 4 (i % 4) + i / 4, so that element r
// of lane j + 16 lk is row 4 r + lk, the FP64 layout.
template <int NP, int LD, int RBK, int JT, int KN>
__device__ __forceinline__ void tile_gemm_bf16_rows(f4 (&acc)[RBK][JT], const double *tile, const s4 *__restrict__ Bq,
                                                    int c0, int r0, int li, int lk)
{
    const s4 *bq = Bq + (long long)lk * NP + 16 * c0 + li;
    const double *arow = tile + (16 * r0 + 4 * (li % 4) + li / 4) * LD + 4 * lk;
#pragma unroll
    for (int ks = 0; ks < KN / 16; ++ks) {
        s4 bv[JT];
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) bv[jt] = bq[(long long)4 * ks * NP + 16 * jt];
#pragma unroll
        for (int rb = 0; rb < RBK; ++rb) {
            const double *ap = arow + 16 * rb * LD + 16 * ks;
            const s4 av = {bf16_bits(ap[0]), bf16_bits(ap[1]), bf16_bits(ap[2]), bf16_bits(ap[3])};
#pragma unroll
            for (int jt = 0; jt < JT; ++jt)
                acc[rb][jt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv[jt], acc[rb][jt], 0, 0, 0);
        }
    }
}

// LDS of one diag tile.
template <int NJ, int RT, bool GENERAL, bool ADJ>
struct DiagShared {
    static constexpr int LD = 16 * NJ + 2;  // row stride (doubles)
    double tile[RT * LD];
    double row_sig[RT], row_zl[RT], row_zr[RT], row_d0[RT], row_dN[RT];
    long long row_col[RT];  // global offset of the row's column start (element start)
    double row_u0[ADJ ? RT : 1], row_uN[ADJ ? RT : 1];
    double part0[4][(GENERAL || ADJ) ? RT : 1], partN[4][(GENERAL || ADJ) ? RT : 1];
};

// Tile `bx` (RT rows) of the diagonalised mode solve; 256 threads.  Starts
// with a barrier so that back-to-back tiles can reuse the LDS.  EO: the
// even/odd eigenbasis (a.eo; one element with Dirichlet faces), whose
// contractions run over half the k range.
template <int NJ, int RT, bool GENERAL, bool ADJ, bool EO = !GENERAL && !ADJ>
__device__ __forceinline__ void diag_tile(const DiagArgs &a, int bx, DiagShared<NJ, RT, GENERAL, ADJ> &sh)
{
    static_assert(!(EO && GENERAL), "the even/odd basis serves one element with Dirichlet faces");
    // EO adjoint: q = y / (mu - sigma) is corrected before the end-value
    // dots use it, so acc holds q (not y) from there on
    constexpr bool QIN = EO && ADJ;
    constexpr int NP = 16 * NJ;      // padded interior size
    constexpr int LD = NP + 2;       // LDS row stride (doubles)
    constexpr int H = NP / 2;        // EO: the odd coordinates' offset
    constexpr int KN = EO ? H : NP;  // contraction length
    // The four waves split the output's NJ column tiles WC ways and its RBK
    // row blocks WR ways; each wave computes JT column tiles of RBW row blocks.
    constexpr int WC = NJ % 4 == 0 ? 4 : 2, WR = 4 / WC;
    constexpr int JT = NJ / WC;
    constexpr int RBK = RT / 16;
    constexpr int RBW = RBK / WR;
    static_assert(NJ % WC == 0 && RBK % WR == 0, "the waves split the tile evenly");
    double *tile = sh.tile;
    auto &row_sig = sh.row_sig;
    auto &row_zl = sh.row_zl;
    auto &row_zr = sh.row_zr;
    auto &row_d0 = sh.row_d0;
    auto &row_dN = sh.row_dN;
    auto &row_col = sh.row_col;
    auto &row_u0 = sh.row_u0;
    auto &row_uN = sh.row_uN;
    auto &part0 = sh.part0;
    auto &partN = sh.partN;
    __syncthreads();

    const int tid = threadIdx.x, lane = tid % WAVE, w = uniform(tid / WAVE);
    const int c0 = (w % WC) * JT, r0 = (w / WC) * RBW;  // the wave's first column tile, row block
    const int n = a.n, N = a.N, E = a.ne, K = a.K, M = a.M;
    const int ho = n / 2, he = n - ho;                // EO: even and odd coordinates
    const int rows_per_fc = 2 * E;                    // re/im x elements
    const int fc_per_tile = RT / rows_per_fc;
    const int fc0 = bx * fc_per_tile;
    const int nfc = a.nx * M;
    // EO: a wave's column tiles lie in one half, so its contractions run
    // over that half's k range only (the eigenbasis is block diagonal)
    const int kb = EO && 16 * c0 >= H ? H : 0;

    // ---- row bookkeeping: row = (fc_local*2 + plane)*E + e ----
    for (int row = tid; row < RT; row += 256) {
        const int fl = row / rows_per_fc, rem = row % rows_per_fc;
        const int plane = rem / E, e = rem % E;
        const int fc = fc0 + fl;
        double sig = 0.0;
        long long col = -1;
        if (fl < fc_per_tile && fc < nfc) {
            const int i = fc / M, m = fc % M;
            const int iw = i <= a.nx / 2 ? i : a.nx - i;
            sig = a.shift[iw * M + m];
            col = (long long)i * a.xs + (long long)m * K + (long long)e * (N - 1) + (plane ? a.plane_off : 0);
        }
        row_sig[row] = sig;
        row_col[row] = col;
        if constexpr (ADJ) {
            row_u0[row] = col >= 0 ? a.spec[col] : 0.0;
            row_uN[row] = col >= 0 ? a.spec[col + N - 1] : 0.0;
        }
    }
    __syncthreads();
    // ---- stage F^T: tile[row][k] = f interior of the row's element ----
    // EO: the folded interior, even coordinate i (k = i < H) f_i + f_{n-1-i}
    // (f_i at the centre of an odd n), odd coordinate i (k = H + i)
    // f_i - f_{n-1-i}.  Unrolled in chunks, all loads issued before any LDS
    // store (a loop with one load per trip waits out the memory latency on
    // every trip).
    {
        constexpr int NIT = RT * NP / 256;  // tile entries per thread
        constexpr int CH = EO ? (NIT % 8 == 0 ? 8 : NIT)
                              : (NIT % 16 == 0 ? 16 : NIT % 12 == 0 ? 12 : NIT % 8 == 0 ? 8 : NIT);
        static_assert(NIT % CH == 0 && RT * NP % 256 == 0, "whole chunks");
#pragma unroll
        for (int c0 = 0; c0 < NIT; c0 += CH) {
            double v[CH];
#pragma unroll
            for (int it = 0; it < CH; ++it) {
                const int idx = tid + 256 * (c0 + it);
                const int row = idx / NP, k = idx % NP;
                const long long col = row_col[row];
                if constexpr (EO) {
                    const bool even = k < H;
                    const int i = even ? k : k - H;
                    double x0 = 0.0, x1 = 0.0;
                    if (col >= 0 && i < (even ? he : ho)) {
                        x0 = a.spec[col + 1 + i];
                        if (i < ho) x1 = a.spec[col + n - i];
                    }
                    v[it] = even ? x0 + x1 : x0 - x1;
                } else {
                    v[it] = (col >= 0 && k < n) ? a.spec[col + 1 + k] : 0.0;
                }
            }
#pragma unroll
            for (int it = 0; it < CH; ++it) {
                const int idx = tid + 256 * (c0 + it);
                tile[(idx / NP) * LD + idx % NP] = v[it];
            }
        }
    }
    __syncthreads();

    // ---- Y^T = F^T Vinv^T ----
    d4 acc[RBW][JT];
#pragma unroll
    for (int rb = 0; rb < RBW; ++rb)
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) acc[rb][jt] = d4{0.0, 0.0, 0.0, 0.0};
    const int li = lane % 16, lk = lane / 16;
    tile_gemm<NP, LD, RBW, JT, false, KN>(acc, tile + kb, a.VinvT + (long long)kb * NP, c0, r0, li, lk);
    __syncthreads();  // every wave has read the F tile
    if constexpr (QIN) {
        // transposed first-order correction, q -= (M^T q) / (mu - sigma)
        // (a.eoM holds M^T's operand): q into the tile, the bf16
        // contraction over the other half, the update in registers
        double iq[RBW][JT][4];
#pragma unroll
        for (int rb = 0; rb < RBW; ++rb)
#pragma unroll
            for (int jt = 0; jt < JT; ++jt) {
                const int j = 16 * (c0 + jt) + li;
                const bool jv = j < H ? j < he : j - H < ho;
                const double mu = jv ? a.mu[j] : 0.0;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * (r0 + rb) + 4 * r + lk;
                    iq[rb][jt][r] = jv ? 1.0 / (mu - row_sig[row]) : 0.0;
                    acc[rb][jt][r] *= iq[rb][jt][r];
                    tile[row * LD + j] = acc[rb][jt][r];
                }
            }
        __syncthreads();
        f4 cacc[RBW][JT];
#pragma unroll
        for (int rb = 0; rb < RBW; ++rb)
#pragma unroll
            for (int jt = 0; jt < JT; ++jt) cacc[rb][jt] = f4{0.0f, 0.0f, 0.0f, 0.0f};
        const int ko = H - kb;
        tile_gemm_bf16_rows<NP, LD, RBW, JT, H>(cacc, tile + ko, reinterpret_cast<const s4 *>(a.eoM) + ko / 4 * NP,
                                                c0, r0, li, lk);
#pragma unroll
        for (int rb = 0; rb < RBW; ++rb)
#pragma unroll
            for (int jt = 0; jt < JT; ++jt)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    acc[rb][jt][r] = fma(-(double)cacc[rb][jt][r], iq[rb][jt][r], acc[rb][jt][r]);
        // the tile is next written after the barrier that follows the dots
    }

    // ---- shift, interface sweep, assemble Z^T in the tile ----
    // q = y / (mu - sigma) stays in registers; GENERAL needs r0|rN . q per
    // row (end-point derivatives of the particular solution), ADJ needs
    // w0|wN . q (the adjoint load on zl, zr).  Each wave reduces its columns
    // over the 16 lanes of a row, then the four waves' partials are summed.
    // One row block at a time, so that only its four partial pairs are live.
    double inv[RBW][JT][4];
    const double *dv0 = ADJ ? a.w0 : a.r0, *dvN = ADJ ? a.wN : a.rN;
#pragma unroll
    for (int rb = 0; rb < RBW; ++rb) {
        double p0[4] = {0.0, 0.0, 0.0, 0.0}, pN[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) {
            const int j = 16 * (c0 + jt) + li;
            const bool jv = EO ? (j < H ? j < he : j - H < ho) : j < n;
            const double mu = jv ? a.mu[j] : 0.0;
            const double v0 = (GENERAL || ADJ) && jv ? dv0[j] : 0.0;
            const double vN = (GENERAL || ADJ) && jv ? dvN[j] : 0.0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * (r0 + rb) + 4 * r + lk;
                const double d = mu - row_sig[row];
                const double iv = jv ? 1.0 / d : 0.0;
                inv[rb][jt][r] = iv;
                if (GENERAL || ADJ) {
                    const double q = QIN ? acc[rb][jt][r] : acc[rb][jt][r] * iv;
                    p0[r] = fma(v0, q, p0[r]);
                    pN[r] = fma(vN, q, pN[r]);
                }
            }
        }
        if constexpr (GENERAL || ADJ) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
#pragma unroll
                for (int o = 1; o < 16; o <<= 1) {
                    p0[r] += __shfl_xor(p0[r], o);
                    pN[r] += __shfl_xor(pN[r], o);
                }
                if (li == 0) {
                    const int row = 16 * (r0 + rb) + 4 * r + lk;
                    part0[w][row] = p0[r];
                    partN[w][row] = pN[r];
                }
            }
        }
    }
    if constexpr (GENERAL || ADJ) {
        __syncthreads();
        for (int row = tid; row < RT; row += 256) {
            double s0 = 0.0, sN = 0.0;
            const int wbase = (row / 16 / RBW) * WC;  // the waves that cover the row
#pragma unroll
            for (int c = 0; c < WC; ++c) {
                s0 += part0[wbase + c][row];
                sN += partN[wbase + c][row];
            }
            row_d0[row] = s0;
            row_dN[row] = sN;
        }
    }
    if constexpr (ADJ) {
        __syncthreads();
        // transposed interface system per real column: zbar -> gbar and the
        // adjoints P0bar, PNbar of each element's end-point derivatives
        // (Robin faces), or the transposed cyclic closure (periodic z, one
        // element).  Rows that belong to no column stay zero.
        for (int row = tid; row < RT; row += 256) {
            const int sc = row / E;
            if (sc >= 2 * fc_per_tile || fc0 + (sc >> 1) >= nfc) {
                row_zl[row] = 0.0;
                row_zr[row] = 0.0;
            }
        }
        for (int sc = tid; sc < 2 * fc_per_tile; sc += 256) {
            const int fl = sc >> 1, plane = sc & 1;
            const int fc = fc0 + fl;
            if (fc >= nfc) continue;
            const int i = fc / M, m = fc % M;
            const int iw = i <= a.nx / 2 ? i : a.nx - i;
            const int mode = iw * M + m;
            const double L0 = a.hom[4 * mode + 0], LN = a.hom[4 * mode + 1];
            const double R0 = a.hom[4 * mode + 2], RN = a.hom[4 * mode + 3];
            const int r0w = sc * E;  // the column's element rows
            if (a.periodic && E > 1) {
                periodic_sweep_adjoint(E, L0, LN, R0, RN, row_u0 + r0w, row_uN + r0w, row_d0 + r0w, row_dN + r0w,
                                       row_zl + r0w, row_zr + r0w);
                continue;
            }
            if (a.periodic) {
                // u_0 = u_{N-1} = z = (dN - d0) / bp (periodic_sweep, one
                // element): zbar collects both end values and the load
                // through ul + ur; d0bar = -zbar / bp, dNbar = zbar / bp
                const double bp = (L0 + R0) - (LN + RN);
                const double zb = (row_u0[r0w] + row_uN[r0w]) + (row_d0[r0w] + row_dN[r0w]);
                row_zl[r0w] = -zb / bp;
                row_zr[r0w] = zb / bp;
                continue;
            }
            double *gre = plane ? a.gbar_im : a.gbar_re;
            if (E == 1) {
                const double b = -a.alpha_l * L0 + a.beta_l, c = -a.alpha_l * R0;
                const double lb = a.alpha_r * LN;
                const double br = a.alpha_r * RN + a.beta_r - (c / b) * lb;
                const double zlb = row_u0[r0w] + row_d0[r0w];
                const double zrb = row_uN[r0w] + row_dN[r0w] - (c / b) * zlb;
                const double rhsrb = zrb / br;
                const double rhsb = zlb / b - (lb / b) * rhsrb;
                row_zl[r0w] = a.alpha_l * rhsb;  // reused: P0bar, PNbar per row
                row_zr[r0w] = -a.alpha_r * rhsrb;
                gre[fc] = rhsb;
                gre[(long long)nfc + fc] = rhsrb;
                continue;
            }
            double glb, grb;
            robin_sweep_adjoint(E, L0, LN, R0, RN, a.alpha_l, a.beta_l, a.alpha_r, a.beta_r, row_u0 + r0w,
                                row_uN + r0w, row_d0 + r0w, row_dN + r0w, row_zl + r0w, row_zr + r0w, &glb, &grb);
            gre[fc] = glb;
            gre[(long long)nfc + fc] = grb;
        }
    } else if (GENERAL) {
        __syncthreads();
        // sweep over the elements of each real spectral column
        for (int sc = tid; sc < 2 * fc_per_tile; sc += 256) {
            const int fl = sc >> 1, plane = sc & 1;
            const int fc = fc0 + fl;
            if (fc >= nfc) continue;
            const int i = fc / M, m = fc % M;
            const int iw = i <= a.nx / 2 ? i : a.nx - i;
            const int mode = iw * M + m;
            const double d0ul = a.hom[4 * mode + 0], dNul = a.hom[4 * mode + 1];
            const double d0ur = a.hom[4 * mode + 2], dNur = a.hom[4 * mode + 3];
            const int r0w = (fl * 2 + plane) * E;
            if (a.periodic) {
                periodic_sweep(E, d0ul, dNul, d0ur, dNur, row_d0 + r0w, row_dN + r0w, row_zl + r0w,
                               row_zr + r0w, 1, a.closure ? a.closure + (plane ? nfc : 0) + fc : nullptr);
                continue;
            }
            const double *bsrc = plane ? a.bc_im : a.bc_re;
            robin_sweep(E, d0ul, dNul, d0ur, dNur, bsrc[fc], bsrc[(long long)nfc + fc], a.alpha_l, a.beta_l,
                        a.alpha_r, a.beta_r, row_d0 + r0w, row_dN + r0w, row_zl + r0w, row_zr + r0w);
        }
    } else if constexpr (!ADJ) {
        // one element, Dirichlet data on both faces: zl = g_l, zr = g_r
        for (int row = tid; row < RT; row += 256) {
            const int fl = row >> 1, plane = row & 1;
            const int fc = fc0 + fl;
            double zl = 0.0, zr = 0.0;
            if (fl < fc_per_tile && fc < nfc) {
                const double *bsrc = plane ? a.bc_im : a.bc_re;
                zl = bsrc[fc] / a.beta_l;
                zr = bsrc[(long long)nfc + fc] / a.beta_r;
            }
            row_zl[row] = zl;
            row_zr[row] = zr;
        }
    }
    __syncthreads();
    const double *zv0 = ADJ ? a.r0 : a.w0, *zvN = ADJ ? a.rN : a.wN;
#pragma unroll
    for (int rb = 0; rb < RBW; ++rb)
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) {
            const int j = 16 * (c0 + jt) + li;
            const bool jv = EO ? (j < H ? j < he : j - H < ho) : j < n;
            const double w0 = jv ? zv0[j] : 0.0, wN = jv ? zvN[j] : 0.0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * (r0 + rb) + 4 * r + lk;
                const double h = fma(row_zl[row], w0, row_zr[row] * wN);
                tile[row * LD + j] =
                    QIN ? fma(h, inv[rb][jt][r], acc[rb][jt][r]) : (acc[rb][jt][r] + h) * inv[rb][jt][r];
            }
        }
    __syncthreads();
    if constexpr (EO && !ADJ) {
        // first-order correction from the symmetrised operator to the
        // reference's: z -= (M z) / (mu - sigma), M z from the other half's
        // coordinates (M maps even to odd and back), bf16 operands with FP32
        // sums (M z is about 1e-12 |z|)
        f4 cacc[RBW][JT];
#pragma unroll
        for (int rb = 0; rb < RBW; ++rb)
#pragma unroll
            for (int jt = 0; jt < JT; ++jt) cacc[rb][jt] = f4{0.0f, 0.0f, 0.0f, 0.0f};
        const int ko = H - kb;
        tile_gemm_bf16_rows<NP, LD, RBW, JT, H>(cacc, tile + ko, reinterpret_cast<const s4 *>(a.eoM) + ko / 4 * NP,
                                                c0, r0, li, lk);
        __syncthreads();  // every wave has read the Z tile
#pragma unroll
        for (int rb = 0; rb < RBW; ++rb)
#pragma unroll
            for (int jt = 0; jt < JT; ++jt) {
                const int j = 16 * (c0 + jt) + li;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int idx = (16 * (r0 + rb) + 4 * r + lk) * LD + j;
                    tile[idx] = fma(-(double)cacc[rb][jt][r], inv[rb][jt][r], tile[idx]);
                }
            }
        __syncthreads();
    }

    // ---- U^T = Z^T V^T, stored along z (EO: unfolded first) ----
#pragma unroll
    for (int rb = 0; rb < RBW; ++rb)
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) acc[rb][jt] = d4{0.0, 0.0, 0.0, 0.0};
    tile_gemm<NP, LD, RBW, JT, false, KN>(acc, tile + kb, a.VT + (long long)kb * NP, c0, r0, li, lk);
    if constexpr (EO) {
        __syncthreads();  // every wave has read the Z tile
#pragma unroll
        for (int rb = 0; rb < RBW; ++rb)
#pragma unroll
            for (int jt = 0; jt < JT; ++jt) {
                const int k = 16 * (c0 + jt) + li;
#pragma unroll
                for (int r = 0; r < 4; ++r) tile[(16 * (r0 + rb) + 4 * r + lk) * LD + k] = acc[rb][jt][r];
            }
        __syncthreads();
        // u_i = e_i + o_i, u_{n-1-i} = e_i - o_i (i < n/2), the centre of an odd n e_i
        for (int idx = tid; idx < RT * NP; idx += 256) {
            const int row = idx / NP, k = idx % NP;
            const long long col = row_col[row];
            if (col < 0 || k >= n) continue;
            const int i = k < n - 1 - k ? k : n - 1 - k;
            const double e = tile[row * LD + i];
            const double o = 2 * k + 1 == n ? 0.0 : tile[row * LD + H + i];
            a.spec[col + 1 + k] = k <= n - 1 - k ? e + o : e - o;
        }
    } else {
#pragma unroll
        for (int rb = 0; rb < RBW; ++rb)
#pragma unroll
            for (int jt = 0; jt < JT; ++jt) {
                const int k = 16 * (c0 + jt) + li;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * (r0 + rb) + 4 * r + lk;
                    const long long col = row_col[row];
                    if (col >= 0 && k < n) a.spec[col + 1 + k] = acc[rb][jt][r];
                }
            }
    }
    // element end values (the adjoint load has none: f is not used there)
    for (int row = tid; row < RT; row += 256) {
        const long long col = row_col[row];
        if (col >= 0) {
            a.spec[col] = ADJ ? 0.0 : row_zl[row];
            a.spec[col + N - 1] = ADJ ? 0.0 : row_zr[row];
        }
    }
}

}  // namespace
}  // namespace hz
