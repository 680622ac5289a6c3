// lu.hip -- the reference algorithm per (|kx|,|ky|) mode: one work-group per
// mode, the whole augmented system resident in registers.
//
// For mode (iw, jw) the work-group
//   1. assembles A = Dzz_int - sigma I (or the caller's operator) and the
//      right-hand sides: the two homogeneous ones (-Dzz[1..N-2][0],
//      -Dzz[1..N-2][N-1], NonPeriodic3D_MPI.cpp:211-219) and the interior
//      spectrum of every element for each real column of the mode
//      (:188-201, :248-325);
//   2. eliminates with the reference's pivot rule (LU_Decomposition.cpp:7-72):
//      the candidate at position i is |a_ik| / s_i where s_i is the max-norm
//      of the ORIGINAL row at position i (the scale is never exchanged with
//      the rows), the strictly-greater test makes the first maximum in
//      position order win, and |scaled pivot| <= 1e-12 is singular.
//      Carrying the right-hand sides through the elimination is LUSolve's
//      forward phase (:86-95);
//   3. back-substitutes the right-hand sides (:97-107);
//   4. forms the end-point derivatives Dz[0|N-1].col (GLLQuadrature.cpp:234-245),
//      runs the element-interface sweep with the Robin end rows
//      (NonPeriodic3D_MPI.cpp:327-580) and assembles u = p + zl ul + zr ur
//      (:582-587) straight into the spectrum.
//
// Exchanges are implicit: rows never move, every physical row carries its
// current position, and the pivot row is found wherever it lives.  The mode
// kernel (eliminate_modes) uses FMA and exact exchanges, so its results
// agree with the reference to round-off.  The standalone factor and solve
// kernels (the drop-in LUDecomposition / LUSolve) round every product and
// difference separately, in the reference's order, apply the reference's
// a+b-b exchange arithmetic, and reproduce it bit for bit.
//
// Layout: 64*TC threads.  Wave c (of TC) owns matrix columns c + TC*b
// (b < CB) and right-hand sides c + TC*t (t < NR); lane r owns physical rows
// r + 64*a (a < RB).  A column lives in one wave, so the pivot search is a
// wave reduction (DPP and readlane, no LDS).
#include "hz_internal.h"
#include "lu_slots.h"
#include "sweep.h"

namespace hz {
namespace {

constexpr double kEps = 1.000000000000E-12;  // include/General.h:15
constexpr int WAVE = 64;
// Right-hand sides: ul, ur, then per element one per real column of the
// mode (2 per complex column; 2 complex columns when 0 < |kx| < nx/2, which
// needs nx >= 4).  At most 4 x 64 (the NR = 64 instance); the LDS arrays
// indexed by right-hand side are sized with room to spare.
constexpr int kMaxRhs = 2 + 4 * 64;

__host__ __device__ __forceinline__ int rhs_per_element(int nx) { return nx >= 4 ? 4 : 2; }

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// a - l*p with the product rounded before the difference
// (LU_Decomposition.cpp:56-57 `A[l][m] -= A[l][k]*A[k][m]`).
__device__ __forceinline__ double sub_product_rounded(double a, double l, double p)
{
#pragma clang fp contract(off)
    return a - l * p;
}

// Publish buffers of the mode kernel's elimination: a wave is at most four
// positions behind the latest published one, so 8 buffers are never
// overwritten while a wave may still read them.
constexpr int kPub = 8;
// Bound on a wave's wait for a published position (s_sleep trips of 64
// cycles, ~2 ms): past it the mode reports HZ_ERR_HIP instead of hanging.
constexpr int kPubSpin = 1 << 16;

template <int RB, int CB, int NR, int TC>
struct Smem {
    static constexpr int ROWS = WAVE * RB;
    double scale[ROWS];        // s by position, LU_Decomposition.cpp:20-29
    double lbuf[kPub][ROWS];   // multipliers by physical row
    double part[TC][ROWS];
    alignas(16) double prow[TC][CB];   // per-wave staging of the pivot row
    alignas(16) double krow[TC][CB];   // ... and of the row it is exchanged with
    alignas(16) double prowy[TC][NR];
    alignas(16) double krowy[TC][NR];
    int info[kPub][3];         // {npivot[k], physical pivot row, physical row at position k}
    int flag[kPub];            // k + 1 once position k is published in buffer k % kPub
    int phys_of_pos[ROWS];
    double udiag[ROWS];        // U diagonal by position
    double det;
    int singular;  // 1: a pivot failed the reference's test; 2: a wait ran to its bound
};

// Elimination core.  A[a][b]: physical row lane + 64a, matrix column
// c + TC*b.  Y[a][t]: same row, right-hand side c + TC*t (nrhs of them).
// pos[a]: current position of the row.  With out != nullptr the factors are
// written in the reference's storage (multipliers at their positions, U rows
// in position order) and the pivots to piv_out.  EXACT: rank-1 updates
// without fused multiply-adds (the reference's rounding).
template <int RB, int CB, int NR, int TC, bool EXACT = false>
__device__ __forceinline__ void eliminate(double (&A)[RB][CB], double (&Y)[RB][NR], int (&pos)[RB], int n,
                                          int nrhs, Smem<RB, CB, NR, TC> &s, double *out, int *piv_out)
{
    const int lane = threadIdx.x % WAVE, c = uniform(threadIdx.x / WAVE);
#pragma unroll
    for (int a = 0; a < RB; ++a) {
        double m = 0.0;
#pragma unroll
        for (int b = 0; b < CB; ++b)
            if (c + TC * b < n) m = fmax(m, fabs(A[a][b]));
        s.part[c][lane + WAVE * a] = m;
        pos[a] = lane + WAVE * a;
    }
    if (threadIdx.x == 0) {
        s.det = 1.0;
        s.singular = 0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < WAVE * RB; i += WAVE * TC) {
        double m = s.part[0][i];
#pragma unroll
        for (int cc = 1; cc < TC; ++cc) m = fmax(m, s.part[cc][i]);
        s.scale[i] = m;
    }
    __syncthreads();

    for (int k = 0; k < n - 1; ++k) {
        const int buf = k & 1;
        const int ck = k % TC, bk = k / TC;  // uniform
        // ---------------- phase A: pivot search in the wave owning column k
        if (c == ck) {
            double colv[RB];
#pragma unroll
            for (int a = 0; a < RB; ++a) colv[a] = 0.0;
#define HZ_GETCOL(b)                                                            \
    if constexpr ((b) < CB) {                                                   \
        _Pragma("unroll") for (int a = 0; a < RB; ++a) colv[a] = A[a][(b)];     \
    }
            HZ_SLOT_AT(bk, HZ_GETCOL)
#undef HZ_GETCOL
            double best = -2.0;
            int bpos = 0x7fffffff, brow = 0, mine = -1;
#pragma unroll
            for (int a = 0; a < RB; ++a) {
                const int row = lane + WAVE * a;
                if (row < n && pos[a] >= k) {
                    if (pos[a] == k) mine = row;
                    double ratio = fabs(colv[a]) / s.scale[pos[a]];
                    if (!(ratio >= 0.0)) ratio = -1.0;  // NaN scale: lowest priority
                    if (ratio > best || (ratio == best && pos[a] < bpos)) {
                        best = ratio;
                        bpos = pos[a];
                        brow = row;
                    }
                }
            }
#pragma unroll
            for (int m = 1; m < WAVE; m <<= 1) {
                const double ob = __shfl_xor(best, m);
                const int op = __shfl_xor(bpos, m);
                const int orw = __shfl_xor(brow, m);
                if (ob > best || (ob == best && op < bpos)) {
                    best = ob;
                    bpos = op;
                    brow = orw;
                }
            }
            const int ppos = uniform(bpos), ph = uniform(brow);
            const int phk = uniform(__builtin_amdgcn_readlane(mine, __builtin_ctzll(__ballot(mine >= 0))));
            double own = colv[0], ownk = colv[0];
#pragma unroll
            for (int a = 1; a < RB; ++a) {
                if (a == ph / WAVE) own = colv[a];
                if (a == phk / WAVE) ownk = colv[a];
            }
            const double bkv = __shfl(own, ph % WAVE);    // pivot row, column k
            const double akv = __shfl(ownk, phk % WAVE);  // row at position k, column k
            double piv = bkv, akswap = akv;
            if (ppos != k) {  // Swap(A[k][k], A[p][k])
                const double sum = akv + bkv;
                akswap = sum - bkv;
                piv = sum - akswap;
            }
#pragma unroll
            for (int a = 0; a < RB; ++a) {
                const int row = lane + WAVE * a;
                int np = pos[a];
                if (row == ph) np = k;
                else if (np == k) np = ppos;
                const bool active = row < n && np > k;
                const double v = (row == phk && ppos != k) ? akswap : colv[a];
                const double l = active ? v / piv : 0.0;
                s.lbuf[buf][row] = l;
                if (out) {
                    if (active) out[(long long)np * n + k] = l;
                    if (row == ph) out[(long long)k * n + k] = piv;
                }
            }
            if (lane == 0) {
                s.udiag[k] = piv;
                s.info[buf][0] = ppos;
                s.info[buf][1] = ph;
                s.info[buf][2] = phk;
                s.phys_of_pos[k] = ph;
                if (piv_out) piv_out[k] = ppos;
                double det = s.det;
                if (ppos != k) det = -det;
                s.det = det * piv;
                if (fabs(best) <= kEps || best < 0.0) s.singular = 1;
            }
        }
        __syncthreads();
        // ---------------- phase C: exchange positions, rank-1 update
        const int ppos = uniform(s.info[buf][0]), ph = uniform(s.info[buf][1]);
        const int phk = uniform(s.info[buf][2]);
        const bool swapped = ppos != k;
        double l[RB];
#pragma unroll
        for (int a = 0; a < RB; ++a) {
            const int row = lane + WAVE * a;
            if (row == ph) pos[a] = k;
            else if (pos[a] == k) pos[a] = ppos;
            l[a] = s.lbuf[buf][row];
        }
        const int b0 = (c > ck) ? bk : bk + 1;  // first slot of this wave right of k
        const int phslot = ph / WAVE, phlane = ph % WAVE;
        const int kslot = phk / WAVE, klane = phk % WAVE;
        // stage the pivot row (and, on an exchange, the row it replaces)
#pragma unroll
        for (int aa = 0; aa < RB; ++aa) {
            if (aa == phslot && lane == phlane) {
#define HZ_STAGE(b)                                                             \
    if constexpr ((b) < CB) s.prow[c][(b)] = A[aa][(b)];
                HZ_SLOTS_FROM(b0, HZ_STAGE)
#undef HZ_STAGE
#pragma unroll
                for (int t = 0; t < NR; ++t) s.prowy[c][t] = Y[aa][t];
            }
            if (swapped && aa == kslot && lane == klane) {
#define HZ_STAGEK(b)                                                            \
    if constexpr ((b) < CB) s.krow[c][(b)] = A[aa][(b)];
                HZ_SLOTS_FROM(b0, HZ_STAGEK)
#undef HZ_STAGEK
#pragma unroll
                for (int t = 0; t < NR; ++t) s.krowy[c][t] = Y[aa][t];
            }
        }
        wave_sync();
        if (swapped) {
            // The reference exchanges rows k and p with a+b-b arithmetic; the
            // two owner lanes apply the same roundings to their rows.
            if (lane == phlane || lane == klane) {
#define HZ_XCHG(b)                                                              \
    if constexpr ((b) < CB) {                                                   \
        const double bv = s.prow[c][(b)], av = s.krow[c][(b)];                  \
        const double sum = av + bv;                                             \
        const double kv = sum - bv; /* LUDecomposition: Swap(A[k][j], A[p][j]) */ \
        const double pv = sum - kv;                                             \
        _Pragma("unroll") for (int a = 0; a < RB; ++a) {                        \
            if (a == kslot && lane == klane) A[a][(b)] = kv;                    \
            if (a == phslot && lane == phlane) A[a][(b)] = pv;                  \
        }                                                                       \
        if (lane == phlane) s.prow[c][(b)] = pv;                                \
    }
                HZ_SLOTS_FROM(b0, HZ_XCHG)
#undef HZ_XCHG
#pragma unroll
                for (int t = 0; t < NR; ++t) {
                    const double bv = s.prowy[c][t], av = s.krowy[c][t];
                    const double sum = av + bv;
                    const double pv = sum - av;  // LUSolve: Swap(b[p], b[k])
                    const double kv = sum - pv;
#pragma unroll
                    for (int a = 0; a < RB; ++a) {
                        if (a == kslot && lane == klane) Y[a][t] = kv;
                        if (a == phslot && lane == phlane) Y[a][t] = pv;
                    }
                    if (lane == phlane) s.prowy[c][t] = pv;
                }
            }
            wave_sync();
        }
        if (out && lane == 0) {
            for (int b = b0; b < CB; ++b) {
                const int j = c + TC * b;
                if (j < n) out[(long long)k * n + j] = s.prow[c][b];
            }
        }
#define HZ_UPDATE(b)                                                            \
    if constexpr ((b) < CB) {                                                   \
        const double pv = s.prow[c][(b)];                                       \
        _Pragma("unroll") for (int a = 0; a < RB; ++a)                          \
            A[a][(b)] = EXACT ? sub_product_rounded(A[a][(b)], l[a], pv) : fma(-l[a], pv, A[a][(b)]); \
    }
        HZ_SLOTS_FROM(b0, HZ_UPDATE)
#undef HZ_UPDATE
        if (nrhs > 0) {
#pragma unroll
            for (int t = 0; t < NR; ++t) {
                const double pv = s.prowy[c][t];
#pragma unroll
                for (int a = 0; a < RB; ++a) Y[a][t] = fma(-l[a], pv, Y[a][t]);
            }
        }
        wave_sync();
    }
    // last position: its row is the last pivot (not tested by the reference)
    __syncthreads();
    {
        const int k = n - 1, ck = k % TC, bk = k / TC;
        if (c == ck) {
#pragma unroll
            for (int a = 0; a < RB; ++a) {
                const int row = lane + WAVE * a;
                if (row < n && pos[a] == k) {
                    double v = 0.0;
#define HZ_GETV(b)                                                              \
    if constexpr ((b) < CB) v = A[a][(b)];
                    HZ_SLOT_AT(bk, HZ_GETV)
#undef HZ_GETV
                    s.udiag[k] = v;
                    s.det = s.det * v;
                    s.phys_of_pos[k] = row;
                    if (out) out[(long long)k * n + k] = v;
                }
            }
        }
    }
    __syncthreads();
}

// ---- the mode kernel's elimination ------------------------------------------
// Lane exchanges without LDS: DPP moves within rows of 16 lanes, readlane
// across them.
template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)dpp_i32<CTRL>((int)(unsigned)b);
    const unsigned hi = (unsigned)dpp_i32<CTRL>((int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// v of lane l (l uniform)
__device__ __forceinline__ double lane_f64(double v, int l)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// DPP controls: quad_perm [1,0,3,2] and [2,3,0,1] (lane xor 1, xor 2), then
// the half-row and row mirrors, which leave every lane with the reduction
// over its row of 16 lanes; the four rows are then combined by readlane.
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141, kDppMirror = 0x140;

// Maximum of v over the wave (uniform result); v is not NaN.
__device__ __forceinline__ double wave_max_f64(double v)
{
    v = fmax(v, dpp_f64<kDppXor1>(v));
    v = fmax(v, dpp_f64<kDppXor2>(v));
    v = fmax(v, dpp_f64<kDppHalfMirror>(v));
    v = fmax(v, dpp_f64<kDppMirror>(v));
    return fmax(fmax(lane_f64(v, 0), lane_f64(v, 16)), fmax(lane_f64(v, 32), lane_f64(v, 48)));
}

// Minimum of v over the wave (uniform result).
__device__ __forceinline__ int wave_min_i32(int v)
{
    v = min(v, dpp_i32<kDppXor1>(v));
    v = min(v, dpp_i32<kDppXor2>(v));
    v = min(v, dpp_i32<kDppHalfMirror>(v));
    v = min(v, dpp_i32<kDppMirror>(v));
    return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

// Pivot of position k (LU_Decomposition.cpp:32-49), run by the wave that
// owns matrix column k with colv = that column brought up to date: the
// candidate at position i >= k is |a_ik| / s_i (s: the original max-norm
// by position), the first maximum in position order wins, and a scaled
// pivot <= 1e-12 is singular.  Publishes in buffer `buf` the exchanged
// position and the pivot's physical row, the multipliers a_ik / pivot of
// every physical row still to be eliminated (0 for the others), the pivot
// value and the singular flag, then raises the buffer's flag to k + 1.
template <int RB, int CB, int NR, int TC>
__device__ __forceinline__ void pivot_publish(const double (&colv)[RB], const int (&pos)[RB], const double (&scl)[RB],
                                              int k, int n, Smem<RB, CB, NR, TC> &s, int buf)
{
    const int lane = threadIdx.x % WAVE;
    double best = -2.0, bval = 0.0;
    int bpos = 0x7fffffff, brow = 0;
#pragma unroll
    for (int a = 0; a < RB; ++a) {
        const int row = lane + WAVE * a;
        if (row < n && pos[a] >= k) {
            double ratio = fabs(colv[a]) / scl[a];
            if (!(ratio >= 0.0)) ratio = -1.0;  // NaN: lowest priority
            if (ratio > best || (ratio == best && pos[a] < bpos)) {
                best = ratio;
                bpos = pos[a];
                brow = row;
                bval = colv[a];
            }
        }
    }
    const double g = wave_max_f64(best);
    const bool top = best == g;
    unsigned long long m = __ballot(top);
    if (__popcll(m) > 1) {  // several rows share the maximum: the first position wins
        const int mp = wave_min_i32(top ? bpos : 0x7fffffff);
        m = __ballot(top && bpos == mp);
    }
    const int w = __builtin_ctzll(m);
    const int ppos = uniform(__builtin_amdgcn_readlane(bpos, w));
    const int ph = uniform(__builtin_amdgcn_readlane(brow, w));
    const double piv = lane_f64(bval, w);
#pragma unroll
    for (int a = 0; a < RB; ++a) {
        const int row = lane + WAVE * a;
        const int np = row == ph ? k : (pos[a] == k ? ppos : pos[a]);
        s.lbuf[buf][row] = (row < n && np > k) ? colv[a] / piv : 0.0;
    }
    if (lane == 0) {
        s.info[buf][0] = ppos;
        s.info[buf][1] = ph;
        s.udiag[k] = piv;
        s.phys_of_pos[k] = ph;
        if (g <= kEps) atomicMax(&s.singular, 1);  // includes a NaN column (g = -1)
        // the wave's multiplier and info stores complete before the flag
        __hip_atomic_store(&s.flag[buf], k + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

// Wait until position k is published (its buffer's flag reads k + 1); a
// wait past kPubSpin trips marks the mode as failed and goes on.
template <int RB, int CB, int NR, int TC>
__device__ __forceinline__ void await_position(Smem<RB, CB, NR, TC> &s, int k)
{
    const int b = k % kPub;
    int spins = 0;
    while (uniform(__hip_atomic_load(&s.flag[b], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP)) != k + 1) {
        if (++spins > kPubSpin) {
            if (threadIdx.x % WAVE == 0) atomicMax(&s.singular, 2);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// Elimination of the mode kernel: the reference's pivot sequence applied to
// the matrix and the right-hand sides (LUDecomposition and the forward phase
// of LUSolve).  Exchanges are implicit -- rows keep their registers and only
// their positions change -- and exact: the reference's a+b-b exchange
// arithmetic is not reproduced here (a rounding-level difference; the
// standalone factor kernel keeps it).
//
// Position k's pivot is published (multipliers, exchange, flag) by the wave
// that owns matrix column k.  Every wave runs the positions in order at its
// own pace, with no work-group barrier: it waits for position k's flag,
// applies the exchange to its row positions, and -- if it owns column k+1 --
// first brings that column up to date (the pivot row's entry is read from
// its lane) and publishes position k+1; then it stages the pivot row's
// entries of its own columns in LDS (a wave-private broadcast) and updates
// its columns right of k and the right-hand sides.  The chain from one
// publication to the next is one column update and one pivot search; the
// rank-1 updates of the other waves overlap it.  A wave can be at most four
// positions behind the latest publication (each publication needs its owner
// to have reached the previous position), so kPub = 8 buffers suffice.
// Leaves pos, the U rows in A, Y = L^-1 P b, s.udiag, s.phys_of_pos and
// s.singular.
template <int RB, int CB, int NR, int TC>
__device__ __forceinline__ void eliminate_modes(double (&A)[RB][CB], double (&Y)[RB][NR], int (&pos)[RB], int n,
                                                Smem<RB, CB, NR, TC> &s)
{
    const int lane = threadIdx.x % WAVE, c = uniform(threadIdx.x / WAVE);
    // s by position (positions start as the physical rows), LU_Decomposition.cpp:20-29
#pragma unroll
    for (int a = 0; a < RB; ++a) {
        double m = 0.0;
#pragma unroll
        for (int b = 0; b < CB; ++b)
            if (c + TC * b < n) m = fmax(m, fabs(A[a][b]));
        s.part[c][lane + WAVE * a] = m;
        pos[a] = lane + WAVE * a;
    }
    if (threadIdx.x == 0) s.singular = 0;
    if (threadIdx.x < kPub) s.flag[threadIdx.x] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < WAVE * RB; i += WAVE * TC) {
        double m = s.part[0][i];
#pragma unroll
        for (int cc = 1; cc < TC; ++cc) m = fmax(m, s.part[cc][i]);
        s.scale[i] = m;
    }
    __syncthreads();
    double scl[RB];  // s of each row's current position
#pragma unroll
    for (int a = 0; a < RB; ++a) scl[a] = s.scale[lane + WAVE * a];
    if (n > 1 && c == 0) {
        double colv[RB];
#pragma unroll
        for (int a = 0; a < RB; ++a) colv[a] = A[a][0];
        pivot_publish<RB, CB, NR, TC>(colv, pos, scl, 0, n, s, 0);
    }

    for (int k = 0; k < n - 1; ++k) {
        const int buf = k % kPub;
        const int ck = k % TC, bk = k / TC;
        await_position(s, k);
        const int ppos = uniform(s.info[buf][0]), ph = uniform(s.info[buf][1]);
        const int pa = ph / WAVE, pl = ph % WAVE;
        double l[RB];
#pragma unroll
        for (int a = 0; a < RB; ++a) l[a] = s.lbuf[buf][lane + WAVE * a];
        {
            // the displaced row takes position ppos and its scale, which is
            // the pivot row's current scale (scl follows the positions)
            double ps = scl[0];
#pragma unroll
            for (int a = 1; a < RB; ++a)
                if (a == pa) ps = scl[a];
            ps = lane_f64(ps, pl);
#pragma unroll
            for (int a = 0; a < RB; ++a) {
                const int row = lane + WAVE * a;
                if (row == ph) {
                    pos[a] = k;
                } else if (pos[a] == k) {
                    pos[a] = ppos;
                    scl[a] = ps;
                }
            }
        }
        const int k1 = k + 1;
        if (k1 < n - 1 && c == k1 % TC) {
            // look-ahead: column k+1 brought up to date and position k+1's
            // pivot published before this wave's other work
            const int b1 = k1 / TC;
            double colv[RB];
#pragma unroll
            for (int a = 0; a < RB; ++a) colv[a] = 0.0;
#define HZ_COLNEXT(b)                                                           \
    if constexpr ((b) < CB) {                                                   \
        _Pragma("unroll") for (int a = 0; a < RB; ++a) colv[a] = A[a][(b)];     \
    }
            HZ_SLOT_AT(b1, HZ_COLNEXT)
#undef HZ_COLNEXT
            double pown = colv[0];
#pragma unroll
            for (int a = 1; a < RB; ++a)
                if (a == pa) pown = colv[a];
            const double pv = lane_f64(pown, pl);
#pragma unroll
            for (int a = 0; a < RB; ++a) colv[a] = fma(-l[a], pv, colv[a]);
            pivot_publish<RB, CB, NR, TC>(colv, pos, scl, k1, n, s, k1 % kPub);
        }
        // the pivot row's entries of this wave's columns and right-hand sides
        if (lane == pl) {
#pragma unroll
            for (int a = 0; a < RB; ++a)
                if (a == pa) {
#pragma unroll
                    for (int b = 0; b < CB; ++b) s.prow[c][b] = A[a][b];
#pragma unroll
                    for (int t = 0; t < NR; ++t) s.prowy[c][t] = Y[a][t];
                }
        }
        wave_sync();
        // Rank-1 update in chunks of UC slots, each chunk's pivot-row values
        // loaded together before its FMAs; chunks wholly left of column k are
        // skipped, and the slots left of k inside the first chunk are updated
        // too (their rows either have zero multipliers -- rows already
        // pivoted -- or hold eliminated entries that are never read again).
        const int b0 = (c > ck) ? bk : bk + 1;  // first slot of this wave right of column k
        constexpr int UC = RB == 2 ? 4 : (CB < 8 ? CB : 8);
        static_assert(CB % UC == 0, "whole update chunks");
#pragma unroll
        for (int ch = 0; ch < CB / UC; ++ch) {
            if (UC * (ch + 1) > b0) {  // uniform
                double pv[UC];
#pragma unroll
                for (int j = 0; j < UC; ++j) pv[j] = s.prow[c][UC * ch + j];
#pragma unroll
                for (int j = 0; j < UC; ++j)
#pragma unroll
                    for (int a = 0; a < RB; ++a) A[a][UC * ch + j] = fma(-l[a], pv[j], A[a][UC * ch + j]);
            }
        }
#pragma unroll
        for (int t = 0; t < NR; ++t) {
            const double pv = s.prowy[c][t];
#pragma unroll
            for (int a = 0; a < RB; ++a) Y[a][t] = fma(-l[a], pv, Y[a][t]);
        }
        // the next position's staging rewrites prow: this wave's reads of it are done
        wave_sync();
    }
    __syncthreads();
    // last position: its row is the last pivot (not tested by the reference)
    {
        const int k = n - 1, ck = k % TC, bk = k / TC;
        if (c == ck) {
#pragma unroll
            for (int a = 0; a < RB; ++a) {
                const int row = lane + WAVE * a;
                if (row < n && pos[a] == k) {
                    double v = 0.0;
#define HZ_GETV(b)                                                              \
    if constexpr ((b) < CB) v = A[a][(b)];
                    HZ_SLOT_AT(bk, HZ_GETV)
#undef HZ_GETV
                    s.udiag[k] = v;
                    s.phys_of_pos[k] = row;
                }
            }
        }
    }
    __syncthreads();
}

// Work-groups per CU the mode kernel is compiled for: three for the
// 128-row instance with few right-hand sides (168 VGPRs, a few spilled in
// the assembly phase; measured 14% faster than two at 256^2 x 129), the
// register file's own limit otherwise.
template <int RB, int NR>
constexpr int lu_modes_min_wgs() { return RB == 2 && NR <= 2 ? 3 : 1; }

// ---------------------------------------------------------------------------
template <int RB, int CB, int NR, int TC>
__global__ __launch_bounds__(WAVE *TC, (lu_modes_min_wgs<RB, NR>())) void lu_modes_kernel(const LuModeArgs a)
{
    constexpr int ROWS = WAVE * RB;
    constexpr int BQ = 2 * TC;  // back-substitution block of U columns
    __shared__ Smem<RB, CB, NR, TC> sm;
    __shared__ double ubuf[2][BQ][ROWS];
    __shared__ double sh_ul[ROWS], sh_ur[ROWS];
    __shared__ double dots[2][kMaxRhs];
    __shared__ double zeta[2][kMaxRhs];

    const int lane = threadIdx.x % WAVE, c = uniform(threadIdx.x / WAVE);
    const int n = a.n, N = a.N, E = a.ne, K = a.K, M = a.M;
    const int mode = blockIdx.x;
    const int iw = mode / M, jw = mode % M;
    const int xi0 = iw, xi1 = a.nx - iw, nxi = (iw != 0 && iw != a.nx / 2) ? 2 : 1;
    const int nc = 2 * nxi;
    const int CPE = rhs_per_element(a.nx);  // right-hand sides per element
    const int nrhs = 2 + CPE * E;

    // ---- assemble (physical row i = position i) ----
    double A[RB][CB], Y[RB][NR];
    int pos[RB];
    const double sig = a.shift ? a.shift[mode] : 0.0;
#pragma unroll
    for (int aa = 0; aa < RB; ++aa) {
        const int i = lane + WAVE * aa;
#pragma unroll
        for (int b = 0; b < CB; ++b) {
            const int j = c + TC * b;
            double v = 0.0;
            if (i < n && j < n) {
                if (a.op_full) {
                    v = a.op_full[((long long)mode * n + i) * n + j];
                } else {
                    v = a.DzzT[(long long)j * n + i];
                    if (i == j) v = a.op_diag ? a.op_diag[(long long)mode * n + i] : v - sig;
                }
            }
            A[aa][b] = v;
        }
#pragma unroll
        for (int t = 0; t < NR; ++t) {
            const int r = c + TC * t;  // right-hand-side index
            double v = 0.0;
            if (i < n && r < nrhs) {
                if (r == 0) {
                    v = -a.Dzz[(long long)(i + 1) * N];
                } else if (r == 1) {
                    v = -a.Dzz[(long long)(i + 1) * N + (N - 1)];
                } else {
                    const int e = (r - 2) / CPE, cc = (r - 2) % CPE;
                    if (cc < nc) {
                        const int xi = (cc >> 1) ? xi1 : xi0;
                        const long long col = (long long)xi * a.xs + (long long)jw * K;
                        const double *src = (cc & 1) ? a.spec_im : a.spec_re;
                        v = src[col + (long long)e * (N - 1) + 1 + i];
                    }
                }
            }
            Y[aa][t] = v;
        }
    }

    eliminate_modes<RB, CB, NR, TC>(A, Y, pos, n, sm);

    // ---- back substitution, U staged through LDS in blocks of BQ columns ----
    // Within a block the positions are unrolled, so the loads of U, of the
    // pivot rows and of the diagonal issue ahead of the division chain.
    const int nblk = (n + BQ - 1) / BQ;
    int buf = 0;
    for (int blk = nblk - 1; blk >= 0; --blk) {
        const int q0 = blk * BQ;
        // wave c owns the block's columns q0 + c + TC*tq, slots blk*BQ/TC + tq
        switch (blk) {
#define HZ_UBLK(B)                                                              \
    case (B):                                                                   \
        if constexpr (((B) + 1) * (BQ / TC) <= CB) {                            \
            _Pragma("unroll") for (int tq = 0; tq < BQ / TC; ++tq)              \
            _Pragma("unroll") for (int aa = 0; aa < RB; ++aa)                   \
                ubuf[buf][c + TC * tq][lane + WAVE * aa] = A[aa][(B) * (BQ / TC) + tq]; \
        }                                                                       \
        break;
            HZ_UBLK(0) HZ_UBLK(1) HZ_UBLK(2) HZ_UBLK(3) HZ_UBLK(4) HZ_UBLK(5) HZ_UBLK(6) HZ_UBLK(7)
            HZ_UBLK(8) HZ_UBLK(9) HZ_UBLK(10) HZ_UBLK(11) HZ_UBLK(12) HZ_UBLK(13) HZ_UBLK(14) HZ_UBLK(15)
#undef HZ_UBLK
        default: break;
        }
        static_assert(CB / (BQ / TC) <= 16, "at most 16 back-substitution blocks");
        __syncthreads();
#pragma unroll
        for (int qq = BQ - 1; qq >= 0; --qq) {
            const int q = q0 + qq;
            if (q >= n) continue;  // uniform
            const int ph = uniform(sm.phys_of_pos[q]);
            const double uqq = sm.udiag[q];
            double u[RB];
#pragma unroll
            for (int aa = 0; aa < RB; ++aa) u[aa] = ubuf[buf][qq][lane + WAVE * aa];
#pragma unroll
            for (int t = 0; t < NR; ++t) {
                if (c + TC * t >= nrhs) continue;  // uniform
                double own = Y[0][t];
#pragma unroll
                for (int aa = 1; aa < RB; ++aa)
                    if (aa == ph / WAVE) own = Y[aa][t];
                const double xq = lane_f64(own, ph % WAVE) / uqq;
#pragma unroll
                for (int aa = 0; aa < RB; ++aa) {
                    const int row = lane + WAVE * aa;
                    double v = Y[aa][t];
                    if (pos[aa] < q && row < n) v = fma(-u[aa], xq, v);
                    if (row == ph) v = xq;
                    Y[aa][t] = v;
                }
            }
        }
        buf ^= 1;
    }
    __syncthreads();

    // ---- end-point derivatives of every right-hand side ----
#pragma unroll
    for (int t = 0; t < NR; ++t) {
        const int r = c + TC * t;
        if (r >= nrhs) continue;  // uniform
        double d0 = 0.0, dN = 0.0;
#pragma unroll
        for (int aa = 0; aa < RB; ++aa) {
            const int row = lane + WAVE * aa;
            if (row < n) {
                const int q = pos[aa];
                d0 = fma(a.Dz[q + 1], Y[aa][t], d0);
                dN = fma(a.Dz[(long long)(N - 1) * N + q + 1], Y[aa][t], dN);
                if (r == 0) sh_ul[q] = Y[aa][t];
                if (r == 1) sh_ur[q] = Y[aa][t];
            }
        }
#pragma unroll
        for (int m = 1; m < WAVE; m <<= 1) {
            d0 += __shfl_xor(d0, m);
            dN += __shfl_xor(dN, m);
        }
        if (r == 0) {  // ul has u(0) = 1
            d0 += a.Dz[0];
            dN += a.Dz[(long long)(N - 1) * N];
        } else if (r == 1) {  // ur has u(N-1) = 1
            d0 += a.Dz[N - 1];
            dN += a.Dz[(long long)(N - 1) * N + N - 1];
        }
        if (lane == 0) {
            dots[0][r] = d0;
            dots[1][r] = dN;
        }
    }
    __syncthreads();

    // ---- element-interface sweep, one thread per real column ----
    if (threadIdx.x < nc && a.periodic) {
        const int cc = threadIdx.x;
        const long long fc = (long long)((cc >> 1) ? xi1 : xi0) * M + jw;
        periodic_sweep(E, dots[0][0], dots[1][0], dots[0][1], dots[1][1], &dots[0][2 + cc], &dots[1][2 + cc],
                       &zeta[0][cc], &zeta[1][cc], CPE,
                       a.closure ? a.closure + ((cc & 1) ? (long long)a.nx * M : 0) + fc : nullptr);
    } else if (threadIdx.x < nc) {
        const int cc = threadIdx.x;
        const long long bcol = (long long)((cc >> 1) ? xi1 : xi0) * M + jw;
        const double *bsrc = (cc & 1) ? a.bc_im : a.bc_re;
        const double gl = bsrc[bcol], gr = bsrc[(long long)a.nx * M + bcol];
        const double d0ul = dots[0][0], dNul = dots[1][0], d0ur = dots[0][1], dNur = dots[1][1];
        // forward elimination over the elements; b and rhs kept in zeta
        double bprev = -a.alpha_l * d0ul + a.beta_l;
        double cprev = -a.alpha_l * d0ur;
        double rprev = gl + a.alpha_l * dots[0][2 + cc];
        zeta[0][cc] = bprev;
        zeta[1][cc] = rprev;
        for (int e = 1; e < E; ++e) {
            const double lb = dNul;
            const double be = dNur - d0ul - (cprev / bprev) * lb;
            const double re = dots[0][2 + CPE * e + cc] - dots[1][2 + CPE * (e - 1) + cc] - (rprev / bprev) * lb;
            bprev = be;
            cprev = -d0ur;
            rprev = re;
            zeta[0][CPE * e + cc] = be;
            zeta[1][CPE * e + cc] = re;
        }
        const double lb = a.alpha_r * dNul;
        const double br = a.alpha_r * dNur + a.beta_r - (cprev / bprev) * lb;
        const double rhsr = gr - a.alpha_r * dots[1][2 + CPE * (E - 1) + cc] - (rprev / bprev) * lb;
        double zr = rhsr / br;
        for (int e = E - 1; e >= 0; --e) {
            const double be = zeta[0][CPE * e + cc], re = zeta[1][CPE * e + cc];
            const double ce = (e == 0) ? -a.alpha_l * d0ur : -d0ur;
            const double zl = (re - ce * zr) / be;
            zeta[0][CPE * e + cc] = zl;
            zeta[1][CPE * e + cc] = zr;
            zr = zl;
        }
    }
    __syncthreads();

    // ---- assembly straight into the spectrum ----
#pragma unroll
    for (int t = 0; t < NR; ++t) {
        const int r = c + TC * t;
        if (r < 2 || r >= nrhs) continue;
        const int e = (r - 2) / CPE, cc = (r - 2) % CPE;
        if (cc >= nc) continue;
        const double zl = zeta[0][CPE * e + cc], zr = zeta[1][CPE * e + cc];
        const int xi = (cc >> 1) ? xi1 : xi0;
        const long long col = (long long)xi * a.xs + (long long)jw * K + (long long)e * (N - 1);
        double *dst = (cc & 1) ? a.spec_im : a.spec_re;
#pragma unroll
        for (int aa = 0; aa < RB; ++aa) {
            const int row = lane + WAVE * aa;
            if (row < n) {
                const int q = pos[aa];
                dst[col + 1 + q] = Y[aa][t] + (zl * sh_ul[q] + zr * sh_ur[q]);
            }
        }
        if (lane == 0) {
            dst[col] = zl;
            dst[col + N - 1] = zr;
        }
    }
    if (threadIdx.x == 0) a.status[mode] = sm.singular == 2 ? HZ_ERR_HIP : sm.singular ? HZ_ERR_SINGULAR : 0;
}

// Batched factorisation in the reference's storage (no right-hand sides).
template <int RB, int CB, int TC>
__global__ __launch_bounds__(WAVE *TC) void lu_factor_kernel(const LuFactorArgs a)
{
    __shared__ Smem<RB, CB, 1, TC> sm;
    const int lane = threadIdx.x % WAVE, c = uniform(threadIdx.x / WAVE);
    const int n = a.n;
    double *Am = a.A + (long long)blockIdx.x * n * n;
    double A[RB][CB], Y[RB][1];
    int pos[RB];
#pragma unroll
    for (int aa = 0; aa < RB; ++aa) {
        Y[aa][0] = 0.0;
#pragma unroll
        for (int b = 0; b < CB; ++b) {
            const int i = lane + WAVE * aa, j = c + TC * b;
            A[aa][b] = (i < n && j < n) ? Am[(long long)i * n + j] : 0.0;
        }
    }
    __syncthreads();  // every read of Am before the factors overwrite it
    eliminate<RB, CB, 1, TC, true>(A, Y, pos, n, 0, sm, Am,
                                   a.piv ? a.piv + (long long)blockIdx.x * (n - 1) : nullptr);
    if (threadIdx.x == 0) {
        if (a.det) a.det[blockIdx.x] = sm.det;
        if (a.status) a.status[blockIdx.x] = sm.singular ? HZ_ERR_SINGULAR : 0;
    }
}

// LU_Decomposition.cpp:78-108 for many systems, one thread per system (a
// utility for callers of the reference's LUSolve; not on the solve path),
// every product and difference rounded separately as in the reference.
__global__ void lu_solve_kernel(const double *LU, const int *piv, double *b, int n, int count)
{
#pragma clang fp contract(off)
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

template <int RB, int CB, int NR>
int launch_modes_t(const LuModeArgs &a, int modes, hipStream_t s)
{
    hipLaunchKernelGGL((lu_modes_kernel<RB, CB, NR, 4>), dim3(modes), dim3(WAVE * 4), 0, s, a);
    HZ_HIP(hipGetLastError());
    return HZ_OK;
}

template <int RB, int CB>
int launch_factor_t(const LuFactorArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL((lu_factor_kernel<RB, CB, 4>), dim3(a.count), dim3(WAVE * 4), 0, s, a);
    HZ_HIP(hipGetLastError());
    return HZ_OK;
}

template <int RB, int CB>
int launch_modes_nr(const LuModeArgs &a, int modes, hipStream_t s)
{
    const int nr = (2 + rhs_per_element(a.nx) * a.ne + 3) / 4;  // right-hand sides per wave
    if (nr <= 2) return launch_modes_t<RB, CB, 2>(a, modes, s);
    if (nr <= 8) return launch_modes_t<RB, CB, 8>(a, modes, s);
    if constexpr (RB == 1) {
        if (nr <= 32) return launch_modes_t<1, CB, 32>(a, modes, s);
        if (nr <= 64) return launch_modes_t<1, CB, 64>(a, modes, s);
    }
    return set_error(HZ_ERR_UNSUPPORTED, "PLAIN_LU form: too many elements (2 + c*ne right-hand sides, c = 4 "
                                         "when nx >= 4 else 2, must be <= 256, or <= 32 when N-2 > 64)");
}

}  // namespace

int lu_modes_check(int nx, int n, int ne)
{
    if (n > 128) return set_error(HZ_ERR_UNSUPPORTED, "PLAIN_LU form: N-2 must be <= 128");
    const int nrhs = 2 + rhs_per_element(nx) * ne;
    if (nrhs > 4 * (n <= 64 ? 64 : 8))  // 4 waves x NR of the largest instance
        return set_error(HZ_ERR_UNSUPPORTED, "PLAIN_LU form: too many elements (at most 63, or 127 when nx < 4; "
                                             "7, or 15 when nx < 4, for N-2 > 64)");
    return HZ_OK;
}

int lu_modes_shapes(int nx, int n, int ne, int *cols, int max)
{
    // column capacity 4*CB of the four instances: 64 rows (RB = 1) up to 64
    // columns, then 128 rows and columns, whose right-hand sides stop at 8
    // per wave; an instance holds every smaller n (masked columns, padding rows)
    static const int caps[] = {16, 32, 64, 128};
    const int nr = (2 + rhs_per_element(nx) * ne + 3) / 4;
    int k = 0;
    for (int cap : caps) {
        if (cap < n || nr > (cap == 128 ? 8 : 64)) continue;
        if (k < max) cols[k] = cap;
        ++k;
    }
    return k;
}

int launch_lu_modes(const LuModeArgs &a, int modes, hipStream_t s)
{
    if (int st = lu_modes_check(a.nx, a.n, a.ne)) return st;
    const int cols = a.cols ? a.cols : a.n <= 16 ? 16 : a.n <= 32 ? 32 : a.n <= 64 ? 64 : 128;
    int c[4];
    const int k = lu_modes_shapes(a.nx, a.n, a.ne, c, 4);
    bool held = false;
    for (int i = 0; i < k && i < 4; ++i) held = held || c[i] == cols;
    if (!held) return set_error(HZ_ERR_ARG, "PLAIN_LU form: the mode-kernel shape does not hold this plan");
    switch (cols) {
    case 16: return launch_modes_nr<1, 4>(a, modes, s);
    case 32: return launch_modes_nr<1, 8>(a, modes, s);
    case 64: return launch_modes_nr<1, 16>(a, modes, s);
    default: return launch_modes_nr<2, 32>(a, modes, s);
    }
}

int launch_lu_factor(const LuFactorArgs &a, hipStream_t s)
{
    if (a.count <= 0) return HZ_OK;
    const int n = a.n;
    if (n <= 32) return launch_factor_t<1, 8>(a, s);
    if (n <= 64) return launch_factor_t<1, 16>(a, s);
    if (n <= 128) return launch_factor_t<2, 32>(a, s);
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
