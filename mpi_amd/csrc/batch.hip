// batch.hip -- many independent DIAG- or MIXED-form problems of uneven size (nx, ny,
// N, ne, lambda, faces) solved together.
//
// Every problem is the five-stage pipeline of a solve:
//   0 R2C along y (+ faces)   1 C2C along x (+ faces)   2 mode solve (diagonalised, or MIXED)
//   3 C2C^-1 along x          4 C2R along y
// split into tasks: a tile of spectral columns of the mode solve, or for a
// transform pass one uniform-base tile of the standalone transform kernels
// (staged variant) or one whole line batch with all its z-lane blocks
// (fused variant).
//
// Fused variant (built, switched off -- see kVariants): ONE persistent
// launch.  Tasks are numbered stage-major
// (every problem's stage 0, then every problem's stage 1, ...) and handed out
// by an atomic counter.  A task of stage s waits until its problem's stage
// s-1 has completed all its tasks (a per-problem counter, released with an
// agent-scope fence after the task's stores, acquired before the task's
// loads, so the cross-XCD L2s are written back / invalidated).  Everything a
// task waits for was handed out earlier to a running work-group, so the wait
// always ends; a spin limit still bounds it and raises an error flag.
// Staged variant: one launch per stage and task kind (the launch boundary
// is the dependency), the same task bodies.
//
// The autotuner times the enabled variants (mode-solve tile height 32 or 64
// rows) on the first solve of a batch shape, keeps the fastest, and captures
// it in a HIP graph that later solves replay.
#include "fft_body.h"
#include "mixed_body.h"

#include <array>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace hz {

constexpr int kStages = 5;

struct BatchProblem {
    FftArgs f[4];    // passes 0..3 of xy_pass_args
    DiagArgs d;      // DIAG plans
    MixedArgs m;     // MIXED plans
    int ly, lx;      // transform lengths along y, x
    int ntask[kStages];
    int dkind;       // 2 * (NP / 32 - 1) + general (NP = 32, 64, 96, 128), + 8 for a MIXED plan
};

namespace {

// ---- task bodies -------------------------------------------------------------
template <int KIND>
__device__ __noinline__ void fft_rows(const FftArgs &a, int L, int batch, double *smem)
{
    const int lanes = a.lanes + a.xlanes;
    switch (L) {
    case 16:
        for (int lb = 0; lb * 64 < lanes; ++lb) fft_tile<16, 4, 4, 64, KIND>(a, lb, batch, smem);
        break;
    case 32:
        for (int lb = 0; lb * 32 < lanes; ++lb) fft_tile<32, 8, 4, 32, KIND>(a, lb, batch, smem);
        break;
    case 64:
        for (int lb = 0; lb * 32 < lanes; ++lb) fft_tile<64, 8, 8, 32, KIND>(a, lb, batch, smem);
        break;
    case 128:
        for (int lb = 0; lb * 16 < lanes; ++lb) fft_tile<128, 16, 8, 16, KIND>(a, lb, batch, smem);
        break;
    case 256:
        for (int lb = 0; lb * 16 < lanes; ++lb) fft_tile<256, 16, 16, 16, KIND>(a, lb, batch, smem);
        break;
    default:
        break;
    }
}

template <int NJ, int RT, bool GENERAL>
__device__ __forceinline__ void diag_run_t(const DiagArgs &a, int tile, double *smem)
{
    diag_tile<NJ, RT, GENERAL, false>(a, tile, *reinterpret_cast<DiagShared<NJ, RT, GENERAL, false> *>(smem));
}

template <int RT>
__device__ __noinline__ void diag_run(const DiagArgs &a, int kind, int tile, double *smem)
{
    switch (kind) {
    case 0: diag_run_t<2, RT, false>(a, tile, smem); break;
    case 1: diag_run_t<2, RT, true>(a, tile, smem); break;
    case 2: diag_run_t<4, RT, false>(a, tile, smem); break;
    case 3: diag_run_t<4, RT, true>(a, tile, smem); break;
    case 4: diag_run_t<6, RT, false>(a, tile, smem); break;
    case 5: diag_run_t<6, RT, true>(a, tile, smem); break;
    case 6: diag_run_t<8, RT, false>(a, tile, smem); break;
    case 7: diag_run_t<8, RT, true>(a, tile, smem); break;
    default: break;
    }
}

template <int RT>
__device__ void run_task(const BatchProblem &pr, int s, int local, double *smem)
{
    switch (s) {
    case 0: fft_rows<FFT_R2C>(pr.f[0], pr.ly, local, smem); break;
    case 1: fft_rows<FFT_C2C_FWD>(pr.f[1], pr.lx, local, smem); break;
    case 2: diag_run<RT>(pr.d, pr.dkind, local, smem); break;
    case 3: fft_rows<FFT_C2C_INV>(pr.f[2], pr.lx, local, smem); break;
    case 4: fft_rows<FFT_C2R>(pr.f[3], pr.ly, local, smem); break;
    default: break;
    }
}

// ctl layout: [0] task counter, [1] error flag, [2..3] unused,
// [4 + p*kStages + s] completed tasks of stage s of problem p.
__global__ void batch_reset_kernel(int *ctl, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) ctl[i] = 0;
}

// Problem of task `rel` of stage s: off[s*(nprob+1) + p] is the first task
// of problem p within stage s.
__device__ __forceinline__ int find_problem(const int *__restrict__ off, int nprob, int s, int rel, int *local)
{
    const int *o = off + s * (nprob + 1);
    int lo = 0, hi = nprob;  // o[lo] <= rel < o[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (o[mid] <= rel) lo = mid;
        else hi = mid;
    }
    *local = rel - o[lo];
    return lo;
}

// Fused persistent launch.  base[s]: first task of stage s (base[kStages] =
// total); tasks are taken from the counter ctl[0].
template <int RT>
__global__ __launch_bounds__(256, 2) void batch_fused_kernel(const BatchProblem *__restrict__ P, int nprob,
                                                             const int *__restrict__ off,
                                                             const int *__restrict__ base, int *ctl, int spin_limit)
{
    extern __shared__ double smem[];
    __shared__ int s_stage, s_p, s_local;
    int *done = ctl + 4;
    for (;;) {
        if (threadIdx.x == 0) {
            int s = -1, p = 0, local = 0;
            const int t = atomicAdd(&ctl[0], 1);
            if (t < base[kStages]) {
                s = 0;
                while (t >= base[s + 1]) ++s;
                p = find_problem(off, nprob, s, t - base[s], &local);
                if (s > 0) {
                    const int need = P[p].ntask[s - 1];
                    int *cnt = &done[p * kStages + s - 1];
                    int spins = 0;
                    while (__hip_atomic_load(cnt, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < need) {
                        __builtin_amdgcn_s_sleep(4);
                        if (++spins > spin_limit) {
                            atomicOr(&ctl[1], 1);
                            break;
                        }
                    }
                }
            }
            s_stage = s;
            s_p = p;
            s_local = local;
        }
        __syncthreads();
        const int s = s_stage;
        if (s < 0) break;
        // every wave orders its loads after the wait (cross-XCD caches)
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        const int p = s_p;
        run_task<RT>(P[p], s, s_local, smem);
        // every wave's stores are complete and written back before the signal
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (threadIdx.x == 0)
            __hip_atomic_fetch_add(&done[p * kStages + s], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Staged launches: one launch per stage and kind of task (transform length,
// or mode-solve variant), each work-group one task from a (problem, task)
// list; the launch boundaries order the stages, and every kernel inlines only
// its own body, so it gets its own register budget.
// One uniform-base tile per work-group (fft_tile_uniform: a lane block of a
// line batch, or W line batches of a face), entries in XCD-aware order.
template <int L, int R1, int R2, int W, int KIND>
__global__ __launch_bounds__(256) void batch_fft_list(const BatchProblem *__restrict__ P, const int2 *__restrict__ list,
                                                      int pass)
{
    __shared__ double tile[(L + 2) * W];
    const int2 t = list[xcd_logical(blockIdx.x, gridDim.x)];
    const FftArgs &a = P[t.x].f[pass];
    // The lists count both face tiles in the forward passes; a table filled
    // without faces simply ends here instead of reading null face pointers.
    const long long tiles = uniform_main_tiles(a.nbatch, a.lanes, a.in_batch_stride, a.out_batch_stride, W) +
                            (long long)a.xlanes * ((a.nbatch + W - 1) / W);
    if (t.y >= tiles) return;
    fft_tile_uniform<L, R1, R2, W, KIND>(a, t.y, tile);
}

// Length-1 transform passes (2-D and 1-D problems): 256 (batch, lane)
// copies per list entry.
template <int KIND>
__global__ __launch_bounds__(256) void batch_unit_list(const BatchProblem *__restrict__ P, const int2 *__restrict__ list,
                                                       int pass)
{
    const int2 t = list[blockIdx.x];
    unit_line_element<KIND>(P[t.x].f[pass], (long long)t.y * 256 + threadIdx.x);
}

template <int NJ, int RT, bool GENERAL>
__global__ __launch_bounds__(256) void batch_diag_list(const BatchProblem *__restrict__ P, const int2 *__restrict__ list)
{
    __shared__ DiagShared<NJ, RT, GENERAL, false> sh;
    const int2 t = list[blockIdx.x];
    diag_tile<NJ, RT, GENERAL, false>(P[t.x].d, t.y, sh);
}

template <int NJ, int RT, bool GENERAL>
__global__ __launch_bounds__(256, 2) void batch_mixed_list(const BatchProblem *__restrict__ P,
                                                           const int2 *__restrict__ list)
{
    __shared__ MixedShared<NJ, RT, GENERAL> sh;
    const int2 t = list[blockIdx.x];
    mixed_tile<NJ, RT, GENERAL>(P[t.x].m, t.y, sh);
}

template <int RT>
constexpr size_t batch_lds()
{
    constexpr size_t fft = sizeof(double) * (256 + 2) * 16;  // largest transform tile (L = 256, W = 16)
    constexpr size_t dg = sizeof(DiagShared<8, RT, true, false>);
    return fft > dg ? fft : dg;
}

}  // namespace
}  // namespace hz

using namespace hz;

// ---- host side ---------------------------------------------------------------
namespace {

struct Variant {
    int rt;       // diag tile rows
    int fused;    // 1 persistent, 0 one launch per stage
    int wgs;      // persistent work-groups per CU
    int enabled;  // offered to the autotuner and to hz_batch_force_variant
};

// The fused persistent variants are switched off: their first run on the
// GPU did not finish (its dependency waits ran to the spin limit, i.e. the
// completion counters written by work-groups on other XCDs were not seen),
// and the cause is not established.  Only the staged variants run.  (The
// fused task bodies also predate the length-1 passes of 2-D and 1-D plans:
// fft_rows has no L = 1 case, so they would need one before being enabled.)
const Variant kVariants[] = {{32, 1, 2, 0}, {32, 1, 3, 0}, {32, 1, 4, 0}, {64, 1, 2, 0}, {32, 0, 0, 1}, {64, 0, 0, 1}};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);
constexpr int kSpinLimit = 1 << 16;  // ~0.1 s of 2-us probes, then the error flag

std::mutex g_tune_mu;
std::map<std::string, int> g_tuned;  // batch shape -> variant

}  // namespace

struct hz_batch {
    std::vector<hz_plan *> plans;
    int n = 0, device = 0, cus = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr, t0 = nullptr, t1 = nullptr;
    BatchProblem *d_probs = nullptr, *h_probs = nullptr;
    int *d_tab[2] = {nullptr, nullptr};  // per tile height: off then base
    std::vector<int> h_tab[2];
    int *d_ctl = nullptr;
    int nctl = 0;
    // staged variant: per tile height, task lists grouped by (stage, kind)
    int2 *d_list[2] = {nullptr, nullptr};
    std::vector<std::array<int, 3>> groups[2];  // {stage, kind, first}, count in next first
    std::vector<const void *> last;
    std::vector<unsigned> gens;  // each plan's generation when the tables were built
    std::string shape;
    int variant = -1;
    double variant_ms[kNumVariants] = {};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

int rt_index(int rt) { return rt == 64 ? 1 : 0; }

// off/base tables of the batch for diag tiles of `rt` rows
std::vector<int> task_table(const hz_batch *b, int rt, std::vector<std::array<int, kStages>> *counts)
{
    const int n = b->n;
    std::vector<int> tab((size_t)kStages * (n + 1) + kStages + 1, 0);
    counts->assign(n, {});
    for (int p = 0; p < n; ++p) {
        const hz_plan *pl = b->plans[p];
        // MIXED tiles have a fixed height per padded size
        const int rows = pl->d.form == HZ_FORM_MIXED ? mixed_tile_rows(diag_padded(pl->n)) : rt;
        const int fc_per_tile = rows / (2 * pl->ne);
        const int nfc = pl->nx * pl->M;
        (*counts)[p] = {pl->nx, pl->M, (nfc + fc_per_tile - 1) / fc_per_tile, pl->M, pl->nx};
    }
    int total = 0;
    int *base = tab.data() + (size_t)kStages * (n + 1);
    for (int s = 0; s < kStages; ++s) {
        base[s] = total;
        int *o = tab.data() + (size_t)s * (n + 1);
        int acc = 0;
        for (int p = 0; p < n; ++p) {
            o[p] = acc;
            acc += (*counts)[p][s];
        }
        o[n] = acc;
        total += acc;
    }
    base[kStages] = total;
    return tab;
}

int fill_problems(hz_batch *b, const double *const *f, const double *const *gl, const double *const *gr,
                  double *const *u, int rt)
{
    std::vector<std::array<int, kStages>> counts;
    (void)task_table(b, rt, &counts);
    for (int p = 0; p < b->n; ++p) {
        hz_plan *pl = b->plans[p];
        BatchProblem &q = b->h_probs[p];
        q = BatchProblem{};
        const long long xs = pl->xs;
        q.f[0] = xy_pass_args(pl, 0, f[p], pl->spec.re, pl->spec.im, nullptr, pl->K, gl[p], gr[p], nullptr, nullptr,
                              xs);
        q.f[1] = xy_pass_args(pl, 1, nullptr, pl->spec.re, pl->spec.im, nullptr, pl->K, gl[p], gr[p], nullptr,
                              nullptr, xs);
        q.f[2] = xy_pass_args(pl, 2, nullptr, pl->spec.re, pl->spec.im, nullptr, pl->K, nullptr, nullptr, nullptr,
                              nullptr, xs);
        q.f[3] = xy_pass_args(pl, 3, nullptr, pl->spec.re, pl->spec.im, u[p], pl->K, nullptr, nullptr, nullptr,
                              nullptr, xs);
        if (pl->d.form == HZ_FORM_MIXED)
            q.m = mixed_args(pl);
        else
            q.d = diag_args(pl);
        q.ly = pl->ny;
        q.lx = pl->nx;
        for (int s = 0; s < kStages; ++s) q.ntask[s] = counts[p][s];
        const bool general = pl->ne > 1 || pl->d.alpha_l != 0.0 || pl->d.alpha_r != 0.0 || pl->d.periodic_z;
        q.dkind = 2 * (diag_padded(pl->n) / 32 - 1) + (general ? 1 : 0) + (pl->d.form == HZ_FORM_MIXED ? 8 : 0);
    }
    return HZ_OK;
}

// z lanes per work-group of the batch's transform tiles of length L.
int batch_lanes(int L) { return L <= 16 ? 64 : L <= 64 ? 32 : 16; }

// Kind of a transform task: log2 of its length; of a mode-solve task: dkind.
int task_kind(const BatchProblem &q, int s) { return s == 2 ? q.dkind : (s == 0 || s == 4 ? q.ly : q.lx); }

template <int KIND>
void launch_fft_list(hipStream_t s, const BatchProblem *P, const int2 *list, int count, int L, int pass)
{
    switch (L) {
    case 1: hipLaunchKernelGGL((batch_unit_list<KIND>), dim3(count), dim3(256), 0, s, P, list, pass); break;
    case 16: hipLaunchKernelGGL((batch_fft_list<16, 4, 4, 64, KIND>), dim3(count), dim3(256), 0, s, P, list, pass); break;
    case 32: hipLaunchKernelGGL((batch_fft_list<32, 8, 4, 32, KIND>), dim3(count), dim3(256), 0, s, P, list, pass); break;
    case 64: hipLaunchKernelGGL((batch_fft_list<64, 8, 8, 32, KIND>), dim3(count), dim3(256), 0, s, P, list, pass); break;
    case 128: hipLaunchKernelGGL((batch_fft_list<128, 16, 8, 16, KIND>), dim3(count), dim3(256), 0, s, P, list, pass); break;
    default: hipLaunchKernelGGL((batch_fft_list<256, 16, 16, 16, KIND>), dim3(count), dim3(256), 0, s, P, list, pass); break;
    }
}

void launch_mixed_list(hipStream_t s, const BatchProblem *P, const int2 *list, int count, int kind)
{
    const dim3 g(count), b(256);
    switch (kind) {  // MIXED tile heights: mixed_tile_rows
    case 0: hipLaunchKernelGGL((batch_mixed_list<2, 32, false>), g, b, 0, s, P, list); break;
    case 1: hipLaunchKernelGGL((batch_mixed_list<2, 32, true>), g, b, 0, s, P, list); break;
    case 2: hipLaunchKernelGGL((batch_mixed_list<4, 16, false>), g, b, 0, s, P, list); break;
    case 3: hipLaunchKernelGGL((batch_mixed_list<4, 16, true>), g, b, 0, s, P, list); break;
    case 4: hipLaunchKernelGGL((batch_mixed_list<6, 32, false>), g, b, 0, s, P, list); break;
    case 5: hipLaunchKernelGGL((batch_mixed_list<6, 32, true>), g, b, 0, s, P, list); break;
    case 6: hipLaunchKernelGGL((batch_mixed_list<8, 16, false>), g, b, 0, s, P, list); break;
    default: hipLaunchKernelGGL((batch_mixed_list<8, 16, true>), g, b, 0, s, P, list); break;
    }
}

template <int RT>
void launch_diag_list(hipStream_t s, const BatchProblem *P, const int2 *list, int count, int kind)
{
    const dim3 g(count), b(256);
    if (kind >= 8) {
        launch_mixed_list(s, P, list, count, kind - 8);
        return;
    }
    switch (kind) {
    case 0: hipLaunchKernelGGL((batch_diag_list<2, RT, false>), g, b, 0, s, P, list); break;
    case 1: hipLaunchKernelGGL((batch_diag_list<2, RT, true>), g, b, 0, s, P, list); break;
    case 2: hipLaunchKernelGGL((batch_diag_list<4, RT, false>), g, b, 0, s, P, list); break;
    case 3: hipLaunchKernelGGL((batch_diag_list<4, RT, true>), g, b, 0, s, P, list); break;
    case 4: hipLaunchKernelGGL((batch_diag_list<6, RT, false>), g, b, 0, s, P, list); break;
    case 5: hipLaunchKernelGGL((batch_diag_list<6, RT, true>), g, b, 0, s, P, list); break;
    case 6: hipLaunchKernelGGL((batch_diag_list<8, RT, false>), g, b, 0, s, P, list); break;
    default: hipLaunchKernelGGL((batch_diag_list<8, RT, true>), g, b, 0, s, P, list); break;
    }
}

template <int RT>
int enqueue_t(hz_batch *b, const Variant &v, hipStream_t s)
{
    const int r = rt_index(RT);
    const int *tab = b->d_tab[r];
    const int *off = tab, *base = tab + (size_t)kStages * (b->n + 1);
    constexpr size_t lds = batch_lds<RT>();
    static bool attr_set = false;  // opt in to more than 64 KiB of dynamic LDS once
    if (!attr_set) {
        HZ_HIP(hipFuncSetAttribute((const void *)batch_fused_kernel<RT>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        attr_set = true;
    }
    if (v.fused) {
        hipLaunchKernelGGL(batch_reset_kernel, dim3(8), dim3(256), 0, s, b->d_ctl, b->nctl);
        hipLaunchKernelGGL((batch_fused_kernel<RT>), dim3(v.wgs * b->cus), dim3(256), lds, s, b->d_probs, b->n, off,
                           base, b->d_ctl, kSpinLimit);
    } else {
        const auto &g = b->groups[r];
        for (size_t i = 0; i + 1 < g.size(); ++i) {
            const int st = g[i][0], kind = g[i][1], first = g[i][2], count = g[i + 1][2] - first;
            if (count == 0) continue;  // e.g. only in-place length-1 passes (identities) of this kind
            const int2 *list = b->d_list[r] + first;
            const int pass = st < 2 ? st : st - 1;
            switch (st) {
            case 0: launch_fft_list<FFT_R2C>(s, b->d_probs, list, count, kind, pass); break;
            case 1: launch_fft_list<FFT_C2C_FWD>(s, b->d_probs, list, count, kind, pass); break;
            case 2: launch_diag_list<RT>(s, b->d_probs, list, count, kind); break;
            case 3: launch_fft_list<FFT_C2C_INV>(s, b->d_probs, list, count, kind, pass); break;
            default: launch_fft_list<FFT_C2R>(s, b->d_probs, list, count, kind, pass); break;
            }
        }
    }
    HZ_HIP(hipGetLastError());
    return HZ_OK;
}

// Task lists of the staged variant for tile height rt (problem kinds known
// from the filled problem table).
int build_lists(hz_batch *b, int rt)
{
    const int r = rt_index(rt);
    std::vector<std::array<int, kStages>> counts;
    (void)task_table(b, rt, &counts);
    std::map<std::pair<int, int>, std::vector<int2>> by;
    for (int p = 0; p < b->n; ++p)
        for (int st = 0; st < kStages; ++st) {
            auto &v = by[{st, task_kind(b->h_probs[p], st)}];
            int ntask = counts[p][st];
            if (st != 2) {  // transform passes: one entry per uniform-base tile
                // The table is filled with null buffers here, so the face
                // count is set from what hz_batch_solve always passes: both
                // faces in the forward passes (its gl, gr are never null for
                // a problem with faces), none in the inverse ones, none for
                // periodic z.  Face offsets are below nx*M.
                const FftArgs &a = b->h_probs[p].f[st < 2 ? st : st - 1];
                const bool has_faces = !b->plans[p]->d.periodic_z;
                const int L = task_kind(b->h_probs[p], st), W = batch_lanes(L);
                const int faces = st < 2 && has_faces ? 2 : 0;
                if (L == 1) {  // copies, 256 per task; the in-place x passes are identities
                    const long long total = (long long)(a.lanes + faces) * a.nbatch;
                    ntask = (st == 1 || st == 3) ? 0 : (int)((total + 255) / 256);
                    for (int t = 0; t < ntask; ++t) v.push_back(make_int2(p, t));
                    continue;
                }
                const long long lim = 1LL << 29;  // element offsets whose byte offsets fit 32 bits
                const long long m = (W - 1) + (long long)(L - 1) * std::max(a.in_point_stride, a.out_point_stride);
                if (m >= lim) return set_error(HZ_ERR_UNSUPPORTED, "batch: problem too large for a tile");
                ntask = (int)uniform_main_tiles(a.nbatch, a.lanes, a.in_batch_stride, a.out_batch_stride, W) +
                        faces * ((a.nbatch + W - 1) / W);
            }
            for (int t = 0; t < ntask; ++t) v.push_back(make_int2(p, t));
        }
    std::vector<int2> all;
    b->groups[r].clear();
    for (auto &kv : by) {  // ordered by stage, then kind
        b->groups[r].push_back({kv.first.first, kv.first.second, (int)all.size()});
        all.insert(all.end(), kv.second.begin(), kv.second.end());
    }
    b->groups[r].push_back({kStages, 0, (int)all.size()});
    (void)hipFree(b->d_list[r]);
    b->d_list[r] = nullptr;
    if (hipMalloc((void **)&b->d_list[r], sizeof(int2) * all.size()) != hipSuccess)
        return set_error(HZ_ERR_ALLOC, "batch allocation failed");
    HZ_HIP(hipMemcpy(b->d_list[r], all.data(), sizeof(int2) * all.size(), hipMemcpyHostToDevice));
    return HZ_OK;
}

int enqueue(hz_batch *b, int vi, hipStream_t s)
{
    const Variant &v = kVariants[vi];
    return v.rt == 64 ? enqueue_t<64>(b, v, s) : enqueue_t<32>(b, v, s);
}

// Problem tables depend on the variant's tile height through the diag task
// counts: refill and upload them for variant vi.
int upload_problems(hz_batch *b, const double *const *f, const double *const *gl, const double *const *gr,
                    double *const *u, int vi)
{
    HZ_HIP(hipStreamSynchronize(b->stream));  // the pinned table may still be in flight
    fill_problems(b, f, gl, gr, u, kVariants[vi].rt);
    HZ_HIP(hipMemcpyAsync(b->d_probs, b->h_probs, sizeof(BatchProblem) * b->n, hipMemcpyHostToDevice, b->stream));
    return HZ_OK;
}

// Autotuning key of a batch: every plan's size and sweep kind.
std::string batch_shape(const std::vector<hz_plan *> &plans)
{
    std::string shape;
    for (const hz_plan *pl : plans) {
        const bool general = pl->ne > 1 || pl->d.alpha_l != 0.0 || pl->d.alpha_r != 0.0 || pl->d.periodic_z;
        shape += std::to_string(pl->nx) + "x" + std::to_string(pl->ny) + "x" + std::to_string(pl->N) + "e" +
                 std::to_string(pl->ne) + (pl->d.form == HZ_FORM_MIXED ? "m" + std::to_string(pl->sweeps) : "") +
                 (pl->d.periodic_z ? "p;" : general ? "g;" : ";");
    }
    return shape;
}

// Task lists of both tile heights from the plans as they are now (the
// mode-solve kind of a problem follows its Robin coefficients); the
// recorded graph and the uploaded table are dropped, so the next solve
// re-reads every plan.
int rebuild_lists(hz_batch *b)
{
    std::vector<const double *> nil(b->n, nullptr);
    std::vector<double *> nilo(b->n, nullptr);
    for (int r = 0; r < 2; ++r) {
        fill_problems(b, nil.data(), nil.data(), nil.data(), nilo.data(), r ? 64 : 32);
        if (int st = build_lists(b, r ? 64 : 32)) return st;
    }
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->last.clear();
    b->gens.resize(b->n);
    for (int p = 0; p < b->n; ++p) b->gens[p] = b->plans[p]->generation;
    b->shape = batch_shape(b->plans);
    return HZ_OK;
}

void free_batch(hz_batch *b)
{
    if (!b) return;
    // work enqueued by hz_batch_solve may still read the tables below
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (hipEvent_t e : {b->ev_in, b->ev_out, b->t0, b->t1})
        if (e) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    (void)hipFree(b->d_probs);
    (void)hipHostFree(b->h_probs);
    for (int *t : b->d_tab) (void)hipFree(t);
    for (int2 *l : b->d_list) (void)hipFree(l);
    (void)hipFree(b->d_ctl);
    delete b;
}

}  // namespace

extern "C" {

int hz_batch_create(hz_plan *const *plans, int count, hz_batch **batch)
{
    if (!plans || !batch || count < 1) return set_error(HZ_ERR_ARG, "bad argument");
    *batch = nullptr;
    int dev = 0;
    HZ_HIP(hipGetDevice(&dev));
    for (int p = 0; p < count; ++p) {
        const hz_plan *pl = plans[p];
        if (!pl) return set_error(HZ_ERR_ARG, "null plan");
        for (int q = 0; q < p; ++q)
            if (plans[q] == pl) return set_error(HZ_ERR_ARG, "a plan appears twice in the batch");
        if (pl->d.form != HZ_FORM_DIAG && pl->d.form != HZ_FORM_MIXED)
            return set_error(HZ_ERR_UNSUPPORTED, "batch: DIAG- or MIXED-form plans only");
        if (pl->device != dev) return set_error(HZ_ERR_ARG, "batch: plans on another device");
        if (pl->n > 128) return set_error(HZ_ERR_UNSUPPORTED, "batch: N-2 must be <= 128");
        if ((pl->nx != 1 && (pl->nx < 16 || pl->nx > 256)) || (pl->ny != 1 && (pl->ny < 16 || pl->ny > 256)))
            return set_error(HZ_ERR_UNSUPPORTED, "batch: nx, ny must be 1 or in [16, 256]");
        if (2 * pl->ne > 32) return set_error(HZ_ERR_UNSUPPORTED, "batch: at most 16 elements");
    }
    hz_batch *b = new hz_batch();
    b->plans.assign(plans, plans + count);
    b->n = count;
    b->device = dev;
    b->shape = batch_shape(b->plans);
    b->nctl = 4 + kStages * count;
    int st = HZ_OK;
    auto fail = [&](int code) {
        free_batch(b);
        return code;
    };
    if (hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || b->cus < 1)
        return fail(set_error(HZ_ERR_HIP, "cannot read the CU count"));
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming) != hipSuccess ||
        hipEventCreate(&b->t0) != hipSuccess || hipEventCreate(&b->t1) != hipSuccess)
        return fail(set_error(HZ_ERR_HIP, "stream/event creation failed"));
    if (hipMalloc((void **)&b->d_probs, sizeof(BatchProblem) * count) != hipSuccess ||
        hipHostMalloc((void **)&b->h_probs, sizeof(BatchProblem) * count) != hipSuccess ||
        hipMalloc((void **)&b->d_ctl, sizeof(int) * b->nctl) != hipSuccess)
        return fail(set_error(HZ_ERR_ALLOC, "batch allocation failed"));
    // the staged variants never write ctl; hz_batch_status reads its error
    // flag whatever ran, so it starts cleared (outside any graph capture)
    if (hipMemset(b->d_ctl, 0, sizeof(int) * b->nctl) != hipSuccess)
        return fail(set_error(HZ_ERR_HIP, "batch control clear failed"));
    for (int r = 0; r < 2; ++r) {
        std::vector<std::array<int, kStages>> counts;
        b->h_tab[r] = task_table(b, r ? 64 : 32, &counts);
        if (hipMalloc((void **)&b->d_tab[r], sizeof(int) * b->h_tab[r].size()) != hipSuccess)
            return fail(set_error(HZ_ERR_ALLOC, "batch allocation failed"));
        if (hipMemcpy(b->d_tab[r], b->h_tab[r].data(), sizeof(int) * b->h_tab[r].size(), hipMemcpyHostToDevice) !=
            hipSuccess)
            return fail(set_error(HZ_ERR_HIP, "batch table upload failed"));
    }
    // kinds only depend on the plans: the lists are built from a table
    // filled with null buffers
    if ((st = rebuild_lists(b))) return fail(st);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(b->shape);
        if (it != g_tuned.end()) b->variant = it->second;
    }
    *batch = b;
    return HZ_OK;
}

int hz_batch_solve(hz_batch *b, const double *const *f, const double *const *gl, const double *const *gr,
                   double *const *u, hz_stream_t stream)
{
    if (!b || !f || !gl || !gr || !u) return set_error(HZ_ERR_ARG, "null argument");
    for (int p = 0; p < b->n; ++p) {
        if (!f[p] || !u[p]) return set_error(HZ_ERR_ARG, "null buffer");
        if (b->plans[p]->d.periodic_z ? (gl[p] || gr[p]) : (!gl[p] || !gr[p]))
            return set_error(HZ_ERR_ARG, "face data: required without periodic z, NULL with it");
    }
    const hipStream_t cs = (hipStream_t)stream;
    int st;
    // a plan's lambda or Robin coefficients changed since the tables were
    // built (hz_plan_set_lambda / hz_plan_set_robin): rebuild them
    for (int p = 0; p < b->n; ++p)
        if (b->plans[p]->generation != b->gens[p]) {
            HZ_HIP(hipStreamSynchronize(b->stream));  // a replay may still read the old tables
            if ((st = rebuild_lists(b))) return st;
            break;
        }
    HZ_HIP(hipEventRecord(b->ev_in, cs));
    HZ_HIP(hipStreamWaitEvent(b->stream, b->ev_in, 0));
    std::vector<const void *> ptrs;
    for (int p = 0; p < b->n; ++p) {
        ptrs.push_back(f[p]);
        ptrs.push_back(gl[p]);
        ptrs.push_back(gr[p]);
        ptrs.push_back(u[p]);
    }
    if (b->variant < 0) {
        // autotune: time every variant on these buffers (each run is a full solve)
        double best = 1e300;
        int bi = 0;
        for (int vi = 0; vi < kNumVariants; ++vi) {
            if (!kVariants[vi].enabled) continue;
            if ((st = upload_problems(b, f, gl, gr, u, vi))) return st;
            if ((st = enqueue(b, vi, b->stream))) return st;  // warm-up
            HZ_HIP(hipEventRecord(b->t0, b->stream));
            for (int r = 0; r < 3; ++r)
                if ((st = enqueue(b, vi, b->stream))) return st;
            HZ_HIP(hipEventRecord(b->t1, b->stream));
            HZ_HIP(hipEventSynchronize(b->t1));
            float ms = 0.0f;
            HZ_HIP(hipEventElapsedTime(&ms, b->t0, b->t1));
            b->variant_ms[vi] = ms / 3.0;
            if (ms < best) {
                best = ms;
                bi = vi;
            }
        }
        b->variant = bi;
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[b->shape] = bi;
    }
    if (ptrs != b->last) {
        if ((st = upload_problems(b, f, gl, gr, u, b->variant))) return st;
        b->last = ptrs;
    }
    if (!b->exec) {
        HZ_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        st = enqueue(b, b->variant, b->stream);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (st) return st;
        if (e != hipSuccess) return set_error(HZ_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        b->graph = g;
        HZ_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    HZ_HIP(hipGraphLaunch(b->exec, b->stream));
    HZ_HIP(hipEventRecord(b->ev_out, b->stream));
    HZ_HIP(hipStreamWaitEvent(cs, b->ev_out, 0));
    return HZ_OK;
}

int hz_batch_status(hz_batch *b)
{
    if (!b) return set_error(HZ_ERR_ARG, "null argument");
    HZ_HIP(hipStreamSynchronize(b->stream));
    int flags[2] = {0, 0};
    HZ_HIP(hipMemcpy(flags, b->d_ctl, sizeof(flags), hipMemcpyDeviceToHost));
    if (flags[1]) return set_error(HZ_ERR_HIP, "batch: a dependency wait hit its spin limit");
    return HZ_OK;
}

int hz_batch_tuning(const hz_batch *b, int *variant, int *fused, int *tile_rows, int *wgs_per_cu,
                    double *variant_ms, int max_variants)
{
    if (!b) return set_error(HZ_ERR_ARG, "null argument");
    const int vi = b->variant;
    if (variant) *variant = vi;
    if (vi >= 0) {
        if (fused) *fused = kVariants[vi].fused;
        if (tile_rows) *tile_rows = kVariants[vi].rt;
        if (wgs_per_cu) *wgs_per_cu = kVariants[vi].wgs;
    }
    if (variant_ms)
        for (int i = 0; i < kNumVariants && i < max_variants; ++i) variant_ms[i] = b->variant_ms[i];
    return HZ_OK;
}

int hz_batch_force_variant(hz_batch *b, int variant)
{
    if (!b || variant < -1 || variant >= kNumVariants) return set_error(HZ_ERR_ARG, "bad variant");
    if (variant >= 0 && !kVariants[variant].enabled)
        return set_error(HZ_ERR_UNSUPPORTED, "batch: the fused persistent variants are switched off");
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->variant = variant;
    b->last.clear();  // the problem table depends on the variant's tile height
    return HZ_OK;
}

int hz_batch_destroy(hz_batch *b)
{
    free_batch(b);
    return HZ_OK;
}

int hz_solve_batched(hz_plan *const *plans, int count, const double *const *f, const double *const *gl,
                     const double *const *gr, double *const *u, hz_stream_t stream)
{
    hz_batch *b = nullptr;
    int st = hz_batch_create(plans, count, &b);
    if (st) return st;
    st = hz_batch_solve(b, f, gl, gr, u, stream);
    if (!st) st = hz_batch_status(b);
    if (!st) HZ_HIP(hipStreamSynchronize((hipStream_t)stream));
    free_batch(b);
    return st;
}

}  // extern "C"
