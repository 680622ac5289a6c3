/*
 * hz.h -- C ABI of the MI355X-native DirectPseudoSpectral Helmholtz solver.
 *
 * Solves  d2u/dx2 + d2u/dy2 + d2u/dz2 - lambda u = f  on a box that is
 * periodic (Fourier) in x and y and discretised in z by E spectral elements
 * of N Gauss-Lobatto-Legendre points, with Robin data on the two z faces:
 *     beta_l u - alpha_l du/dz = g_l   at the first z node,
 *     beta_r u + alpha_r du/dz = g_r   at the last z node.
 *
 * It replaces the C++ interface of gpritam/MPI,
 * HelmholtzSolvers/DirectPseudoSpectral/include/NonPeriodic3D_MPI.h:33-47
 * (CreatePDEPlan / DestroyPDEPlan / SolvePDE / FirstDerivativeZ /
 * SecondDerivativeZ) and LU_Decomposition.h:8-9, with an explicit plan
 * handle instead of process globals.  The C++ drop-in with the reference's
 * names and globals (hz_compat.h) is built on top of this header.
 *
 * Conventions
 *  - All arrays passed to hz_solve* / hz_adjoint / hz_derivative_z are DEVICE
 *    pointers (HIP), float64, row-major with z fastest:
 *        f, u          [nx][ny][nz_total],   nz_total = ne*(n-1)+1
 *        g_l, g_r      [nx][ny]
 *  - Every entry point returns 0 on success or a positive HZ_ERR_* code;
 *    hz_last_error_string() describes the last failure of the calling thread.
 *  - Work is enqueued on the caller's stream (NULL = default stream); the
 *    call does not synchronise unless the entry point says so.
 *  - A plan is not thread-safe: concurrent calls on one plan race on its
 *    workspace.  Different plans are independent.
 */
#ifndef HZ_H
#define HZ_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hz_plan hz_plan;
typedef struct ihipStream_t *hz_stream_t; /* == hipStream_t */

enum {
    HZ_OK = 0,
    HZ_ERR_ARG = 1,          /* invalid argument or descriptor */
    HZ_ERR_NOT_POW2 = 2,     /* Fourier_Quadrature.cpp:1106-1114 */
    HZ_ERR_N_TOO_SMALL = 3,  /* GLLQuadrature.cpp:82-83 */
    HZ_ERR_SINGULAR = 4,     /* LU_Decomposition.cpp:49-50 */
    HZ_ERR_ALLOC = 5,        /* General.h:311-312 */
    HZ_ERR_HIP = 6,          /* a HIP runtime call failed */
    HZ_ERR_UNSUPPORTED = 7,  /* combination not provided by this form */
    HZ_ERR_NO_DEVICE = 8
};

/* Solution methods ("forms").
 *  PLAIN_LU  : the reference algorithm on the GPU -- every solve assembles and
 *              factors each (|kx|,|ky|) operator Dzz_int - sigma I with the
 *              reference's scaled partial pivoting (LU_Decomposition.cpp:7-72),
 *              solves particular + homogeneous columns, runs the element
 *              interface sweep and assembles.  FP64 throughout.
 *  DIAG      : z-diagonalised form -- Dzz_int = V diag(mu) V^-1 is computed
 *              once per plan; a solve is two FP64 MFMA contractions per
 *              column batch plus an O(n) shift per column.  No factorisation.
 *              One element with Dirichlet faces uses the even/odd split of
 *              the mirror-symmetric GLL operator (hz_host_eigen_split): half
 *              the contraction work, plus a bf16 first-order correction for
 *              the operator's rounding-level asymmetry, so the result keeps
 *              the full basis's accuracy.
 *  MIXED     : FP32 solves through the plan's eigenbasis (two FP32 matrix-core
 *              contractions per column batch) refined with FP64 residuals of
 *              the GLL operator (one FP64 matrix-core contraction per sweep,
 *              refine_sweeps sweeps); no per-mode operators are stored.
 *              Transforms stay FP64.
 */
enum { HZ_FORM_PLAIN_LU = 0, HZ_FORM_DIAG = 1, HZ_FORM_MIXED = 2 };

typedef struct {
    int nx, ny;              /* Fourier points in x, y: powers of two >= 1.
                                nx = 1 is the 2-D solver of
                                NonPeriodic2D_MPI.cpp (its periodic x is
                                this y, its GLL y this z); nx = ny = 1 the
                                1-D solver of NonPeriodic1D_MPI.cpp       */
    int n;                   /* GLL points per element (reference N), >= 3 */
    int ne;                  /* elements in z (reference size*Ns), >= 1    */
    double xl, xr, yl, yr;   /* periodic box                                */
    double zl, zr;           /* z extent; only Lz = zr - zl enters the operator */
    double lambda;           /* Helmholtz shift (reference Lambda)          */
    double alpha_l, beta_l, alpha_r, beta_r; /* Robin coefficients          */
    int form;                /* HZ_FORM_*                                   */
    int refine_sweeps;       /* MIXED only: refinement sweeps (0 = default 2,
                                < 0 = none, <= 16)                          */
    double mixed_fp64_cond;  /* not used (kept for the layout): the MIXED
                                refinement converges for every shift, so no
                                mode needs an FP64 operator               */
    int periodic_z;          /* 1: z periodic as in Periodic3D_MPI.cpp (u and
                                u_z continuous across z = zr -> zl); the Robin
                                coefficients and face data are not used.
                                All three forms.                           */
} hz_plan_desc;

/* Replaces CreatePDEPlan (NonPeriodic3D_MPI.cpp:40-134): GLL nodes, scaled
 * Dz/Dzz, the operator list and device workspace.  Must be called with the
 * target device current. */
int hz_plan_create(const hz_plan_desc *desc, hz_plan **plan);
/* Replaces DestroyPDEPlan (NonPeriodic3D_MPI.cpp:139-170). */
int hz_plan_destroy(hz_plan *plan);

/* Plan geometry.  z (host, nz_total entries) follows the reference exactly:
 * element e spans [e*Lz/ne, (e+1)*Lz/ne], i.e. z starts at 0, not zl
 * (NonPeriodic3D_MPI.cpp:111-120). Dz, Dzz (host, n*n, row-major) are the
 * element matrices scaled by 2ne/Lz and 4ne^2/Lz^2 (:122-131). */
int hz_plan_nz_total(const hz_plan *plan, int *nz_total);
int hz_plan_z_nodes(const hz_plan *plan, double *z);
int hz_plan_operators(const hz_plan *plan, double *Dz, double *Dzz);
int hz_plan_num_modes(const hz_plan *plan, int *num_modes); /* (nx/2+1)(ny/2+1) */
/* Changes lambda (and with it every operator) for later solves. */
int hz_plan_set_lambda(hz_plan *plan, double lambda);
/* Changes the Robin coefficients of the two faces for later solves. */
int hz_plan_set_robin(hz_plan *plan, double alpha_l, double beta_l, double alpha_r, double beta_r);

/* Replaces SolvePDE (NonPeriodic3D_MPI.cpp:237-591; with periodic_z,
 * Periodic3D_MPI.cpp:233-863) for the standard operator of the reference
 * test callback (TestNonPeriodic3D_MPI.cpp:71-79):
 * Dzz_int - (lambda + 4 pi^2 iw^2/Lx^2 + 4 pi^2 jw^2/Ly^2) I.
 * g_l, g_r may be NULL for a periodic-z plan (they are not used).
 * Inputs are not modified (the reference destroys them). */
int hz_solve(hz_plan *plan, const double *f, const double *g_l, const double *g_r,
             double *u, hz_stream_t stream);

/* SolvePDE with caller-assembled operators (the reference's callback
 * contract, NonPeriodic3D_MPI.cpp:248-253): mode m = iw*(ny/2+1) + jw.
 *  op_diag != NULL, op_full == NULL : operator m = Dzz_int with its diagonal
 *     replaced by op_diag[m*(n-2) .. +n-2)  (device, float64)
 *  op_full != NULL : operator m = op_full[m*(n-2)^2 ..] row-major (device).
 * Runs the reference-algorithm (PLAIN_LU) kernel whatever the plan's form;
 * needs N - 2 <= 128. */
int hz_solve_operators(hz_plan *plan, const double *f, const double *g_l,
                       const double *g_r, const double *op_diag,
                       const double *op_full, double *u, hz_stream_t stream);

/* Status of the last solve's factorisations: after the stream has finished,
 * returns HZ_ERR_SINGULAR if any operator failed the reference's pivot test
 * (|scaled pivot| <= 1e-12).  Synchronises the stream. */
int hz_solve_status(hz_plan *plan, hz_stream_t stream);

/* Adjoint (transpose) of the solve map (f, g_l, g_r) -> u at fixed lambda:
 * given ubar = dL/du (device, [nx][ny][nz_total]), writes
 *   fbar  = dL/df      [nx][ny][nz_total]  (zero on element end nodes)
 *   glbar = dL/dg_l, grbar = dL/dg_r   [nx][ny]
 * and, if u != NULL (the forward solution) and lambda_bar != NULL (host),
 *   *lambda_bar = dL/dlambda = sum over nodes of fbar * u.
 * PLAIN_LU, DIAG and MIXED forms, any number of elements; with periodic z
 * there is no face data and glbar, grbar must be NULL. */
int hz_adjoint(hz_plan *plan, const double *ubar, const double *u, double *fbar,
               double *glbar, double *grbar, double *lambda_bar, hz_stream_t stream);

/* No counterpart in the reference; for callers that cannot allocate (code
 * under HIP graph capture, serving loops that forbid hipMalloc): hz_adjoint
 * that allocates nothing and does not synchronise.  dL/dlambda, when
 * lambda_bar_dev != NULL (then u is required), is written to that DEVICE
 * word on `stream`; the reduction's partial sums live in a static device
 * array of 1024 doubles (8 KiB per GPU, kept until the process ends), enough
 * for any field size.  The plan's diagonalised operators must exist: DIAG and
 * MIXED plans have them from creation, a PLAIN_LU plan after its first
 * hz_adjoint call (HZ_ERR_UNSUPPORTED before).  The plan's device must be
 * current.  Calls that overlap on one GPU (two streams at once) race on the
 * shared array.  Other arguments and results as hz_adjoint. */
int hz_adjoint_noalloc(hz_plan *plan, const double *ubar, const double *u, double *fbar, double *glbar,
                       double *grbar, double *lambda_bar_dev, hz_stream_t stream);

/* No counterpart in the reference: the explicit right-hand side of a
 * semi-implicit Allen-Cahn step, f[i] = a3 u[i]^3 + a1 u[i] for i < n
 * (device pointers; f may equal u), computed as u * (a3 u u + a1) in one
 * pass on the caller's stream.  mpi_amd/allen_cahn.py calls it with
 * a3 = 1/eps^2, a1 = -(1/eps^2 + lambda). */
int hz_cubic_rhs(const double *u, double *f, long long n, double a3, double a1, hz_stream_t stream);
/* hz_solve with the right-hand side f = a3 u^3 + a1 u formed inside the
 * forward transform's loads (the operations of hz_cubic_rhs, so the result
 * equals hz_cubic_rhs followed by hz_solve bit for bit), without the
 * intermediate field: a semi-implicit Allen-Cahn step in one call.  u is
 * read only by the first pass, so out may be u.  No counterpart in the
 * reference. */
int hz_solve_cubic(hz_plan *plan, const double *u, double a3, double a1, const double *g_l,
                   const double *g_r, double *out, hz_stream_t stream);
/* One backward step through hz_solve_cubic's right-hand side, fused: with
 * fbar = dL/df of the step u_in -> u_out (from hz_adjoint),
 *   ubar = fbar (3 a3 u_in^2 + a1)          (dL/du_in; ubar may be fbar)
 *   *acc += sum fbar (b1 (u_out - u_in) + b3 (u_in^3 - u_in))
 * (acc a device double; a parameter's gradient through the step, e.g. the
 * Allen-Cahn eps with b1 = dlambda/deps, b3 = da3/deps).  Fields of the
 * plan's size, device pointers; the sum runs in a fixed order (the same on
 * every run).  No counterpart in the reference. */
int hz_cubic_adjoint(hz_plan *plan, const double *fbar, const double *u_in, const double *u_out, double a3,
                     double a1, double b1, double b3, double *ubar, double *acc, hz_stream_t stream);

/* Replaces FirstDerivativeZ / SecondDerivativeZ (NonPeriodic3D_MPI.cpp:632-683):
 * out = Dz u (order 1) or Dzz u (order 2) applied to every element column.
 * Here u and out use the element layout of the reference's FieldVariable
 * arrays, [ne][nx][ny][n] (z fastest, interface nodes stored in both
 * neighbouring elements), device pointers. */
int hz_derivative_z(hz_plan *plan, int order, const double *u, double *out,
                    hz_stream_t stream);

/* Fourier derivatives along x (axis 0) or y (axis 1), order 1 or 2, of a
 * stack of `lanes` planes u[nx][ny][lanes] -> out (device pointers): the
 * 3-D FirstDerivativeFourierX/Y and SecondDerivativeFourierX/Y of
 * Fourier_Quadrature.cpp:1494-1575, :1721-1800 (wavenumber k = j for
 * j < L/2, j - L otherwise), which the reference's FirstDerivativeX/Y and
 * SecondDerivativeX/Y of the Helmholtz solver leave empty
 * (NonPeriodic3D_MPI.cpp:596-627). */
int hz_derivative_xy(hz_plan *plan, int axis, int order, const double *u, double *out, int lanes,
                     hz_stream_t stream);

/* Replaces FourierTransformXY (Fourier_Quadrature.cpp:1040-1155) for a
 * stack of `lanes` planes: isign = +1 maps real in[nx][ny][lanes] to the half
 * spectrum re/im[nx][ny/2+1][lanes] (unnormalised, exp(-2 pi i jk/L), x index
 * in FFT order); isign = -1 maps re/im (overwritten) back to real
 * out[nx][ny][lanes] with the 1/(nx ny) factor.  Device pointers. */
int hz_fourier_xy(hz_plan *plan, int isign, const double *in, double *re, double *im,
                  double *out, int lanes, hz_stream_t stream);

/* Batched LU in the reference's storage: for `count` row-major n x n
 * matrices A (device), factor in place with LU_Decomposition.cpp:7-72's
 * pivot rule; piv[count*(n-1)] (device int32) and det[count] (device, may
 * be NULL) as the reference's npivot and determinant; status[count]
 * (device int32, may be NULL) = 0 or HZ_ERR_SINGULAR. */
int hz_lu_factor_batched(double *A, int n, int count, int *piv, double *det,
                         int *status, hz_stream_t stream);
/* LU_Decomposition.cpp:78-108 for `count` systems: b[count*n] overwritten. */
int hz_lu_solve_batched(const double *LU, const int *piv, double *b, int n, int count,
                        hz_stream_t stream);

/* Many independent problems of uneven size solved together (no
 * counterpart in the reference, which solves one problem per process).
 * plans[p] are DIAG- or MIXED-form plans, mixed freely (distinct; nx, ny 1 or in [16, 256]; N - 2 <= 128;
 * ne <= 16; Robin or periodic z); f[p], g_l[p], g_r[p], u[p] are device
 * arrays in each plan's layout (g_l[p], g_r[p] NULL for a periodic plan).
 * Two ways to run a batch, each with mode-solve tiles of 32 or 64 rows:
 * staged (variants 4, 5: one launch per stage and kind of task, recorded in
 * a HIP graph that later solves replay, re-reading the buffer table when the
 * pointers change), merged (variants 8, 9: one launch per stage over every
 * task kind of the stage, also recorded in a graph) and fused (variants 6, 7:
 * ONE cooperative launch per solve, the five stages separated by grid
 * barriers; built but switched off and refused, since two of its GPU runs did
 * not finish).  The first solve of a batch shape times every usable variant
 * and keeps the fastest.  Variants 0-3 are refused (an earlier fused design,
 * removed).
 * Enqueued on `stream`; a plan in a batch must not be used by another solve
 * while the batch runs. */
typedef struct hz_batch hz_batch;
int hz_batch_create(hz_plan *const *plans, int count, hz_batch **batch);
int hz_batch_solve(hz_batch *batch, const double *const *f, const double *const *gl,
                   const double *const *gr, double *const *u, hz_stream_t stream);
/* Synchronises the batch and reports HZ_ERR_HIP if the fused kernel ran
 * without a cooperative launch (it then does no work). */
int hz_batch_status(hz_batch *batch);
/* The autotuner's choice (variant index; fused = 0 staged per task kind,
 * 1 fused, 2 merged per stage; tile rows; the fused launch's work-groups per
 * CU) and the measured ms of every variant (0 where not timed). */
int hz_batch_tuning(const hz_batch *batch, int *variant, int *fused, int *tile_rows,
                    int *wgs_per_cu, double *variant_ms, int max_variants);
/* Pins the variant (4, 5, 8 or 9; 0-3 and 6-7 are refused; -1 = autotune
 * again on the next solve); drops the recorded graph.  For tests and tuning
 * studies. */
int hz_batch_force_variant(hz_batch *batch, int variant);
int hz_batch_destroy(hz_batch *batch);
/* One-shot form of the above: create the batch, solve, wait for it,
 * destroy it.  Returns when u[] is complete.
 * Every call autotunes and records its graph again, so a caller that solves
 * the same set repeatedly keeps an hz_batch instead. */
int hz_solve_batched(hz_plan *const *plans, int count, const double *const *f, const double *const *gl,
                     const double *const *gr, double *const *u, hz_stream_t stream);

/* Refinement history of the last MIXED solve (synchronises the device):
 * rel_residual[s] = max over spectral columns and elements of
 * |b - A x|_max / |b|_max of the iterate before refinement sweep s + 1,
 * i.e. after the FP32 solve (s = 0) and after s sweeps (s = 1..S-1), with
 * S = refine_sweeps (the residual after the last sweep is not formed: it
 * would cost one more FP64 contraction); *sweeps = number of values written
 * (S, at most max_sweeps).  No counterpart in the reference (its LU is FP64
 * only). */
int hz_mixed_history(const hz_plan *plan, double *rel_residual, int max_sweeps, int *sweeps);
/* Plan-time autotuning results: the mode kernel's shape -- tile rows of the
 * diagonalised kernel (DIAG plans), the column capacity (16, 32, 64 or 128)
 * of the reference-algorithm kernel's instance (PLAIN_LU plans; the LU
 * shapes of LU_Decomposition.cpp:7-72's port), 0 for MIXED -- with the
 * measured ms of each candidate (0 for a candidate that does not apply, and
 * none when only one candidate holds the plan), and the
 * z lanes per work-group chosen for each of the four transform passes
 * (R2C y, C2C x, C2C^-1 x, C2R y; 0 = the length's only shape) with the
 * winner's ms.  Any pointer may be NULL.  No counterpart in the reference. */
int hz_plan_tuning(const hz_plan *plan, int *diag_tile_rows, double *diag_ms, int max_diag,
                   int *fft_lanes, double *fft_ms);
/* The mode kernel's candidate shapes (DIAG tile rows, or PLAIN_LU column
 * capacities), in the order of hz_plan_tuning's diag_ms (*count = their
 * number, at most `max` written), and a way to pin one of them (a DIAG tile
 * must hold one column's 2*ne element rows).  For tests and tuning studies;
 * no counterpart in the reference. */
int hz_plan_diag_candidates(const hz_plan *plan, int *tile_rows, int max, int *count);
int hz_plan_set_diag_tile(hz_plan *plan, int tile_rows);
/* Runs the plan-time autotuning again (transform lane widths; DIAG tile
 * rows or PLAIN_LU instance), e.g. once the GPU has reached its clocks
 * under load.  The plan's device must be current.  Synchronises the device
 * first (the tuners time on the plan's own spectrum and face buffers, which
 * a solve in flight would share), allocates a scratch field for the
 * transform timings, clears the last solve's statuses and periodic
 * closures, and leaves the new choices and times for hz_plan_tuning.
 * Results of later solves are unchanged: every DIAG tile height and
 * PLAIN_LU instance gives the same output bit for bit, and a transform's
 * lane width sets how many lines one work-group takes, not any line's
 * arithmetic (the radix split depends on the length only).  No counterpart
 * in the reference. */
int hz_plan_retune(hz_plan *plan);
/* Periodic-z plans: the closure coefficient of the cyclic element-interface
 * system per real spectral column of the last solve (the reference's
 * alphaPDE, Periodic3D_MPI.cpp:698), half-spectrum order:
 * alpha[plane][i * (ny/2 + 1) + m], plane 0 = real, 1 = imaginary part;
 * zeros with one element (the reference computes none).  Synchronises. */
int hz_plan_periodic_closure(const hz_plan *plan, double *alpha);
/* Number of modes of a MIXED plan that keep an FP64 operator: 0 (every mode
 * runs the FP32 eigenbasis solve; kept for callers of the earlier form). */
int hz_mixed_fp64_modes(const hz_plan *plan, int *fp64_modes);

/* Host-only utilities (no device needed): the plan-time setup that
 * hz_plan_create performs, exposed so it can be checked on any machine.
 * hz_host_plan_operators: GLL nodes on [-1,1] (x[n]), the z array
 * (z[ne*(n-1)+1]) and the scaled Dz, Dzz (n*n each); any pointer may be NULL.
 * hz_host_eigen: Dzz_int = V diag(mu) V^-1 for the (n-2)x(n-2) interior
 * operator of the same plan (row-major V, Vinv; mu ascending).
 * hz_host_eigen_split: the even/odd eigenbasis that the DIAG and adjoint
 * kernels use.  With m = n-2 and the fold
 *   F x = (x_i + x_{m-1-i}, i < m/2 [+ the centre x_{m/2} when m is odd] |
 *          x_i - x_{m-1-i}, i < m/2)
 * (even coordinates at 0.., odd ones at np/2..), the symmetrised operator
 * (A + JAJ)/2, J the reversal, equals F^-1 W diag(mu) Winv F with W, Winv
 * block diagonal (np x np, row-major, zero padding) and mu[np] (zero
 * padding); np even with np/2 >= ceil(m/2).  *asym = max|A - JAJ| / max|A|
 * (the rounding of the GLL nodes).  Pointers may be NULL. */
int hz_host_plan_operators(int n, int ne, double lz, double *x, double *z, double *Dz,
                           double *Dzz);
int hz_host_eigen(int n, int ne, double lz, double *mu, double *V, double *Vinv);
int hz_host_eigen_split(int n, int ne, double lz, int np, double *mu, double *W, double *Winv,
                        double *asym);

/* Stage timing with HIP events on the solve stream (measurement hook).
 * While enabled, every solve records an event before and after each of its
 * kernel stages; hz_plan_profile_read synchronises, returns the mean
 * milliseconds per stage over the recorded solves and clears the record.
 * Stage names (one per slot) are written to `names` as a ';'-joined list. */
int hz_plan_profile(hz_plan *plan, int enable);
int hz_plan_profile_read(hz_plan *plan, double *stage_ms, int max_stages, int *nstages,
                         int *nsolves, char *names, int names_len);

/* Attainable ceilings of the current device, measured on the spot (about
 * 10 ms): the FP64 matrix rate of v_mfma_f64_16x16x4f64 at full occupancy
 * with no memory traffic (TFLOP/s), and the rate of a device-to-device copy
 * of copy_bytes per buffer (GB/s, bytes read + written).  Either output
 * may be NULL.  Allocates scratch buffers; synchronises.  A measurement
 * hook with no counterpart in the reference. */
int hz_probe_ceilings(double *fp64_mfma_tflops, double *copy_gbps, size_t copy_bytes);

const char *hz_last_error_string(void);
const char *hz_version(void);

#ifdef __cplusplus
}
#endif
#endif /* HZ_H */
