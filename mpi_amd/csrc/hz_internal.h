// hz_internal.h -- plan object and internal interfaces of libhz.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/hz.h"

namespace hz {

// ---- host-side plan setup (gll.cpp, eig.cpp) ------------------------------
double legendre(double x, int n, int *status);
int gll_nodes(int N, std::vector<double> &x);
int gll_weights(const std::vector<double> &x, std::vector<double> &w);
int gll_derivative_matrices(const std::vector<double> &x, std::vector<double> &D,
                            std::vector<double> &D2);
int plan_operators(int N, int ne, double Lz, std::vector<double> &z, std::vector<double> &Dz,
                   std::vector<double> &Dzz);
double mode_shift(double lambda, int iw, int jw, double Lx, double Ly);
int eigen_decompose(const std::vector<double> &A, int n, std::vector<double> &mu,
                    std::vector<double> &V, std::vector<double> &Vinv);
int eigen_decompose_split(const std::vector<double> &A, int n, int NP, std::vector<double> &mu,
                          std::vector<double> &W, std::vector<double> &Winv, double *asym);

// ---- error reporting --------------------------------------------------------
int set_error(int code, const std::string &msg);
#define HZ_HIP(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            return ::hz::set_error(HZ_ERR_HIP, std::string(#call) + ": " +           \
                                                   hipGetErrorString(e_));            \
    } while (0)

// ---- spectral data layout ----------------------------------------------------
// Physical fields: [nx][ny][K] (K = nz_total, z fastest).
// Half spectrum (R2C along y then C2C along x), split real/imaginary planes:
//   re[(i*M + m)*K + k], im[...]     i < nx (FFT order), m < M = ny/2+1
// Boundary spectra: [side][nx][M] split planes, side 0 = g_l, 1 = g_r.
struct Spectrum {
    double *re = nullptr, *im = nullptr;
};

// Arguments of the FFT line kernels (fft.hip).
enum FftKind { FFT_R2C = 0, FFT_C2R = 1, FFT_C2C_FWD = 2, FFT_C2C_INV = 3 };
struct FftArgs {
    const double *in_re, *in_im;   // in_im unused for R2C
    double *out_re, *out_im;       // out_im unused for C2R
    long long in_batch_stride, in_point_stride;
    long long out_batch_stride, out_point_stride;
    int lanes;                     // contiguous lanes (K)
    int nbatch;
    const double2 *tw;             // tw[t] = (cos 2 pi t/L, sin 2 pi t/L)
    // Up to two extra single-lane fields transformed in the same launch
    // (the boundary faces): lane `lanes + s` reads xin_*[s] and writes
    // xout_*[s] with the x_* strides.
    int xlanes;
    const double *xin_re[2], *xin_im[2];
    double *xout_re[2], *xout_im[2];
    long long x_in_batch_stride, x_in_point_stride, x_out_batch_stride, x_out_point_stride;
    int w;                         // z lanes per work-group (0 = default for L)
    // R2C only: the main lanes read x (pre_c3 x^2 + pre_c1) for the input x
    int pre_cubic;
    double pre_c3, pre_c1;
};
int launch_fft(int L, FftKind kind, const FftArgs &a, hipStream_t s);
// Lane widths offered for length L (autotuned per plan); returns the count.
int fft_width_choices(int L, int *w, int max);

// Arguments of the plain-LU mode kernel (lu.hip).
struct LuModeArgs {
    int nx, ny, M, K, N, n, ne;
    long long xs;            // spectrum x-row stride (>= M * K)
    double *closure;         // periodic z: [re|im][nx * M] closure coefficients, or null
    const double *DzzT;      // interior operator transposed: DzzT[j*n + i] = Dzz[i+1][j+1]
    const double *Dzz;       // full N x N
    const double *Dz;        // full N x N
    const double *shift;     // [modes] or null
    const double *op_diag;   // [modes][n] or null
    const double *op_full;   // [modes][n][n] or null
    double *spec_re, *spec_im;
    const double *bc_re, *bc_im;
    double alpha_l, beta_l, alpha_r, beta_r;
    int *status;             // [modes]
    int periodic;            // z periodic: cyclic interface system, no face data
    int cols;                // instance column capacity (lu_modes_shapes); 0 = the smallest
};
int launch_lu_modes(const LuModeArgs &a, int modes, hipStream_t s);
// Column capacities (16, 32, 64, 128) of the mode-kernel instances that hold
// (nx, N-2, ne), smallest first; returns their number (at most `max` written).
int lu_modes_shapes(int nx, int n, int ne, int *cols, int max);
// Whether the plain-LU mode kernel has an instance for (nx, N-2, ne); sets
// the error text when not.
int lu_modes_check(int nx, int n, int ne);

// Arguments of the z-diagonalised mode kernel (diag.hip).
struct DiagArgs {
    int nx, M, K, N, n, ne;
    long long xs;               // spectrum x-row stride: column (i, m) starts at i * xs + m * K
    double *spec;               // real plane; imaginary plane at spec + plane_off
    long long plane_off;
    const double *bc_re, *bc_im;
    const double *shift;        // [modes]
    const double *mu, *w0, *wN, *r0, *rN;  // [n], or [NP] with eo
    const double *VinvT, *VT;   // [NP][NP], NP = diag_padded(n): n rounded up to 32 (256 above 128)
    const double *hom;          // [modes][4]: Dz0.ul, DzN.ul, Dz0.ur, DzN.ur
    double alpha_l, beta_l, alpha_r, beta_r;
    int adjoint;                // transposed map; VinvT <- V, VT <- Vinv (row-major)
    int eo;                     // even/odd eigenbasis (one element, Dirichlet faces): the
                                // vectors padded to NP, the operands block diagonal
    const short *eoM;           // eo: the correction to the reference's operator (M, or M^T for
                                // the adjoint), bf16 [NP/4][NP][4]
    double *gbar_re, *gbar_im;  // adjoint only: [side][nx][M] face gradients
    int rt;                     // tile rows (spectral columns x planes x elements); 0 = default
    int periodic;               // z periodic: cyclic interface system, no face data
    double *closure;            // periodic z: [re|im][nx * M] closure coefficients, or null
};
int diag_padded(int n);
// Tile heights offered for padded size NP (autotuned per plan).
int diag_tile_choices(int NP, int *rt, int max);
int launch_diag_modes(const DiagArgs &a, hipStream_t s);


// MIXED form (mixed.hip, mixed_body.h): FP32 eigenbasis solves refined with
// FP64 residuals, both on the matrix cores.
struct MixedSetupArgs {
    int n, NP, modes;
    const double *shift, *mu, *w0, *wN;
    const double *VT;           // [NP][NP]: VT[k][i] = V[i][k]
    double *UL, *UR;            // [modes][n] homogeneous solutions (interior)
};
struct MixedArgs {
    int nx, M, K, N, n, ne, sweeps;
    long long xs;               // spectrum x-row stride (>= M * K)
    double *spec;               // real plane; imaginary plane at spec + plane_off
    long long plane_off;
    const double *bc_re, *bc_im, *shift;
    const double *mu;           // [NP] even/odd eigenvalues of the symmetrised Dzz_int (zero padded)
    const float *VinvT32, *VT32;  // [NP][NP]: Winv^T, (D W)^T of the even/odd basis in FP32
    const double *DzzTP;        // [NP][NP]: DzzTP[k][j] = Dzz[j+1][k+1] (zero padded)
    const double *Dzz, *Dz;     // [N][N] scaled element matrices
    const double *UL, *UR, *hom;
    double alpha_l, beta_l, alpha_r, beta_r;
    double *hist;               // [sweeps][hist_rows]: |r| / |b| per row before each sweep, or null
    long long hist_rows;        // nx * M * 2 * ne
    int periodic;               // z periodic: cyclic interface system, no face data
    double *closure;            // periodic z: [re|im][nx * M] closure coefficients, or null
};
int launch_mixed_setup(const MixedSetupArgs &a, hipStream_t s);
int launch_mixed_modes(const MixedArgs &a, hipStream_t s);
// Tile rows of the MIXED mode kernel for padded size NP.
int mixed_tile_rows(int NP);

struct LuFactorArgs {
    double *A;
    int n, count;
    int *piv;
    double *det;
    int *status;
};
int launch_lu_factor(const LuFactorArgs &a, hipStream_t s);
int launch_lu_solve(const double *LU, const int *piv, double *b, int n, int count, hipStream_t s);
int launch_derivative(const double *D, int N, long long ncols, const double *u, double *out, hipStream_t s);
// Multiplies the half spectrum re/im[A][B][lanes] by (2 pi i k/len)^order along
// axis A (axis_a) or B of length L (Fourier derivative).
int launch_spectral_derivative(double *re, double *im, long long A, long long B, long long lanes, int axis_a, int L,
                               double len, int order, hipStream_t s);
// sum_i a[i]*b[i] into *out (device or pinned host memory), deterministic two-pass reduction;
// partial >= kDotPartials doubles (the first pass runs at most that many work-groups)
constexpr int kDotPartials = 1024;
int launch_dot(const double *a, const double *b, long long n, double *partial, double *out, hipStream_t s);
// ubar = fbar (3 a3 u_in^2 + a1); *acc += sum fbar (b1 (u_out - u_in) + b3 (u_in^3 - u_in))
int launch_cubic_adjoint(const double *fbar, const double *u_in, const double *u_out, long long n, double a3,
                         double a1, double b1, double b3, double *ubar, double *partial, double *acc, hipStream_t s);
// n pseudo-random doubles in [-0.5, 0.5) (a hash of the index and seed):
// timing data for the plan-time tuners.
int launch_fill_noise(double *p, size_t n, unsigned long long seed, hipStream_t s);

// hz_adjoint on a plan whose diagonalised operators exist, with dL/dlambda
// (when lambda_bar_dev is not null) written to device memory through the
// caller's reduction scratch `partial` (kDotPartials doubles): allocates
// nothing and does not synchronise.
int adjoint_prebuilt(hz_plan *p, const double *ubar, const double *u, double *fbar, double *glbar, double *grbar,
                     double *lambda_bar_dev, double *partial, hipStream_t s);

}  // namespace hz

struct hz_plan;
namespace hz {
// The four line-transform passes of a solve on `p`'s grid:
//   0 R2C along y (in -> re/im; gl, gr -> boundary spectra when given),
//   1 C2C along x forward (re/im and, with gl/gr, the boundary spectra),
//   2 C2C along x inverse (boundary spectra too when gl_out/gr_out given),
//   3 C2R along y (re/im -> out; boundary spectra -> gl_out, gr_out).
// xs: x-row stride of the re/im half-spectrum buffers (0 = M * lanes, the
// unpadded layout of the public transform entry points).
FftArgs xy_pass_args(const hz_plan *p, int pass, const double *in, double *re, double *im, double *out,
                     int lanes, const double *gl, const double *gr, double *gl_out, double *gr_out,
                     long long xs = 0);
// x-row stride of a plan's own spectrum: M * K rounded up to 16 doubles, so
// that every 16-lane row segment of the merged x passes starts on a
// 128-byte line.
inline long long spectrum_xs(int M, int K) { return ((long long)M * K + 15) / 16 * 16; }
// Arguments of the forward diagonalised mode solve on `p`'s workspace.
DiagArgs diag_args(const hz_plan *p);
bool diag_general(const hz_plan *p);  // several elements, Robin faces or periodic z
void diag_basis(const hz_plan *p, DiagArgs &a, bool adjoint);
// Arguments of the MIXED mode solve on `p`'s workspace.
MixedArgs mixed_args(const hz_plan *p);
}  // namespace hz

// ---- the plan ----------------------------------------------------------------
struct hz_plan {
    hz_plan_desc d{};
    int device = 0;
    // bumped by every change of lambda or the Robin coefficients, so that a
    // batch holding copies of them (and the kernel kinds they select) knows
    // to rebuild its tables
    unsigned generation = 0;
    int nx = 0, ny = 0, N = 0, n = 0, ne = 0, K = 0, M = 0, modes = 0;
    long long xs = 0;  // x-row stride of spec (spectrum_xs(M, K))
    double Lx = 0, Ly = 0, Lz = 0;
    std::vector<double> z, Dz, Dzz, shift;
    // device
    double *d_Dz = nullptr, *d_Dzz = nullptr, *d_DzzT = nullptr, *d_shift = nullptr;
    double2 *d_twx = nullptr, *d_twy = nullptr;
    hz::Spectrum spec, bc;
    int *d_status = nullptr;
    double *d_closure = nullptr;  // periodic z: closure coefficient per real spectral column
    // z-diagonalised form
    int NP = 0;
    // full eigenbasis (n entries, V and Vinv n x n; the device operands
    // padded to NP x NP)
    std::vector<double> mu, V, Vinv, w0, wN, r0, rN;
    double *d_mu = nullptr, *d_w0 = nullptr, *d_wN = nullptr, *d_r0 = nullptr, *d_rN = nullptr;
    double *d_VinvT = nullptr, *d_VT = nullptr, *d_hom = nullptr;
    double *d_Vrow = nullptr, *d_Vinvrow = nullptr;  // adjoint operands
    // even/odd eigenbasis (one element with Dirichlet faces): mu, w0, wN,
    // r0, rN padded to NP each; operands NP x NP
    std::vector<double> eo_vec, eo_W, eo_Wi;  // eo_W, eo_Wi: W, Winv (MIXED's FP32 copies)
    double *d_eo_vec = nullptr, *d_eo_B1 = nullptr, *d_eo_B2 = nullptr, *d_eo_hom = nullptr;
    double *d_eo_B1a = nullptr, *d_eo_B2a = nullptr;  // adjoint operands
    short *d_eo_M = nullptr, *d_eo_Ma = nullptr;      // first-order correction (bf16), forward and adjoint
    int diag_rt = 0;                                  // autotuned tile rows
    int lu_cols = 0;                                  // autotuned PLAIN instance (0 = smallest)
    std::vector<double> lu_cols_ms;                   // tuning times per LU instance
    int fft_w[4] = {0, 0, 0, 0};                      // autotuned lanes per pass
    double fft_w_ms[4] = {0, 0, 0, 0};                // winner's ms per pass
    std::vector<double> diag_rt_ms;                   // tuning times per choice
    double *d_partial = nullptr;                      // reduction scratch
    double *h_lambda_bar = nullptr;                   // pinned host word the dL/dlambda reduction writes
    double *d_xspec = nullptr;                        // half spectrum of the x/y derivatives
    size_t xspec_n = 0;
    // MIXED form
    int sweeps = 0;
    float *d_VinvT32 = nullptr, *d_VT32 = nullptr;
    double *d_DzzTP = nullptr, *d_UL = nullptr, *d_UR = nullptr, *d_hist = nullptr;
    // stage profiling
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;      // events in use, per solve: nst+1
    std::vector<std::vector<hipEvent_t>> ev_solves;
    std::vector<std::string> stage_names;
};
