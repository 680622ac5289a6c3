// compat_core.h -- the engine shared by the drop-in solvers
// (Helmholtz3D.cpp, Helmholtz2D.cpp, Helmholtz1D.cpp, each built periodic and not):
// error handling, the operator recorder behind LUDecomposition, and one GPU
// plan with pinned staging in the plan's layout.  Internal to the drop-in
// libraries; the reference-named functions (LUDecomposition, LUSolve,
// ErrorMessage, CreateFFTPlan, DestroyFFTPlan) are defined in
// compat_core.cpp.
#pragma once

#include "../../include/hz.h"

#include <hip/hip_runtime.h>

#include <vector>

namespace hzc {

void hip_check(hipError_t e, const char *what);
void hz_check(int code);  // ErrorMessage(hz_last_error_string()) on failure

// While a SolvePDE runs the caller's ConstructOperator callbacks,
// LUDecomposition on the solver's Operator is recorded here instead of
// factored: the diagonal alone while every off-diagonal entry equals the
// interior of the second-derivative matrix (the usual shifted operator), the
// whole matrix from the first callback that differs.
struct Recorder {
    bool on = false, full = false;
    int mode = -1, modes = 0, n = 0;
    double **op = nullptr;   // the solver's Operator
    double **dxx = nullptr;  // its N x N second-derivative matrix
    std::vector<double> diag, ops;
    void begin(int modes_, int n_, double **op_, double **dxx_);
    bool take(double **A, int n_);  // false: not a matrix being recorded
};
Recorder &recorder();

// ConstructFundamentalSolutions of the reference solvers
// (NonPeriodic3D_MPI.cpp:184-232 and its periodic, 2-D and 1-D twins): for
// each of `count` element columns, zero end values and the interior solve of
// rhs[c][1..N-2] with the factored operator (A, piv from LUDecomposition),
// all columns in one batched GPU solve (the reference's LUSolve rounding).
// The homogeneous solutions the reference also forms go to its internal
// workspace, which is not part of its interface; SolvePDE forms its own.
void particular_solutions(double *const *cols, const double *const *rhs, int count, int N, double **A, int *piv);

// Size contract of the reference's FFT plan (Fourier_Quadrature.cpp:16-69):
// -1 when none was created.
int fft_plan_size();

// One GPU plan, with pinned host staging in its layout: h_f holds f (and
// then u) as [nx][ny][K]; h_g the faces [2][nx][ny].
struct Solver {
    hz_plan *plan = nullptr;
    hz_plan_desc desc{};  // as created (form: the one the plan runs)
    int K = 0, modes = 0, nx = 0, ny = 0, N = 0, ne = 0;
    bool periodic = false;
    double *d_f = nullptr, *d_gl = nullptr, *d_gr = nullptr, *d_u = nullptr, *d_ops = nullptr;
    double *h_f = nullptr, *h_g = nullptr;
    // A PLAIN_LU request the reference-algorithm kernel cannot hold (N - 2 >
    // 128, or too many elements) gets a DIAG plan: its solves then need the
    // standard operator (below), which is what the helpers need anyway.
    void create(const hz_plan_desc &d);
    void destroy();
    // Solve with the recorded operators: h_f / h_g in, u back into h_f.
    // When every recorded operator is the standard one, Dzz_int - sigma I
    // with sigma = lambda + 4 pi^2 iw^2/Lx^2 + 4 pi^2 jw^2/Ly^2 to the bit
    // (TestNonPeriodic3D_MPI.cpp:71-79), a DIAG or MIXED plan solves with its
    // own form; any other operator goes to the reference-algorithm kernel.
    void solve_recorded(double lambda, double al, double bl, double ar, double br);
    bool recorded_standard(double lambda) const;
    // Element-array derivative along z: host [ne][nx][ny][N], in place.
    void derivative_z(std::vector<double> &elements, int order);
    // Fourier derivative along axis 0 (x) or 1 (y) of a host stack
    // [nx][ny][lanes], in place.
    void derivative_xy(std::vector<double> &stack, int lanes, int axis, int order);
    // Periodic z: the last solve's closure coefficients in the reference's
    // full-spectrum layout alpha[i][2 m + c] (i < nx, m < ny, c = re/im),
    // the upper half from conjugate symmetry (re kept, im negated).
    void closure_full(std::vector<double> &alpha);
};

}  // namespace hzc
