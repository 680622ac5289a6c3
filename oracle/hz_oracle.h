/*
 * hz_oracle.h -- CPU restatement of the DirectPseudoSpectral 3-D Helmholtz
 * solver of gpritam/MPI (HelmholtzSolvers/DirectPseudoSpectral).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle and the
 * CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library (mpi_amd/) never links
 * or calls it.
 *
 * Every function restates the reference algorithm in its original operation
 * order so that, compiled with -ffp-contract=off on x86-64, results are
 * bit-identical to the reference.  Citations are relative to
 * HelmholtzSolvers/DirectPseudoSpectral/.
 *
 * Matrices are row-major, flat.  Error returns: 0 = ok, nonzero = the
 * reference's ErrorMessage condition (see hzo_error_string()).
 */
#ifndef HZ_ORACLE_H
#define HZ_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    HZO_OK = 0,
    HZO_ERR_SINGULAR = 1,      /* LU_Decomposition.cpp:49-50 */
    HZO_ERR_NOT_POW2 = 2,      /* Fourier_Quadrature.cpp:1106-1114 */
    HZO_ERR_LEGENDRE_ARG = 3,  /* GLLQuadrature.cpp:11-12 */
    HZO_ERR_N_TOO_SMALL = 4,   /* GLLQuadrature.cpp:82-83 */
    HZO_ERR_ALLOC = 5,         /* General.h:311-312 */
    HZO_ERR_ARG = 6
};

const char *hzo_error_string(int code);

/* GLLQuadrature.cpp:7-40 */
double hzo_legendre(double x, int n, int *err);
/* GLLQuadrature.cpp:78-128 with SymmetricEigenSystems.cpp:313-416 */
int hzo_gll_points(double *x, int N);
/* GLLQuadrature.cpp:135-151 */
int hzo_gll_weights(double *w, const double *x, int N);
/* GLLQuadrature.cpp:198-213 */
int hzo_first_derivative_matrix(double *D, const double *x, int N);
/* GLLQuadrature.cpp:250-262 */
void hzo_second_derivative_matrix(double *Dxx, const double *D, int N);
/* NonPeriodic3D_MPI.cpp:103-131: z grid and scaled Dz, Dzz for one rank.
 * z has Ns*(N-1)+1 entries (rank-local). */
int hzo_plan_operators(int N, int Ns, int size, int rank, double Lz,
                       double *z, double *Dz, double *Dzz);

/* LU_Decomposition.cpp:7-72 (in place, LINPACK storage). */
int hzo_lu_decompose(double *A, int *npivot, double *determinant, int n);
/* LU_Decomposition.cpp:78-108 */
void hzo_lu_solve(const double *A, const int *npivot, double *b, int n);

/* Fourier_Quadrature.cpp:1097-1155: in-place 2-D transform of every z-plane
 * of A, laid out [2Nx][2Ny][Nz] (Nz fastest).  isign = +1 forward, -1 inverse.
 * With Nz = 1 this is also the 2-D version (:1040-1091). */
int hzo_fft_xy(double *A, int isign, int Nx, int Ny, int Nz);

/* The reference operator callback (TestNonPeriodic3D_MPI.cpp:71-79): the
 * diagonal shift lambda + 4 pi^2 iw^2/Lx^2 + 4 pi^2 jw^2/Ly^2 evaluated in
 * the same operation order. */
double hzo_shift(double lambda, int iw, int jw, double Lx, double Ly);

typedef struct {
    int Nx, Ny;        /* Fourier points (powers of two) */
    int Nz;            /* requested z points (reference meaning) */
    int N;             /* GLL points per element */
    int size;          /* number of simulated MPI ranks */
    double xl, xr, yl, yr, zl, zr;
    double alphal, betal, alphar, betar, lambda;
} hzo_desc;

/* Number of elements per rank and total z points, NonPeriodic3D_MPI.cpp:70 */
int hzo_elements_per_rank(const hzo_desc *d);
int hzo_total_z_points(const hzo_desc *d);
/* Global z array (ranks concatenated, interfaces once): Ktot entries. */
int hzo_z_grid(const hzo_desc *d, double *z);

/* Full SolvePDE (NonPeriodic3D_MPI.cpp:237-591) with `size` ranks simulated
 * in one process, the MPI point-to-point carries replayed in rank order.
 * f  : [Nx][Ny][Ktot] physical right-hand side (k fastest)
 * gl : [Nx][Ny] left face data, gr : [Nx][Ny] right face data
 * u  : [Nx][Ny][Ktot] output.  Ktot = size*Ns*(N-1)+1. */
int hzo_solve(const hzo_desc *d, const double *f, const double *gl,
              const double *gr, double *u);

/* FirstDerivativeZ / SecondDerivativeZ, NonPeriodic3D_MPI.cpp:632-683, on
 * the reference's element arrays: u, out [size*Ns][Nx][Ny][N] (element e of
 * rank r at e = r*Ns + kk, interface nodes stored in both neighbours),
 * out = Dz u (order 1) or Dzz u (order 2) per element column, summed over
 * p = 0..N-1 in order. */
int hzo_derivative_z(const hzo_desc *d, int order, const double *u, double *out);
/* ConstructFundamentalSolutions, NonPeriodic3D_MPI.cpp:184-232, for one
 * column: U[kk][0..N) (kk < Ns) = 0 at the ends and the interior solve of
 * rhs[kk][1..N-2] with the factored operator (LU, npivot from
 * hzo_lu_decompose, order N-2); ul, ur (N each, may be NULL) = the
 * homogeneous solutions with -Dzz[1..N-2][0] and -Dzz[1..N-2][N-1]. */
int hzo_fundamental_solutions(int N, int Ns, const double *LU, const int *npivot, const double *Dzz,
                              const double *rhs, double *U, double *ul, double *ur);
/* Periodic3D_MPI.cpp SolvePDE at size = 1: u[Nx][Ny][Ktot] from f (z periodic). */
int hzo_solve_periodic(const hzo_desc *d, const double *f, double *u);
/* The same, also writing alphaPDE[Nx][2 Ny] (the closure coefficient per
 * spectral column, Periodic3D_MPI.cpp:698) when there is more than one element. */
int hzo_solve_periodic_alpha(const hzo_desc *d, const double *f, double *u, double *alpha);

#ifdef __cplusplus
}
#endif
#endif
