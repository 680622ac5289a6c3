/*
 * hz_oracle.c -- CPU restatement of gpritam/MPI's DirectPseudoSpectral 3-D
 * Helmholtz solver.  TEST INFRASTRUCTURE ONLY (correctness oracle and CPU
 * baseline); the product never links it.  See hz_oracle.h.
 *
 * Build with -O2 -ffp-contract=off (no FMA contraction) so that every
 * floating-point operation rounds exactly as the reference's plain IEEE
 * double code on x86-64.  Citations are relative to
 * HelmholtzSolvers/DirectPseudoSpectral/.
 */
#include "hz_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define HZO_PI 3.1415926535897932      /* include/General.h:14 */
#define HZO_EPSILON 1.000000000000E-12 /* include/General.h:15 */

const char *hzo_error_string(int code)
{
    switch (code) {
    case HZO_OK: return "ok";
    case HZO_ERR_SINGULAR: return "ERROR: Matrix is singular!";
    case HZO_ERR_NOT_POW2: return "N has to be 2^n for some integer n.";
    case HZO_ERR_LEGENDRE_ARG:
        return "Absolute value of the argument of the Legendre polynomial cannot be more than 1!";
    case HZO_ERR_N_TOO_SMALL: return "Number of points has to be more than or equal to 2.";
    case HZO_ERR_ALLOC: return "Error: Memory can not be allocated!";
    default: return "invalid argument";
    }
}

static double absd(double m) { return m >= 0.0 ? m : -m; }  /* General.h Absolute */

/* General.h Swap: the reference exchanges two values arithmetically
 * (a = a+b; b = a-b; a = a-b), which rounds when the magnitudes differ.
 * The rounding is part of the reference's results, so it is kept. */
static void swap_arith(double *a, double *b)
{
    *a = *a + *b;
    *b = *a - *b;
    *a = *a - *b;
}

/* General.cpp:87-91 */
static double pythag(double a, double b)
{
    return (absd(a) > absd(b) ? absd(a) * sqrt(1.0 + b * b / (a * a))
                              : (absd(b) == 0.0 ? 0.0 : absd(b) * sqrt(1.0 + a * a / (b * b))));
}

/* General.h:1099-1103 */
static double sign2(double a, double b) { return (b >= 0.0 ? absd(a) : -absd(a)); }

/* GLLQuadrature.cpp:7-40 */
double hzo_legendre(double x, int n, int *err)
{
    if (absd(x) > (1.0 + HZO_EPSILON)) {
        if (err) *err = HZO_ERR_LEGENDRE_ARG;
        return 0.0;
    }
    if (x > 1.0) x = 1.0;
    if (x < -1.0) x = -1.0;
    if (n < 0) {
        if (err) *err = HZO_ERR_ARG;
        return 0.0;
    }
    if (n == 0) return 1.0;
    if (n == 1) return x;
    double P0 = 1.0, P1 = x, Pn = 0.0;
    for (int i = 2; i <= n; ++i) {
        Pn = ((2.0 * i - 1.0) * x * P1 - (i - 1.0) * P0) / i;
        P0 = P1;
        P1 = Pn;
    }
    return Pn;
}

/* SymmetricEigenSystems.cpp:313-416: implicit-shift QL on a symmetric
 * tridiagonal matrix (eigenvalues only: the eigenvector rotations of the
 * reference do not feed back into the eigenvalues), then the descending
 * selection sort of :398-415. */
static void ql_tridiagonal(double *Ad, double *As, int N)
{
    int i, l, m;
    double s, r, p, g, f, d, c, b;
    for (l = 0; l < N; ++l) {
        int iteration = 0;
        (void)iteration;
        do {
            for (m = l; m < (N - 1); ++m) {
                d = absd(Ad[m]) + absd(Ad[m + 1]);
                if ((double)(absd(As[m]) + d) == d) break;
            }
            if (m != l) {
                iteration++;
                g = 0.5 * (Ad[l + 1] - Ad[l]) / As[l];
                r = pythag(g, 1.0);
                g = Ad[m] - Ad[l] + As[l] / (g + sign2(r, g));
                s = c = 1.0;
                p = 0.0;
                for (i = (m - 1); i >= l; i--) {
                    f = s * As[i];
                    b = c * As[i];
                    As[i + 1] = (r = pythag(f, g));
                    if (r == 0.0) {
                        Ad[i + 1] -= p;
                        As[m] = 0.0;
                        break;
                    }
                    s = (f / r);
                    c = (g / r);
                    g = (Ad[i + 1] - p);
                    r = (Ad[i] - g) * s + 2.0 * c * b;
                    Ad[i + 1] = g + (p = s * r);
                    g = (c * r - b);
                }
                if ((r == 0.0) && (i >= l)) continue;
                Ad[l] -= p;
                As[l] = g;
                As[m] = 0.0;
            }
        } while (m != l);
    }
    int k;
    for (i = 0; i < (N - 1); ++i) {
        p = Ad[k = i];
        for (int j = i; j < N; ++j)
            if (Ad[j] >= p) p = Ad[k = j];
        if (k != i) {
            Ad[k] = Ad[i];
            Ad[i] = p;
        }
    }
}

/* GLLQuadrature.cpp:78-128: GLL nodes as eigenvalues of the
 * Lobatto-modified Jacobi matrix, returned in ascending order. */
int hzo_gll_points(double *x, int N)
{
    if (N < 2) return HZO_ERR_N_TOO_SMALL;
    if (N == 2) {
        x[0] = -1.0;
        x[1] = 1.0;
        return HZO_OK;
    }
    double *As = (double *)calloc((size_t)N, sizeof(double));
    if (!As) return HZO_ERR_ALLOC;
    for (int i = 0; i < (N - 1); ++i) x[i] = 0.0;
    for (int i = 0; i < (N - 2); ++i)
        As[i] = sqrt((i + 1.0) * (i + 1.0) / (4.0 * (i + 1.0) * (i + 1.0) - 1.0));
    double sign = (N % 2 == 0 ? 1.0 : -1.0);
    double a = -sign * (N - 1.0) / (2.0 * N - 3.0);
    double b = sign;
    double c = (N - 1.0) / (2.0 * N - 3.0);
    double d = 1.0;
    double e = sign * (N - 1.0) / (2.0 * N - 3.0);
    double f = (N - 1.0) / (2.0 * N - 3.0);
    double alpha = (d * e - b * f) / (a * d - b * c);
    double beta = (a * f - e * c) / (a * d - b * c);
    x[N - 1] = sqrt(alpha);
    As[N - 2] = sqrt(beta);
    ql_tridiagonal(x, As, N);
    int M = (N % 2 == 0 ? N / 2 : (N - 1) / 2);
    for (int i = 0; i < M; ++i) swap_arith(&x[i], &x[N - i - 1]);
    free(As);
    return HZO_OK;
}

/* GLLQuadrature.cpp:135-151 */
int hzo_gll_weights(double *w, const double *x, int N)
{
    int points = (N % 2 == 0 ? N / 2 : (N + 1) / 2);
    double numerator = 2.0 / ((N - 1.0) * N), Pn;
    int err = 0;
    for (int i = 0; i < points; ++i) {
        Pn = hzo_legendre(x[i], N - 1, &err);
        w[i] = numerator / (Pn * Pn);
        w[N - i - 1] = w[i];
    }
    return err;
}

/* GLLQuadrature.cpp:198-213 */
int hzo_first_derivative_matrix(double *D, const double *x, int N)
{
    int err = 0;
    D[0] = -0.25 * N * (N - 1);
    D[(size_t)(N - 1) * N + (N - 1)] = 0.25 * N * (N - 1);
    for (int i = 1; i < (N - 1); ++i) D[(size_t)i * N + i] = 0.0;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            if (i != j)
                D[(size_t)i * N + j] = hzo_legendre(x[i], N - 1, &err) /
                                       (hzo_legendre(x[j], N - 1, &err) * (x[i] - x[j]));
    return err;
}

/* GLLQuadrature.cpp:250-262 (naive i-j-k triple loop, k innermost) */
void hzo_second_derivative_matrix(double *Dxx, const double *D, int N)
{
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) Dxx[(size_t)i * N + j] = 0.0;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k)
                Dxx[(size_t)i * N + j] += D[(size_t)i * N + k] * D[(size_t)k * N + j];
}

/* NonPeriodic3D_MPI.cpp:103-131 */
int hzo_plan_operators(int N, int Ns, int size, int rank, double Lz,
                       double *z, double *Dz, double *Dzz)
{
    double Lzp = Lz / size;
    double *zlocal = (double *)calloc((size_t)N, sizeof(double));
    if (!zlocal) return HZO_ERR_ALLOC;
    int err = hzo_gll_points(zlocal, N);
    if (err) { free(zlocal); return err; }
    if (z) {
        for (int kk = 0; kk < Ns; ++kk) {
            /* The element bounds shadow the global zl and start at 0
             * (NonPeriodic3D_MPI.cpp:113-116); kept as in the reference. */
            double zle = rank * Lzp + kk * Lzp / Ns;
            double zre = zle + Lzp / Ns;
            for (int k = 0; k < N; ++k)
                z[kk * (N - 1) + k] = 0.5 * (zle + zre + (Lzp / Ns) * zlocal[k]);
        }
    }
    err = hzo_first_derivative_matrix(Dz, zlocal, N);
    if (err) { free(zlocal); return err; }
    hzo_second_derivative_matrix(Dzz, Dz, N);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) Dz[(size_t)i * N + j] *= ((2.0 * Ns) / Lzp);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) Dzz[(size_t)i * N + j] *= ((4.0 * Ns * Ns) / (Lzp * Lzp));
    free(zlocal);
    return HZO_OK;
}

/* LU_Decomposition.cpp:7-72: scaled partial pivoting.  The row scale s[p]
 * is computed once from the original rows and is indexed by position (it
 * is not swapped with the rows); swaps touch columns >= k only; the last
 * pivot is not tested. */
int hzo_lu_decompose(double *A, int *npivot, double *determinant, int n)
{
    double *s = (double *)calloc((size_t)n, sizeof(double));
    double *c = (double *)calloc((size_t)n, sizeof(double));
    if (!s || !c) { free(s); free(c); return HZO_ERR_ALLOC; }
    int status = HZO_OK;
    *determinant = 1.0;
#define A_(i, j) A[(size_t)(i) * n + (j)]
    for (int p = 0; p < n; ++p) {
        s[p] = absd(A_(p, 0));
        for (int q = 1; q < n; ++q)
            if (absd(A_(p, q)) > s[p]) s[p] = absd(A_(p, q));
    }
    for (int k = 0; k < (n - 1); ++k) {
        c[k] = absd(A_(k, k)) / s[k];
        int index = k;
        for (int i = k + 1; i < n; ++i) {
            if ((absd(A_(i, k)) / s[i]) > c[k]) {
                index = i;
                c[k] = (absd(A_(i, k)) / s[i]);
            }
        }
        npivot[k] = index;
        if (absd(c[k]) <= HZO_EPSILON) { status = HZO_ERR_SINGULAR; break; }
        if (index != k) {
            *determinant = -*determinant;
            for (int j = k; j < n; ++j) swap_arith(&A_(k, j), &A_(index, j));
        }
        for (int l = k + 1; l < n; ++l) {
            A_(l, k) = A_(l, k) / A_(k, k);
            for (int m = k + 1; m < n; ++m) A_(l, m) -= A_(l, k) * A_(k, m);
        }
        *determinant = *determinant * A_(k, k);
    }
    if (status == HZO_OK) *determinant = *determinant * A_(n - 1, n - 1);
#undef A_
    free(s);
    free(c);
    return status;
}

/* LU_Decomposition.cpp:78-108 */
void hzo_lu_solve(const double *A, const int *npivot, double *b, int n)
{
    for (int k = 0; k < (n - 1); ++k) {
        int index = npivot[k];
        if (index != k) swap_arith(&b[index], &b[k]);
        for (int i = k + 1; i < n; ++i) b[i] -= A[(size_t)i * n + k] * b[k];
    }
    b[n - 1] /= A[(size_t)(n - 1) * n + (n - 1)];
    for (int p = n - 2; p >= 0; p--) {
        double temp = 0;
        for (int q = p + 1; q < n; ++q) temp += A[(size_t)p * n + q] * b[q];
        b[p] = (b[p] - temp) / A[(size_t)p * n + p];
    }
}

/* One in-place radix-2 complex FFT of nn points stored interleaved
 * (re, im) at base[t*stride], t = 0..2nn-1: bit reversal then the
 * Danielson-Lanczos passes with the trigonometric recurrence of
 * Fourier_Quadrature.cpp:756-805 (and :862-910 for y). */
static void fft_strided(double *base, size_t stride, int nn, int isign)
{
    int mmax = 2, m, jj = 1, istep;
    double wtemp, wr, wi, wpr, wpi, tempr, tempi, theta;
#define E_(t) base[(size_t)(t) * stride]
    for (int ii = 1; ii < 2 * nn; ii += 2) {
        if (jj > ii) {
            swap_arith(&E_(jj - 1), &E_(ii - 1));
            swap_arith(&E_(jj), &E_(ii));
        }
        m = nn;
        while (m >= 2 && jj > m) {
            jj -= m;
            m /= 2;
        }
        jj += m;
    }
    while (2 * nn > mmax) {
        istep = mmax * 2;
        theta = isign * (-(2.0 * HZO_PI) / mmax);
        wtemp = sin(0.5 * theta);
        wpr = -2.0 * wtemp * wtemp;
        wpi = sin(theta);
        wr = 1.0;
        wi = 0.0;
        for (int mm = 1; mm < mmax; mm += 2) {
            for (int ii = mm; ii <= 2 * nn; ii += istep) {
                jj = ii + mmax;
                tempr = wr * E_(jj - 1) - wi * E_(jj);
                tempi = wr * E_(jj) + wi * E_(jj - 1);
                E_(jj - 1) = E_(ii - 1) - tempr;
                E_(jj) = E_(ii) - tempi;
                E_(ii - 1) += tempr;
                E_(ii) += tempi;
            }
            wtemp = wr;
            wr += (wr * wpr - wi * wpi);
            wi += (wi * wpr + wtemp * wpi);
        }
        mmax = istep;
    }
#undef E_
}

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* Fourier_Quadrature.cpp:719-818 (3-D FourierTransformX) on A[2Nx][2Ny][Nz],
 * transforming rows j < Ny. */
static void fourier_x(double *A, int isign, int Nx, int Ny, int Nz, int real, double *tmp)
{
    const size_t sx = (size_t)2 * Ny * Nz; /* stride of the first index */
#define A3(i, j, k) A[(size_t)(i) * sx + (size_t)(j) * Nz + (k)]
    if (isign == 1) {
        if (real)
            for (int i = 0; i < Nx; ++i)
                for (int j = 0; j < Ny; ++j)
                    for (int k = 0; k < Nz; ++k) A3(Nx + i, j, k) = 0.0;
        /* RealComplexX(...,1), :318-334 */
        for (int k = 0; k < Nz; ++k)
            for (int j = 0; j < Ny; ++j) {
                for (int i = 0; i < Nx; ++i) {
                    tmp[2 * i] = A3(i, j, k);
                    tmp[2 * i + 1] = A3(Nx + i, j, k);
                }
                for (int i = 0; i < 2 * Nx; ++i) A3(i, j, k) = tmp[i];
            }
    }
    for (int j = 0; j < Ny; ++j)
        for (int k = 0; k < Nz; ++k) fft_strided(&A3(0, j, k), sx, Nx, isign);
    if (isign == -1) {
        for (int i = 0; i < 2 * Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                for (int k = 0; k < Nz; ++k) A3(i, j, k) /= Nx;
        /* RealComplexX(...,0), :301-316 */
        for (int k = 0; k < Nz; ++k)
            for (int j = 0; j < Ny; ++j) {
                for (int i = 0; i < Nx; ++i) {
                    tmp[i] = A3(2 * i, j, k);
                    tmp[Nx + i] = A3(2 * i + 1, j, k);
                }
                for (int i = 0; i < 2 * Nx; ++i) A3(i, j, k) = tmp[i];
            }
    }
#undef A3
}

/* Fourier_Quadrature.cpp:827-926 (3-D FourierTransformY) on A[..][2Ny][Nz],
 * transforming rows i < Nx. */
static void fourier_y(double *A, int isign, int Nx, int Ny, int Nz, int real, double *tmp)
{
    const size_t sx = (size_t)2 * Ny * Nz;
#define A3(i, j, k) A[(size_t)(i) * sx + (size_t)(j) * Nz + (k)]
    if (isign == 1) {
        if (real)
            for (int i = 0; i < Nx; ++i)
                for (int j = 0; j < Ny; ++j)
                    for (int k = 0; k < Nz; ++k) A3(i, Ny + j, k) = 0.0;
        /* RealComplexY(...,1), :343-358 */
        for (int k = 0; k < Nz; ++k)
            for (int i = 0; i < Nx; ++i) {
                for (int j = 0; j < Ny; ++j) {
                    tmp[2 * j] = A3(i, j, k);
                    tmp[2 * j + 1] = A3(i, Ny + j, k);
                }
                for (int j = 0; j < 2 * Ny; ++j) A3(i, j, k) = tmp[j];
            }
    }
    for (int i = 0; i < Nx; ++i)
        for (int k = 0; k < Nz; ++k) fft_strided(&A3(i, 0, k), (size_t)Nz, Ny, isign);
    if (isign == -1) {
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < 2 * Ny; ++j)
                for (int k = 0; k < Nz; ++k) A3(i, j, k) /= Ny;
        /* RealComplexY(...,0), :326-341 */
        for (int k = 0; k < Nz; ++k)
            for (int i = 0; i < Nx; ++i) {
                for (int j = 0; j < Ny; ++j) {
                    tmp[j] = A3(i, 2 * j, k);
                    tmp[Ny + j] = A3(i, 2 * j + 1, k);
                }
                for (int j = 0; j < 2 * Ny; ++j) A3(i, j, k) = tmp[j];
            }
    }
#undef A3
}

/* Fourier_Quadrature.cpp:1097-1155 */
int hzo_fft_xy(double *A, int isign, int Nx, int Ny, int Nz)
{
    if (!is_pow2(Nx) || !is_pow2(Ny)) return HZO_ERR_NOT_POW2;
    int L = Nx > Ny ? Nx : Ny;
    double *tmp = (double *)malloc(sizeof(double) * 2 * (size_t)L);
    if (!tmp) return HZO_ERR_ALLOC;
    const size_t sx = (size_t)2 * Ny * Nz;
#define A3(i, j, k) A[(size_t)(i) * sx + (size_t)(j) * Nz + (k)]
    if (isign == 1) {
        fourier_x(A, 1, Nx, Ny, Nz, 1, tmp);
        for (int k = 0; k < Nz; ++k)
            for (int i = 0; i < Nx; ++i)
                for (int j = 0; j < Ny; ++j) {
                    A3(i, j, k) = A3(2 * i, j, k);
                    A3(i, Ny + j, k) = A3(2 * i + 1, j, k);
                }
        fourier_y(A, 1, Nx, Ny, Nz, 0, tmp);
    } else {
        fourier_y(A, -1, Nx, Ny, Nz, 0, tmp);
        for (int i = Nx - 1; i >= 0; i--)
            for (int k = 0; k < Nz; ++k)
                for (int j = 0; j < Ny; ++j) {
                    A3(2 * i, j, k) = A3(i, j, k);
                    A3(2 * i + 1, j, k) = A3(i, Ny + j, k);
                }
        fourier_x(A, -1, Nx, Ny, Nz, 1, tmp);
    }
#undef A3
    free(tmp);
    return HZO_OK;
}

/* TestNonPeriodic3D_MPI.cpp:76 */
double hzo_shift(double lambda, int iw, int jw, double Lx, double Ly)
{
    return (lambda + 4.0 * HZO_PI * HZO_PI * iw * iw / (Lx * Lx) +
            4.0 * HZO_PI * HZO_PI * jw * jw / (Ly * Ly));
}

int hzo_elements_per_rank(const hzo_desc *d)
{
    return (int)ceil(((double)(d->Nz)) / (d->size * d->N)); /* NonPeriodic3D_MPI.cpp:70 */
}

int hzo_total_z_points(const hzo_desc *d)
{
    return d->size * hzo_elements_per_rank(d) * (d->N - 1) + 1;
}

int hzo_z_grid(const hzo_desc *d, double *z)
{
    int Ns = hzo_elements_per_rank(d), N = d->N;
    double *zr = (double *)malloc(sizeof(double) * (size_t)(Ns * (N - 1) + 1));
    double *Dz = (double *)malloc(sizeof(double) * (size_t)N * N);
    double *Dzz = (double *)malloc(sizeof(double) * (size_t)N * N);
    int err = (!zr || !Dz || !Dzz) ? HZO_ERR_ALLOC : HZO_OK;
    for (int r = 0; r < d->size && !err; ++r) {
        err = hzo_plan_operators(N, Ns, d->size, r, d->zr - d->zl, zr, Dz, Dzz);
        for (int k = 0; k < Ns * (N - 1) + 1 && !err; ++k) z[r * Ns * (N - 1) + k] = zr[k];
    }
    free(zr);
    free(Dz);
    free(Dzz);
    return err;
}

/* NonPeriodic3D_MPI.cpp:175-179 -> GLLQuadrature.cpp:234-245 */
static double first_derivative(int k, const double *h, const double *Dz, int N)
{
    double sum = 0.0;
    for (int i = 0; i < N; ++i) sum += Dz[(size_t)k * N + i] * h[i];
    return sum;
}

/* Per-rank state of the reference solver (NonPeriodic3D_MPI.cpp:14-35). */
typedef struct {
    double *ul, *ur;                      /* [Ns][Nx][NY][N] */
    double *zetal, *zetar, *b, *c, *rhs;  /* [Ns][Nx][NY] */
    double *br, *rhsr;                    /* [Nx][NY] */
    double *rhsPDE;                       /* [Ns][2Nx][NY][N] */
    double *U;                            /* [Ns][2Nx][NY][N] */
    double *leftBC, *rightBC;             /* [2Nx][NY] */
    double *Dz, *Dzz, *Op;
    int *npivot;
} rank_state;

static void free_rank(rank_state *s)
{
    free(s->ul); free(s->ur); free(s->zetal); free(s->zetar); free(s->b);
    free(s->c); free(s->rhs); free(s->br); free(s->rhsr); free(s->rhsPDE);
    free(s->U); free(s->leftBC); free(s->rightBC); free(s->Dz); free(s->Dzz);
    free(s->Op); free(s->npivot);
}

/* NonPeriodic3D_MPI.cpp:184-232, with U/S element arrays of one rank. */
static void fundamental_solutions(rank_state *s, int i, int j, int Nx, int NY, int N, int Ns)
{
    const int n = N - 2;
    const size_t col = ((size_t)i * NY + j) * N;          /* offset in [2Nx][NY][N] */
    const size_t el = (size_t)2 * Nx * NY * N;             /* element stride of U/rhsPDE */
    const size_t sel = (size_t)Nx * NY * N;                /* element stride of ul/ur */
    for (int kk = 0; kk < Ns; ++kk) {
        double *u = s->U + kk * el + col;
        const double *r = s->rhsPDE + kk * el + col;
        u[0] = 0.0;
        u[N - 1] = 0.0;
        for (int k = 0; k < n; ++k) u[1 + k] = r[k + 1];
        hzo_lu_solve(s->Op, s->npivot, u + 1, n);
    }
    double *ul = s->ul + col, *ur = s->ur + col;
    ul[0] = 1.0; ul[N - 1] = 0.0;
    ur[0] = 0.0; ur[N - 1] = 1.0;
    for (int k = 0; k < n; ++k) ul[1 + k] = -s->Dzz[(size_t)(k + 1) * N + 0];
    hzo_lu_solve(s->Op, s->npivot, ul + 1, n);
    for (int k = 0; k < n; ++k) ur[1 + k] = -s->Dzz[(size_t)(k + 1) * N + (N - 1)];
    hzo_lu_solve(s->Op, s->npivot, ur + 1, n);
    for (int kk = 1; kk < Ns; ++kk)
        for (int k = 0; k < N; ++k) {
            s->ul[kk * sel + col + k] = ul[k];
            s->ur[kk * sel + col + k] = ur[k];
        }
}

int hzo_solve(const hzo_desc *d, const double *f, const double *gl,
              const double *gr, double *u)
{
    const int Nx = d->Nx, Ny = d->Ny, N = d->N, P = d->size;
    if (!is_pow2(Nx) || !is_pow2(Ny)) return HZO_ERR_NOT_POW2;
    if (N < 3 || P < 1) return HZO_ERR_ARG;
    const int NY = 2 * Ny, n = N - 2;
    const int Ns = hzo_elements_per_rank(d);
    const int Ktot = P * Ns * (N - 1) + 1;
    const double Lx = d->xr - d->xl, Ly = d->yr - d->yl, Lz = d->zr - d->zl;
    const size_t el = (size_t)2 * Nx * NY * N, sel = (size_t)Nx * NY * N, pl = (size_t)Nx * NY;
    int status = HZO_OK;

    rank_state *R = (rank_state *)calloc((size_t)P, sizeof(rank_state));
    if (!R) return HZO_ERR_ALLOC;
    for (int r = 0; r < P; ++r) {
        rank_state *s = &R[r];
        s->ul = calloc(Ns * sel, 8); s->ur = calloc(Ns * sel, 8);
        s->zetal = calloc(Ns * pl, 8); s->zetar = calloc(Ns * pl, 8);
        s->b = calloc(Ns * pl, 8); s->c = calloc(Ns * pl, 8); s->rhs = calloc(Ns * pl, 8);
        s->br = calloc(pl, 8); s->rhsr = calloc(pl, 8);
        s->rhsPDE = calloc(Ns * el, 8); s->U = calloc(Ns * el, 8);
        s->leftBC = calloc(2 * pl, 8); s->rightBC = calloc(2 * pl, 8);
        s->Dz = calloc((size_t)N * N, 8); s->Dzz = calloc((size_t)N * N, 8);
        s->Op = calloc((size_t)n * n, 8); s->npivot = calloc((size_t)n, sizeof(int));
        if (!s->ul || !s->ur || !s->zetal || !s->zetar || !s->b || !s->c || !s->rhs ||
            !s->br || !s->rhsr || !s->rhsPDE || !s->U || !s->leftBC || !s->rightBC ||
            !s->Dz || !s->Dzz || !s->Op || !s->npivot) { status = HZO_ERR_ALLOC; goto done; }
        status = hzo_plan_operators(N, Ns, P, r, Lz, NULL, s->Dz, s->Dzz);
        if (status) goto done;
        /* Inputs in the reference layout (TestNonPeriodic3D_MPI.cpp:116-128). */
        for (int kk = 0; kk < Ns; ++kk)
            for (int i = 0; i < Nx; ++i)
                for (int j = 0; j < Ny; ++j)
                    for (int k = 0; k < N; ++k)
                        s->rhsPDE[kk * el + ((size_t)i * NY + j) * N + k] =
                            f[((size_t)i * Ny + j) * Ktot + (size_t)(r * Ns + kk) * (N - 1) + k];
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j) {
                s->leftBC[(size_t)i * NY + j] = gl[(size_t)i * Ny + j];
                s->rightBC[(size_t)i * NY + j] = gr[(size_t)i * Ny + j];
            }
    }

    for (int r = 0; r < P; ++r) {
        rank_state *s = &R[r];
        /* Step 1, :241-245 */
        for (int kk = 0; kk < Ns; ++kk) hzo_fft_xy(s->rhsPDE + kk * el, 1, Nx, Ny, N);
        hzo_fft_xy(s->leftBC, 1, Nx, Ny, 1);
        hzo_fft_xy(s->rightBC, 1, Nx, Ny, 1);
        /* Step 2, :248-325 */
        for (int iw = 0; iw <= Nx / 2 && !status; iw++)
            for (int jw = 0; jw <= Ny / 2 && !status; jw++) {
                double shift = hzo_shift(d->lambda, iw, jw, Lx, Ly);
                for (int a = 0; a < n; ++a)
                    for (int b = 0; b < n; ++b)
                        s->Op[(size_t)a * n + b] =
                            s->Dzz[(size_t)(a + 1) * N + (b + 1)] - (a == b ? shift : 0.0);
                double det;
                status = hzo_lu_decompose(s->Op, s->npivot, &det, n);
                if (status) break;
                int is[2], js[4], ni, nj;
                if (iw == 0 || iw == Nx / 2) { is[0] = iw; ni = 1; }
                else { is[0] = iw; is[1] = Nx - iw; ni = 2; }
                if (jw == 0) { js[0] = 0; js[1] = 1; nj = 2; }
                else if (jw == Ny / 2) { js[0] = Ny; js[1] = Ny + 1; nj = 2; }
                else { js[0] = 2 * jw; js[1] = 2 * jw + 1; js[2] = 2 * (Ny - jw); js[3] = 2 * (Ny - jw) + 1; nj = 4; }
                for (int a = 0; a < ni; ++a)
                    for (int b = 0; b < nj; ++b) fundamental_solutions(s, is[a], js[b], Nx, NY, N, Ns);
            }
        if (status) goto done;
    }

    /* Step 3, forward sweep :327-517, ranks in pipeline order. */
    for (int r = 0; r < P; ++r) {
        rank_state *s = &R[r];
#define PL(arr, kk, i, j) arr[(size_t)(kk) * pl + (size_t)(i) * NY + (j)]
#define COLU(kk, i, j) (s->U + (size_t)(kk) * el + ((size_t)(i) * NY + (j)) * N)
#define COLL(arr, kk, i, j) (arr + (size_t)(kk) * sel + ((size_t)(i) * NY + (j)) * N)
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < NY; ++j) {
                if (r == 0) {
                    PL(s->b, 0, i, j) = -d->alphal * first_derivative(0, COLL(s->ul, 0, i, j), s->Dz, N) + d->betal;
                    PL(s->c, 0, i, j) = -d->alphal * first_derivative(0, COLL(s->ur, 0, i, j), s->Dz, N);
                    PL(s->rhs, 0, i, j) = s->leftBC[(size_t)i * NY + j] + d->alphal * first_derivative(0, COLU(0, i, j), s->Dz, N);
                } else {
                    /* carries received from rank r-1 (MPI_Recv :388-389) */
                    PL(s->b, 0, i, j) = -first_derivative(0, COLL(s->ul, 0, i, j), s->Dz, N) + PL(s->zetal, 0, i, j);
                    PL(s->c, 0, i, j) = -first_derivative(0, COLL(s->ur, 0, i, j), s->Dz, N);
                    PL(s->rhs, 0, i, j) = first_derivative(0, COLU(0, i, j), s->Dz, N) + PL(s->zetar, 0, i, j);
                }
            }
        for (int kk = 1; kk < Ns; ++kk) {
            for (int i = 0; i < Nx; ++i)
                for (int j = 0; j < NY; ++j)
                    s->leftBC[(size_t)i * NY + j] = first_derivative(N - 1, COLL(s->ul, kk - 1, i, j), s->Dz, N);
            for (int i = 0; i < Nx; ++i)
                for (int j = 0; j < NY; ++j) {
                    double lb = s->leftBC[(size_t)i * NY + j];
                    PL(s->b, kk, i, j) = first_derivative(N - 1, COLL(s->ur, kk - 1, i, j), s->Dz, N) -
                                         first_derivative(0, COLL(s->ul, kk, i, j), s->Dz, N) -
                                         (PL(s->c, kk - 1, i, j) / PL(s->b, kk - 1, i, j)) * lb;
                    PL(s->c, kk, i, j) = -first_derivative(0, COLL(s->ur, kk, i, j), s->Dz, N);
                    PL(s->rhs, kk, i, j) = first_derivative(0, COLU(kk, i, j), s->Dz, N) -
                                           first_derivative(N - 1, COLU(kk - 1, i, j), s->Dz, N) -
                                           (PL(s->rhs, kk - 1, i, j) / PL(s->b, kk - 1, i, j)) * lb;
                }
        }
        const int e = Ns - 1;
        if (r == P - 1) {
            for (int i = 0; i < Nx; ++i)
                for (int j = 0; j < NY; ++j)
                    s->leftBC[(size_t)i * NY + j] = d->alphar * first_derivative(N - 1, COLL(s->ul, e, i, j), s->Dz, N);
            for (int i = 0; i < Nx; ++i)
                for (int j = 0; j < NY; ++j) {
                    double lb = s->leftBC[(size_t)i * NY + j];
                    s->br[(size_t)i * NY + j] = d->alphar * first_derivative(N - 1, COLL(s->ur, e, i, j), s->Dz, N) + d->betar -
                                                (PL(s->c, e, i, j) / PL(s->b, e, i, j)) * lb;
                    s->rhsr[(size_t)i * NY + j] = s->rightBC[(size_t)i * NY + j] -
                                                  d->alphar * first_derivative(N - 1, COLU(e, i, j), s->Dz, N) -
                                                  (PL(s->rhs, e, i, j) / PL(s->b, e, i, j)) * lb;
                }
        } else {
            rank_state *t = &R[r + 1];
            for (int i = 0; i < Nx; ++i)
                for (int j = 0; j < NY; ++j)
                    s->leftBC[(size_t)i * NY + j] = first_derivative(N - 1, COLL(s->ul, e, i, j), s->Dz, N);
            for (int i = 0; i < Nx; ++i)
                for (int j = 0; j < NY; ++j) {
                    double lb = s->leftBC[(size_t)i * NY + j];
                    PL(s->zetal, e, i, j) = first_derivative(N - 1, COLL(s->ur, e, i, j), s->Dz, N) -
                                            (PL(s->c, e, i, j) / PL(s->b, e, i, j)) * lb;
                    PL(s->zetar, e, i, j) = -first_derivative(N - 1, COLU(e, i, j), s->Dz, N) -
                                            (PL(s->rhs, e, i, j) / PL(s->b, e, i, j)) * lb;
                }
            /* MPI_Send :432-433 / :513-514 -> rank r+1's S[0].zetal/zetar */
            memcpy(t->zetal, s->zetal + (size_t)e * pl, pl * 8);
            memcpy(t->zetar, s->zetar + (size_t)e * pl, pl * 8);
        }
    }

    /* Step 4, backward sweep :519-580, last rank first. */
    for (int r = P - 1; r >= 0; --r) {
        rank_state *s = &R[r];
        for (int kk = Ns - 1; kk >= 0; --kk)
            for (int i = 0; i < Nx; ++i)
                for (int j = 0; j < NY; ++j) {
                    if (kk == Ns - 1)
                        PL(s->zetar, kk, i, j) = (r == P - 1) ? s->rhsr[(size_t)i * NY + j] / s->br[(size_t)i * NY + j]
                                                              : PL(R[r + 1].zetal, 0, i, j);
                    else
                        PL(s->zetar, kk, i, j) = PL(s->zetal, kk + 1, i, j);
                    PL(s->zetal, kk, i, j) = (PL(s->rhs, kk, i, j) - PL(s->c, kk, i, j) * PL(s->zetar, kk, i, j)) /
                                             PL(s->b, kk, i, j);
                }
    }

    /* Step 5 :582-587 and Step 6 :589-590 */
    for (int r = 0; r < P; ++r) {
        rank_state *s = &R[r];
        for (int kk = 0; kk < Ns; ++kk)
            for (int i = 0; i < Nx; ++i)
                for (int j = 0; j < NY; ++j) {
                    double *uc = COLU(kk, i, j);
                    const double *ulc = COLL(s->ul, kk, i, j), *urc = COLL(s->ur, kk, i, j);
                    double zl = PL(s->zetal, kk, i, j), zr = PL(s->zetar, kk, i, j);
                    for (int k = 0; k < N; ++k) uc[k] += zl * ulc[k] + zr * urc[k];
                }
        for (int kk = 0; kk < Ns; ++kk) hzo_fft_xy(s->U + kk * el, -1, Nx, Ny, N);
        for (int kk = 0; kk < Ns; ++kk)
            for (int i = 0; i < Nx; ++i)
                for (int j = 0; j < Ny; ++j)
                    for (int k = 0; k < N; ++k)
                        u[((size_t)i * Ny + j) * Ktot + (size_t)(r * Ns + kk) * (N - 1) + k] =
                            s->U[kk * el + ((size_t)i * NY + j) * N + k];
    }
#undef PL
#undef COLU
#undef COLL
done:
    for (int r = 0; r < P; ++r) free_rank(&R[r]);
    free(R);
    return status;
}

/* Periodic3D_MPI.cpp:233-863 for one rank (size = 1): the same transforms,
 * operators and fundamental solutions (:184-232), then the cyclic
 * element-interface system.  With Ns = 1 (:323-333, :655-664) the single
 * interface value is rhs0/b; with Ns > 1 a forward sweep carries two right-
 * hand sides, the particular one (rhs0) and the periodic coupling column
 * (rhs1) (:336-376), the backward sweep gives zeta and eta (:666-694), and the
 * closure alpha = -(zl + zr)/(1 + el + er) of the last element combines them
 * (:696-711).  Assembly and the inverse transform as in the non-periodic
 * solver (:854-862).  desc->size must be 1; the Robin fields are unused. */
static int solve_periodic(const hzo_desc *d, const double *f, double *u, double *alpha_out);

int hzo_solve_periodic(const hzo_desc *d, const double *f, double *u) { return solve_periodic(d, f, u, NULL); }

/* As hzo_solve_periodic, also returning the closure coefficients alphaPDE
 * (Periodic3D_MPI.cpp:698, [Nx][2 Ny] in the reference's spectral layout;
 * untouched with one element, where the reference computes none). */
int hzo_solve_periodic_alpha(const hzo_desc *d, const double *f, double *u, double *alpha)
{
    return solve_periodic(d, f, u, alpha);
}

static int solve_periodic(const hzo_desc *d, const double *f, double *u, double *alpha_out)
{
    const int Nx = d->Nx, Ny = d->Ny, N = d->N;
    if (!is_pow2(Nx) || !is_pow2(Ny)) return HZO_ERR_NOT_POW2;
    if (N < 3 || d->size != 1) return HZO_ERR_ARG;
    const int NY = 2 * Ny, n = N - 2;
    const int Ns = hzo_elements_per_rank(d);
    const int Ktot = Ns * (N - 1) + 1;
    const double Lx = d->xr - d->xl, Ly = d->yr - d->yl, Lz = d->zr - d->zl;
    const size_t el = (size_t)2 * Nx * NY * N, sel = (size_t)Nx * NY * N, pl = (size_t)Nx * NY;
    int status = HZO_OK;
    rank_state st;
    memset(&st, 0, sizeof(st));
    rank_state *s = &st;
    double *rhs1 = NULL, *etal = NULL, *etar = NULL;
    s->ul = calloc(Ns * sel, 8); s->ur = calloc(Ns * sel, 8);
    s->zetal = calloc(Ns * pl, 8); s->zetar = calloc(Ns * pl, 8);
    s->b = calloc(Ns * pl, 8); s->c = calloc(Ns * pl, 8); s->rhs = calloc(Ns * pl, 8);
    s->rhsPDE = calloc(Ns * el, 8); s->U = calloc(Ns * el, 8);
    s->Dz = calloc((size_t)N * N, 8); s->Dzz = calloc((size_t)N * N, 8);
    s->Op = calloc((size_t)n * n, 8); s->npivot = calloc((size_t)n, sizeof(int));
    rhs1 = calloc(Ns * pl, 8); etal = calloc(Ns * pl, 8); etar = calloc(Ns * pl, 8);
    if (!s->ul || !s->ur || !s->zetal || !s->zetar || !s->b || !s->c || !s->rhs || !s->rhsPDE || !s->U ||
        !s->Dz || !s->Dzz || !s->Op || !s->npivot || !rhs1 || !etal || !etar) { status = HZO_ERR_ALLOC; goto done; }
    status = hzo_plan_operators(N, Ns, 1, 0, Lz, NULL, s->Dz, s->Dzz);
    if (status) goto done;
    for (int kk = 0; kk < Ns; ++kk)
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                for (int k = 0; k < N; ++k)
                    s->rhsPDE[kk * el + ((size_t)i * NY + j) * N + k] =
                        f[((size_t)i * Ny + j) * Ktot + (size_t)kk * (N - 1) + k];
    /* Step 1 (:236-238) and Step 2 (:240-318) */
    for (int kk = 0; kk < Ns; ++kk) hzo_fft_xy(s->rhsPDE + kk * el, 1, Nx, Ny, N);
    for (int iw = 0; iw <= Nx / 2 && !status; iw++)
        for (int jw = 0; jw <= Ny / 2 && !status; jw++) {
            double shift = hzo_shift(d->lambda, iw, jw, Lx, Ly);
            for (int a = 0; a < n; ++a)
                for (int b = 0; b < n; ++b)
                    s->Op[(size_t)a * n + b] = s->Dzz[(size_t)(a + 1) * N + (b + 1)] - (a == b ? shift : 0.0);
            double det;
            status = hzo_lu_decompose(s->Op, s->npivot, &det, n);
            if (status) break;
            int is[2], js[4], ni, nj;
            if (iw == 0 || iw == Nx / 2) { is[0] = iw; ni = 1; }
            else { is[0] = iw; is[1] = Nx - iw; ni = 2; }
            if (jw == 0) { js[0] = 0; js[1] = 1; nj = 2; }
            else if (jw == Ny / 2) { js[0] = Ny; js[1] = Ny + 1; nj = 2; }
            else { js[0] = 2 * jw; js[1] = 2 * jw + 1; js[2] = 2 * (Ny - jw); js[3] = 2 * (Ny - jw) + 1; nj = 4; }
            for (int a = 0; a < ni; ++a)
                for (int b = 0; b < nj; ++b) fundamental_solutions(s, is[a], js[b], Nx, NY, N, Ns);
        }
    if (status) goto done;
#define PL(arr, kk, i, j) arr[(size_t)(kk) * pl + (size_t)(i) * NY + (j)]
#define COLU(kk, i, j) (s->U + (size_t)(kk) * el + ((size_t)(i) * NY + (j)) * N)
#define COLL(arr, kk, i, j) (arr + (size_t)(kk) * sel + ((size_t)(i) * NY + (j)) * N)
#define FD(k, h) first_derivative((k), (h), s->Dz, N)
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < NY; ++j) {
            if (Ns == 1) {
                /* :323-333 and :655-664 */
                double b = (FD(0, COLL(s->ul, 0, i, j)) + FD(0, COLL(s->ur, 0, i, j))) -
                           (FD(N - 1, COLL(s->ul, 0, i, j)) + FD(N - 1, COLL(s->ur, 0, i, j)));
                double r0 = FD(N - 1, COLU(0, i, j)) - FD(0, COLU(0, i, j));
                PL(s->zetal, 0, i, j) = r0 / b;
                PL(s->zetar, 0, i, j) = PL(s->zetal, 0, i, j);
                continue;
            }
            /* forward sweep :336-376 */
            const int e = Ns - 1;
            PL(s->b, 0, i, j) = FD(N - 1, COLL(s->ur, e, i, j)) - FD(N - 1, COLL(s->ul, e, i, j)) -
                                FD(0, COLL(s->ul, 0, i, j));
            PL(s->c, 0, i, j) = -FD(0, COLL(s->ur, 0, i, j));
            PL(s->rhs, 0, i, j) = FD(0, COLU(0, i, j)) - FD(N - 1, COLU(e, i, j));
            PL(rhs1, 0, i, j) = FD(N - 1, COLL(s->ul, e, i, j));
            for (int kk = 1; kk < Ns - 1; ++kk) {
                const double lb = FD(N - 1, COLL(s->ul, kk - 1, i, j));
                PL(s->b, kk, i, j) = FD(N - 1, COLL(s->ur, kk - 1, i, j)) - FD(0, COLL(s->ul, kk, i, j)) -
                                     (PL(s->c, kk - 1, i, j) / PL(s->b, kk - 1, i, j)) * lb;
                PL(s->c, kk, i, j) = -FD(0, COLL(s->ur, kk, i, j));
                PL(s->rhs, kk, i, j) = FD(0, COLU(kk, i, j)) - FD(N - 1, COLU(kk - 1, i, j)) -
                                       (PL(s->rhs, kk - 1, i, j) / PL(s->b, kk - 1, i, j)) * lb;
                PL(rhs1, kk, i, j) = -(PL(rhs1, kk - 1, i, j) / PL(s->b, kk - 1, i, j)) * lb;
            }
            {
                const double lb = FD(N - 1, COLL(s->ul, e - 1, i, j));
                PL(s->b, e, i, j) = FD(N - 1, COLL(s->ur, e - 1, i, j)) - FD(0, COLL(s->ul, e, i, j)) +
                                    FD(0, COLL(s->ur, e, i, j)) - (PL(s->c, e - 1, i, j) / PL(s->b, e - 1, i, j)) * lb;
                PL(s->rhs, e, i, j) = FD(0, COLU(e, i, j)) - FD(N - 1, COLU(e - 1, i, j)) -
                                      (PL(s->rhs, e - 1, i, j) / PL(s->b, e - 1, i, j)) * lb;
                PL(rhs1, e, i, j) = -FD(0, COLL(s->ur, e, i, j)) - (PL(rhs1, e - 1, i, j) / PL(s->b, e - 1, i, j)) * lb;
            }
            /* backward sweep :666-694 */
            PL(s->zetal, e, i, j) = PL(s->rhs, e, i, j) / PL(s->b, e, i, j);
            PL(etal, e, i, j) = PL(rhs1, e, i, j) / PL(s->b, e, i, j);
            for (int kk = Ns - 2; kk >= 0; --kk) {
                PL(s->zetar, kk, i, j) = PL(s->zetal, kk + 1, i, j);
                PL(etar, kk, i, j) = PL(etal, kk + 1, i, j);
                PL(s->zetal, kk, i, j) = (PL(s->rhs, kk, i, j) - PL(s->c, kk, i, j) * PL(s->zetar, kk, i, j)) /
                                         PL(s->b, kk, i, j);
                PL(etal, kk, i, j) = (PL(rhs1, kk, i, j) - PL(s->c, kk, i, j) * PL(etar, kk, i, j)) / PL(s->b, kk, i, j);
            }
            PL(s->zetar, e, i, j) = PL(s->zetal, 0, i, j);
            PL(etar, e, i, j) = PL(etal, 0, i, j);
            /* closure :698-711 */
            const double alpha = -(PL(s->zetal, e, i, j) + PL(s->zetar, e, i, j)) /
                                 (1.0 + PL(etal, e, i, j) + PL(etar, e, i, j));
            if (alpha_out) alpha_out[(size_t)i * NY + j] = alpha;
            for (int kk = 0; kk < Ns; ++kk) {
                PL(s->zetal, kk, i, j) += alpha * PL(etal, kk, i, j);
                PL(s->zetar, kk, i, j) += alpha * PL(etar, kk, i, j);
            }
        }
    /* Step 5 :854-859, Step 6 :861-862 */
    for (int kk = 0; kk < Ns; ++kk)
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < NY; ++j) {
                double *uc = COLU(kk, i, j);
                const double *ulc = COLL(s->ul, kk, i, j), *urc = COLL(s->ur, kk, i, j);
                const double zl = PL(s->zetal, kk, i, j), zr = PL(s->zetar, kk, i, j);
                for (int k = 0; k < N; ++k) uc[k] += zl * ulc[k] + zr * urc[k];
            }
    for (int kk = 0; kk < Ns; ++kk) hzo_fft_xy(s->U + kk * el, -1, Nx, Ny, N);
    for (int kk = 0; kk < Ns; ++kk)
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                for (int k = 0; k < N; ++k)
                    u[((size_t)i * Ny + j) * Ktot + (size_t)kk * (N - 1) + k] = s->U[kk * el + ((size_t)i * NY + j) * N + k];
#undef PL
#undef COLU
#undef COLL
#undef FD
done:
    free_rank(s);
    free(rhs1);
    free(etal);
    free(etar);
    return status;
}

/* NonPeriodic3D_MPI.cpp:632-683 on the global layout. */
int hzo_derivative_z(const hzo_desc *d, int order, const double *u, double *out)
{
    const int N = d->N, P = d->size, Ns = hzo_elements_per_rank(d);
    const int E = P * Ns;
    const double Lz = d->zr - d->zl;
    if (order != 1 && order != 2) return HZO_ERR_ARG;
    double *Dz = malloc(sizeof(double) * (size_t)N * N), *Dzz = malloc(sizeof(double) * (size_t)N * N);
    if (!Dz || !Dzz) { free(Dz); free(Dzz); return HZO_ERR_ALLOC; }
    /* every rank scales by the same 2 Ns / Lzp (NonPeriodic3D_MPI.cpp:125-131) */
    int err = hzo_plan_operators(N, Ns, P, 0, Lz, NULL, Dz, Dzz);
    const double *M = order == 1 ? Dz : Dzz;
    const size_t ncol = (size_t)E * d->Nx * d->Ny;
    for (size_t c = 0; c < ncol && !err; ++c) {
        const double *uc = u + c * N;
        double *oc = out + c * N;
        for (int k = 0; k < N; ++k) {
            double a = 0.0;
            for (int p = 0; p < N; ++p) a += M[(size_t)k * N + p] * uc[p];
            oc[k] = a;
        }
    }
    free(Dz);
    free(Dzz);
    return err;
}

int hzo_fundamental_solutions(int N, int Ns, const double *LU, const int *npivot, const double *Dzz,
                              const double *rhs, double *U, double *ul, double *ur)
{
    const int n = N - 2;
    if (N < 3 || Ns < 1) return HZO_ERR_ARG;
    for (int kk = 0; kk < Ns; ++kk) {
        double *u = U + (size_t)kk * N;
        u[0] = 0.0;
        u[N - 1] = 0.0;
        for (int k = 0; k < n; ++k) u[1 + k] = rhs[(size_t)kk * N + k + 1];
        hzo_lu_solve(LU, npivot, u + 1, n);
    }
    if (ul) {
        ul[0] = 1.0;
        ul[N - 1] = 0.0;
        for (int k = 0; k < n; ++k) ul[1 + k] = -Dzz[(size_t)(k + 1) * N + 0];
        hzo_lu_solve(LU, npivot, ul + 1, n);
    }
    if (ur) {
        ur[0] = 0.0;
        ur[N - 1] = 1.0;
        for (int k = 0; k < n; ++k) ur[1 + k] = -Dzz[(size_t)(k + 1) * N + (N - 1)];
        hzo_lu_solve(LU, npivot, ur + 1, n);
    }
    return HZO_OK;
}
