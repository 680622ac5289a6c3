// gll.cpp -- plan-time GLL setup on the host: nodes, weights, first and
// second differentiation matrices, in the reference's operation order so the
// device operators are bit-identical to the reference's (compile with
// -ffp-contract=off).  Citations: HelmholtzSolvers/DirectPseudoSpectral/.
#include "hz_internal.h"

#include <cmath>
#include <vector>

namespace hz {

static const double kEpsilon = 1.000000000000E-12;  // include/General.h:15

static inline double absd(double m) { return m >= 0.0 ? m : -m; }

// The reference exchanges values arithmetically (include/General.h Swap);
// the rounding of a+b-b is part of its node values.
static inline void swap_arith(double &a, double &b)
{
    a = a + b;
    b = a - b;
    a = a - b;
}

// General.cpp:87-91
static double pythag(double a, double b)
{
    return (absd(a) > absd(b) ? absd(a) * std::sqrt(1.0 + b * b / (a * a))
                              : (absd(b) == 0.0 ? 0.0 : absd(b) * std::sqrt(1.0 + a * a / (b * b))));
}

// GLLQuadrature.cpp:7-40 (argument check and clamp, three-term recurrence)
double legendre(double x, int n, int *status)
{
    if (absd(x) > (1.0 + kEpsilon)) {
        if (status) *status = HZ_ERR_ARG;
        return 0.0;
    }
    if (x > 1.0) x = 1.0;
    if (x < -1.0) x = -1.0;
    if (n == 0) return 1.0;
    if (n == 1) return x;
    double p0 = 1.0, p1 = x, pn = 0.0;
    for (int i = 2; i <= n; ++i) {
        pn = ((2.0 * i - 1.0) * x * p1 - (i - 1.0) * p0) / i;
        p0 = p1;
        p1 = pn;
    }
    return pn;
}

// SymmetricEigenSystems.cpp:313-416: implicit QL on a symmetric tridiagonal
// matrix, eigenvalues only, followed by the reference's descending sort.
static void tridiagonal_ql(std::vector<double> &d, std::vector<double> &e)
{
    const int n = (int)d.size();
    int i, l, m;
    double s, r, p, g, f, dd, c, b;
    for (l = 0; l < n; ++l) {
        do {
            for (m = l; m < n - 1; ++m) {
                dd = absd(d[m]) + absd(d[m + 1]);
                if ((double)(absd(e[m]) + dd) == dd) break;
            }
            if (m != l) {
                g = 0.5 * (d[l + 1] - d[l]) / e[l];
                r = pythag(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? absd(r) : -absd(r)));
                s = c = 1.0;
                p = 0.0;
                for (i = m - 1; i >= l; i--) {
                    f = s * e[i];
                    b = c * e[i];
                    e[i + 1] = (r = pythag(f, g));
                    if (r == 0.0) {
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    d[i + 1] = g + (p = s * r);
                    g = c * r - b;
                }
                if (r == 0.0 && i >= l) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
    for (i = 0; i < n - 1; ++i) {
        int k = i;
        p = d[i];
        for (int j = i; j < n; ++j)
            if (d[j] >= p) p = d[k = j];
        if (k != i) {
            d[k] = d[i];
            d[i] = p;
        }
    }
}

// GLLQuadrature.cpp:78-128: nodes = eigenvalues of the Lobatto-modified
// Jacobi matrix, reversed into ascending order.
int gll_nodes(int N, std::vector<double> &x)
{
    if (N < 2) return HZ_ERR_N_TOO_SMALL;
    x.assign(N, 0.0);
    if (N == 2) {
        x[0] = -1.0;
        x[1] = 1.0;
        return HZ_OK;
    }
    std::vector<double> e(N, 0.0);
    for (int i = 0; i < N - 2; ++i)
        e[i] = std::sqrt((i + 1.0) * (i + 1.0) / (4.0 * (i + 1.0) * (i + 1.0) - 1.0));
    const double sign = (N % 2 == 0 ? 1.0 : -1.0);
    const double a = -sign * (N - 1.0) / (2.0 * N - 3.0), b = sign;
    const double c = (N - 1.0) / (2.0 * N - 3.0), d = 1.0;
    const double ee = sign * (N - 1.0) / (2.0 * N - 3.0), f = (N - 1.0) / (2.0 * N - 3.0);
    const double alpha = (d * ee - b * f) / (a * d - b * c);
    const double beta = (a * f - ee * c) / (a * d - b * c);
    x[N - 1] = std::sqrt(alpha);
    e[N - 2] = std::sqrt(beta);
    tridiagonal_ql(x, e);
    const int M = (N % 2 == 0 ? N / 2 : (N - 1) / 2);
    for (int i = 0; i < M; ++i) swap_arith(x[i], x[N - i - 1]);
    return HZ_OK;
}

// GLLQuadrature.cpp:135-151
int gll_weights(const std::vector<double> &x, std::vector<double> &w)
{
    const int N = (int)x.size();
    w.assign(N, 0.0);
    const int points = (N % 2 == 0 ? N / 2 : (N + 1) / 2);
    const double numerator = 2.0 / ((N - 1.0) * N);
    int status = HZ_OK;
    for (int i = 0; i < points; ++i) {
        const double pn = legendre(x[i], N - 1, &status);
        w[i] = numerator / (pn * pn);
        w[N - i - 1] = w[i];
    }
    return status;
}

// GLLQuadrature.cpp:198-213 and :250-262 (row-major N x N)
int gll_derivative_matrices(const std::vector<double> &x, std::vector<double> &D,
                            std::vector<double> &D2)
{
    const int N = (int)x.size();
    int status = HZ_OK;
    D.assign((size_t)N * N, 0.0);
    D2.assign((size_t)N * N, 0.0);
    D[0] = -0.25 * N * (N - 1);
    D[(size_t)(N - 1) * N + (N - 1)] = 0.25 * N * (N - 1);
    std::vector<double> P(N);
    for (int i = 0; i < N; ++i) P[i] = legendre(x[i], N - 1, &status);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            if (i != j) D[(size_t)i * N + j] = P[i] / (P[j] * (x[i] - x[j]));
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            double acc = 0.0;
            for (int k = 0; k < N; ++k) acc += D[(size_t)i * N + k] * D[(size_t)k * N + j];
            D2[(size_t)i * N + j] = acc;
        }
    return status;
}

// NonPeriodic3D_MPI.cpp:103-131 for one process holding all `ne` elements:
// z from 0 (the reference's element bounds ignore zl), Dz and Dzz scaled.
int plan_operators(int N, int ne, double Lz, std::vector<double> &z, std::vector<double> &Dz,
                   std::vector<double> &Dzz)
{
    std::vector<double> xi;
    int st = gll_nodes(N, xi);
    if (st) return st;
    st = gll_derivative_matrices(xi, Dz, Dzz);
    if (st) return st;
    const double Lzp = Lz;  // one rank holds the whole z extent
    const int Ns = ne;
    z.assign((size_t)Ns * (N - 1) + 1, 0.0);
    for (int kk = 0; kk < Ns; ++kk) {
        const double zle = 0 * Lzp + kk * Lzp / Ns;
        const double zre = zle + Lzp / Ns;
        for (int k = 0; k < N; ++k) z[(size_t)kk * (N - 1) + k] = 0.5 * (zle + zre + (Lzp / Ns) * xi[k]);
    }
    for (auto &v : Dz) v *= ((2.0 * Ns) / Lzp);
    for (auto &v : Dzz) v *= ((4.0 * Ns * Ns) / (Lzp * Lzp));
    return HZ_OK;
}

// TestNonPeriodic3D_MPI.cpp:76: the shift of mode (iw, jw), same order.
double mode_shift(double lambda, int iw, int jw, double Lx, double Ly)
{
    const double PI = 3.1415926535897932;  // include/General.h:14
    return lambda + 4.0 * PI * PI * iw * iw / (Lx * Lx) + 4.0 * PI * PI * jw * jw / (Ly * Ly);
}

}  // namespace hz
