// eig.cpp -- plan-time eigendecomposition of the interior GLL second-
// derivative operator for the z-diagonalised form:
//     Dzz_int = V diag(mu) V^{-1}
// The eigenvalues of the Dirichlet GLL collocation operator are real,
// negative and distinct, and V is well conditioned (about 3 at n = 127), so
// a real algorithm suffices:
//   1. Householder reduction to upper Hessenberg form,
//   2. implicit single-shift QR with Wilkinson shifts and deflation
//      (eigenvalues),
//   3. inverse iteration on the original matrix (eigenvectors),
//   4. V^{-1} by partial-pivot LU of V.
// Everything runs in long double (x87 extended) and is rounded to double at
// the end, so the factors carry no more than one rounding.
#include "hz_internal.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

namespace hz {

typedef long double ld;

static void hessenberg(std::vector<ld> &H, int n)
{
    std::vector<ld> v(n);
    for (int k = 0; k < n - 2; ++k) {
        ld alpha = 0;
        for (int i = k + 1; i < n; ++i) alpha += H[i * n + k] * H[i * n + k];
        alpha = std::sqrt(alpha);
        if (alpha == 0) continue;
        if (H[(k + 1) * n + k] > 0) alpha = -alpha;
        for (int i = 0; i < n; ++i) v[i] = 0;
        v[k + 1] = H[(k + 1) * n + k] - alpha;
        for (int i = k + 2; i < n; ++i) v[i] = H[i * n + k];
        ld vnorm2 = 0;
        for (int i = k + 1; i < n; ++i) vnorm2 += v[i] * v[i];
        if (vnorm2 == 0) continue;
        // H = (I - 2vv^T/v^Tv) H (I - 2vv^T/v^Tv)
        for (int j = 0; j < n; ++j) {
            ld s = 0;
            for (int i = k + 1; i < n; ++i) s += v[i] * H[i * n + j];
            s = 2 * s / vnorm2;
            for (int i = k + 1; i < n; ++i) H[i * n + j] -= s * v[i];
        }
        for (int i = 0; i < n; ++i) {
            ld s = 0;
            for (int j = k + 1; j < n; ++j) s += H[i * n + j] * v[j];
            s = 2 * s / vnorm2;
            for (int j = k + 1; j < n; ++j) H[i * n + j] -= s * v[j];
        }
        for (int i = k + 2; i < n; ++i) H[i * n + k] = 0;
    }
}

// Eigenvalues of an upper Hessenberg matrix with real spectrum.
static int hessenberg_qr(std::vector<ld> &H, int n, std::vector<ld> &mu)
{
    const ld eps = std::numeric_limits<ld>::epsilon();
    int hi = n - 1, iter = 0;
    mu.assign(n, 0);
    while (hi >= 0) {
        if (hi == 0) {
            mu[0] = H[0];
            break;
        }
        int lo = hi;
        while (lo > 0) {
            ld s = std::fabs(H[(lo - 1) * n + lo - 1]) + std::fabs(H[lo * n + lo]);
            if (std::fabs(H[lo * n + lo - 1]) <= eps * s) {
                H[lo * n + lo - 1] = 0;
                break;
            }
            --lo;
        }
        if (lo == hi) {
            mu[hi] = H[hi * n + hi];
            --hi;
            iter = 0;
            continue;
        }
        if (++iter > 200) return HZ_ERR_UNSUPPORTED;
        // Wilkinson shift from the trailing 2x2 block.
        const ld a = H[(hi - 1) * n + hi - 1], b = H[(hi - 1) * n + hi];
        const ld c = H[hi * n + hi - 1], d = H[hi * n + hi];
        const ld tr = a + d, det = a * d - b * c, disc = tr * tr / 4 - det;
        ld shift = d;
        if (disc >= 0) {
            const ld r = std::sqrt(disc);
            const ld l1 = tr / 2 + r, l2 = tr / 2 - r;
            shift = (std::fabs(l1 - d) < std::fabs(l2 - d)) ? l1 : l2;
        }
        if (iter % 11 == 10) shift += std::fabs(H[hi * n + hi - 1]);  // exceptional shift
        // Implicit single-shift QR sweep on rows/cols lo..hi (bulge chase).
        ld x = H[lo * n + lo] - shift, y = H[(lo + 1) * n + lo];
        for (int k = lo; k < hi; ++k) {
            ld r = std::hypot(x, y);
            ld cs = (r == 0) ? 1 : x / r, sn = (r == 0) ? 0 : y / r;
            const int j0 = (k > lo) ? k - 1 : lo;
            for (int j = j0; j < n; ++j) {
                ld t1 = H[k * n + j], t2 = H[(k + 1) * n + j];
                H[k * n + j] = cs * t1 + sn * t2;
                H[(k + 1) * n + j] = -sn * t1 + cs * t2;
            }
            const int i1 = std::min(k + 2, hi);
            for (int i = 0; i <= i1; ++i) {
                ld t1 = H[i * n + k], t2 = H[i * n + k + 1];
                H[i * n + k] = cs * t1 + sn * t2;
                H[i * n + k + 1] = -sn * t1 + cs * t2;
            }
            if (k > lo) H[(k + 1) * n + k - 1] = 0;
            if (k + 1 < hi) {
                x = H[(k + 1) * n + k];
                y = H[(k + 2) * n + k];
            }
        }
    }
    return HZ_OK;
}

// Partial-pivot LU in place (long double); returns false on a zero pivot.
static bool lu_ld(std::vector<ld> &A, int n, std::vector<int> &piv)
{
    piv.resize(n);
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(A[i * n + k]) > std::fabs(A[p * n + k])) p = i;
        piv[k] = p;
        if (A[p * n + k] == 0) return false;
        if (p != k)
            for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[p * n + j]);
        for (int i = k + 1; i < n; ++i) {
            A[i * n + k] /= A[k * n + k];
            const ld l = A[i * n + k];
            for (int j = k + 1; j < n; ++j) A[i * n + j] -= l * A[k * n + j];
        }
    }
    return true;
}

static void lu_solve_ld(const std::vector<ld> &A, int n, const std::vector<int> &piv, ld *b)
{
    // rows were exchanged whole during the factorisation: permute, then solve
    for (int k = 0; k < n; ++k) std::swap(b[k], b[piv[k]]);
    for (int k = 0; k < n; ++k)
        for (int i = k + 1; i < n; ++i) b[i] -= A[i * n + k] * b[k];
    for (int i = n - 1; i >= 0; --i) {
        ld s = b[i];
        for (int j = i + 1; j < n; ++j) s -= A[i * n + j] * b[j];
        b[i] = s / A[i * n + i];
    }
}

// A (n x n, row-major double) -> mu (ascending), V (columns = eigenvectors,
// unit 2-norm, largest component positive), Vinv.
int eigen_decompose(const std::vector<double> &A, int n, std::vector<double> &mu,
                    std::vector<double> &V, std::vector<double> &Vinv)
{
    std::vector<ld> H(A.begin(), A.end());
    hessenberg(H, n);
    std::vector<ld> ev;
    int st = hessenberg_qr(H, n, ev);
    if (st) return st;
    std::sort(ev.begin(), ev.end());
    ld anorm = 0;
    for (double a : A) anorm = std::max(anorm, (ld)std::fabs(a));
    std::vector<ld> Vl((size_t)n * n), M((size_t)n * n), x(n);
    std::vector<int> piv;
    for (int e = 0; e < n; ++e) {
        ld shift = ev[e];
        for (int attempt = 0; attempt < 4; ++attempt) {
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) M[i * n + j] = (ld)A[(size_t)i * n + j] - (i == j ? shift : 0);
            if (lu_ld(M, n, piv)) break;
            shift += anorm * 64 * std::numeric_limits<ld>::epsilon();
        }
        for (int i = 0; i < n; ++i) x[i] = 1 + (ld)i / n;
        for (int it = 0; it < 3; ++it) {
            lu_solve_ld(M, n, piv, x.data());
            ld nrm = 0;
            for (int i = 0; i < n; ++i) nrm += x[i] * x[i];
            nrm = std::sqrt(nrm);
            for (int i = 0; i < n; ++i) x[i] /= nrm;
        }
        int imax = 0;
        for (int i = 1; i < n; ++i)
            if (std::fabs(x[i]) > std::fabs(x[imax])) imax = i;
        const ld sg = x[imax] < 0 ? -1 : 1;
        for (int i = 0; i < n; ++i) Vl[(size_t)i * n + e] = sg * x[i];
    }
    // Inverse of V.
    std::vector<ld> W(Vl);
    if (!lu_ld(W, n, piv)) return HZ_ERR_UNSUPPORTED;
    std::vector<ld> col(n);
    Vinv.assign((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < n; ++i) col[i] = (i == j);
        lu_solve_ld(W, n, piv, col.data());
        for (int i = 0; i < n; ++i) Vinv[(size_t)i * n + j] = (double)col[i];
    }
    V.resize((size_t)n * n);
    for (size_t i = 0; i < V.size(); ++i) V[i] = (double)Vl[i];
    mu.resize(n);
    for (int i = 0; i < n; ++i) mu[i] = (double)ev[i];
    return HZ_OK;
}

// Even/odd split.  The GLL nodes are symmetric about the element centre, so
// the interior operator A commutes with the reversal (J x)_i = x_{n-1-i} up
// to the rounding of the nodes (|A - JAJ| is about 7e-13 |A| at n = 127).
// The fold F x = (x_i + x_{n-1-i} | x_i - x_{n-1-i}), i < n/2, with the
// centre entry of an odd n kept in the even part, takes the symmetrised
// As = (A + JAJ)/2 to block-diagonal form F As F^-1 = diag(Ae, Ao), Ae of
// size he = n - n/2 and Ao of size ho = n/2, and each block is diagonalised
// on its own: a contraction with the eigenbasis then costs he^2 + ho^2
// multiply-adds instead of n^2.  Padded layout (NP even, NP/2 >= he): even
// coordinate i at index i, odd coordinate i at index NP/2 + i; W = diag(Ve,
// Vo) and Winv = diag(Ve^-1, Vo^-1) (NP x NP, row-major, zero outside the
// blocks), mu the eigenvalues (zero in the padding), so that
//   As = F^-1 W diag(mu) Winv F,  F^-1 y = ((y_e + y_o)/2 | (y_e - y_o)/2),
// the centre entry of an odd n taken from y_e.  *asym = max|A - JAJ| / max|A|.
int eigen_decompose_split(const std::vector<double> &A, int n, int NP, std::vector<double> &mu,
                          std::vector<double> &W, std::vector<double> &Winv, double *asym)
{
    const int m = n / 2, he = n - m, ho = m, H = NP / 2;
    if (NP % 2 || H < he) return HZ_ERR_ARG;
    ld amax = 0, dmax = 0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            amax = std::max(amax, (ld)std::fabs(A[(size_t)i * n + j]));
            dmax = std::max(dmax, (ld)std::fabs((ld)A[(size_t)i * n + j] -
                                                (ld)A[(size_t)(n - 1 - i) * n + (n - 1 - j)]));
        }
    if (asym) *asym = amax > 0 ? (double)(dmax / amax) : 0.0;
    // B = F As F^-1 column by column: F^-1 of a unit vector has one or two
    // entries, As x costs O(n), the fold O(n).
    auto as = [&](int r, int c) {
        return ((ld)A[(size_t)r * n + c] + (ld)A[(size_t)(n - 1 - r) * n + (n - 1 - c)]) / 2;
    };
    std::vector<ld> B((size_t)NP * NP, 0), y(n);
    for (int c = 0; c < NP; ++c) {
        const bool even = c < H;
        const int j = even ? c : c - H;
        if (j >= (even ? he : ho)) continue;
        for (int r = 0; r < n; ++r) {
            if (even && 2 * j + 1 == n) {
                y[r] = as(r, j);
            } else {
                const ld s = even ? 1 : -1;
                y[r] = (as(r, j) + s * as(r, n - 1 - j)) / 2;
            }
        }
        for (int i = 0; i < he; ++i) B[(size_t)i * NP + c] = 2 * i + 1 == n ? y[i] : y[i] + y[n - 1 - i];
        for (int i = 0; i < ho; ++i) B[(size_t)(H + i) * NP + c] = y[i] - y[n - 1 - i];
    }
    ld bmax = 0, off = 0;
    for (int r = 0; r < NP; ++r)
        for (int c = 0; c < NP; ++c) {
            const ld v = std::fabs(B[(size_t)r * NP + c]);
            bmax = std::max(bmax, v);
            if ((r < H) != (c < H)) off = std::max(off, v);
        }
    if (off > bmax * 1e-12L) return HZ_ERR_UNSUPPORTED;  // not centrosymmetric
    mu.assign(NP, 0.0);
    W.assign((size_t)NP * NP, 0.0);
    Winv.assign((size_t)NP * NP, 0.0);
    for (int half = 0; half < 2; ++half) {
        const int b = half ? ho : he, o = half ? H : 0;
        if (!b) continue;
        std::vector<double> blk((size_t)b * b), ev, V, Vi;
        for (int i = 0; i < b; ++i)
            for (int j = 0; j < b; ++j) blk[(size_t)i * b + j] = (double)B[(size_t)(o + i) * NP + (o + j)];
        const int st = eigen_decompose(blk, b, ev, V, Vi);
        if (st) return st;
        for (int i = 0; i < b; ++i) {
            mu[o + i] = ev[i];
            for (int j = 0; j < b; ++j) {
                W[(size_t)(o + i) * NP + (o + j)] = V[(size_t)i * b + j];
                Winv[(size_t)(o + i) * NP + (o + j)] = Vi[(size_t)i * b + j];
            }
        }
    }
    return HZ_OK;
}

}  // namespace hz
