// sweep.h -- element-interface systems shared by the mode kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace hz {
namespace {

// Element-interface system of one real column with z periodic
// (Periodic3D_MPI.cpp:320-376, :652-711 at one rank).  Element e of every
// array is at [e * st].  d0, dN: end-point derivatives of the element's
// particular solution; d0ul, dNul, d0ur, dNur those of the homogeneous
// solutions, the same in every element.  On return zl, zr hold each
// element's end values.
//   E = 1: one value with u(0) = u(L) and matching u_z.
//   E > 1: forward sweep with the particular (r0) and the periodic-coupling
//   (r1) right-hand sides, backward sweep (zeta, eta), then the closure
//   alpha = -(zeta_l + zeta_r)/(1 + eta_l + eta_r) of the last element.
// zl, zr carry b and r0 through the forward sweep and d0, dN the backward
// values (their inputs are dead by then); r1 is recomputed in the backward
// pass instead of stored (the same operations, so the same values).
// alpha_out (may be null) receives the closure coefficient (E > 1 only, as
// the reference's alphaPDE, Periodic3D_MPI.cpp:698).
__device__ __forceinline__ void periodic_sweep(int E, double d0ul, double dNul, double d0ur, double dNur, double *d0,
                                               double *dN, double *zl, double *zr, int st, double *alpha_out = nullptr)
{
    if (E == 1) {
        const double b = (d0ul + d0ur) - (dNul + dNur);
        const double z = (dN[0] - d0[0]) / b;
        zl[0] = z;
        zr[0] = z;
        return;
    }
    const double c = -d0ur, lb = dNul;
    double bp = dNur - dNul - d0ul;
    double r0p = d0[0] - dN[(E - 1) * st];
    double r1p = dNul;
    zl[0] = bp;
    zr[0] = r0p;
    for (int e = 1; e < E; ++e) {
        const bool last = e == E - 1;  // the last element's rows carry the wrap terms
        const double be = dNur - d0ul + (last ? d0ur : 0.0) - (c / bp) * lb;
        const double re = d0[e * st] - dN[(e - 1) * st] - (r0p / bp) * lb;
        const double qe = (last ? -d0ur : 0.0) - (r1p / bp) * lb;
        bp = be;
        r0p = re;
        r1p = qe;
        zl[e * st] = be;
        zr[e * st] = re;
    }
    double zeta = zr[(E - 1) * st] / zl[(E - 1) * st];
    double eta = r1p / zl[(E - 1) * st];
    d0[(E - 1) * st] = zeta;
    dN[(E - 1) * st] = eta;
    for (int e = E - 2; e >= 0; --e) {
        double r1e = dNul;
        for (int k = 1; k <= e; ++k) r1e = -(r1e / zl[(k - 1) * st]) * lb;
        const double zle = (zr[e * st] - c * zeta) / zl[e * st];
        const double ele = (r1e - c * eta) / zl[e * st];
        zeta = zle;
        eta = ele;
        d0[e * st] = zle;
        dN[e * st] = ele;
    }
    const double alpha = -(d0[(E - 1) * st] + d0[0]) / (1.0 + dN[(E - 1) * st] + dN[0]);
    if (alpha_out) *alpha_out = alpha;
    for (int e = 0; e < E; ++e) {
        const int n1 = e + 1 < E ? e + 1 : 0;
        zl[e * st] = d0[e * st] + alpha * dN[e * st];
        zr[e * st] = d0[n1 * st] + alpha * dN[n1 * st];
    }
}

// Element-interface system of one real column with Robin faces
// beta_l u - alpha_l u_z = gl, beta_r u + alpha_r u_z = gr
// (NonPeriodic3D_MPI.cpp:327-587 at one rank): forward elimination over the
// elements, then back substitution.  d0, dN, zl, zr as in periodic_sweep
// (element e at [e]); zl, zr carry b and the right-hand side through the
// forward sweep.
__device__ __forceinline__ void robin_sweep(int E, double d0ul, double dNul, double d0ur, double dNur, double gl,
                                            double gr, double al, double bl, double ar, double br,
                                            const double *d0, const double *dN, double *zl, double *zr)
{
    double bprev = -al * d0ul + bl;
    double cprev = -al * d0ur;
    double rprev = gl + al * d0[0];
    zl[0] = bprev;
    zr[0] = rprev;
    for (int e = 1; e < E; ++e) {
        const double lb = dNul;
        const double be = dNur - d0ul - (cprev / bprev) * lb;
        const double re = d0[e] - dN[e - 1] - (rprev / bprev) * lb;
        bprev = be;
        cprev = -d0ur;
        rprev = re;
        zl[e] = be;
        zr[e] = re;
    }
    const double lb = ar * dNul;
    const double bE = ar * dNur + br - (cprev / bprev) * lb;
    const double rhsr = gr - ar * dN[E - 1] - (rprev / bprev) * lb;
    double z = rhsr / bE;
    for (int e = E - 1; e >= 0; --e) {
        const double be = zl[e], re = zr[e];
        const double ce = (e == 0) ? -al * d0ur : -d0ur;
        const double zle = (re - ce * z) / be;
        zl[e] = zle;
        zr[e] = z;
        z = zle;
    }
}

// Transpose of robin_sweep's interface system T zeta = r (unknowns zeta_0
// .. zeta_E: the face and interface values; rows: left Robin, E - 1 slope
// matching conditions, right Robin), for the adjoint of a column of E > 1
// elements.  Inputs per element e: u0[e], uN[e] the adjoints at its first
// and last node (a shared interface node is read from u0 of the element to
// its right), w0y[e], wNy[e] the loads (w0 / (mu - sigma)) . y_e and
// (wN / (mu - sigma)) . y_e through ul, ur.  zetabar_i collects them,
// lambda = T^-T zetabar, and since r_0 = g_l + alpha_l d0_0,
// r_e = d0_e - dN_(e-1), r_E = g_r - alpha_r dN_(E-1):
//   p0b[e] = d0bar_e, pNb[e] = dNbar_e, *glb = lambda_0, *grb = lambda_E.
// p0b, pNb carry the eliminated diagonal and right-hand side until the back
// substitution overwrites them.
__device__ __forceinline__ void robin_sweep_adjoint(int E, double d0ul, double dNul, double d0ur, double dNur,
                                                    double al, double bl, double ar, double br, const double *u0,
                                                    const double *uN, const double *w0y, const double *wNy,
                                                    double *p0b, double *pNb, double *glb, double *grb)
{
    // T^T: diagonal d_i, sub-diagonal a_i (coefficient of lambda_(i-1)),
    // super-diagonal s_i (coefficient of lambda_(i+1))
    auto diag = [&](int i) { return i == 0 ? -al * d0ul + bl : i == E ? ar * dNur + br : dNur - d0ul; };
    auto sub = [&](int i) { return i == 1 ? -al * d0ur : -d0ur; };
    auto sup = [&](int i) { return i == E - 1 ? ar * dNul : dNul; };
    auto zbar = [&](int i) {
        return i == 0 ? u0[0] + w0y[0] : i == E ? uN[E - 1] + wNy[E - 1] : u0[i] + w0y[i] + wNy[i - 1];
    };
    double dp = diag(0), yp = zbar(0);
    p0b[0] = dp;
    pNb[0] = yp;
    for (int i = 1; i <= E; ++i) {
        const double m = sub(i) / dp;
        dp = diag(i) - m * sup(i - 1);
        yp = zbar(i) - m * yp;
        if (i < E) {
            p0b[i] = dp;
            pNb[i] = yp;
        }
    }
    const double lamE = yp / dp;
    double next = lamE;
    for (int i = E - 1; i >= 0; --i) {
        const double lam = (pNb[i] - sup(i) * next) / p0b[i];
        pNb[i] = i == E - 1 ? -ar * lamE : -next;
        p0b[i] = i == 0 ? al * lam : lam;
        next = lam;
    }
    *glb = next;
    *grb = lamE;
}

// Transpose of periodic_sweep's cyclic interface system, E > 1 elements.
// The interface values zeta_0 .. zeta_(E-1) (zeta_0 at both ends of the
// column) solve the circulant tridiagonal system
//   zeta_(e-1) dNul + zeta_e (dNur - d0ul) - zeta_(e+1) d0ur = d0_e - dN_(e-1)
// (indices mod E: slope matching at every interface).  Inputs per element
// as in robin_sweep_adjoint (u0, uN, w0y, wNy; the column's first and last
// nodes both carry zeta_0); lambda = C^-T zetabar is solved directly (E = 2)
// or with the Sherman-Morrison correction of a tridiagonal solve, and
//   p0b[e] = d0bar_e = lambda_e,  pNb[e] = dNbar_e = -lambda_((e+1) mod E).
// u0 and pNb serve as scratch once their inputs are consumed.
__device__ __forceinline__ void periodic_sweep_adjoint(int E, double d0ul, double dNul, double d0ur, double dNur,
                                                       double *u0, const double *uN, const double *w0y,
                                                       const double *wNy, double *p0b, double *pNb)
{
    for (int e = 0; e < E; ++e)  // zetabar
        pNb[e] = e == 0 ? (u0[0] + uN[E - 1]) + (w0y[0] + wNy[E - 1]) : u0[e] + (w0y[e] + wNy[e - 1]);
    // C^T: sub-diagonal a (coefficient of lambda_(i-1)), diagonal b,
    // super-diagonal c (coefficient of lambda_(i+1)), wrapping around
    const double a = -d0ur, b = dNur - d0ul, c = dNul;
    if (E == 2) {
        const double o = a + c, det = b * b - o * o;
        const double l0 = (b * pNb[0] - o * pNb[1]) / det, l1 = (b * pNb[1] - o * pNb[0]) / det;
        p0b[0] = l0;
        p0b[1] = l1;
        pNb[0] = -l1;
        pNb[1] = -l0;
        return;
    }
    // tridiagonal part with the corners folded into the first and last
    // diagonal entries, two right-hand sides: zetabar (y) and u = (g, 0, .., 0, c) (z)
    const double g = -b, beta = a;  // corners: row 0 holds a * lambda_(E-1), row E-1 holds c * lambda_0
    double cp = 0.0, yp = 0.0, zp = 0.0;
    for (int i = 0; i < E; ++i) {
        const double bi = i == 0 ? b - g : i == E - 1 ? b - c * beta / g : b;
        const double den = i == 0 ? bi : bi - a * cp;
        const double ui = i == 0 ? g : i == E - 1 ? c : 0.0;
        yp = (pNb[i] - (i ? a * yp : 0.0)) / den;
        zp = (ui - (i ? a * zp : 0.0)) / den;
        cp = c / den;
        p0b[i] = cp;
        pNb[i] = yp;
        u0[i] = zp;
    }
    double ynext = pNb[E - 1], znext = u0[E - 1];
    for (int i = E - 2; i >= 0; --i) {
        ynext = pNb[i] - p0b[i] * ynext;
        znext = u0[i] - p0b[i] * znext;
        pNb[i] = ynext;
        u0[i] = znext;
    }
    const double f = (pNb[0] + beta * pNb[E - 1] / g) / (1.0 + u0[0] + beta * u0[E - 1] / g);
    for (int e = 0; e < E; ++e) p0b[e] = pNb[e] - f * u0[e];
    for (int e = 0; e < E; ++e) pNb[e] = -p0b[(e + 1) % E];
}

}  // namespace
}  // namespace hz
