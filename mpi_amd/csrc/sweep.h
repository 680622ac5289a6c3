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
__device__ __forceinline__ void periodic_sweep(int E, double d0ul, double dNul, double d0ur, double dNur, double *d0,
                                               double *dN, double *zl, double *zr, int st)
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
    for (int e = 0; e < E; ++e) {
        const int n1 = e + 1 < E ? e + 1 : 0;
        zl[e * st] = d0[e * st] + alpha * dN[e * st];
        zr[e * st] = d0[n1 * st] + alpha * dN[n1 * st];
    }
}

}  // namespace
}  // namespace hz
