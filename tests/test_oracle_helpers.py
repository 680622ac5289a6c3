"""The oracle's z derivatives, fundamental solutions, LU solve and periodic
closure coefficients against the reference's own outputs (golden vectors of
tests/golden/make_golden.py --helpers / --palpha, written by gpritam/MPI's
DirectPseudoSpectral code compiled from its sources), plus checks that
follow from the mathematics alone.  Bit-exact except where noted."""
import math

import numpy as np
import pytest

from golden_util import (fundsol_inputs, load, lusolve_inputs, palpha_inputs, sha, tecplot_inputs, unhex,
                         zderiv_inputs)
from oracle import pyoracle as po

ZD = load("zderiv.json")
FS = load("fundsol.json")
LS = load("lusolve.json")
TP = load("tecplot.json")
PA = load("palpha.json")


def _zdesc(rec):
    return po.make_desc(rec["Nx"], rec["Ny"], rec["Nz"], rec["N"], size=rec["size"], zl=rec["zl"], zr=rec["zr"])


@pytest.mark.parametrize("key", sorted(ZD))
def test_derivative_z_bitexact(key):
    """FirstDerivativeZ / SecondDerivativeZ (NonPeriodic3D_MPI.cpp:632-683) on
    element arrays, one and two ranks, N up to 257."""
    rec = ZD[key]
    u = zderiv_inputs(rec)
    d = _zdesc(rec)
    for order, k in ((1, "sha_d1"), (2, "sha_d2")):
        out = po.derivative_z(d, u, order)
        assert sha(out) == rec[k], (key, order)
        assert np.array_equal(out.ravel(), unhex(rec["d1" if order == 1 else "d2"]))


@pytest.mark.parametrize("N,ne", [(10, 3), (33, 2), (65, 1)])
def test_derivative_z_mathematics(N, ne):
    """Independent of the reference: on every element the scaled Dz and Dzz
    differentiate polynomials of degree < N exactly (to the matrices'
    round-off), with dz = (Lz/ne)/2 dxi."""
    zl, zr = -1.0, 1.0
    d = po.make_desc(2, 2, ne * N, N, size=1, zl=zl, zr=zr)
    z = po.z_grid(d)
    u = np.zeros((ne, 2, 2, N))
    want1 = np.zeros_like(u)
    want2 = np.zeros_like(u)
    for deg in (0, 1, 2, N // 2, N - 1):
        for e in range(ne):
            ze = z[e * (N - 1):e * (N - 1) + N]
            c = 0.5 * (ze[0] + ze[-1])
            t = (ze - c)
            u[e] = t ** deg
            want1[e] = deg * t ** max(deg - 1, 0) if deg else 0.0
            want2[e] = deg * (deg - 1) * t ** max(deg - 2, 0) if deg > 1 else 0.0
        s1 = max(1.0, np.abs(want1).max())
        s2 = max(1.0, np.abs(want2).max())
        assert np.abs(po.derivative_z(d, u, 1) - want1).max() <= 1e-9 * N ** 2 * s1
        assert np.abs(po.derivative_z(d, u, 2) - want2).max() <= 1e-9 * N ** 4 * s2


def _fs_operator(rec):
    """The test callback's operator of mode (iw, jw) with the reference's
    scaled Dzz, factored by the oracle's LUDecomposition."""
    xl, xr, yl, yr, zl, zr = rec["bounds"]
    N, Ns = rec["N"], rec["Ns"]
    _, _, Dzz = po.plan_operators(N, Ns, 1, 0, zr - zl)
    shift = po.shift(rec["lam"], rec["iw"], rec["jw"], xr - xl, yr - yl)
    A = Dzz[1:-1, 1:-1] - shift * np.eye(N - 2)
    LU, piv, _ = po.lu(A)
    return A, LU, piv, Dzz


@pytest.mark.parametrize("key", sorted(FS))
def test_fundamental_solutions_bitexact(key):
    """ConstructFundamentalSolutions (NonPeriodic3D_MPI.cpp:184-232) after the
    test callback (TestNonPeriodic3D_MPI.cpp:71-79) factored the operator."""
    rec = FS[key]
    rhs = fundsol_inputs(rec)
    A, LU, piv, Dzz = _fs_operator(rec)
    U, ul, ur = po.fundamental_solutions(rec["N"], LU, piv, Dzz, rhs)
    assert np.array_equal(U.ravel(), unhex(rec["U"]))
    # mathematics: zero ends, interior solves A x = rhs interior; the
    # homogeneous solutions solve A x = -Dzz[1..N-2][0 or N-1]
    assert not U[:, 0].any() and not U[:, -1].any()
    for kk in range(rec["Ns"]):
        r = A @ U[kk, 1:-1] - rhs[kk, 1:-1]
        assert np.abs(r).max() <= 1e-10 * np.abs(A).max() * np.abs(U).max() * rec["N"]
    for h, col in ((ul, 0), (ur, -1)):
        r = A @ h[1:-1] + Dzz[1:-1, col]
        assert np.abs(r).max() <= 1e-10 * np.abs(A).max() * rec["N"]


@pytest.mark.parametrize("key", sorted(LS))
def test_lu_solve_bitexact(key):
    """LUDecomposition then LUSolve (LU_Decomposition.cpp:7-108)."""
    rec = LS[key]
    A, b = lusolve_inputs(key, rec)
    LU, piv, _ = po.lu(A)
    x = po.lu_solve(LU, piv, b)
    assert np.array_equal(x, unhex(rec["x"]))
    # mathematics: the residual of the solved system
    assert np.abs(A @ x - b).max() <= 1e-9 * np.abs(A).max() * max(1.0, np.abs(x).max()) * rec["n"]


@pytest.mark.parametrize("key", sorted(TP))
def test_tecplot_fixture_layout(key):
    """The reference's Tecplot file (NonPeriodic3D_MPI.cpp:713-741) holds the
    written arrays: the header, one line per (kk, k, j, i) in that loop
    order, X, Y, Z on the plan's grid, U to 8 decimals and the exact
    function's value."""
    rec = TP[key]
    U = tecplot_inputs(rec)
    nx, ny, nz, N = (int(v) for v in rec["args"][:4])
    xl, xr, yl, yr, zl, zr = (float(v) for v in rec["args"][4:10])
    Ns = rec["Ns"]
    lines = rec["text"].split("\n")
    assert lines[0] == 'TITLE = "Helmholtz solver"'
    assert lines[1] == 'Variables = "X","Y","Z","U","Exact"'
    assert lines[2] == f"Zone I = {nx}, J = {ny}, K = {Ns * N}, DATAPACKING=POINT"
    body = [ln for ln in lines[3:] if ln]
    assert len(body) == Ns * N * nx * ny
    d = po.make_desc(nx, ny, nz, N, size=1, zl=zl, zr=zr)
    z = po.z_grid(d)
    t = 0
    for kk in range(Ns):
        for k in range(N):
            for j in range(ny):
                for i in range(nx):
                    X, Y, Z, u, ex = (float(v) for v in body[t].split("\t"))
                    x = xl + i * ((xr - xl) / nx)
                    y = yl + j * ((yr - yl) / ny)
                    zz = z[kk * (N - 1) + k]
                    assert abs(X - x) <= 5e-9 and abs(Y - y) <= 5e-9 and abs(Z - zz) <= 5e-9
                    assert abs(u - U[kk, i, j, k]) <= 5e-9
                    assert abs(ex - (x + 2 * y - 3 * zz + x * y * zz)) <= 5e-9
                    t += 1


def _padesc(rec, size=None):
    xl, xr, yl, yr, zl, zr = rec["bounds"]
    nz = rec["Nz"] if size is None else rec["elements"] * rec["N"]
    return po.make_desc(rec["Nx"], rec["Ny"], nz, rec["N"], size=size or rec["size"], xl=xl, xr=xr, yl=yl, yr=yr,
                        zl=zl, zr=zr, lam=rec["lam"])


@pytest.mark.parametrize("key", sorted(PA))
def test_periodic_closure_against_reference(key):
    """The closure coefficients alphaPDE of the cyclic interface system
    (Periodic3D_MPI.cpp:698): bit-exact with the one-rank reference; the
    multi-rank reference (:738, :804, broadcast at :758, :838) solves the
    same system in another order and agrees to round-off."""
    rec = PA[key]
    f = palpha_inputs(rec)
    want = unhex(rec["alpha"]).reshape(rec["Nx"], 2 * rec["Ny"])
    if rec["size"] == 1:
        u, a = po.solve_periodic(_padesc(rec), f, alpha=True)
        assert sha(u) == rec["sha_u"]
        assert np.array_equal(a, want)
    else:
        d1 = _padesc(rec, size=1)
        assert po.total_z_points(d1) == rec["K"]
        u, a = po.solve_periodic(d1, f, alpha=True)
        assert np.abs(a - want).max() <= 1e-12 * rec["max_abs_alpha"]


def test_periodic_closure_mathematics():
    """Independent of the reference: with the closure coefficient the
    solution is periodic and its slope continuous across the wrap; for one
    element the reference computes none (zeros)."""
    rec = PA["8x16_N9_E3_np1"]
    f = palpha_inputs(rec)
    d = _padesc(rec)
    u, a = po.solve_periodic(d, f, alpha=True)
    assert np.abs(a).max() > 0
    assert np.abs(u[:, :, 0] - u[:, :, -1]).max() <= 1e-12 * np.abs(u).max()
    one = po.make_desc(8, 8, 9, 9, size=1, xl=0, xr=2 * math.pi, yl=0, yr=2 * math.pi, zl=0, zr=1, lam=1.0)
    _, a1 = po.solve_periodic(one, np.ones((8, 8, 9)), alpha=True)
    assert not a1.any()
