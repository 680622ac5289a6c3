"""The oracle (oracle/hz_oracle.c) against the reference's own outputs.

The golden vectors were produced by gpritam/MPI's DirectPseudoSpectral code
compiled from its sources (tests/golden/make_golden.py).  The oracle follows
the reference's operation order, so every comparison here is bit-exact
(SHA-256 of the output bytes), except where noted.
"""
import math

import numpy as np
import pytest

from golden_util import case_inputs, load, sha, unhex
from oracle import pyoracle as po


@pytest.fixture(scope="module")
def gll():
    return load("gll.json")


@pytest.mark.parametrize("N", list(range(2, 41)) + [65, 129, 257])
def test_gll_bitexact(gll, N):
    """GLLQuadrature.cpp:78-262 via SymmetricEigenSystems.cpp:313-416."""
    x, w, D, Dzz = po.gll(N)
    g = gll[str(N)]
    assert sha(x) == g["sha_x"]
    assert sha(w) == g["sha_w"]
    assert sha(D) == g["sha_D"]
    assert sha(Dzz) == g["sha_Dzz"]


@pytest.mark.parametrize("N", [5, 17, 33, 65])
def test_gll_mathematics(N):
    """Independent of the reference: nodes are the roots of (1-x^2)P'_{N-1},
    weights integrate polynomials of degree 2N-3 exactly, D differentiates
    polynomials of degree N-1 exactly (rows of D sum to ~0)."""
    from numpy.polynomial import legendre as L
    x, w, D, Dzz = po.gll(N)
    c = np.zeros(N)
    c[-1] = 1
    roots = np.sort(L.legroots(L.legder(c)))
    assert np.allclose(x[1:-1], roots, atol=1e-14)
    assert abs(x[0] + 1) < 1e-14 and abs(x[-1] - 1) < 1e-14
    for deg in range(0, 2 * N - 2):
        exact = 0.0 if deg % 2 else 2.0 / (deg + 1)
        assert abs(np.dot(w, x ** deg) - exact) < 1e-12 * max(1, N)
    for deg in range(0, N):
        p = x ** deg
        dp = deg * x ** max(deg - 1, 0) if deg else 0 * x
        assert np.allclose(D @ p, dp, atol=1e-9 * N ** 2)
    assert np.allclose(Dzz, D @ D, rtol=0, atol=1e-9 * np.abs(Dzz).max())


def test_lu_bitexact():
    """LU_Decomposition.cpp:7-72: factors, pivots and determinant bit-exact."""
    from golden_util import load as gl
    import tests_lu_cases as cases  # noqa: F401
    fx = gl("lu.json")
    for name, n, A in cases.lu_cases():
        rec = fx[name]
        assert sha(A) == rec["sha_input"], name
        LU, piv, det = po.lu(A)
        assert [int(v) for v in piv] == rec["pivots"], name
        assert sha(LU.ravel()) == rec["sha_factors"], name
        assert det == float.fromhex(rec["det"]), name


def test_lu_solves_linear_system():
    """Independent check: LU + LUSolve solves A x = b to backward-stable accuracy,
    and a plausible mistake (pivot on |a| instead of |a|/s) changes the pivots
    of the row-scaled case."""
    import tests_lu_cases as cases
    for name, n, A in cases.lu_cases():
        LU, piv, det = po.lu(A)
        rng = np.random.default_rng(n)
        xt = rng.standard_normal(n)
        b = A @ xt
        x = po.lu_solve(LU, piv, b)
        res = np.abs(A @ x - b).max() / (np.abs(A).max() * np.abs(x).max())
        assert res < 1e-13, (name, res)
        sign, logdet = np.linalg.slogdet(A)
        if abs(logdet) < 700:  # the reference's determinant is a plain product and overflows beyond
            assert np.sign(det) == sign and abs(math.log(abs(det)) - logdet) < 1e-5 * max(1, abs(logdet))


@pytest.mark.parametrize("key", ["8x4x3", "32x32x5", "16x64x1", "4x8x2"])
def test_fft_bitexact_and_dft(key):
    """Fourier_Quadrature.cpp:1040-1155 bit-exact; forward equals the DFT
    definition (numpy.fft.fft2) and the inverse round-trips."""
    from mpi_amd import workloads as wl
    rec = load("fft.json")[key]
    nx, ny, nz = rec["shape"]
    rng = wl.SplitMix64(77)
    # regenerate in the generator's order of cases
    order = ["8x4x3", "32x32x5", "16x64x1", "4x8x2"]
    for k in order:
        a, b, c = map(int, k.split("x"))
        vals = np.array([rng.normal() for _ in range(a * b * c)]).reshape(a, b, c)
        if k == key:
            break
    A = np.zeros((2 * nx, 2 * ny, nz))
    A[:nx, :ny, :] = vals
    assert sha(A) == rec["sha_input"]
    for isign in (1, -1):
        out = po.fft_xy(A if nz > 1 else A[:, :, 0], isign).reshape(A.shape)
        assert sha(out) == rec[f"sha_out_{isign}"]
    F = po.fft_xy(A if nz > 1 else A[:, :, 0], 1).reshape(A.shape)
    got = F[:nx, 0:2 * ny:2, :] + 1j * F[:nx, 1:2 * ny:2, :]
    assert np.allclose(got, np.fft.fft2(vals, axes=(0, 1)), rtol=0, atol=1e-12 * np.abs(vals).sum())
    back = po.fft_xy(F if nz > 1 else F[:, :, 0], -1).reshape(A.shape)
    assert np.allclose(back[:nx, :ny, :], vals, atol=1e-13)


@pytest.mark.parametrize("name", ["shipped_np1", "shipped_np4", "config1_N33_np1", "config1_N9_np1",
                                  "config1_N9_np4", "config2_N33_np4", "config2_N129_np1"])
def test_solve_bitexact(name):
    """SolvePDE (NonPeriodic3D_MPI.cpp:237-591) with `size` ranks simulated:
    bit-exact against the reference at np=1 and np=4."""
    rec = load(f"solve_{name}.json")
    d = rec["desc"]
    desc = po.make_desc(d["Nx"], d["Ny"], d["Nz"], d["N"], size=d["size"], xl=d["xl"], xr=d["xr"],
                        yl=d["yl"], yr=d["yr"], zl=d["zl"], zr=d["zr"], alphal=d["alphal"],
                        betal=d["betal"], alphar=d["alphar"], betar=d["betar"], lam=d["lam"])
    z = po.z_grid(desc)
    rec, f, gl, gr, ex = case_inputs(name, z)
    u = po.solve(desc, f, gl, gr)
    assert sha(u) == rec["sha_u"]
    if ex is not None:
        assert np.abs(u - ex).max() == rec["err_vs_exact"]


def test_manufactured_solution_and_rank_independence():
    """Independent of the reference: config 1 reproduces the manufactured
    u = sin x cos 2y (1-z^2) e^z to spectral accuracy, and the 4-element
    discretisation agrees whether it is split over 1 or 4 ranks."""
    from mpi_amd import workloads as wl
    out = {}
    for N, size in ((33, 1), (9, 1), (9, 4)):
        desc = po.make_desc(32, 32, 33, N, size=size, xl=0, xr=2 * math.pi, yl=0, yr=2 * math.pi,
                            zl=-1, zr=1, lam=1.0)
        x, y, zeta = wl.grid(32, 32, po.z_grid(desc))
        f, gl, gr, ue = wl.manufactured_config1(x, y, zeta)
        u = po.solve(desc, f, gl, gr)
        out[(N, size)] = u
        tol = 1e-13 if N == 33 else 2e-10
        assert np.abs(u - ue).max() < tol
    assert np.abs(out[(9, 1)] - out[(9, 4)]).max() < 1e-13


def test_robin_residual():
    """Independent of the reference: the solution of a Robin problem satisfies
    the collocation equations (x-y spectral Laplacian + Dzz) at interior nodes
    and the Robin rows at both faces."""
    N, nx, ny = 17, 16, 8
    al, bl, ar, br, lam = 0.7, 1.3, 0.4, 2.0, 3.0
    desc = po.make_desc(nx, ny, N, N, size=1, xl=0, xr=2 * math.pi, yl=0, yr=2 * math.pi, zl=0, zr=2,
                        alphal=al, betal=bl, alphar=ar, betar=br, lam=lam)
    rng = np.random.default_rng(3)
    K = po.total_z_points(desc)
    f = rng.standard_normal((nx, ny, K))
    gl = rng.standard_normal((nx, ny))
    gr = rng.standard_normal((nx, ny))
    u = po.solve(desc, f, gl, gr)
    _, Dz, Dzz = po.plan_operators(N, 1, 1, 0, 2.0)
    kx = np.fft.fftfreq(nx, 1.0 / nx)
    ky = np.fft.fftfreq(ny, 1.0 / ny)
    U = np.fft.fft2(u, axes=(0, 1))
    lap = np.fft.ifft2(-(kx[:, None, None] ** 2 + ky[None, :, None] ** 2) * U, axes=(0, 1)).real
    res = lap + u @ Dzz.T - lam * u - f
    assert np.abs(res[:, :, 1:-1]).max() < 1e-9 * np.abs(f).max()
    du = u @ Dz.T
    assert np.abs(bl * u[:, :, 0] - al * du[:, :, 0] - gl).max() < 1e-10
    assert np.abs(br * u[:, :, -1] + ar * du[:, :, -1] - gr).max() < 1e-10


def _pdesc(d, size=None, nz=None):
    return po.make_desc(d["Nx"], d["Ny"], nz or d["Nz"], d["N"], size=size or d["size"], xl=d["xl"], xr=d["xr"],
                        yl=d["yl"], yr=d["yr"], zl=d["zl"], zr=d["zr"], lam=d["lam"])


@pytest.mark.parametrize("name", ["shipped_np1", "random_64_N33", "random_128_N129", "random_32_N9"])
def test_periodic_solve_bitexact(name):
    """Periodic-z SolvePDE (Periodic3D_MPI.cpp:233-863) at np=1: bit-exact
    against the reference build."""
    from golden_util import periodic_case_inputs
    rec = load(f"psolve_{name}.json")
    desc = _pdesc(rec["desc"])
    rec, f, ex = periodic_case_inputs(name, po.z_grid(desc))
    u = po.solve_periodic(desc, f)
    assert sha(u) == rec["sha_u"]
    if ex is not None:
        assert np.abs(u - ex).max() == rec["err_vs_exact"]


def test_periodic_four_ranks_same_system():
    """The reference at np=4 (7 elements per rank) solves the 28-element
    system; the one-rank restatement of that system agrees to round-off."""
    from golden_util import periodic_case_inputs, sample
    rec = load("psolve_shipped_np4.json")
    d = rec["desc"]
    desc4 = _pdesc(d)
    rec, f, ex = periodic_case_inputs("shipped_np4", po.z_grid(desc4))
    desc1 = _pdesc(d, size=1, nz=280)  # ceil(280/10) = 28 elements on one rank
    assert po.total_z_points(desc1) == po.total_z_points(desc4)
    u = po.solve_periodic(desc1, f)
    got, want = sample(u, rec)
    assert np.abs(got - want).max() <= 1e-12 * rec["max_abs_u"]


def test_periodic_collocation_residual():
    """Independent of the reference: interior collocation rows, u(0) = u(L),
    and continuity of u_z at every element interface including the periodic
    wrap hold for the solution."""
    nx, ny, N, ne, lam = 8, 4, 9, 3, 2.0
    desc = po.make_desc(nx, ny, ne * N, N, size=1, xl=0, xr=2 * math.pi, yl=0, yr=2 * math.pi, zl=0, zr=3, lam=lam)
    K = po.total_z_points(desc)
    assert K == ne * (N - 1) + 1
    rng = np.random.default_rng(5)
    f = rng.standard_normal((nx, ny, K))
    u = po.solve_periodic(desc, f)
    _, Dz, Dzz = po.plan_operators(N, ne, 1, 0, 3.0)
    kx = np.fft.fftfreq(nx, 1.0 / nx)
    ky = np.fft.fftfreq(ny, 1.0 / ny)
    lap = np.fft.ifft2(-(kx[:, None, None] ** 2 + ky[None, :, None] ** 2) * np.fft.fft2(u, axes=(0, 1)), axes=(0, 1)).real
    assert np.abs(u[:, :, 0] - u[:, :, -1]).max() < 1e-12 * np.abs(u).max()
    dl, dr = [], []
    for e in range(ne):
        ue = u[:, :, e * (N - 1):e * (N - 1) + N]
        fe = f[:, :, e * (N - 1):e * (N - 1) + N]
        res = lap[:, :, e * (N - 1):e * (N - 1) + N] + ue @ Dzz.T - lam * ue - fe
        assert np.abs(res[:, :, 1:-1]).max() < 1e-10 * np.abs(f).max()
        du = ue @ Dz.T
        dl.append(du[:, :, 0])
        dr.append(du[:, :, -1])
    scale = max(np.abs(d).max() for d in dl)
    for e in range(ne):
        assert np.abs(dr[e] - dl[(e + 1) % ne]).max() < 1e-9 * scale


def test_fderiv_fixture_convention():
    """The reference's Fourier derivatives (Fourier_Quadrature.cpp:1494-1800)
    are the spectral derivative with wavenumber j for j < L/2 and j - L
    otherwise (the Nyquist point as -L/2): checked on its own output."""
    from mpi_amd import workloads as wl
    from golden_util import unhex
    fx = load("fderiv.json")
    rng = wl.SplitMix64(91)
    for key in ["16x8x3", "64x32x5", "8x64x2"]:
        rec = fx[key]
        nx, ny, nz = rec["shape"]
        F = np.array([rng.normal() for _ in range(nx * ny * nz)]).reshape(nx, ny, nz)
        assert sha(F) == rec["sha_input"]
        for w, (axis, order) in enumerate([(0, 1), (1, 1), (0, 2), (1, 2)]):
            Lp = (nx, ny)[axis]
            k = np.arange(Lp, dtype=np.float64)
            k = np.where(k < Lp // 2, k, k - Lp)
            c = 2 * math.pi * k / rec["L"][axis]
            shape = [1, 1, 1]
            shape[axis] = Lp
            mult = (1j * c) ** order
            want = np.fft.ifft(np.fft.fft(F, axis=axis) * mult.reshape(shape), axis=axis).real
            got = unhex(rec["out"][w] if rec["out"] else rec["head"][w])
            assert np.abs(got - want.ravel()[:len(got)]).max() <= 1e-12 * rec["max_abs"][w]
