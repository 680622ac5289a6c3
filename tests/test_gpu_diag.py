"""DIAG form (z-diagonalised, FP64 MFMA) against the oracle and the reference.

Bar: relative max-norm error <= 1e-12 against the FP64 oracle and the
reference's own output (measured ~6e-14 at 256 x 256 x 129).  On rough random
inputs the bar is max(1e-12, 2 x the rounding spread of the reference
algorithm itself: its restatement built with fused multiply-adds against the
plain build), the bar of the low-dimensional tests.
"""
import math

import numpy as np
import pytest

from golden_util import case_inputs, load, sample
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
TOL = 1e-12


def spread_bar(desc, f, gl, gr, uo):
    """max(TOL, 2 x how far rounding alone moves the reference algorithm on
    this input: the oracle built with fused multiply-adds against the
    oracle)."""
    uf = po.solve(desc, f, gl, gr, fma=True)
    return max(TOL, 2.0 * float(np.abs(uf - uo).max() / np.abs(uo).max()))


def to_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def rel_err(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def solve_case(name, form="diag"):
    from mpi_amd.solver import HelmholtzPlan
    rec = load(f"solve_{name}.json")
    d = rec["desc"]
    desc = po.make_desc(d["Nx"], d["Ny"], d["Nz"], d["N"], size=d["size"], xl=d["xl"], xr=d["xr"], yl=d["yl"],
                        yr=d["yr"], zl=d["zl"], zr=d["zr"], alphal=d["alphal"], betal=d["betal"],
                        alphar=d["alphar"], betar=d["betar"], lam=d["lam"])
    K = po.total_z_points(desc)
    ne = (K - 1) // (d["N"] - 1)
    plan = HelmholtzPlan(d["Nx"], d["Ny"], d["N"], ne, xl=d["xl"], xr=d["xr"], yl=d["yl"], yr=d["yr"],
                         zl=d["zl"], zr=d["zr"], lam=d["lam"], alpha_l=d["alphal"], beta_l=d["betal"],
                         alpha_r=d["alphar"], beta_r=d["betar"], form=form)
    rec, f, gl, gr, ex = case_inputs(name, po.z_grid(desc))
    u = plan.solve(to_dev(f), to_dev(gl), to_dev(gr)).cpu().numpy()
    return rec, desc, f, gl, gr, ex, u


@pytest.mark.parametrize("name", ["config1_N33_np1", "config1_N9_np1", "shipped_np1"])
def test_diag_matches_oracle(cuda, name):
    rec, desc, f, gl, gr, ex, u = solve_case(name)
    desc.size = 1
    uo = po.solve(desc, f, gl, gr)
    print(f"{name}: rel err {rel_err(u, uo):.2e}")
    assert rel_err(u, uo) <= TOL
    if ex is not None:
        assert np.abs(u - ex).max() <= 10 * rec["err_vs_exact"] + 1e-13


def test_diag_config2_matches_reference(cuda):
    rec, desc, f, gl, gr, ex, u = solve_case("config2_N129_np1")
    got, want = sample(u, rec)
    print(f"config2 diag: rel err vs the reference {np.abs(got - want).max() / rec['max_abs_u']:.2e}")
    assert np.abs(got - want).max() <= TOL * rec["max_abs_u"]


def test_diag_config2_multi_element(cuda):
    rec, desc, f, gl, gr, ex, u = solve_case("config2_N33_np4")
    got, want = sample(u, rec)
    assert np.abs(got - want).max() <= TOL * rec["max_abs_u"]


@pytest.mark.parametrize("N,ne,robin", [(17, 3, True), (65, 1, True), (257, 1, False), (130, 1, True),
                                        (130, 1, False), (34, 1, False), (6, 1, False), (5, 1, False)])
def test_diag_general_against_oracle(cuda, N, ne, robin):
    """Robin faces, several elements, and the largest element size (N = 257);
    one Dirichlet element (the even/odd split with its first-order
    correction) with an even interior size (no centre entry: N = 130, 34, 6)
    and the smallest odd one (N = 5, one odd coordinate)."""
    from mpi_amd.solver import HelmholtzPlan
    nx, ny = 16, 8
    al, bl, ar, br = (0.7, 1.3, 0.4, 2.0) if robin else (0.0, 1.0, 0.0, 1.0)
    lam = 3.0
    plan = HelmholtzPlan(nx, ny, N, ne, xl=0, xr=2 * math.pi, yl=0, yr=3.0, zl=0, zr=2, lam=lam,
                         alpha_l=al, beta_l=bl, alpha_r=ar, beta_r=br, form="diag")
    desc = po.make_desc(nx, ny, ne * N, N, size=1, xl=0, xr=2 * math.pi, yl=0, yr=3.0, zl=0, zr=2,
                        alphal=al, betal=bl, alphar=ar, betar=br, lam=lam)
    assert po.total_z_points(desc) == plan.nz
    rng = np.random.default_rng(N + ne)
    f = rng.standard_normal((nx, ny, plan.nz))
    gl = rng.standard_normal((nx, ny))
    gr = rng.standard_normal((nx, ny))
    u = plan.solve(to_dev(f), to_dev(gl), to_dev(gr)).cpu().numpy()
    uo = po.solve(desc, f, gl, gr)
    err, bar = rel_err(u, uo), spread_bar(desc, f, gl, gr, uo)
    print(f"N={N} ne={ne} robin={robin}: rel err {err:.2e}, bar {bar:.2e}")
    assert err <= bar


@pytest.mark.parametrize("N", [33, 129, 257])
def test_diag_plan_autotune(cuda, N):
    """The plan-time tuner picks one of its candidate shapes, and the tuned
    plan solves to the oracle's answer."""
    from mpi_amd.solver import HelmholtzPlan
    plan = HelmholtzPlan(64, 16, N, 1, lam=2.0, form="diag")
    t = plan.tuning()
    print(f"N={N}: {t}")
    assert t["diag_tile_rows"] in set(plan.diag_candidates()) and len(t["choice_ms"]) >= 2
    assert set(t["choices"]) <= set(plan.diag_candidates())
    # nx = 64 offers three lane widths (passes 1, 2), ny = 16 one shape (passes 0, 3)
    assert t["fft_lanes"][0] == 0 and t["fft_lanes"][3] == 0
    assert t["fft_lanes"][1] in {8, 16, 32} and t["fft_lanes"][2] in {8, 16, 32}
    desc = po.make_desc(64, 16, N, N, size=1, xl=0, xr=2 * math.pi, yl=0, yr=2 * math.pi, zl=-1, zr=1, lam=2.0)
    rng = np.random.default_rng(N)
    f = rng.standard_normal((64, 16, plan.nz))
    g = rng.standard_normal((64, 16))
    u = plan.solve(to_dev(f), to_dev(g), to_dev(g)).cpu().numpy()
    uo = po.solve(desc, f, g, g)
    err, bar = rel_err(u, uo), spread_bar(desc, f, g, g, uo)
    print(f"autotuned N={N}: rel err {err:.2e}, bar {bar:.2e}")
    assert err <= bar
    # tuning again (on the plan's own buffers) leaves valid choices and the
    # same solve bit for bit
    plan.retune()
    t2 = plan.tuning()
    print(f"retuned N={N}: {t2}")
    assert t2["diag_tile_rows"] in set(plan.diag_candidates()) and len(t2["choice_ms"]) == len(t["choice_ms"])
    assert t2["fft_lanes"][0] == 0 and t2["fft_lanes"][1] in {8, 16, 32}
    u2 = plan.solve(to_dev(f), to_dev(g), to_dev(g)).cpu().numpy()
    assert np.array_equal(u2, u)
    plan.close()


@pytest.mark.parametrize("form", ["plain", "mixed"])
def test_retune_other_forms(cuda, form):
    """hz_plan_retune on a PLAIN plan re-times its LU instances and the
    transform lane widths, on a MIXED plan the lane widths only; the solve
    stays the same bit for bit."""
    from mpi_amd.solver import HelmholtzPlan
    plan = HelmholtzPlan(64, 32, 33, 1, lam=2.0, form=form)
    rng = np.random.default_rng(5)
    f = to_dev(rng.standard_normal((64, 32, plan.nz)))
    g = to_dev(rng.standard_normal((64, 32)))
    u = plan.solve(f, g, g).cpu().numpy()
    plan.retune()
    t = plan.tuning()
    if form == "mixed":
        assert t["diag_tile_rows"] == 0
    else:  # N - 2 = 31: the LU instances of 32, 64 and 128 columns hold it
        assert plan.diag_candidates() == [32, 64, 128] and t["diag_tile_rows"] in {32, 64, 128}
        assert len(t["choice_ms"]) == 3
    # nx = 64: three lane widths for the x passes; ny = 32: one shape (0) for the y passes
    assert t["fft_lanes"][0] == 0 and t["fft_lanes"][3] == 0
    assert t["fft_lanes"][1] in {8, 16, 32} and t["fft_lanes"][2] in {8, 16, 32}
    assert np.array_equal(plan.solve(f, g, g).cpu().numpy(), u)
    plan.close()


# N-2 is padded to a multiple of 32 for the FP64 MFMA contractions: 15, 31
# -> 32 (the waves split columns and row blocks two ways), 63 -> 64, 95 -> 96
# (two ways again), 127 -> 128.  (ne, robin): the plain Dirichlet tile and the
# general one with the element sweep.
@pytest.mark.parametrize("N", [17, 33, 65, 97, 129])
@pytest.mark.parametrize("ne,robin", [(1, False), (2, True)])
def test_diag_every_tile_shape(cuda, N, ne, robin):
    """Every candidate tile shape gives bit for bit the same solve (each
    output element is the same sequence of operations whatever the tile
    height), and that solve matches the oracle."""
    import torch
    from mpi_amd.solver import HelmholtzPlan
    nx, ny = 16, 8
    al, bl, ar, br = (0.6, 1.2, 0.3, 0.9) if robin else (0.0, 1.0, 0.0, 1.0)
    plan = HelmholtzPlan(nx, ny, N, ne, lam=3.0, alpha_l=al, beta_l=bl, alpha_r=ar, beta_r=br, form="diag")
    rng = np.random.default_rng(N + 7 * ne)
    f = rng.standard_normal((nx, ny, plan.nz))
    gl = rng.standard_normal((nx, ny))
    gr = rng.standard_normal((nx, ny))
    shapes = [rt for rt in plan.diag_candidates() if rt >= 2 * ne]
    assert len(shapes) >= 2
    outs = []
    for rt in shapes:
        plan.set_diag_tile(rt)
        outs.append(plan.solve(to_dev(f), to_dev(gl), to_dev(gr)))
    for rt, u in zip(shapes[1:], outs[1:]):
        assert torch.equal(u, outs[0]), f"tile {rt} differs from tile {shapes[0]}"
    desc = po.make_desc(nx, ny, ne * N, N, size=1, xl=0, xr=2 * math.pi, yl=0, yr=2 * math.pi, zl=-1, zr=1,
                        alphal=al, betal=bl, alphar=ar, betar=br, lam=3.0)
    uo = po.solve(desc, f, gl, gr)
    err, bar = rel_err(outs[0].cpu().numpy(), uo), spread_bar(desc, f, gl, gr, uo)
    print(f"tiles N={N} ne={ne}: rel err {err:.2e}, bar {bar:.2e}")
    assert err <= bar
    plan.close()


def test_diag_tile_pinning_rules(cuda):
    """Only the plan's candidate shapes that hold a column's element rows are
    accepted."""
    from mpi_amd._lib import HzError
    from mpi_amd.solver import HelmholtzPlan
    small = HelmholtzPlan(16, 16, 33, 1, lam=1.0, form="diag")
    assert small.diag_candidates() == [32, 64, 128]
    for bad in (16, 24, -32):
        with pytest.raises(HzError):
            small.set_diag_tile(bad)
    small.close()
    many = HelmholtzPlan(16, 16, 129, 20, lam=1.0, form="diag")
    assert 32 in many.diag_candidates()
    with pytest.raises(HzError):
        many.set_diag_tile(32)  # 40 element rows per column
    many.set_diag_tile(48)
    many.close()


def test_probe_ceilings_are_plausible(cuda):
    """hz_probe_ceilings measures the FP64 MFMA and copy rates the roofline
    fractions are also quoted against: within (0, published peak]."""
    import ctypes
    from mpi_amd import _lib
    lib = _lib.load()
    mfma, copy = ctypes.c_double(0), ctypes.c_double(0)
    _lib.check(lib.hz_probe_ceilings(ctypes.byref(mfma), ctypes.byref(copy), 64 << 20))
    assert 5.0 < mfma.value <= 78.6 * 1.02
    assert 300.0 < copy.value <= 8000.0 * 1.5  # an Infinity-Cache-resident copy can beat HBM
    with pytest.raises(_lib.HzError):
        _lib.check(lib.hz_probe_ceilings(None, None, 1 << 20))


@pytest.mark.parametrize("nx,ny", [(1024, 16), (16, 1024), (512, 64)])
@pytest.mark.parametrize("form", ["diag", "plain"])
def test_largest_transform_lengths_solve(cuda, nx, ny, form):
    """Solves whose transforms run at the longest lengths (1024 and 512
    points: the two-stage radix kernels, on the padded spectrum rows),
    against the oracle."""
    from mpi_amd.solver import HelmholtzPlan
    N = 9
    plan = HelmholtzPlan(nx, ny, N, 1, xl=0, xr=2 * math.pi, yl=0, yr=2 * math.pi, zl=-1, zr=1, lam=2.0,
                         alpha_l=0.3, beta_l=1.0, alpha_r=0.2, beta_r=1.0, form=form)
    rng = np.random.default_rng(nx + 3 * ny)
    f = rng.standard_normal((nx, ny, plan.nz))
    gl = rng.standard_normal((nx, ny))
    gr = rng.standard_normal((nx, ny))
    u = plan.solve(to_dev(f), to_dev(gl), to_dev(gr)).cpu().numpy()
    desc = po.make_desc(nx, ny, N, N, size=1, xl=0, xr=2 * math.pi, yl=0, yr=2 * math.pi, zl=-1, zr=1,
                        alphal=0.3, betal=1.0, alphar=0.2, betar=1.0, lam=2.0)
    uo = po.solve(desc, f, gl, gr)
    err, bar = rel_err(u, uo), spread_bar(desc, f, gl, gr, uo)
    print(f"{nx}x{ny} {form}: rel err {err:.2e}, bar {bar:.2e}")
    assert err <= bar
    plan.close()
