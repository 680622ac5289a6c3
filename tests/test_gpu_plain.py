"""PLAIN_LU form on the GPU against the oracle (the reference algorithm).

Bar: max |u_gpu - u_oracle| / max |u_oracle| <= 1e-12 (FP64), pivots of the
batched factorisation identical to the reference's.
"""
import math

import numpy as np
import pytest

from golden_util import case_inputs, load, sample, sha, unhex
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel_err(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def to_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def make_plan(desc_dict, form="plain", size=1):
    from mpi_amd.solver import HelmholtzPlan
    d = desc_dict
    ne = size * po.lib().hzo_elements_per_rank(__import__("ctypes").byref(
        po.make_desc(d["Nx"], d["Ny"], d["Nz"], d["N"], size=size)))
    return HelmholtzPlan(d["Nx"], d["Ny"], d["N"], ne, xl=d["xl"], xr=d["xr"], yl=d["yl"], yr=d["yr"],
                         zl=d["zl"], zr=d["zr"], lam=d["lam"], alpha_l=d["alphal"], beta_l=d["betal"],
                         alpha_r=d["alphar"], beta_r=d["betar"], form=form)


@pytest.mark.parametrize("nx,ny,lanes", [(32, 32, 33), (8, 4, 3), (256, 256, 129), (16, 64, 5), (2, 2, 1),
                                         (512, 512, 7), (1024, 8, 2), (128, 128, 65), (128, 1024, 17),
                                         (1024, 128, 1), (1, 64, 4), (64, 1, 4), (1, 1, 9)])
def test_fourier_xy_matches_dft(cuda, nx, ny, lanes):
    """Forward half spectrum equals numpy's rfft2 over (x, y); inverse round-trips.
    Every length from 1 to 1024, and lane counts that leave a one-lane
    remainder group (17, 33, 65, 129) or a single lane."""
    from mpi_amd.solver import HelmholtzPlan
    plan = HelmholtzPlan(nx, ny, 5, 1)
    rng = np.random.default_rng(nx * 7 + ny)
    x = rng.standard_normal((nx, ny, lanes))
    re, im = plan.fourier_xy(to_dev(x))
    F = np.fft.rfft2(x, axes=(0, 1))
    got = re.cpu().numpy() + 1j * im.cpu().numpy()
    assert np.abs(got - F).max() <= 1e-13 * np.abs(F).max() * math.log2(nx * ny)
    back = plan.inverse_fourier_xy(re, im).cpu().numpy()
    assert np.abs(back - x).max() <= 1e-14 * math.log2(nx * ny) * np.abs(x).max()


def test_lu_factor_matches_reference_pivots(cuda):
    """Batched factorisation (the drop-in's LUDecomposition): pivots, factors
    and determinant identical to the reference's, bit for bit."""
    import torch
    import tests_lu_cases as cases
    from mpi_amd.solver import lu_factor_batched
    fx = load("lu.json")
    for name, n, A in cases.lu_cases():
        rec = fx[name]
        LU_ref, piv_ref, det_ref = po.lu(A)
        dA = to_dev(A[None])
        piv, det, st = lu_factor_batched(dA)
        torch.cuda.synchronize()
        assert [int(v) for v in piv.cpu().numpy()[0, : n - 1]] == rec["pivots"], name
        LU = dA.cpu().numpy()[0]
        assert sha(LU) == rec["sha_factors"], name
        assert np.array_equal(LU, LU_ref), name
        d = det.cpu().numpy()[0]
        assert d == float.fromhex(rec["det"]) or (np.isnan(d) and np.isnan(det_ref)), name
        assert int(st.cpu().numpy()[0]) == 0


def test_lu_factor_singular_flag(cuda):
    """A matrix with a zero column fails the reference's pivot test (LU_Decomposition.cpp:49-50)."""
    import torch
    from mpi_amd.solver import lu_factor_batched
    A = np.random.default_rng(0).standard_normal((2, 6, 6))
    A[1, :, 2] = 0.0
    piv, det, st = lu_factor_batched(to_dev(A))
    torch.cuda.synchronize()
    assert list(st.cpu().numpy()) == [0, 4]


def _gpu_solve(name, form="plain"):
    rec = load(f"solve_{name}.json")
    d = rec["desc"]
    plan = make_plan(d, form=form, size=d["size"])
    z = plan.z_nodes()
    rec, f, gl, gr, ex = case_inputs(name, po.z_grid(po.make_desc(d["Nx"], d["Ny"], d["Nz"], d["N"],
                                                                  size=d["size"], zl=d["zl"], zr=d["zr"])))
    u = plan.solve(to_dev(f), to_dev(gl), to_dev(gr)).cpu().numpy()
    plan.status()
    return rec, plan, f, gl, gr, ex, u


@pytest.mark.parametrize("name", ["shipped_np1", "config1_N33_np1", "config1_N9_np1"])
def test_plain_solve_matches_oracle(cuda, name):
    rec, plan, f, gl, gr, ex, u = _gpu_solve(name)
    d = rec["desc"]
    desc = po.make_desc(d["Nx"], d["Ny"], d["Nz"], d["N"], size=1, xl=d["xl"], xr=d["xr"], yl=d["yl"],
                        yr=d["yr"], zl=d["zl"], zr=d["zr"], alphal=d["alphal"], betal=d["betal"],
                        alphar=d["alphar"], betar=d["betar"], lam=d["lam"])
    uo = po.solve(desc, f, gl, gr)
    assert sha(uo) == rec["sha_u"]
    assert rel_err(u, uo) <= TOL
    if ex is not None:
        assert np.abs(u - ex).max() <= 2.0 * rec["err_vs_exact"] + 1e-14


def test_plain_solve_config2_matches_reference(cuda):
    """256 x 256 x 129 (N = 129), lambda = 10, the config the headline metric is
    quoted on: the reference's sampled output (4096 points) and max-norm."""
    rec, plan, f, gl, gr, ex, u = _gpu_solve("config2_N129_np1")
    got, want = sample(u, rec)
    assert np.abs(got - want).max() <= TOL * rec["max_abs_u"]
    assert abs(np.abs(u).max() - rec["max_abs_u"]) <= TOL * rec["max_abs_u"]


def test_plain_solve_config2_multi_element(cuda):
    """Same problem with 4 elements of N = 33 (reference at np = 4, Ns = 1)."""
    rec, plan, f, gl, gr, ex, u = _gpu_solve("config2_N33_np4")
    got, want = sample(u, rec)
    assert np.abs(got - want).max() <= 1e-11 * rec["max_abs_u"]


def test_plain_solve_robin_against_oracle(cuda):
    """General Robin coefficients on both faces and 3 elements."""
    from mpi_amd.solver import HelmholtzPlan
    nx, ny, N, ne = 16, 32, 12, 3
    al, bl, ar, br, lam = 0.7, 1.3, 0.4, 2.0, 3.0
    plan = HelmholtzPlan(nx, ny, N, ne, xl=0, xr=2 * math.pi, yl=0, yr=4.0, zl=0, zr=2, lam=lam,
                         alpha_l=al, beta_l=bl, alpha_r=ar, beta_r=br)
    desc = po.make_desc(nx, ny, ne * (N - 1) + 1 - 1, N, size=1, xl=0, xr=2 * math.pi, yl=0, yr=4.0, zl=0, zr=2,
                        alphal=al, betal=bl, alphar=ar, betar=br, lam=lam)
    desc.Nz = ne * N
    assert po.total_z_points(desc) == plan.nz
    rng = np.random.default_rng(5)
    f = rng.standard_normal((nx, ny, plan.nz))
    gl = rng.standard_normal((nx, ny))
    gr = rng.standard_normal((nx, ny))
    u = plan.solve(to_dev(f), to_dev(gl), to_dev(gr)).cpu().numpy()
    uo = po.solve(desc, f, gl, gr)
    assert rel_err(u, uo) <= TOL


def test_plain_solve_operators_callback(cuda):
    """Caller-assembled operators (the callback contract): diagonal and full
    matrices that equal the standard shift reproduce hz_solve exactly in
    structure and to rounding in value."""
    import torch
    from mpi_amd.solver import HelmholtzPlan
    nx, ny, N = 8, 16, 17
    plan = HelmholtzPlan(nx, ny, N, 1, lam=2.5)
    _, Dzz = plan.operators()
    n = N - 2
    rng = np.random.default_rng(9)
    f = rng.standard_normal((nx, ny, plan.nz))
    gl = rng.standard_normal((nx, ny))
    gr = rng.standard_normal((nx, ny))
    u0 = plan.solve(to_dev(f), to_dev(gl), to_dev(gr)).cpu().numpy()
    M = ny // 2 + 1
    diag = np.zeros((plan.num_modes, n))
    full = np.zeros((plan.num_modes, n, n))
    for iw in range(nx // 2 + 1):
        for jw in range(M):
            s = po.shift(2.5, iw, jw, 2 * math.pi, 2 * math.pi)
            diag[iw * M + jw] = Dzz.diagonal()[1:-1] - s
            full[iw * M + jw] = Dzz[1:-1, 1:-1] - s * np.eye(n)
    u1 = plan.solve_operators(to_dev(f), to_dev(gl), to_dev(gr), op_diag=to_dev(diag)).cpu().numpy()
    u2 = plan.solve_operators(to_dev(f), to_dev(gl), to_dev(gr), op_full=to_dev(full)).cpu().numpy()
    torch.cuda.synchronize()
    assert np.array_equal(u0, u1)
    assert np.array_equal(u0, u2)


def test_singular_operator_reported(cuda):
    """An operator whose first column vanishes fails the reference's pivot test
    (LU_Decomposition.cpp:49-50); the solve reports HZ_ERR_SINGULAR."""
    from mpi_amd._lib import HZ_ERR_SINGULAR, HzError
    from mpi_amd.solver import HelmholtzPlan
    plan = HelmholtzPlan(8, 8, 9, 1, lam=1.0)
    n = 7
    z = to_dev(np.zeros((8, 8, plan.nz)))
    g = to_dev(np.zeros((8, 8)))
    full = np.broadcast_to(np.eye(n), (plan.num_modes, n, n)).copy()
    plan.solve_operators(z, g, g, op_full=to_dev(full))
    plan.status()
    full[3, :, 0] = 0.0
    plan.solve_operators(z, g, g, op_full=to_dev(full))
    with pytest.raises(HzError) as e:
        plan.status()
    assert e.value.code == HZ_ERR_SINGULAR


@pytest.mark.parametrize("N,ne,robin", [(9, 1, False), (17, 1, False), (33, 2, True), (65, 1, True), (17, 8, True)])
def test_plain_every_lu_instance(cuda, N, ne, robin):
    """Every instance of the reference-algorithm mode kernel that holds the
    plan (column capacity 16, 32, 64, 128, the plan-time tuner's candidates)
    gives the same solve bit for bit, and that solve matches the oracle
    within max(1e-12, 2 x its rounding spread)."""
    import torch
    from mpi_amd.solver import HelmholtzPlan
    nx, ny, lam = 16, 8, 3.0
    al, bl, ar, br = (0.6, 1.2, 0.3, 0.9) if robin else (0.0, 1.0, 0.0, 1.0)
    plan = HelmholtzPlan(nx, ny, N, ne, xl=0, xr=2 * math.pi, yl=0, yr=3.0, zl=0, zr=2, lam=lam,
                         alpha_l=al, beta_l=bl, alpha_r=ar, beta_r=br, form="plain")
    shapes = plan.diag_candidates()
    t = plan.tuning()
    assert shapes and t["diag_tile_rows"] in shapes
    assert len(t["choice_ms"]) == (len(shapes) if len(shapes) > 1 else 0)
    rng = np.random.default_rng(N + 10 * ne)
    f = rng.standard_normal((nx, ny, plan.nz))
    gl = rng.standard_normal((nx, ny))
    gr = rng.standard_normal((nx, ny))
    outs = []
    for cols in shapes:
        print(f"case N={N} ne={ne} robin={robin} LU instance {cols} columns", flush=True)
        plan.set_diag_tile(cols)
        outs.append(plan.solve(to_dev(f), to_dev(gl), to_dev(gr)))
        plan.status()
    for cols, u in zip(shapes[1:], outs[1:]):
        assert torch.equal(u, outs[0]), f"instance {cols} differs from instance {shapes[0]}"
    desc = po.make_desc(nx, ny, ne * N, N, size=1, xl=0, xr=2 * math.pi, yl=0, yr=3.0, zl=0, zr=2,
                        alphal=al, betal=bl, alphar=ar, betar=br, lam=lam)
    uo = po.solve(desc, f, gl, gr)
    uf = po.solve(desc, f, gl, gr, fma=True)
    bar = max(1e-12, 2.0 * float(np.abs(uf - uo).max() / np.abs(uo).max()))
    err = float(np.abs(outs[0].cpu().numpy() - uo).max() / np.abs(uo).max())
    print(f"  {len(shapes)} instances, rel err {err:.2e}, bar {bar:.2e}")
    assert err <= bar
    plan.close()
