"""Periodic-z solver (Periodic3D_MPI.cpp) on the GPU, PLAIN_LU, DIAG and MIXED
forms, against the reference's own output (golden vectors) and the oracle.

Bars: PLAIN, DIAG and MIXED (three refinement sweeps) <= 1e-12
relative max-norm error.
"""
import numpy as np
import pytest

from golden_util import load, periodic_case_inputs, sample
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
TOL = {"plain": 1e-12, "diag": 1e-12, "mixed": 1e-12}


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def plan_for(d, ne, form):
    from mpi_amd.solver import HelmholtzPlan
    return HelmholtzPlan(d["Nx"], d["Ny"], d["N"], ne, xl=d["xl"], xr=d["xr"], yl=d["yl"], yr=d["yr"],
                         zl=d["zl"], zr=d["zr"], lam=d["lam"], form=form, periodic_z=True)


CASES = [("shipped_np1", "plain"), ("shipped_np1", "diag")] + [
    (name, form) for name in ("random_64_N33", "random_128_N129", "random_32_N9")
    for form in ("plain", "diag", "mixed")]  # MIXED takes at most 8 elements: not the 25-element shipped case


@pytest.mark.parametrize("name,form", CASES)
def test_periodic_matches_reference(cuda, name, form):
    rec = load(f"psolve_{name}.json")
    d = rec["desc"]
    desc = po.make_desc(d["Nx"], d["Ny"], d["Nz"], d["N"], size=1, xl=d["xl"], xr=d["xr"], yl=d["yl"],
                        yr=d["yr"], zl=d["zl"], zr=d["zr"], lam=d["lam"])
    K = po.total_z_points(desc)
    ne = (K - 1) // (d["N"] - 1)
    plan = plan_for(d, ne, form)
    assert plan.nz == K
    rec, f, ex = periodic_case_inputs(name, po.z_grid(desc))
    u = plan.solve(dev(f)).cpu().numpy()
    got, want = sample(u, rec)
    err = float(np.abs(got - want).max() / rec["max_abs_u"])
    print(f"{name} {form}: rel err vs reference {err:.3e}")
    assert err <= TOL[form]
    if ex is not None:
        assert np.abs(u - ex).max() <= 10 * rec["err_vs_exact"] + 1e-13
    if K < 200:
        uo = po.solve_periodic(desc, f)
        assert float(np.abs(u - uo).max() / np.abs(uo).max()) <= TOL[form]
    plan.close()


@pytest.mark.parametrize("form", ["plain", "diag"])
def test_periodic_four_rank_reference(cuda, form):
    """The reference at np=4 (4 x 7 elements of N = 10) against a 28-element plan."""
    rec = load("psolve_shipped_np4.json")
    d = rec["desc"]
    desc4 = po.make_desc(d["Nx"], d["Ny"], d["Nz"], d["N"], size=4, xl=d["xl"], xr=d["xr"], yl=d["yl"],
                         yr=d["yr"], zl=d["zl"], zr=d["zr"], lam=d["lam"])
    plan = plan_for(d, 28, form)
    rec, f, ex = periodic_case_inputs("shipped_np4", po.z_grid(desc4))
    u = plan.solve(dev(f)).cpu().numpy()
    got, want = sample(u, rec)
    assert float(np.abs(got - want).max() / rec["max_abs_u"]) <= TOL[form]
    plan.close()


def test_periodic_rejects_unsupported(cuda):
    import torch
    from mpi_amd._lib import HzError
    from mpi_amd.solver import HelmholtzPlan
    with pytest.raises(HzError):  # MIXED: a column's 2 x 17 element rows exceed its 32-row tile
        HelmholtzPlan(8, 8, 9, 17, form="mixed", periodic_z=True)
    p = HelmholtzPlan(8, 8, 9, 2, form="diag", periodic_z=True)
    fbar, glbar, grbar, _ = p.adjoint(torch.zeros((8, 8, p.nz), dtype=torch.float64, device="cuda"))
    assert glbar is None and grbar is None and float(fbar.abs().max()) == 0.0
    p.close()


@pytest.mark.parametrize("form", ["plain", "diag", "mixed"])
def test_periodic_closure_matches_oracle(cuda, form):
    """The closure coefficients of the cyclic interface system (the
    reference's alphaPDE) from a three-element periodic solve, expanded to the
    reference's full-spectrum layout the way the drop-in does, against the
    oracle's."""
    from mpi_amd.solver import HelmholtzPlan
    nx, ny, N, ne, lam = 8, 16, 9, 3, 1.5
    plan = HelmholtzPlan(nx, ny, N, ne, xl=0, xr=2 * np.pi, yl=0, yr=3.0, zl=0, zr=2, lam=lam, form=form,
                         periodic_z=True)
    desc = po.make_desc(nx, ny, ne * N, N, size=1, xl=0, xr=2 * np.pi, yl=0, yr=3.0, zl=0, zr=2, lam=lam)
    f = np.random.default_rng(77).standard_normal((nx, ny, plan.nz))
    plan.solve(dev(f))
    half = plan.periodic_closure()
    M = ny // 2 + 1
    full = np.zeros((nx, 2 * ny))
    for i in range(nx):
        for m in range(ny):
            low = m < M
            ii, mm = (i, m) if low else ((nx - i) % nx, ny - m)
            full[i, 2 * m] = half[0, ii, mm]
            full[i, 2 * m + 1] = half[1, ii, mm] if low else -half[1, ii, mm]
    _, want = po.solve_periodic(desc, f, alpha=True)
    assert np.abs(want).max() > 0
    assert np.abs(full - want).max() <= 1e-12 * np.abs(want).max()
    one = HelmholtzPlan(8, 8, 9, 1, lam=1.0, form=form, periodic_z=True)
    one.solve(dev(np.ones((8, 8, one.nz))))
    assert not one.periodic_closure().any()  # one element: no closure, as in the reference
    # tuning again runs the mode kernels on noise: the last solve's closures
    # are cleared, and the next solve gives them again bit for bit
    plan.retune()
    assert not plan.periodic_closure().any()
    plan.status()
    plan.solve(dev(f))
    assert np.array_equal(plan.periodic_closure(), half)
    plan.close()
    one.close()
