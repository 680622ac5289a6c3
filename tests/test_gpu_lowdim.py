"""The 2-D and 1-D solvers on the GPU: plans with unit Fourier axes against
the reference's own outputs (NonPeriodic2D_MPI.cpp, NonPeriodic1D_MPI.cpp,
Periodic2D_MPI.cpp, Periodic1D_MPI.cpp; golden vectors from
tests/golden/make_golden.py --lowdim) and, at sizes it finishes quickly, the
oracle.

A 2-D problem (periodic x of Nx points, GLL y) runs as nx = 1, ny = Nx (the
half spectrum along the periodic axis, as the default mapping) and as
nx = Nx, ny = 1 (the full-spectrum x passes, the reference's own layout);
a 1-D problem as nx = ny = 1.  Bars: PLAIN, DIAG and MIXED <= 1e-12
relative max-norm error against the reference, or twice the case's
rounding_spread where that is larger: the reference's own algorithm, rebuilt
with fused multiply-adds, moves by up to 3.2e-11 on the 64-element 1-D
problems (the interface sweep amplifies rounding in its coefficients), so no
implementation with other rounding can be held to 1e-12 there.  The error
against the exact solution stays within 10x what the reference reaches.
"""
import numpy as np
import pytest

from golden_util import load, lowdim_inputs, lowdim_sample, unhex
from test_oracle_lowdim import CASES, as3d, oracle_desc
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
TOL = {"plain": 1e-12, "diag": 1e-12, "mixed": 1e-12}


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def make_plan(rec, form, mapping):
    from mpi_amd.solver import HelmholtzPlan
    p = rec["params"]
    kw = dict(lam=p["lam"], alpha_l=p["al"], beta_l=p["bl"], alpha_r=p["ar"], beta_r=p["br"], form=form,
              periodic_z=rec["periodic"])
    if form == "mixed":
        kw["refine_sweeps"] = 3
    if rec["dim"] == 2:
        nx, ny = (1, p["Nx"]) if mapping == "y" else (p["Nx"], 1)
        px = dict(xl=0.0, xr=1.0, yl=p["xl"], yr=p["xr"]) if mapping == "y" else \
            dict(xl=p["xl"], xr=p["xr"], yl=0.0, yr=1.0)
        return HelmholtzPlan(nx, ny, p["N"], rec["elements"], zl=p["yl"], zr=p["yr"], **px, **kw)
    return HelmholtzPlan(1, 1, p["N"], rec["elements"], xl=0.0, xr=1.0, yl=0.0, yr=1.0, zl=p["xl"], zr=p["xr"],
                         **kw)


def shaped(rec, mapping, f, gl, gr):
    K = rec["K"]
    lines = rec["params"]["Nx"] if rec["dim"] == 2 else 1
    shape = (1, lines) if mapping == "y" or rec["dim"] == 1 else (lines, 1)
    F = np.reshape(f, shape + (K,))
    if gl is None:
        return F, None, None
    return F, np.reshape(gl, shape), np.reshape(gr, shape)


def variants():
    out = []
    for name in CASES:
        rec = load(f"low_{name}.json")
        maps = ["y", "x"] if rec["dim"] == 2 else ["1"]
        forms = ["plain", "diag"]
        if not rec["periodic"] and rec["elements"] <= 4:
            forms.append("mixed")
        for m in maps:
            for form in forms:
                out.append((name, m, form))
    return out


@pytest.mark.parametrize("name,mapping,form", variants())
def test_lowdim_matches_reference(cuda, name, mapping, form):
    rec = load(f"low_{name}.json")
    print(f"case {name} mapping {mapping} form {form}", flush=True)
    f, gl, gr, ex = lowdim_inputs(rec)
    plan = make_plan(rec, form, mapping)
    if rec["params"]["size"] == 1:
        assert np.array_equal(plan.z_nodes(), unhex(rec["grid"]))
    F, GL, GR = shaped(rec, mapping, f, gl, gr)
    u = plan.solve(dev(F), None if GL is None else dev(GL), None if GR is None else dev(GR))
    plan.status()
    u = u.cpu().numpy().reshape(f.shape)
    plan.close()
    got, want = lowdim_sample(u, rec)
    err = float(np.abs(got - want).max() / rec["max_abs_u"])
    bar = max(TOL[form], 2 * rec["rounding_spread"])
    print(f"  rel err vs reference {err:.3e} (bar {bar:.1e})")
    assert err <= bar
    if ex is not None:
        assert np.abs(u - ex).max() <= 10 * rec["err_vs_exact"] + 1e-13
    if rec["K"] * (rec["params"]["Nx"] if rec["dim"] == 2 else 1) <= 40000 and rec["params"]["size"] == 1:
        d = oracle_desc(rec)
        F3, GL3, GR3 = as3d(rec, f, gl, gr)
        uo = (po.solve_periodic(d, F3) if rec["periodic"] else po.solve(d, F3, GL3, GR3)).reshape(f.shape)
        assert float(np.abs(u - uo).max() / np.abs(uo).max()) <= bar


@pytest.mark.parametrize("form", ["plain", "diag"])
def test_lowdim_adjoint_dot_product(cuda, form):
    """<S(f, gl, gr), ubar> = <f, fbar> + <gl, glbar> + <gr, grbar> for a
    one-element 2-D plan (nx = 1) and a 1-D plan: the adjoint is the transpose
    of the forward solve through the unit-axis transforms too."""
    import torch
    from mpi_amd.solver import HelmholtzPlan
    g = torch.Generator(device="cpu").manual_seed(5)
    for nx, ny in ((1, 64), (1, 1)):
        plan = HelmholtzPlan(nx, ny, 33, 1, zl=-1.0, zr=1.0, lam=3.0, alpha_l=0.5, beta_l=1.0, alpha_r=0.2,
                             beta_r=1.5, form=form)
        f = torch.randn((nx, ny, plan.nz), dtype=torch.float64, generator=g)
        gl = torch.randn((nx, ny), dtype=torch.float64, generator=g)
        gr = torch.randn((nx, ny), dtype=torch.float64, generator=g)
        ubar = torch.randn((nx, ny, plan.nz), dtype=torch.float64, generator=g)
        u = plan.solve(f.cuda(), gl.cuda(), gr.cuda()).cpu()
        fbar, glbar, grbar, _ = plan.adjoint(ubar.cuda())
        lhs = float((u * ubar).sum())
        rhs = float((f * fbar.cpu()).sum() + (gl * glbar.cpu()).sum() + (gr * grbar.cpu()).sum())
        plan.close()
        assert abs(lhs - rhs) <= 1e-11 * max(1.0, abs(lhs)), (nx, ny, lhs, rhs)


def test_lowdim_limits(cuda):
    """Element counts beyond one tile / one right-hand-side set are refused
    with an error, not solved wrongly."""
    from mpi_amd._lib import HzError
    from mpi_amd.solver import HelmholtzPlan
    with pytest.raises(HzError):
        HelmholtzPlan(1, 1, 8, 65, form="diag")       # 2*65 rows > the 128-row tile
    with pytest.raises(HzError):
        HelmholtzPlan(1, 1, 8, 128, form="plain")     # 2 + 2*128 right-hand sides > 256
    with pytest.raises(HzError):
        HelmholtzPlan(8, 8, 8, 64, form="plain")      # 2 + 4*64 > 256 with nx >= 4
    p = HelmholtzPlan(1, 1, 8, 127, form="plain")     # the largest PLAIN 1-D plan
    p.close()


@pytest.mark.parametrize("form", ["diag", "mixed", "plain"])
@pytest.mark.parametrize("N", [17, 129])
def test_lowdim_one_dirichlet_element(cuda, form, N):
    """One element with Dirichlet faces (the even/odd split of the DIAG and
    MIXED forms, forward and adjoint) in a 2-D plan, both mappings, and a 1-D
    plan: against the oracle (x a single point) within max(1e-12, 2 x its
    rounding spread), and <S x, ubar> = <x, S^T ubar> for the adjoint."""
    import torch
    from mpi_amd.solver import HelmholtzPlan
    rng = np.random.default_rng(N + len(form))
    lam, K = 2.5, N
    for lines, mapping in ((64, "y"), (64, "x"), (1, "1")):
        print(f"case form={form} N={N} mapping={mapping}", flush=True)
        f = rng.standard_normal((lines, K))
        gl = rng.standard_normal(lines)
        gr = rng.standard_normal(lines)
        d = po.make_desc(1, lines, N, N, size=1, xl=0.0, xr=1.0, yl=0.0, yr=3.0, zl=-1.0, zr=1.0, lam=lam)
        args = (f.reshape(1, lines, K).copy(), gl.reshape(1, lines).copy(), gr.reshape(1, lines).copy())
        uo = po.solve(d, *args).reshape(lines, K)
        uf = po.solve(d, *args, fma=True).reshape(lines, K)
        bar = max(1e-12, 2.0 * float(np.abs(uf - uo).max() / np.abs(uo).max()))
        if mapping == "x":
            shape = (lines, 1)
            plan = HelmholtzPlan(lines, 1, N, 1, xl=0.0, xr=3.0, yl=0.0, yr=1.0, zl=-1.0, zr=1.0, lam=lam, form=form)
        else:
            shape = (1, lines)
            plan = HelmholtzPlan(1, lines, N, 1, xl=0.0, xr=1.0, yl=0.0, yr=3.0, zl=-1.0, zr=1.0, lam=lam, form=form)
        F, GL, GR = dev(f.reshape(shape + (K,))), dev(gl.reshape(shape)), dev(gr.reshape(shape))
        u = plan.solve(F, GL, GR)
        plan.status()
        err = float(np.abs(u.cpu().numpy().reshape(lines, K) - uo).max() / np.abs(uo).max())
        print(f"  rel err {err:.2e}, bar {bar:.2e}")
        assert err <= bar
        ubar = torch.from_numpy(rng.standard_normal(shape + (K,))).cuda()
        fbar, glbar, grbar, _ = plan.adjoint(ubar)
        lhs = float((u * ubar).sum())
        rhs = float((F * fbar).sum() + (GL * glbar).sum() + (GR * grbar).sum())
        plan.close()
        assert abs(lhs - rhs) <= 1e-11 * max(1.0, abs(lhs)), (mapping, lhs, rhs)
