"""MIXED form: FP32 solves through the plan's eigenbasis (FP32 matrix cores)
refined with FP64 residuals (FP64 matrix cores).

Bar: relative max-norm error <= 1e-12 against the FP64 oracle / the
reference's own output (max(1e-12, 2 x the reference algorithm's own rounding
spread) on rough random inputs), with the error and the residual history
stated per refinement sweep.
"""
import math
import os

import numpy as np
import pytest

from golden_util import GOLDEN, case_inputs, load, sample
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
TOL = 1e-12


def spread_bar(desc, f, gl, gr, uo):
    """max(TOL, 2 x how far rounding alone moves the reference algorithm on
    this input: the oracle built with fused multiply-adds against the
    oracle)."""
    uf = po.solve(desc, f, gl, gr, fma=True)
    return max(TOL, 2.0 * float(np.abs(uf - uo).max() / np.abs(uo).max()))


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def rel_err(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def plan_for(d, ne, form="mixed", **kw):
    from mpi_amd.solver import HelmholtzPlan
    return HelmholtzPlan(d["Nx"], d["Ny"], d["N"], ne, xl=d["xl"], xr=d["xr"], yl=d["yl"], yr=d["yr"],
                         zl=d["zl"], zr=d["zr"], lam=d["lam"], alpha_l=d["alphal"], beta_l=d["betal"],
                         alpha_r=d["alphar"], beta_r=d["betar"], form=form, **kw)


def golden_desc(name):
    rec = load(f"solve_{name}.json")
    d = rec["desc"]
    desc = po.make_desc(d["Nx"], d["Ny"], d["Nz"], d["N"], size=d["size"], xl=d["xl"], xr=d["xr"], yl=d["yl"],
                        yr=d["yr"], zl=d["zl"], zr=d["zr"], alphal=d["alphal"], betal=d["betal"],
                        alphar=d["alphar"], betar=d["betar"], lam=d["lam"])
    K = po.total_z_points(desc)
    return rec, d, desc, (K - 1) // (d["N"] - 1)


@pytest.mark.parametrize("name", ["config1_N33_np1", "config1_N9_np1"])
def test_mixed_matches_oracle(cuda, name):
    rec, d, desc, ne = golden_desc(name)
    plan = plan_for(d, ne)
    rec, f, gl, gr, ex = case_inputs(name, po.z_grid(desc))
    u = plan.solve(dev(f), dev(gl), dev(gr)).cpu().numpy()
    uo = po.solve(desc, f, gl, gr)
    err = rel_err(u, uo)
    print(f"{name}: rel err vs oracle {err:.3e}, history {plan.mixed_history().tolist()}")
    assert err <= TOL
    plan.close()


@pytest.mark.parametrize("N,ne", [(17, 3), (65, 1), (130, 1), (257, 1)])
def test_mixed_robin_against_oracle(cuda, N, ne):
    from mpi_amd.solver import HelmholtzPlan
    nx, ny, lam = 16, 8, 3.0
    al, bl, ar, br = 0.7, 1.3, 0.4, 2.0
    plan = HelmholtzPlan(nx, ny, N, ne, xl=0, xr=2 * math.pi, yl=0, yr=3.0, zl=0, zr=2, lam=lam,
                         alpha_l=al, beta_l=bl, alpha_r=ar, beta_r=br, form="mixed")
    desc = po.make_desc(nx, ny, ne * N, N, size=1, xl=0, xr=2 * math.pi, yl=0, yr=3.0, zl=0, zr=2,
                        alphal=al, betal=bl, alphar=ar, betar=br, lam=lam)
    rng = np.random.default_rng(N + ne)
    f = rng.standard_normal((nx, ny, plan.nz))
    gl = rng.standard_normal((nx, ny))
    gr = rng.standard_normal((nx, ny))
    u = plan.solve(dev(f), dev(gl), dev(gr)).cpu().numpy()
    uo = po.solve(desc, f, gl, gr)
    err, bar = rel_err(u, uo), spread_bar(desc, f, gl, gr, uo)
    print(f"N={N} ne={ne}: rel err {err:.3e}, bar {bar:.2e}, fp64 modes {plan.mixed_fp64_modes()}/{plan.num_modes}")
    assert err <= bar
    plan.close()


def test_mixed_config2_matches_reference(cuda):
    rec, d, desc, ne = golden_desc("config2_N129_np1")
    plan = plan_for(d, ne)
    rec, f, gl, gr, _ = case_inputs("config2_N129_np1", po.z_grid(desc))
    u = plan.solve(dev(f), dev(gl), dev(gr)).cpu().numpy()
    got, want = sample(u, rec)
    err = float(np.abs(got - want).max() / rec["max_abs_u"])
    hist = plan.mixed_history()
    print(f"config2 mixed: rel err {err:.3e}, history {hist.tolist()}, "
          f"fp64 modes {plan.mixed_fp64_modes()}/{plan.num_modes}")
    assert err <= TOL
    assert hist[-1] < hist[0]
    plan.close()


def test_mixed_error_falls_per_sweep(cuda):
    """Error against the oracle after 0, 1, 2, 3 refinement sweeps
    (refine_sweeps = -1 asks for the FP32 solve alone)."""
    rec, d, desc, ne = golden_desc("config1_N33_np1")
    rec, f, gl, gr, ex = case_inputs("config1_N33_np1", po.z_grid(desc))
    uo = po.solve(desc, f, gl, gr)
    errs = []
    for s in (-1, 1, 2, 3):
        plan = plan_for(d, ne, refine_sweeps=s)
        u = plan.solve(dev(f), dev(gl), dev(gr)).cpu().numpy()
        errs.append(rel_err(u, uo))
        hist = plan.mixed_history()
        assert len(hist) == max(s, 0)  # the residual before each sweep
        if len(hist) > 1:
            assert hist[1] < 1e-3 * hist[0]
        plan.close()
    print("error per sweep count 0..3:", errs)
    assert errs[0] < 1e-4  # the FP32 solve alone
    assert errs[-1] <= TOL and errs[0] > 1e3 * errs[-1]


@pytest.mark.parametrize("name", ["config2_N129_np1"])
def test_mixed_all_fp32_modes_converge(cuda, name):
    """Every mode in FP32, including the worst-conditioned ones: the
    refinement still reaches the FP64 answer (history printed per sweep)."""
    rec, d, desc, ne = golden_desc(name)
    plan = plan_for(d, ne, mixed_fp64_cond=1e-300, refine_sweeps=4)
    assert plan.mixed_fp64_modes() == 0
    rec, f, gl, gr, _ = case_inputs(name, po.z_grid(desc))
    u = plan.solve(dev(f), dev(gl), dev(gr)).cpu().numpy()
    got, want = sample(u, rec)
    err = float(np.abs(got - want).max() / rec["max_abs_u"])
    print(f"{name} all-FP32: rel err {err:.3e}, history {plan.mixed_history().tolist()}")
    assert err <= TOL
    plan.close()


@pytest.mark.parametrize("N,lam", [(129, 1e-2), (257, 1e-2), (257, 1e3)])
def test_mixed_refinement_contracts_every_mode(cuda, N, lam):
    """The refinement contracts the near-singular constant mode (lambda =
    1e-2) as fast as the others: the FP32 eigenbasis solve's backward error
    does not depend on the shift.  Rough random input, against the oracle."""
    from mpi_amd.solver import HelmholtzPlan
    nx, ny = 16, 8
    plan = HelmholtzPlan(nx, ny, N, 1, lam=lam, form="mixed", refine_sweeps=2)
    assert plan.mixed_fp64_modes() == 0
    desc = po.make_desc(nx, ny, N, N, size=1, xl=0, xr=2 * math.pi, yl=0, yr=2 * math.pi, zl=-1, zr=1, lam=lam)
    rng = np.random.default_rng(N)
    f = rng.standard_normal((nx, ny, plan.nz))
    gl = rng.standard_normal((nx, ny))
    gr = rng.standard_normal((nx, ny))
    u = plan.solve(dev(f), dev(gl), dev(gr)).cpu().numpy()
    uo = po.solve(desc, f, gl, gr)
    err, bar = rel_err(u, uo), spread_bar(desc, f, gl, gr, uo)
    hist = plan.mixed_history()
    print(f"N={N} lam={lam}: rel err {err:.2e} (bar {bar:.2e}), residual before each sweep {hist.tolist()}")
    assert hist[1] < 1e-3 * hist[0]
    assert err <= bar
    plan.close()


@pytest.mark.skipif(not os.path.exists(os.path.join(GOLDEN, "solve_config4_N257_np1.json")),
                    reason="512x512x257 golden vectors not generated")
def test_mixed_config4_matches_reference(cuda):
    """512 x 512 x 257, lambda = 100: the precision workload."""
    import torch
    from mpi_amd import workloads as wl
    rec, d, desc, ne = golden_desc("config4_N257_np1")
    plan = plan_for(d, ne)
    x, y, zeta = wl.grid(512, 512, po.z_grid(desc))
    f, gl, gr = wl.random_problem(x, y, zeta, seed=4, nmodes=256, kmax=32, lmax=64)
    u = plan.solve(dev(f), dev(gl), dev(gr))
    torch.cuda.synchronize()
    got, want = sample(u.cpu().numpy(), rec)
    err = float(np.abs(got - want).max() / rec["max_abs_u"])
    print(f"config4 mixed: rel err {err:.3e}, history {plan.mixed_history().tolist()}, "
          f"fp64 modes {plan.mixed_fp64_modes()}/{plan.num_modes}")
    assert err <= TOL
    plan.close()
