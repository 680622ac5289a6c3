"""Many problems of uneven size in one batch (hz_batch_*).

Every usable variant (launches per stage and task kind, or per stage over every kind, both
tile heights) must reproduce each problem's single-plan DIAG solve, which is
itself pinned to the oracle; a few problems are also checked against the
oracle directly.  Bar: relative max-norm difference <= 1e-13 against the
single-plan solve; against the oracle <= max(1e-12, 2 x the reference
algorithm's own rounding spread: the oracle built with fused multiply-adds
against the plain build) on these rough random inputs.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


# (nx, ny, N, ne, lam, robin)
MIXED_SET = [(16, 32, 17, 1, 0.5, False), (32, 16, 33, 2, 3.0, True), (64, 64, 65, 1, 10.0, False),
             (128, 32, 129, 1, 1e-2, False), (32, 128, 97, 1, 1e3, True), (256, 16, 21, 3, 2.0, False),
             (16, 256, 9, 4, 7.0, True)]


def make_set(spec, forms=None):
    from mpi_amd.solver import HelmholtzPlan
    plans, inputs, descs = [], [], []
    for k, (nx, ny, N, ne, lam, robin) in enumerate(spec):
        al, bl, ar, br = (0.7, 1.3, 0.4, 2.0) if robin else (0.0, 1.0, 0.0, 1.0)
        p = HelmholtzPlan(nx, ny, N, ne, xl=0, xr=2 * math.pi, yl=0, yr=3.0, zl=0, zr=2, lam=lam,
                          alpha_l=al, beta_l=bl, alpha_r=ar, beta_r=br, form=forms[k] if forms else "diag")
        rng = np.random.default_rng(100 + k)
        f = rng.standard_normal((nx, ny, p.nz))
        gl = rng.standard_normal((nx, ny))
        gr = rng.standard_normal((nx, ny))
        plans.append(p)
        inputs.append((f, gl, gr))
        descs.append(po.make_desc(nx, ny, ne * N, N, size=1, xl=0, xr=2 * math.pi, yl=0, yr=3.0, zl=0, zr=2,
                                  alphal=al, betal=bl, alphar=ar, betar=br, lam=lam))
    return plans, inputs, descs


FUSED_OFF = ("variants 0-3, a persistent launch whose tasks waited on per-problem completion counters, were "
             "removed: its first GPU run did not finish and the cause was not established")
# The cooperative fused launch (variants 6 and 7) is refused: variant 6
# reproduced the staged results bit for bit (512 problems; periodic, 2-D and
# 1-D sets), but variant 7's first run did not finish, nor did a run of the
# MIXED-plan batch test below while variant 6 was timed and forced on it; the
# cause is not established.
COOP_OFF = ("fused cooperative launch (variants 6, 7) switched off: two GPU runs did not finish; cause not "
            "established")
STAGED, MERGED = (4, 5), (8, 9)


@pytest.mark.parametrize("variant", [pytest.param(v, marks=pytest.mark.skip(reason=FUSED_OFF)) for v in range(4)]
                         + list(STAGED) + [pytest.param(v, marks=pytest.mark.skip(reason=COOP_OFF)) for v in (6, 7)]
                         + list(MERGED))
def test_batch_variant_matches_single_solves(cuda, variant):
    from mpi_amd.solver import BatchSolver
    plans, inputs, _ = make_set(MIXED_SET)
    fs = [dev(f) for f, _, _ in inputs]
    gls = [dev(g) for _, g, _ in inputs]
    grs = [dev(g) for _, _, g in inputs]
    want = [p.solve(f, gl, gr).cpu().numpy() for p, f, gl, gr in zip(plans, fs, gls, grs)]
    b = BatchSolver(plans)
    b.force_variant(variant)
    us = b.solve(fs, gls, grs)
    b.status()
    for k, (u, w) in enumerate(zip(us, want)):
        e = rel(u.cpu().numpy(), w)
        assert e <= 1e-13, f"problem {k} ({MIXED_SET[k]}): {e:.3e}"
    b.close()


def test_batch_against_oracle(cuda):
    from mpi_amd.solver import BatchSolver
    plans, inputs, descs = make_set(MIXED_SET)
    b = BatchSolver(plans)
    us = b.solve([dev(f) for f, _, _ in inputs], [dev(g) for _, g, _ in inputs], [dev(g) for _, _, g in inputs])
    b.status()
    print("tuning:", b.tuning())
    for k in (0, 1, 4, 6):
        f, gl, gr = inputs[k]
        uo = po.solve(descs[k], f, gl, gr)
        e = rel(us[k].cpu().numpy(), uo)
        bar = max(1e-12, 2 * rel(po.solve(descs[k], f, gl, gr, fma=True), uo))
        print(f"problem {k}: rel err {e:.2e}, bar {bar:.2e}")
        assert e <= bar, f"problem {k}: {e:.3e}"
    b.close()


def test_batch_graph_replay_and_new_buffers(cuda):
    """Replays of the recorded graph, with unrelated GPU work and host
    synchronisation between them, reproduce the first result exactly; new
    output buffers are picked up."""
    import torch
    from mpi_amd.solver import BatchSolver
    plans, inputs, _ = make_set(MIXED_SET[:4])
    fs = [dev(f) for f, _, _ in inputs]
    gls = [dev(g) for _, g, _ in inputs]
    grs = [dev(g) for _, _, g in inputs]
    b = BatchSolver(plans)
    first = [u.clone() for u in b.solve(fs, gls, grs)]
    for rep in range(3):
        junk = torch.randn(1 << 20, device="cuda", dtype=torch.float64)
        junk.mul_(2.0).sum().item()
        torch.cuda.synchronize()
        us = b.solve(fs, gls, grs)
        torch.cuda.synchronize()
        for u, w in zip(us, first):
            assert torch.equal(u, w), f"replay {rep} differs"
    # same shapes, fresh input values in fresh buffers
    fs2 = [f * 2.0 for f in fs]
    us2 = b.solve(fs2, gls, grs)
    for p, u, f, gl, gr in zip(plans, us2, fs2, gls, grs):
        w = p.solve(f, gl, gr)
        assert rel(u.cpu().numpy(), w.cpu().numpy()) <= 1e-13
    b.close()


def test_batch_many_problem_workload(cuda):
    """512 problems: nx, ny in {32, 64, 128}, N odd in [17, 129], lambda
    log-uniform in [1e-2, 1e3]; every one checked against its single solve."""
    from mpi_amd import workloads as wl
    from mpi_amd.solver import BatchSolver, HelmholtzPlan
    specs = wl.batch_specs(512, seed=3)
    plans, fs, gls, grs = [], [], [], []
    for s in specs:
        p = HelmholtzPlan(s["nx"], s["ny"], s["N"], 1, lam=s["lam"], form="diag")
        f, gl, gr = wl.batch_fields(s, p.z_nodes() - 1.0)
        plans.append(p)
        fs.append(dev(f))
        gls.append(dev(gl))
        grs.append(dev(gr))
    b = BatchSolver(plans)
    us = b.solve(fs, gls, grs)
    b.status()
    print("tuning:", b.tuning())
    worst = 0.0
    for p, u, f, gl, gr in zip(plans, us, fs, gls, grs):
        worst = max(worst, rel(u.cpu().numpy(), p.solve(f, gl, gr).cpu().numpy()))
    print(f"worst relative difference over 512 problems: {worst:.3e}")
    assert worst <= 1e-13
    # the staged and merged variants give the same bits over the 512 problems
    import torch
    first = [u.clone() for u in us]
    for v in STAGED + MERGED:
        b.force_variant(v)
        for rep in range(2):
            got = b.solve(fs, gls, grs)
            b.status()
            for k, (u, w) in enumerate(zip(got, first)):
                assert torch.equal(u, w), f"variant {v} solve {rep}: problem {k} differs"
    b.close()


def test_batch_rejects_bad_sets(cuda):
    from mpi_amd._lib import HzError
    from mpi_amd.solver import BatchSolver, HelmholtzPlan
    p = HelmholtzPlan(16, 16, 9, 1, form="diag")
    q = HelmholtzPlan(16, 16, 9, 1, form="plain")
    with pytest.raises(HzError):
        BatchSolver([p, p])
    with pytest.raises(HzError):
        BatchSolver([p, q])
    b = BatchSolver([p])
    for v in (0, 3, 6, 7, 10, -2):  # 0-3 removed, 6-7 switched off; 10, -2 out of range
        with pytest.raises(HzError):
            b.force_variant(v)
    b.close()


def test_one_shot_batched_solve(cuda):
    """hz_solve_batched (create, solve, wait, destroy in one call) equals the
    single-plan solves, like a kept batch."""
    import ctypes
    from mpi_amd._lib import check, load
    plans, inputs, _ = make_set(MIXED_SET)
    fs = [dev(f) for f, _, _ in inputs]
    gls = [dev(g) for _, g, _ in inputs]
    grs = [dev(g) for _, _, g in inputs]
    want = [p.solve(f, gl, gr).cpu().numpy() for p, f, gl, gr in zip(plans, fs, gls, grs)]
    us = [p.empty_field() for p in plans]
    n = len(plans)
    VP = ctypes.c_void_p
    arr = lambda xs: (VP * n)(*[VP(x) for x in xs])  # noqa: E731
    check(load().hz_solve_batched(arr([p.handle.value for p in plans]), n, arr([t.data_ptr() for t in fs]),
                                  arr([t.data_ptr() for t in gls]), arr([t.data_ptr() for t in grs]),
                                  arr([t.data_ptr() for t in us]), None))
    for k, (u, w) in enumerate(zip(us, want)):
        assert rel(u.cpu().numpy(), w) <= 1e-13, k
    for p in plans:
        p.close()


# (nx, ny, N, ne, lam): periodic-z problems mixed into a batch with Robin ones
PERIODIC_SET = [(32, 16, 33, 1, 0.5), (16, 64, 17, 3, 2.0), (64, 32, 65, 2, 7.0), (16, 16, 97, 1, 1e-2)]


def test_batch_with_periodic_problems(cuda):
    """Periodic-z plans (no face data) batched together with Robin/Dirichlet
    ones: every problem equals its single-plan solve, the periodic ones match
    the oracle's periodic solve, and face data for a periodic problem is
    refused."""
    from mpi_amd._lib import HzError
    from mpi_amd.solver import BatchSolver, HelmholtzPlan
    plans, inputs, _ = make_set(MIXED_SET[:3])
    pdescs = []
    for k, (nx, ny, N, ne, lam) in enumerate(PERIODIC_SET):
        p = HelmholtzPlan(nx, ny, N, ne, xl=0, xr=2 * math.pi, yl=0, yr=3.0, zl=0, zr=2, lam=lam, form="diag",
                          periodic_z=True)
        rng = np.random.default_rng(300 + k)
        plans.append(p)
        inputs.append((rng.standard_normal((nx, ny, p.nz)), None, None))
        pdescs.append(po.make_desc(nx, ny, ne * N, N, size=1, xl=0, xr=2 * math.pi, yl=0, yr=3.0, zl=0, zr=2,
                                   lam=lam))
    fs = [dev(f) for f, _, _ in inputs]
    gls = [None if g is None else dev(g) for _, g, _ in inputs]
    grs = [None if g is None else dev(g) for _, _, g in inputs]
    want = [p.solve(f, gl, gr).cpu().numpy() for p, f, gl, gr in zip(plans, fs, gls, grs)]
    b = BatchSolver(plans)
    us = b.solve(fs, gls, grs)
    b.status()
    for k, (u, w) in enumerate(zip(us, want)):
        assert rel(u.cpu().numpy(), w) <= 1e-13, f"problem {k}"
    for v in MERGED:  # periodic tiles (no faces) in the merged launches
        b.force_variant(v)
        got = b.solve(fs, gls, grs)
        b.status()
        for k, (u, w) in enumerate(zip(got, want)):
            assert rel(u.cpu().numpy(), w) <= 1e-13, f"merged variant {v}, problem {k}"
    for k, d in enumerate(pdescs):
        # Bar: 1e-12, or twice the reference algorithm's own rounding spread
        # (the oracle built with FMA contraction) where that is larger: the
        # constant mode of a periodic problem with lambda = 1e-2 is close to
        # singular, and the two oracle builds differ by 5e-10 there.
        f = inputs[3 + k][0]
        uo = po.solve_periodic(d, f)
        spread = rel(po.solve_periodic(d, f, fma=True), uo)
        assert rel(us[3 + k].cpu().numpy(), uo) <= max(1e-12, 2 * spread), f"periodic problem {k}"
    with pytest.raises(HzError):  # face data given for a periodic problem
        b.solve(fs, gls[:3] + [fs[3][:, :, 0].contiguous()] + gls[4:], grs)
    b.close()


def test_batch_with_two_and_one_dimensional_problems(cuda):
    """Plans with a unit Fourier axis (the 2-D solvers: nx = 1; the 1-D ones:
    nx = ny = 1), Robin, multi-element and periodic, batched with a 3-D
    problem: every problem equals its single-plan solve."""
    from mpi_amd.solver import BatchSolver, HelmholtzPlan
    specs = [  # (nx, ny, N, ne, lam, robin, periodic)
        (1, 32, 33, 1, 2.0, False, False), (1, 64, 17, 2, 0.5, True, False), (1, 16, 65, 1, 3.0, False, True),
        (1, 1, 129, 1, 1.5, False, False), (1, 1, 33, 3, 4.0, True, False), (1, 1, 17, 2, 2.5, False, True),
        (32, 16, 33, 1, 1.0, False, False)]
    plans, fs, gls, grs = [], [], [], []
    for k, (nx, ny, N, ne, lam, robin, periodic) in enumerate(specs):
        al, bl, ar, br = (0.7, 1.3, 0.4, 2.0) if robin else (0.0, 1.0, 0.0, 1.0)
        p = HelmholtzPlan(nx, ny, N, ne, xl=0, xr=2 * math.pi, yl=0, yr=3.0, zl=0, zr=2, lam=lam,
                          alpha_l=al, beta_l=bl, alpha_r=ar, beta_r=br, form="diag", periodic_z=periodic)
        rng = np.random.default_rng(500 + k)
        plans.append(p)
        fs.append(dev(rng.standard_normal((nx, ny, p.nz))))
        gls.append(None if periodic else dev(rng.standard_normal((nx, ny))))
        grs.append(None if periodic else dev(rng.standard_normal((nx, ny))))
    want = [p.solve(f, gl, gr).cpu().numpy() for p, f, gl, gr in zip(plans, fs, gls, grs)]
    b = BatchSolver(plans)
    for rep in range(2):  # the tuning solve, then a graph replay
        us = b.solve(fs, gls, grs)
        b.status()
        for k, (u, w) in enumerate(zip(us, want)):
            assert rel(u.cpu().numpy(), w) <= 1e-13, f"solve {rep}, problem {k} {specs[k]}"
    for v in MERGED:  # length-1 passes in the merged launches
        b.force_variant(v)
        us = b.solve(fs, gls, grs)
        b.status()
        for k, (u, w) in enumerate(zip(us, want)):
            assert rel(u.cpu().numpy(), w) <= 1e-13, f"merged variant {v}, problem {k} {specs[k]}"
    b.close()


def test_batch_follows_plan_changes(cuda):
    """hz_plan_set_robin and hz_plan_set_lambda after hz_batch_create apply to
    the batch's later solves (a Dirichlet plan turned Robin switches its
    mode-solve kind), each solve equal to the single-plan solves."""
    from mpi_amd.solver import BatchSolver
    spec = [(16, 32, 17, 1, 0.5, False), (32, 16, 33, 2, 3.0, True), (64, 64, 65, 1, 10.0, False)]
    plans, inputs, _ = make_set(spec)
    fs = [dev(f) for f, _, _ in inputs]
    gls = [dev(g) for _, g, _ in inputs]
    grs = [dev(g) for _, _, g in inputs]
    b = BatchSolver(plans)

    def check_against_single(tag):
        us = b.solve(fs, gls, grs)
        b.status()
        for k, (p, u) in enumerate(zip(plans, us)):
            w = p.solve(fs[k], gls[k], grs[k]).cpu().numpy()
            e = rel(u.cpu().numpy(), w)
            assert e <= 1e-13, f"{tag}: problem {k}: {e:.3e}"
        return [u.cpu().numpy() for u in us]

    before = check_against_single("initial")
    plans[0].set_robin(0.5, 1.0, 0.25, 2.0)  # Dirichlet -> Robin: the general mode kernel
    plans[2].set_lambda(3.0)
    after = check_against_single("after set_robin / set_lambda")
    assert rel(after[0], before[0]) > 1e-6 and rel(after[2], before[2]) > 1e-6
    assert rel(after[1], before[1]) <= 1e-13  # the untouched plan
    plans[0].set_robin(0.0, 1.0, 0.0, 1.0)  # back to Dirichlet
    check_against_single("back to Dirichlet")
    b.close()


def test_batch_with_mixed_plans(cuda):
    """MIXED plans (FP32 eigenbasis solves refined with FP64 residuals) batch
    with DIAG plans: each problem equals its single-plan solve (the same tile
    body) and the oracle to the rounding-spread bar."""
    from mpi_amd.solver import BatchSolver
    forms = ["mixed", "diag", "mixed", "mixed", "mixed", "diag", "mixed"]
    plans, inputs, descs = make_set(MIXED_SET, forms)
    fs = [dev(f) for f, _, _ in inputs]
    gls = [dev(g) for _, g, _ in inputs]
    grs = [dev(g) for _, _, g in inputs]
    want = [p.solve(f, gl, gr).cpu().numpy() for p, f, gl, gr in zip(plans, fs, gls, grs)]
    b = BatchSolver(plans)
    for rep in range(2):  # autotune + capture, then a replay
        us = b.solve(fs, gls, grs)
        b.status()
        for k, (u, w) in enumerate(zip(us, want)):
            e = rel(u.cpu().numpy(), w)
            assert e <= 1e-13, f"rep {rep} problem {k} ({forms[k]}): {e:.3e}"
    for v in MERGED:  # MIXED tiles in the merged launches
        b.force_variant(v)
        got = b.solve(fs, gls, grs)
        b.status()
        for k, (u, w) in enumerate(zip(got, want)):
            e = rel(u.cpu().numpy(), w)
            assert e <= 1e-13, f"merged variant {v} problem {k} ({forms[k]}): {e:.3e}"
    for k in (0, 2, 3, 6):
        f, gl, gr = inputs[k]
        uo = po.solve(descs[k], f, gl, gr)
        e = rel(us[k].cpu().numpy(), uo)
        bar = max(1e-12, 2 * rel(po.solve(descs[k], f, gl, gr, fma=True), uo))
        print(f"mixed problem {k}: rel err {e:.2e}, bar {bar:.2e}")
        assert e <= bar
    b.close()


# one Dirichlet element of even N (the even/odd split with no centre
# coordinate), padded sizes 32, 64, 32 and 128
EVEN_SET = [(16, 32, 18, 1, 0.5, False), (64, 64, 66, 1, 10.0, False), (32, 16, 34, 1, 3.0, False),
            (128, 32, 130, 1, 2.0, False)]


def test_batch_even_element_sizes(cuda):
    """DIAG and MIXED plans of even N in one batch: each problem equals its
    single-plan solve (staged and merged launches, a graph replay) and the
    oracle to the rounding-spread bar."""
    from mpi_amd.solver import BatchSolver
    forms = ["diag", "mixed", "mixed", "diag"]
    plans, inputs, descs = make_set(EVEN_SET, forms)
    fs = [dev(f) for f, _, _ in inputs]
    gls = [dev(g) for _, g, _ in inputs]
    grs = [dev(g) for _, _, g in inputs]
    want = [p.solve(f, gl, gr).cpu().numpy() for p, f, gl, gr in zip(plans, fs, gls, grs)]
    b = BatchSolver(plans)
    for rep in range(2):  # autotune + capture, then a replay
        us = b.solve(fs, gls, grs)
        b.status()
        for k, (u, w) in enumerate(zip(us, want)):
            e = rel(u.cpu().numpy(), w)
            assert e <= 1e-13, f"rep {rep} problem {k} ({EVEN_SET[k]}): {e:.3e}"
    for v in MERGED:
        b.force_variant(v)
        got = b.solve(fs, gls, grs)
        b.status()
        for k, (u, w) in enumerate(zip(got, want)):
            e = rel(u.cpu().numpy(), w)
            assert e <= 1e-13, f"merged variant {v} problem {k}: {e:.3e}"
    for k, (f, gl, gr) in enumerate(inputs):
        uo = po.solve(descs[k], f, gl, gr)
        e = rel(us[k].cpu().numpy(), uo)
        bar = max(1e-12, 2 * rel(po.solve(descs[k], f, gl, gr, fma=True), uo))
        print(f"even problem {k} {EVEN_SET[k]} ({forms[k]}): rel err {e:.2e}, bar {bar:.2e}")
        assert e <= bar
    b.close()
    for p in plans:
        p.close()
