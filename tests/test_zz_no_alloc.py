"""hz_adjoint_noalloc (the adjoint for callers that cannot allocate) against
hz_adjoint through the C ABI: fbar, glbar, grbar and dL/dlambda identical,
on two calls in a row on one GPU, the second with a smaller problem (the
static reduction scratch keeps the first call's partial sums, which must not
leak into the second's result)."""
import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _vp(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _run_pair(plan, seed, periodic):
    import torch
    from mpi_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(seed)
    shape = (plan.nx, plan.ny, plan.nz)
    f = torch.from_numpy(rng.standard_normal(shape)).cuda()
    faces = None if periodic else [torch.from_numpy(rng.standard_normal(shape[:2])).cuda() for _ in range(2)]
    u = plan.solve(f, *(faces or [None, None]))
    ubar = torch.from_numpy(rng.standard_normal(shape)).cuda()
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    outs = []
    for which in ("alloc", "noalloc"):
        fbar = torch.full(shape, float("nan"), dtype=torch.float64, device="cuda")
        gl = gr = None
        if not periodic:
            gl = torch.full(shape[:2], float("nan"), dtype=torch.float64, device="cuda")
            gr = torch.full(shape[:2], float("nan"), dtype=torch.float64, device="cuda")
        if which == "alloc":
            lb = ctypes.c_double(float("nan"))
            _lib.check(lib.hz_adjoint(plan.handle, _vp(ubar), _vp(u), _vp(fbar), _vp(gl), _vp(gr), ctypes.byref(lb),
                                      sp))
            lam_bar = lb.value
        else:
            lbd = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
            _lib.check(lib.hz_adjoint_noalloc(plan.handle, _vp(ubar), _vp(u), _vp(fbar), _vp(gl), _vp(gr), _vp(lbd),
                                              sp))
            torch.cuda.synchronize()
            lam_bar = float(lbd.cpu().numpy()[0])
        outs.append((fbar.cpu().numpy(), None if gl is None else gl.cpu().numpy(),
                     None if gr is None else gr.cpu().numpy(), lam_bar))
    (fa, gla, gra, la), (fb, glb, grb, lb_) = outs
    assert np.array_equal(fa, fb)
    if not periodic:
        assert np.array_equal(gla, glb) and np.array_equal(gra, grb)
    assert la == lb_ and math.isfinite(la)
    return la


@pytest.mark.parametrize("form", ["diag", "plain", "mixed"])
def test_adjoint_noalloc_matches_adjoint_two_calls(cuda, form):
    from mpi_amd.solver import HelmholtzPlan
    big = HelmholtzPlan(64, 32, 33, 2, lam=3.0, alpha_l=0.5, beta_l=1.0, alpha_r=0.2, beta_r=1.0, form=form)
    small = HelmholtzPlan(16, 8, 17, 1, lam=2.0, form=form)
    _run_pair(big, 1, False)    # PLAIN: hz_adjoint's first call builds the operators
    _run_pair(small, 2, False)  # the second call: smaller, over the first call's scratch contents
    big.close()
    small.close()


def test_adjoint_noalloc_periodic_and_refusals(cuda):
    import torch
    from mpi_amd import _lib
    from mpi_amd.solver import HelmholtzPlan
    p = HelmholtzPlan(32, 16, 17, 3, lam=1.5, form="diag", periodic_z=True)
    _run_pair(p, 3, True)
    lib = _lib.load()
    x = torch.zeros((p.nx, p.ny, p.nz), dtype=torch.float64, device="cuda")
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    lbd = torch.zeros(1, dtype=torch.float64, device="cuda")
    # lambda_bar without u, and face gradients for a periodic plan, are refused
    with pytest.raises(_lib.HzError):
        _lib.check(lib.hz_adjoint_noalloc(p.handle, _vp(x), None, _vp(x), None, None, _vp(lbd), sp))
    with pytest.raises(_lib.HzError):
        _lib.check(lib.hz_adjoint_noalloc(p.handle, _vp(x), _vp(x), _vp(x), _vp(lbd), _vp(lbd), None, sp))
    p.close()
    # a PLAIN plan before its first hz_adjoint has no diagonalised operators
    q = HelmholtzPlan(16, 16, 9, 1, lam=1.0, form="plain")
    y = torch.zeros((q.nx, q.ny, q.nz), dtype=torch.float64, device="cuda")
    g = torch.zeros((q.nx, q.ny), dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.HzError):
        _lib.check(lib.hz_adjoint_noalloc(q.handle, _vp(y), _vp(y), _vp(y), _vp(g), _vp(g), _vp(lbd), sp))
    q.close()
