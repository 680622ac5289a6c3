"""Element sizes at both ends of what a plan accepts.  The smallest, N = 3
(one interior node, so the even/odd split has an empty odd half) and N = 4
(one node in each half), for every form, Dirichlet and Robin faces, one to
three elements, against the oracle; and the adjoint against the oracle's
transposed solve map.  The largest each form takes (N - 2 = 128 for PLAIN,
256 for DIAG and MIXED) against the oracle, and one more node refused at plan
creation with HZ_ERR_UNSUPPORTED.

Bar as in the DIAG tests: relative max-norm error <= max(1e-12, 2 x the
rounding spread of the reference algorithm on the same input: the oracle built
with fused multiply-adds against the plain oracle).  The adjoint bar is 1e-12
relative to max |S^T ubar|, as in the other adjoint tests.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
TOL = 1e-12
DIRICHLET = (0.0, 1.0, 0.0, 1.0)
ROBIN = (0.7, 1.3, 0.4, 2.0)


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def rel_err(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def make_plan(nx, ny, N, ne, bc, lam, form):
    from mpi_amd.solver import HelmholtzPlan
    al, bl, ar, br = bc
    return HelmholtzPlan(nx, ny, N, ne, xl=0, xr=2 * math.pi, yl=0, yr=3.0, zl=0, zr=2, lam=lam,
                         alpha_l=al, beta_l=bl, alpha_r=ar, beta_r=br, form=form)


def make_desc(nx, ny, N, ne, bc, lam):
    al, bl, ar, br = bc
    return po.make_desc(nx, ny, ne * N, N, size=1, xl=0, xr=2 * math.pi, yl=0, yr=3.0, zl=0, zr=2,
                        alphal=al, betal=bl, alphar=ar, betar=br, lam=lam)


@pytest.mark.parametrize("form", ["diag", "mixed", "plain"])
@pytest.mark.parametrize("bc", [DIRICHLET, ROBIN], ids=["dirichlet", "robin"])
@pytest.mark.parametrize("N,ne", [(3, 1), (4, 1), (3, 2), (4, 3)])
def test_smallest_elements_against_oracle(cuda, form, bc, N, ne):
    nx, ny, lam = 16, 8, 3.0
    plan = make_plan(nx, ny, N, ne, bc, lam, form)
    desc = make_desc(nx, ny, N, ne, bc, lam)
    assert po.total_z_points(desc) == plan.nz
    rng = np.random.default_rng(100 * N + 10 * ne + (bc == ROBIN))
    f = rng.standard_normal((nx, ny, plan.nz))
    gl = rng.standard_normal((nx, ny))
    gr = rng.standard_normal((nx, ny))
    print(f"case form={form} bc={'robin' if bc == ROBIN else 'dirichlet'} N={N} ne={ne}", flush=True)
    u = plan.solve(dev(f), dev(gl), dev(gr)).cpu().numpy()
    plan.status()
    uo = po.solve(desc, f, gl, gr)
    bar = max(TOL, 2.0 * rel_err(po.solve(desc, f, gl, gr, fma=True), uo))
    err = rel_err(u, uo)
    print(f"rel err {err:.2e}, bar {bar:.2e}")
    assert err <= bar
    plan.close()


@pytest.mark.parametrize("form", ["diag", "mixed", "plain"])
@pytest.mark.parametrize("bc", [DIRICHLET, ROBIN], ids=["dirichlet", "robin"])
@pytest.mark.parametrize("N", [3, 4])
def test_smallest_elements_adjoint(cuda, form, bc, N):
    nx, ny, lam = 4, 8, 3.0
    desc = make_desc(nx, ny, N, 1, bc, lam)
    K = po.total_z_points(desc)
    nf, ng = nx * ny * K, nx * ny
    S = np.zeros((nf, nf + 2 * ng))  # the oracle's solve map, one unit input per column
    for c in range(nf + 2 * ng):
        f, gl, gr = np.zeros(nf), np.zeros(ng), np.zeros(ng)
        (f if c < nf else gl if c < nf + ng else gr)[c if c < nf else (c - nf) % ng] = 1.0
        S[:, c] = po.solve(desc, f.reshape(nx, ny, K), gl.reshape(nx, ny), gr.reshape(nx, ny)).ravel()
    rng = np.random.default_rng(7 * N + (bc == ROBIN))
    ubar = rng.standard_normal((nx, ny, K))
    want = S.T @ ubar.ravel()
    print(f"case adjoint form={form} bc={'robin' if bc == ROBIN else 'dirichlet'} N={N}", flush=True)
    plan = make_plan(nx, ny, N, 1, bc, lam, form)
    fbar, glbar, grbar, _ = plan.adjoint(dev(ubar))
    got = np.concatenate([fbar.cpu().numpy().ravel(), glbar.cpu().numpy().ravel(), grbar.cpu().numpy().ravel()])
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    plan.close()


@pytest.mark.parametrize("form,N", [("plain", 130), ("diag", 258), ("mixed", 258)])
@pytest.mark.parametrize("bc", [DIRICHLET, ROBIN], ids=["dirichlet", "robin"])
def test_largest_elements_against_oracle(cuda, form, N, bc):
    nx, ny, lam = 4, 8, 3.0
    plan = make_plan(nx, ny, N, 1, bc, lam, form)
    desc = make_desc(nx, ny, N, 1, bc, lam)
    rng = np.random.default_rng(N + (bc == ROBIN))
    f = rng.standard_normal((nx, ny, plan.nz))
    gl = rng.standard_normal((nx, ny))
    gr = rng.standard_normal((nx, ny))
    print(f"case form={form} bc={'robin' if bc == ROBIN else 'dirichlet'} N={N}", flush=True)
    u = plan.solve(dev(f), dev(gl), dev(gr)).cpu().numpy()
    plan.status()
    uo = po.solve(desc, f, gl, gr)
    bar = max(TOL, 2.0 * rel_err(po.solve(desc, f, gl, gr, fma=True), uo))
    err = rel_err(u, uo)
    print(f"rel err {err:.2e}, bar {bar:.2e}")
    assert err <= bar
    plan.close()


@pytest.mark.parametrize("form,N", [("plain", 131), ("diag", 259), ("mixed", 259)])
def test_elements_past_the_limit_refused(cuda, form, N):
    from mpi_amd import _lib
    with pytest.raises(_lib.HzError) as e:
        make_plan(4, 8, N, 1, DIRICHLET, 3.0, form)
    assert e.value.code == _lib.HZ_ERR_UNSUPPORTED
