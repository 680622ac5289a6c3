"""The C-ABI library without a device: it loads, exports every symbol that
include/hz.h declares, and its host-side plan setup is bit-identical to the
reference (golden vectors) -- no compute calls on a GPU."""
import ctypes
import json
import math
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from golden_util import load, sha

ROOT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..")
HEADER = os.path.join(ROOT, "include", "hz.h")


def declared_symbols():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(hz_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from mpi_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(_lib.SIGNATURES), set(syms) ^ set(_lib.SIGNATURES)
    _lib.load()


@pytest.mark.parametrize("N,ne", [(3, 1), (10, 25), (33, 1), (9, 4), (129, 1), (257, 1)])
def test_host_plan_setup_bitexact(N, ne):
    """hz_plan_create's GLL setup equals the reference's bit for bit (the
    reference's D, Dzz for one element; scaling as NonPeriodic3D_MPI.cpp:125-131)."""
    from mpi_amd import _lib
    from oracle import pyoracle as po
    lib = _lib.load()
    P = ctypes.POINTER(ctypes.c_double)
    x = np.zeros(N)
    z = np.zeros(ne * (N - 1) + 1)
    Dz = np.zeros((N, N))
    Dzz = np.zeros((N, N))
    _lib.check(lib.hz_host_plan_operators(N, ne, 2.0, x.ctypes.data_as(P), z.ctypes.data_as(P),
                                          Dz.ctypes.data_as(P), Dzz.ctypes.data_as(P)))
    g = load("gll.json")
    if str(N) in g:
        assert sha(x) == g[str(N)]["sha_x"]
    zo, Dzo, Dzzo = po.plan_operators(N, ne, 1, 0, 2.0)
    assert np.array_equal(z, zo) and np.array_equal(Dz, Dzo) and np.array_equal(Dzz, Dzzo)


@pytest.mark.parametrize("N", [9, 33, 129])
def test_host_eigendecomposition(N):
    """Dzz_int = V diag(mu) V^-1 to round-off; eigenvalues real, negative, and
    the lowest one tends to -(pi/2)^2 * (2/Lz)^2 (Dirichlet Laplacian on [-1,1])."""
    from mpi_amd import _lib
    lib = _lib.load()
    P = ctypes.POINTER(ctypes.c_double)
    m = N - 2
    mu = np.zeros(m)
    V = np.zeros((m, m))
    Vi = np.zeros((m, m))
    _lib.check(lib.hz_host_eigen(N, 1, 2.0, mu.ctypes.data_as(P), V.ctypes.data_as(P), Vi.ctypes.data_as(P)))
    Dzz = np.zeros((N, N))
    _lib.check(lib.hz_host_plan_operators(N, 1, 2.0, None, None, None, Dzz.ctypes.data_as(P)))
    A = Dzz[1:-1, 1:-1]
    assert np.all(mu < 0) and np.all(np.diff(mu) > 0)
    assert np.abs(A @ V - V * mu).max() <= 1e-15 * np.abs(mu).max() * m
    assert np.abs(V @ Vi - np.eye(m)).max() <= 1e-14
    assert abs(mu[-1] + (math.pi / 2) ** 2) < 1e-6



@pytest.mark.parametrize("N,NP", [(9, 32), (10, 32), (33, 32), (34, 64), (129, 128), (257, 256)])
def test_host_even_odd_eigendecomposition(N, NP):
    """The DIAG kernel's even/odd eigenbasis: F^-1 W diag(mu) Winv F is the
    symmetrised interior operator (A + JAJ)/2 to round-off, which is A up to
    the rounding of the GLL nodes; W and Winv are block diagonal and inverse
    to each other; the eigenvalues are those of A.  And the solve it gives,
    F^-1 W (mu - s)^-1 Winv F f, is numpy's LU solve of A - s to 1e-13
    (relative to max|u|) for the shifts a 256^2 plan spans."""
    from mpi_amd import _lib
    lib = _lib.load()
    P = ctypes.POINTER(ctypes.c_double)
    m = N - 2
    mu = np.zeros(NP)
    W = np.zeros((NP, NP))
    Wi = np.zeros((NP, NP))
    asym = ctypes.c_double()
    _lib.check(lib.hz_host_eigen_split(N, 1, 2.0 * math.pi, NP, mu.ctypes.data_as(P), W.ctypes.data_as(P),
                                       Wi.ctypes.data_as(P), ctypes.byref(asym)))
    Dzz = np.zeros((N, N))
    _lib.check(lib.hz_host_plan_operators(N, 1, 2.0 * math.pi, None, None, None, Dzz.ctypes.data_as(P)))
    A = Dzz[1:-1, 1:-1]
    J = np.eye(m)[::-1]
    As = (A + J @ A @ J) / 2
    assert asym.value == np.abs(A - J @ A @ J).max() / np.abs(A).max() and asym.value < 1e-11
    H, h = NP // 2, m // 2
    he = m - h
    F = np.zeros((NP, m))
    for i in range(h):
        F[i, i] = F[i, m - 1 - i] = 1.0
        F[H + i, i], F[H + i, m - 1 - i] = 1.0, -1.0
    if m % 2:
        F[h, h] = 1.0
    Fi = np.linalg.pinv(F)  # exact left inverse: halves of sums and differences
    valid = np.zeros(NP, bool)
    valid[:he] = True
    valid[H:H + h] = True
    assert not mu[~valid].any() and not W[~valid].any() and not W[:, ~valid].any() and not Wi[~valid].any()
    assert not W[:H, H:].any() and not W[H:, :H].any() and not Wi[:H, H:].any() and not Wi[H:, :H].any()
    assert np.abs(W @ Wi - np.diag(valid.astype(float))).max() <= 1e-13
    R = Fi @ W @ np.diag(mu) @ Wi @ F
    assert np.abs(R - As).max() <= 1e-13 * np.abs(A).max()
    assert np.allclose(np.sort(mu[valid]), np.sort(np.linalg.eigvals(A).real), rtol=1e-10, atol=0)
    z = np.linspace(0.0, 2.0 * math.pi, m + 2)[1:-1]
    f = np.sin(3 * z) * np.exp(np.cos(z))
    for s in (10.0, 12.0, 1e3, 10.0 + 2 * 128.0 ** 2):
        inv = np.where(valid, 1.0 / np.where(valid, mu - s, 1.0), 0.0)
        u = Fi @ (W @ (inv * (Wi @ (F @ f))))
        ul = np.linalg.solve(A - s * np.eye(m), f)
        assert np.abs(u - ul).max() <= 1e-13 * np.abs(ul).max() * max(1.0, m / 64)


def test_bench_refuses_more_than_one_gpu():
    """bench.py --gpus 2 exits with status 2 and one JSON line naming the
    reason, before it touches a device."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


@pytest.mark.parametrize("change,code", [
    ({"nx": 96}, "HZ_ERR_NOT_POW2"), ({"ny": 3}, "HZ_ERR_NOT_POW2"), ({"nx": 0}, "HZ_ERR_NOT_POW2"),
    ({"n": 2}, "HZ_ERR_N_TOO_SMALL"), ({"nx": 2048}, "HZ_ERR_UNSUPPORTED"), ({"ne": 0}, "HZ_ERR_ARG"),
    ({"zr": -1.0}, "HZ_ERR_ARG"), ({"xr": 0.0}, "HZ_ERR_ARG"), ({"form": 3}, "HZ_ERR_ARG"),
    ({"form": -1}, "HZ_ERR_ARG"), ({}, "HZ_ERR_NO_DEVICE")])
def test_plan_create_refusals_before_the_device(change, code):
    """hz_plan_create checks the descriptor before it looks for a device: the
    reference's own refusals (non-power-of-two Fourier sizes,
    Fourier_Quadrature.cpp:1106-1114; N < 3, GLLQuadrature.cpp:82-83) and the
    argument errors come back as their codes, a valid descriptor on a host
    without a GPU as HZ_ERR_NO_DEVICE, and no plan is returned."""
    import torch
    from mpi_amd import _lib
    if not change and torch.cuda.device_count() > 0:
        pytest.skip("a HIP device is present: the valid descriptor creates a plan")
    lib = _lib.load()
    fields = dict(nx=64, ny=32, n=33, ne=1, xl=0.0, xr=2 * math.pi, yl=0.0, yr=2 * math.pi, zl=-1.0, zr=1.0,
                  lam=1.0, alpha_l=0.0, beta_l=1.0, alpha_r=0.0, beta_r=1.0, form=1, refine_sweeps=2,
                  mixed_fp64_cond=0.0, periodic_z=0)
    fields.update(change)
    desc = _lib.PlanDesc(**fields)
    handle = ctypes.c_void_p(12345)
    rc = lib.hz_plan_create(ctypes.byref(desc), ctypes.byref(handle))
    assert rc == getattr(_lib, code)
    assert handle.value is None
