"""The 2-D and 1-D solvers (NonPeriodic2D_MPI.cpp, NonPeriodic1D_MPI.cpp,
Periodic2D_MPI.cpp, Periodic1D_MPI.cpp) against the oracle run with unit
Fourier axes.

The reference's 2-D solver transforms its periodic x and keeps a GLL axis y;
its 1-D solver has no transform.  Both are the 3-D algorithm with one or two
fewer transform axes, so the oracle with nx = 1 (the reference's x as this
y, its y as this z) or nx = ny = 1 must reproduce their outputs.  It does, bit
for bit, at one rank and at four (the golden vectors come from the
reference's own sources, tests/golden/make_golden.py --lowdim).  The plan's
host-side grid (the library's hz_host_plan_operators) must equal the
reference's grid bit for bit as well.
"""
import ctypes
import glob
import os

import numpy as np
import pytest

from golden_util import GOLDEN, load, lowdim_inputs, lowdim_sample, sha, unhex
from oracle import pyoracle as po

CASES = sorted(os.path.basename(p)[4:-5] for p in glob.glob(os.path.join(GOLDEN, "low_*.json")))


def oracle_desc(rec):
    """The 3-D descriptor of a 2-D / 1-D case: x is a single point."""
    p = rec["params"]
    robin = dict(alphal=p["al"], betal=p["bl"], alphar=p["ar"], betar=p["br"], lam=p["lam"])
    if rec["dim"] == 2:
        return po.make_desc(1, p["Nx"], p["Ny"], p["N"], size=p["size"], xl=0.0, xr=1.0, yl=p["xl"], yr=p["xr"],
                            zl=p["yl"], zr=p["yr"], **robin)
    return po.make_desc(1, 1, p["Nx"], p["N"], size=p["size"], xl=0.0, xr=1.0, yl=0.0, yr=1.0,
                        zl=p["xl"], zr=p["xr"], **robin)


def as3d(rec, f, gl, gr):
    K = rec["K"]
    lines = rec["params"]["Nx"] if rec["dim"] == 2 else 1
    F = np.ascontiguousarray(f.reshape(1, lines, K))
    if gl is None:
        return F, None, None
    return F, np.reshape(gl, (1, lines)).copy(), np.reshape(gr, (1, lines)).copy()


def test_cases_present():
    assert len(CASES) >= 12
    kinds = {(load(f"low_{c}.json")["dim"], load(f"low_{c}.json")["periodic"]) for c in CASES}
    assert kinds == {(2, False), (2, True), (1, False), (1, True)}


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference(name):
    rec = load(f"low_{name}.json")
    f, gl, gr, ex = lowdim_inputs(rec)
    d = oracle_desc(rec)
    assert np.array_equal(po.z_grid(d), unhex(rec["grid"]))
    F, GL, GR = as3d(rec, f, gl, gr)
    if rec["periodic"]:
        u = po.solve_periodic(d, F)
    else:
        u = po.solve(d, F, GL, GR)
    u = u.reshape(f.shape)
    got, want = lowdim_sample(u, rec)
    assert np.array_equal(got, want), f"max diff {np.abs(got - want).max():.3e}"
    if "u" not in rec:
        assert sha(u) == rec["sha_u"]
    if ex is not None:  # the discretisation error the reference reports
        assert np.abs(u - ex).max() == pytest.approx(rec["err_vs_exact"], rel=1e-12, abs=0)


@pytest.mark.parametrize("name", CASES)
def test_library_grid_matches_reference(name):
    """The GPU plan's z array comes from hz_host_plan_operators (z from 0, the
    reference's quirk); it must be the reference's grid exactly."""
    from mpi_amd import _lib
    rec = load(f"low_{name}.json")
    p = rec["params"]
    ne, N = rec["elements"], p["N"]
    lz = (p["yr"] - p["yl"]) if rec["dim"] == 2 else (p["xr"] - p["xl"])
    z = np.zeros(ne * (N - 1) + 1)
    P = ctypes.POINTER(ctypes.c_double)
    _lib.check(_lib.load().hz_host_plan_operators(N, ne, lz, None, z.ctypes.data_as(P), None, None))
    if p["size"] == 1:
        assert np.array_equal(z, unhex(rec["grid"]))
    else:  # the reference builds each rank's elements from (L/P)/Ns: equal up to rounding
        assert np.abs(z - unhex(rec["grid"])).max() <= 4 * np.finfo(float).eps * lz


@pytest.mark.parametrize("name", CASES)
def test_rounding_spread_reproduces(name):
    """The fixture's rounding_spread (the oracle built with fused multiply-adds
    against the reference's output) is what that build gives here.  It sets
    the parity bar of the GPU tests where the reference's algorithm itself
    moves by more than 1e-12 under a change of rounding (many elements: the
    interface sweep amplifies rounding in its own coefficients, which input
    noise never touches)."""
    rec = load(f"low_{name}.json")
    f, gl, gr, _ = lowdim_inputs(rec)
    d = oracle_desc(rec)
    F, GL, GR = as3d(rec, f, gl, gr)
    exact_build = (po.solve_periodic(d, F) if rec["periodic"] else po.solve(d, F, GL, GR)).reshape(f.shape)
    fused = (po.solve_periodic(d, F, fma=True) if rec["periodic"] else
             po.solve(d, F, GL, GR, fma=True)).reshape(f.shape)
    spread = float(np.abs(fused - exact_build).max() / np.abs(exact_build).max())
    assert spread > 0.0  # the fused build really rounds differently
    assert spread == pytest.approx(rec["rounding_spread"], rel=1e-6)


def test_oracle_planted_mistake_is_caught():
    """A wrong shift (lambda off by one ulp-scale relative amount) in the
    oracle's input must break bit-exactness: the comparison is not vacuous."""
    rec = load("low_1d_random_N33.json")
    f, gl, gr, _ = lowdim_inputs(rec)
    d = oracle_desc(rec)
    d.lam = d.lam * (1 + 1e-12)
    F, GL, GR = as3d(rec, f, gl, gr)
    u = po.solve(d, F, GL, GR).reshape(f.shape)
    got, want = lowdim_sample(u, rec)
    assert not np.array_equal(got, want)
