"""The solver's helper entry points on the GPU against the reference's own
outputs (tests/golden/make_golden.py --helpers / --palpha):

- hz_derivative_z (FirstDerivativeZ / SecondDerivativeZ,
  NonPeriodic3D_MPI.cpp:632-683): bit-exact, every fixture;
- hz_lu_factor_batched + hz_lu_solve_batched (LUDecomposition / LUSolve,
  LU_Decomposition.cpp:7-108): bit-exact;
- the drop-in's FirstDerivativeZ, SecondDerivativeZ,
  ConstructFundamentalSolutions, LUSolve and WriteFile, driven by
  mpi_amd/compat/TestDropinHelpers with the fixture driver's own command
  line: outputs equal byte for byte (the Tecplot file included);
- hz_plan_periodic_closure (alphaPDE, Periodic3D_MPI.cpp:698, :738, :804)
  within max(1e-12, 2 x the reference algorithm's own rounding spread) of
  max |alpha| for PLAIN, DIAG and MIXED plans (the spread: the oracle built
  with fused multiply-adds against the plain build; it is 1e-11 where the
  constant mode of a lambda = 1e-2 problem is nearly singular).
"""
import os
import subprocess

import numpy as np
import pytest

from golden_util import (fundsol_inputs, load, lusolve_inputs, palpha_inputs, sha, tecplot_inputs, unhex,
                         zderiv_inputs)

pytestmark = pytest.mark.gpu

ZD = load("zderiv.json")
FS = load("fundsol.json")
LS = load("lusolve.json")
TP = load("tecplot.json")
PA = load("palpha.json")
EXE = os.path.join(os.path.dirname(__file__), "..", "mpi_amd", "compat", "TestDropinHelpers")


def to_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


@pytest.mark.parametrize("key", sorted(ZD))
def test_derivative_z_matches_reference(cuda, key):
    import torch
    from mpi_amd.solver import HelmholtzPlan
    rec = ZD[key]
    u = zderiv_inputs(rec)
    plan = HelmholtzPlan(rec["Nx"], rec["Ny"], rec["N"], rec["elements"], xl=0.0, xr=1.0, yl=0.0, yr=1.0,
                         zl=rec["zl"], zr=rec["zr"], form="diag")
    du = to_dev(u)
    for order, k in ((1, "d1"), (2, "d2")):
        out = plan.derivative_z(du, order=order, out=torch.empty_like(du)).cpu().numpy()
        assert sha(out) == rec["sha_" + k], (key, order, float(np.abs(out.ravel() - unhex(rec[k])).max()))
    plan.close()


@pytest.mark.parametrize("key", sorted(LS))
def test_lu_solve_batched_matches_reference(cuda, key):
    import torch
    from mpi_amd.solver import lu_factor_batched, lu_solve_batched
    rec = LS[key]
    A, b = lusolve_inputs(key, rec)
    # several copies in one launch, so the batch indexing is exercised too
    count = 3
    dA = to_dev(np.repeat(A[None], count, axis=0))
    piv, _, st = lu_factor_batched(dA)
    db = to_dev(np.repeat(b[None], count, axis=0))
    lu_solve_batched(dA, piv, db)
    torch.cuda.synchronize()
    assert not st.cpu().numpy().any()
    x = db.cpu().numpy()
    want = unhex(rec["x"])
    for c in range(count):
        assert np.array_equal(x[c], want), (key, c, float(np.abs(x[c] - want).max()))


def _run(args, cwd=None):
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=120, cwd=cwd)
    assert r.returncode == 0, r.stdout + r.stderr
    return r


@pytest.mark.parametrize("key", [k for k in sorted(ZD) if ZD[k]["size"] == 1])
def test_dropin_derivative_z(cuda, tmp_path, key):
    rec = ZD[key]
    u = zderiv_inputs(rec)
    inp, out = tmp_path / "u.bin", tmp_path / "d.bin"
    u.tofile(inp)
    _run(["zderiv", str(rec["Nx"]), str(rec["Ny"]), str(rec["Nz"]), str(rec["N"]), repr(rec["zl"]), repr(rec["zr"]),
          str(inp), str(out)])
    got = np.fromfile(str(out) + ".rank0").reshape(2, -1)
    assert sha(got[0].reshape(u.shape)) == rec["sha_d1"]
    assert sha(got[1].reshape(u.shape)) == rec["sha_d2"]


@pytest.mark.parametrize("key", sorted(FS))
def test_dropin_fundamental_solutions(cuda, tmp_path, key):
    rec = FS[key]
    rhs = fundsol_inputs(rec)
    inp, out = tmp_path / "r.bin", tmp_path / "U.bin"
    rhs.tofile(inp)
    _run(["fundsol", *rec["args"], str(inp), str(out)])
    got = np.fromfile(out)
    want = unhex(rec["U"])
    assert np.array_equal(got, want), float(np.abs(got - want).max())


def test_dropin_lu_solve(cuda, tmp_path):
    names = sorted(LS)
    for key in names:
        rec = LS[key]
        A, b = lusolve_inputs(key, rec)
        inp, out = tmp_path / f"{key}.in", tmp_path / f"{key}.out"
        np.concatenate([A.ravel(), b]).tofile(inp)
        _run(["lusolve", str(rec["n"]), "1", str(inp), str(out)])
        assert np.array_equal(np.fromfile(out), unhex(rec["x"])), key


@pytest.mark.parametrize("key", sorted(TP))
def test_dropin_writefile_tecplot(cuda, tmp_path, key):
    import hashlib
    rec = TP[key]
    U = tecplot_inputs(rec)
    (tmp_path / "Output").mkdir()
    inp = tmp_path / "u.bin"
    U.tofile(inp)
    _run(["write", *rec["args"], str(inp)], cwd=tmp_path)
    data = (tmp_path / "Output" / "Solution-0000.tec").read_bytes()
    assert hashlib.sha256(data).hexdigest() == rec["sha_file"]
    assert data.decode("ascii") == rec["text"]


@pytest.mark.parametrize("form", ["plain", "diag", "mixed"])
@pytest.mark.parametrize("key", sorted(PA))
def test_periodic_closure_matches_reference(cuda, key, form):
    from mpi_amd.solver import HelmholtzPlan
    from oracle import pyoracle as po
    rec = PA[key]
    if form == "mixed" and rec["elements"] > 8:
        pytest.skip("MIXED plans take at most 8 elements")
    f = palpha_inputs(rec)
    xl, xr, yl, yr, zl, zr = rec["bounds"]
    nx, ny, N = rec["Nx"], rec["Ny"], rec["N"]
    d1 = po.make_desc(nx, ny, rec["elements"] * N, N, size=1, xl=xl, xr=xr, yl=yl, yr=yr, zl=zl, zr=zr,
                      lam=rec["lam"])
    _, a_plain = po.solve_periodic(d1, f, alpha=True)
    _, a_fma = po.solve_periodic(d1, f, fma=True, alpha=True)
    bar = max(1e-12, 2.0 * float(np.abs(a_fma - a_plain).max()) / rec["max_abs_alpha"])
    plan = HelmholtzPlan(nx, ny, N, rec["elements"], xl=xl, xr=xr, yl=yl, yr=yr, zl=zl, zr=zr, lam=rec["lam"],
                         form=form, periodic_z=True)
    assert plan.nz == rec["K"]
    plan.solve(to_dev(f))
    half = plan.periodic_closure()
    M = ny // 2 + 1
    full = np.zeros((nx, 2 * ny))
    for i in range(nx):
        for m in range(ny):
            low = m < M
            ii, mm = (i, m) if low else ((nx - i) % nx, ny - m)
            full[i, 2 * m] = half[0, ii, mm]
            full[i, 2 * m + 1] = half[1, ii, mm] if low else -half[1, ii, mm]
    want = unhex(rec["alpha"]).reshape(nx, 2 * ny)
    err = float(np.abs(full - want).max() / rec["max_abs_alpha"])
    print(f"{key} {form}: alpha rel err {err:.2e}, bar {bar:.2e}")
    assert err <= bar
    plan.close()
