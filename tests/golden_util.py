"""Helpers that read the golden vectors written by tests/golden/make_golden.py."""
import hashlib
import json
import math
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLDEN, name)) as fp:
        return json.load(fp)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def unhex(lst):
    return np.array([float.fromhex(v) for v in lst])


def shipped_inputs(desc, z):
    """Inputs of the reference's test program (TestNonPeriodic3D_MPI.cpp:26-128)."""
    PI = 3.1415926535897932
    x = np.arange(desc["Nx"]) / desc["Nx"]
    y = np.arange(desc["Ny"]) / desc["Ny"]
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    f = -(64 * PI * PI + 36 * PI * PI + 16 * PI * PI + 1.0) * np.sin(8 * PI * X) * np.cos(6 * PI * Y) * np.cos(4 * PI * Z)
    XX, YY = np.meshgrid(x, y, indexing="ij")
    gl = (1.0 * 4.0 * PI * np.sin(0.0) / 1.0 + 0.0 * np.cos(0.0)) * np.sin(8 * PI * XX) * np.cos(6 * PI * YY)
    gr = (-0.0 * 4.0 * PI * np.sin(4 * PI) / 1.0 + 1.0 * np.cos(4 * PI)) * np.sin(8 * PI * XX) * np.cos(6 * PI * YY)
    ex = np.sin(8 * PI * X) * np.cos(6 * PI * Y) * np.cos(4 * PI * Z)
    return f, gl, gr, ex


def case_inputs(name, z):
    """Regenerate the inputs of golden solve case `name` on the z grid `z`."""
    from mpi_amd import workloads as wl
    rec = load(f"solve_{name}.json")
    d = rec["desc"]
    if name.startswith("shipped"):
        f, gl, gr, ex = shipped_inputs(d, z)
    elif name.startswith("config1"):
        x, y, zeta = wl.grid(d["Nx"], d["Ny"], z)
        f, gl, gr, ex = wl.manufactured_config1(x, y, zeta, 1.0)
    else:
        x, y, zeta = wl.grid(d["Nx"], d["Ny"], z)
        f, gl, gr = wl.random_problem(x, y, zeta, seed=2, nmodes=64, kmax=32, lmax=64)
        ex = None
    assert sha(f) == rec["sha_f"] and sha(gl) == rec["sha_gl"] and sha(gr) == rec["sha_gr"], \
        "input generator drifted from the golden vectors"
    return rec, f, gl, gr, ex


def sample(u, rec):
    idx = rec["sample_idx"]
    return np.array([u[i, j, k] for i, j, k in idx]), unhex(rec["sample_u"])


def periodic_case_inputs(name, z):
    """Regenerate the inputs of periodic golden case `name` on the z grid `z`."""
    from mpi_amd import workloads as wl
    rec = load(f"psolve_{name}.json")
    d = rec["desc"]
    ex = None
    if name.startswith("shipped"):
        PI = 3.1415926535897932
        x = np.arange(d["Nx"]) / d["Nx"]
        y = np.arange(d["Ny"]) / d["Ny"]
        X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
        f = -(64 * PI * PI + 36 * PI * PI + 16 * PI * PI + d["lam"]) * np.sin(8 * PI * X) * np.cos(6 * PI * Y) * \
            np.cos(4 * PI * Z)
        ex = np.sin(8 * PI * X) * np.cos(6 * PI * Y) * np.cos(4 * PI * Z)
    else:
        seed = int(rec["generator"].split("seed=")[1].split(",")[0])
        kmax = int(rec["generator"].split("kmax=")[1].split(",")[0])
        lmax = int(rec["generator"].split("lmax=")[1].split(")")[0])
        x, y, zeta = wl.grid(d["Nx"], d["Ny"], z, xl=d["xl"], xr=d["xr"], yl=d["yl"], yr=d["yr"], zl=d["zl"])
        f, _, _ = wl.random_problem(x, y, zeta, seed=seed, nmodes=32, kmax=kmax, lmax=lmax)
    f = np.ascontiguousarray(f)
    assert sha(f) == rec["sha_f"], "input generator drifted from the golden vectors"
    return rec, f, ex


# ---------------------------------------------------------------------------
# 2-D and 1-D solvers (NonPeriodic2D/1D_MPI.cpp, Periodic2D/1D_MPI.cpp)

def lowdim_inputs(rec):
    """Regenerate the inputs of a 2-D / 1-D golden case from its record.

    Returns (f, gl, gr, exact): 2-D arrays are [Nx][K] (the reference's
    periodic x, then its GLL axis), 1-D ones [K]; faces are [Nx] / scalars
    (None for the periodic solvers); exact is None for random cases."""
    from mpi_amd import workloads as wl
    PI = 3.1415926535897932
    p, g = rec["params"], rec["gen"]
    grid = unhex(rec["grid"])
    dim, periodic = rec["dim"], rec["periodic"]
    xl, xr = p["xl"], p["xr"]
    Lx = xr - xl
    lam = p["lam"]
    al, bl, ar, br = p["al"], p["bl"], p["ar"], p["br"]
    gl = gr = ex = None
    if g["kind"] == "shipped" and dim == 2:
        yl, yr = p["yl"], p["yr"]
        Ly = yr - yl
        x = xl + np.arange(p["Nx"]) * (Lx / p["Nx"])
        X, Y = np.meshgrid(x, grid, indexing="ij")
        sx = np.sin(8.0 * PI * (X - xl) / Lx)
        f = -(64.0 * PI * PI / (Lx * Lx) + 36.0 * PI * PI / (Ly * Ly) + lam) * sx * np.cos(6.0 * PI * (Y - yl) / Ly)
        ex = sx * np.cos(6.0 * PI * (Y - yl) / Ly)
        if not periodic:
            s = np.sin(8.0 * PI * (x - xl) / Lx)
            gl = (al * 6.0 * PI * np.sin(6.0 * PI * (yl - yl) / Ly) / Ly + bl * np.cos(6.0 * PI * (yl - yl) / Ly)) * s
            gr = (-ar * 6.0 * PI * np.sin(6.0 * PI * (yr - yl) / Ly) / Ly + br * np.cos(6.0 * PI * (yr - yl) / Ly)) * s
    elif g["kind"] == "shipped":
        if periodic:
            f = -(64.0 * PI * PI / (Lx * Lx) + lam) * np.sin(8.0 * PI * (grid - xl) / Lx)
            ex = np.sin(8.0 * PI * (grid - xl) / Lx)
        else:
            f = -(4.0 * PI * PI / (Lx * Lx) + lam) * np.sin(2.0 * PI * (grid - xl) / Lx)
            ex = np.sin(2.0 * PI * (grid - xl) / Lx)
            gl = np.float64(-al * 2.0 * PI * math.cos(2.0 * PI * (xl - xl) / Lx) / Lx + bl * math.sin(0.0))
            gr = np.float64(ar * 2.0 * PI * math.cos(2.0 * PI * (xr - xl) / Lx) / Lx +
                            br * math.sin(2.0 * PI * (xr - xl) / Lx))
    else:  # random smooth modes; the GLL axis mapped to [-1, 1]
        zeta = grid + g["zl"]
        if dim == 2:
            x = xl + np.arange(p["Nx"]) * (Lx / p["Nx"])
            f3, gl3, gr3 = wl.random_problem(np.array([0.0]), x, zeta, seed=g["seed"], nmodes=16,
                                             kmax=g["kmax"], lmax=g["lmax"])
            f, gl, gr = f3[0], gl3[0], gr3[0]
        else:
            f3, gl3, gr3 = wl.random_problem(np.array([0.0]), np.array([0.0]), zeta, seed=g["seed"], nmodes=16,
                                             kmax=0, lmax=g["lmax"])
            f, gl, gr = f3[0, 0], gl3[0, 0], gr3[0, 0]
        if periodic:
            gl = gr = None
    f = np.ascontiguousarray(f, dtype=np.float64)
    if gl is not None:
        gl = np.ascontiguousarray(gl, dtype=np.float64)
        gr = np.ascontiguousarray(gr, dtype=np.float64)
    assert sha(f) == rec["sha_f"], "input generator drifted from the golden vectors"
    if gl is not None:
        assert sha(gl) == rec["sha_gl"] and sha(gr) == rec["sha_gr"], "face generator drifted"
    return f, gl, gr, ex


def lowdim_reference_u(rec):
    """The reference's solution: full array when stored, else None."""
    if "u" in rec:
        shape = (rec["params"]["Nx"], rec["K"]) if rec["dim"] == 2 else (rec["K"],)
        return unhex(rec["u"]).reshape(shape)
    return None


def lowdim_sample(u, rec):
    """u at the record's sample points (2-D: (i, k) pairs) and the reference's values there."""
    if "u" in rec:
        ref = lowdim_reference_u(rec)
        return np.asarray(u).ravel(), ref.ravel()
    idx = rec["sample_idx"]
    return np.array([u[i, k] for i, k in idx]), unhex(rec["sample_u"])


# ---------------------------------------------------------------------------
# Helper fixtures: zderiv.json, fundsol.json, lusolve.json, tecplot.json,
# palpha.json (tests/golden/make_golden.py --helpers / --palpha).  Their
# inputs are seeded uniform fields.

def zderiv_inputs(rec):
    """Element arrays u[E][Nx][Ny][N] of a zderiv.json case."""
    from mpi_amd import workloads as wl
    u = wl.uniform_field((rec["elements"], rec["Nx"], rec["Ny"], rec["N"]), -1.0, 1.0, rec["seed"])
    assert sha(u) == rec["sha_u"], "input generator drifted from the golden vectors"
    return u


def fundsol_inputs(rec):
    """rhsPDE[Ns][N] of the column of a fundsol.json case."""
    from mpi_amd import workloads as wl
    rhs = wl.uniform_field((rec["Ns"], rec["N"]), -1.0, 1.0, rec["seed"])
    assert sha(rhs) == rec["sha_rhs"], "input generator drifted from the golden vectors"
    return rhs


def lusolve_inputs(name, rec):
    """(A, b) of a lusolve.json case."""
    import tests_lu_cases
    from mpi_amd import workloads as wl
    A = dict((nm, a) for nm, _, a in tests_lu_cases.lu_cases())[name]
    b = wl.uniform_field((rec["n"],), -1.0, 1.0, rec["seed"])
    assert sha(A) == rec["sha_input"] and sha(b) == rec["sha_b"], "input generator drifted"
    return A, b


def tecplot_inputs(rec):
    """Element arrays U[Ns][Nx][Ny][N] written by a tecplot.json case."""
    from mpi_amd import workloads as wl
    nx, ny, _, N = (int(v) for v in rec["args"][:4])
    U = wl.uniform_field((rec["Ns"], nx, ny, N), -2.0, 2.0, rec["seed"])
    assert sha(U) == rec["sha_u"], "input generator drifted from the golden vectors"
    return U


def palpha_inputs(rec):
    """f[Nx][Ny][K] of a palpha.json case."""
    from mpi_amd import workloads as wl
    f = wl.uniform_field((rec["Nx"], rec["Ny"], rec["K"]), -1.0, 1.0, rec["seed"])
    assert sha(f) == rec["sha_f"], "input generator drifted from the golden vectors"
    return f
