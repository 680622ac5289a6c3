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
