"""Writes the golden vectors under tests/golden/ by running gpritam/MPI's own
DirectPseudoSpectral code, compiled from the reference sources by
``make -C oracle ref`` into oracle/_ref/ref_driver (this container only; the
reference and that binary never reach the GPU machine).

Every fixture stores inputs by their generator (seeds and parameters, plus a
SHA-256 of the generated input bytes) and the reference's outputs either in
full (small cases, hex floats) or as a SHA-256 of the exact output bytes
plus sampled values and norms (large cases).

    python tests/golden/make_golden.py              # all but the largest case
    python tests/golden/make_golden.py --config4     # 512 x 512 x 257 (minutes)
    python tests/golden/make_golden.py --periodic    # the periodic-z solver
    python tests/golden/make_golden.py --fderiv      # Fourier derivatives
    python tests/golden/make_golden.py --lowdim      # the 2-D and 1-D solvers
"""
import hashlib
import json
import math
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)
from mpi_amd import workloads as wl  # noqa: E402
from oracle import pyoracle as po  # noqa: E402  (for z grids / sizes only)
sys.path.insert(0, os.path.join(ROOT, "tests"))
from tests_lu_cases import lu_cases  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
REFP = os.path.join(ROOT, "oracle", "_ref", "ref_driver_periodic")
MPIEXEC = "/opt/conda/bin/mpiexec"
TMP = tempfile.mkdtemp(prefix="golden_")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def hexs(a):
    return [float(v).hex() for v in np.asarray(a, dtype=np.float64).ravel()]


def run(args, **kw):
    env = dict(os.environ, PATH="/opt/conda/bin:" + os.environ["PATH"])
    return subprocess.run(args, check=True, capture_output=True, text=True, env=env, **kw)


def gll_fixture():
    out = {}
    for N in list(range(2, 41)) + [65, 129, 257]:
        path = os.path.join(TMP, "g.bin")
        run([REF, "gll", str(N), path])
        a = np.fromfile(path)
        x, w, D, Dzz = a[:N], a[N:2 * N], a[2 * N:2 * N + N * N], a[2 * N + N * N:]
        e = {"sha_x": sha(x), "sha_w": sha(w), "sha_D": sha(D), "sha_Dzz": sha(Dzz),
             "x": hexs(x), "w": hexs(w), "D_row0": hexs(D[:N]), "Dzz_diag": hexs(Dzz.reshape(N, N).diagonal())}
        out[str(N)] = e
    return out


def lu_fixture():
    out = {}
    for name, n, A in lu_cases():
        inp = os.path.join(TMP, "lu_in.bin")
        outp = os.path.join(TMP, "lu_out.bin")
        A.tofile(inp)
        run([REF, "lu", str(n), "1", inp, outp])
        raw = np.fromfile(outp, dtype=np.uint8)
        dd = raw[: (n * n + 1) * 8].view(np.float64)
        piv = raw[(n * n + 1) * 8:].view(np.int32)[: n - 1]
        out[name] = {"n": n, "sha_input": sha(A), "sha_factors": sha(dd[: n * n]), "det": float(dd[-1]).hex(),
                     "pivots": [int(v) for v in piv], "factors_head": hexs(dd[: min(n * n, 64)])}
    return out


def fft_fixture():
    out = {}
    rng = wl.SplitMix64(77)
    for nx, ny, nz in ((8, 4, 3), (32, 32, 5), (16, 64, 1), (4, 8, 2)):
        A = np.zeros((2 * nx, 2 * ny, nz))
        A[:nx, :ny, :] = np.array([rng.normal() for _ in range(nx * ny * nz)]).reshape(nx, ny, nz)
        inp = os.path.join(TMP, "f_in.bin")
        outp = os.path.join(TMP, "f_out.bin")
        A.tofile(inp)
        rec = {"shape": [nx, ny, nz], "sha_input": sha(A)}
        for isign in (1, -1):
            run([REF, "fft", str(nx), str(ny), str(nz), str(isign), inp, outp])
            r = np.fromfile(outp)
            rec[f"sha_out_{isign}"] = sha(r)
            rec[f"head_{isign}"] = hexs(r[:32])
        out[f"{nx}x{ny}x{nz}"] = rec
    return out


def fderiv_fixture():
    """Fourier derivatives along x and y of 3-D arrays (Fourier_Quadrature.cpp:1494-1575, :1721-1800)."""
    out = {}
    rng = wl.SplitMix64(91)
    for nx, ny, nz, lx, ly in ((16, 8, 3, 2 * math.pi, 3.0), (64, 32, 5, 1.0, 2.0), (8, 64, 2, 5.0, 2 * math.pi)):
        F = np.array([rng.normal() for _ in range(nx * ny * nz)]).reshape(nx, ny, nz)
        inp = os.path.join(TMP, "d_in.bin")
        outp = os.path.join(TMP, "d_out.bin")
        F.tofile(inp)
        run([REF, "fderiv", str(nx), str(ny), str(nz), repr(lx), repr(ly), inp, outp])
        r = np.fromfile(outp).reshape(4, nx, ny, nz)
        out[f"{nx}x{ny}x{nz}"] = {"shape": [nx, ny, nz], "L": [lx, ly], "sha_input": sha(F),
                                  "sha_out": [sha(r[w]) for w in range(4)],
                                  "out": [hexs(r[w].ravel()) for w in range(4)] if F.size <= 4096 else None,
                                  "head": [hexs(r[w].ravel()[:64]) for w in range(4)],
                                  "max_abs": [float(np.abs(r[w]).max()) for w in range(4)]}
    return out


def ref_solve(desc, f, gl, gr):
    K = po.total_z_points(desc)
    inp = os.path.join(TMP, "s_in.bin")
    outp = os.path.join(TMP, "s_out.bin")
    np.concatenate([f.ravel(), gl.ravel(), gr.ravel()]).tofile(inp)
    args = [MPIEXEC, "-np", str(desc.size), REF, "solve", str(desc.Nx), str(desc.Ny), str(desc.Nz), str(desc.N)]
    args += [repr(v) for v in (desc.xl, desc.xr, desc.yl, desc.yr, desc.zl, desc.zr,
                               desc.alphal, desc.betal, desc.alphar, desc.betar, desc.lam)]
    args += [inp, outp]
    res = run(args)
    Ns = (K - 1) // (desc.size * (desc.N - 1))
    N = desc.N
    u = np.zeros((desc.Nx, desc.Ny, K))
    for r in range(desc.size):
        a = np.fromfile(f"{outp}.rank{r}")
        blk = a[: Ns * desc.Nx * desc.Ny * N].reshape(Ns, desc.Nx, desc.Ny, N)
        for kk in range(Ns):
            g = (r * Ns + kk) * (N - 1)
            u[:, :, g:g + N] = blk[kk]
    secs = float(res.stdout.split()[1])
    return u, secs


def ref_solve_periodic(desc, f):
    """Periodic3D_MPI SolvePDE under mpiexec -np desc.size."""
    K = po.total_z_points(desc)
    inp = os.path.join(TMP, "p_in.bin")
    outp = os.path.join(TMP, "p_out.bin")
    np.ascontiguousarray(f).ravel().tofile(inp)
    args = [MPIEXEC, "-np", str(desc.size), REFP, "psolve", str(desc.Nx), str(desc.Ny), str(desc.Nz), str(desc.N)]
    args += [repr(v) for v in (desc.xl, desc.xr, desc.yl, desc.yr, desc.zl, desc.zr, desc.lam)]
    args += [inp, outp]
    res = run(args)
    Ns = (K - 1) // (desc.size * (desc.N - 1))
    N = desc.N
    u = np.zeros((desc.Nx, desc.Ny, K))
    for r in range(desc.size):
        a = np.fromfile(f"{outp}.rank{r}")
        blk = a[: Ns * desc.Nx * desc.Ny * N].reshape(Ns, desc.Nx, desc.Ny, N)
        for kk in range(Ns):
            g = (r * Ns + kk) * (N - 1)
            u[:, :, g:g + N] = blk[kk]
    return u, float(res.stdout.split()[1])


LOWREF = {(2, False): "ref_driver_2d", (1, False): "ref_driver_1d",
          (2, True): "ref_driver_p2d", (1, True): "ref_driver_p1d"}


def ref_solve_low(dim, periodic, p, grid_only=False, f=None, gl=None, gr=None):
    """NonPeriodic/Periodic 2D/1D_MPI SolvePDE under mpiexec -np p['size'].
    Returns (u, grid, seconds); u is [Nx][K] (2-D) or [K] (1-D)."""
    exe = os.path.join(ROOT, "oracle", "_ref", LOWREF[(dim, periodic)])
    N, size = p["N"], p["size"]
    along = p["Ny"] if dim == 2 else p["Nx"]
    Ns = -(-along // (size * N))
    K = size * Ns * (N - 1) + 1
    lines = p["Nx"] if dim == 2 else 1
    inp = os.path.join(TMP, "l_in.bin")
    outp = os.path.join(TMP, "l_out.bin")
    if grid_only:
        f = np.zeros((lines, K))
        gl = gr = np.zeros(lines)
    parts = [np.asarray(f, dtype=np.float64).ravel()]
    if not periodic:
        parts += [np.atleast_1d(gl).astype(np.float64).ravel(), np.atleast_1d(gr).astype(np.float64).ravel()]
    np.concatenate(parts).tofile(inp)
    args = [MPIEXEC, "-np", str(size), exe, "solve", str(p["Nx"])]
    if dim == 2:
        args += [str(p["Ny"])]
    args += [str(N), repr(p["xl"]), repr(p["xr"])]
    if dim == 2:
        args += [repr(p["yl"]), repr(p["yr"])]
    args += [repr(p[k]) for k in ("al", "bl", "ar", "br", "lam")] + [inp, outp]
    res = run(args)
    u = np.zeros((lines, K))
    grid = np.zeros(K)
    for r in range(size):
        a = np.fromfile(f"{outp}.rank{r}")
        blk = a[: Ns * lines * N].reshape(Ns, lines, N)
        loc = a[Ns * lines * N:]
        for kk in range(Ns):
            gidx = (r * Ns + kk) * (N - 1)
            u[:, gidx:gidx + N] = blk[kk]
            grid[gidx:gidx + N] = loc[kk * (N - 1):kk * (N - 1) + N]
    if dim == 1:
        u = u[0]
    return u, grid, float(res.stdout.split()[1])


def lowdim_case(name, dim, periodic, p, gen):
    """One 2-D / 1-D golden case: the reference's grid first (a zero solve),
    then the inputs regenerated by tests/golden_util.lowdim_inputs and the
    reference's solution of them."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import golden_util as gu
    _, grid, _ = ref_solve_low(dim, periodic, p, grid_only=True)
    K = grid.size
    rec = {"dim": dim, "periodic": periodic, "params": p, "gen": gen, "K": K,
           "elements": (K - 1) // (p["N"] - 1), "grid": hexs(grid)}
    # inputs first with placeholder hashes, then the real ones
    rec["sha_f"] = rec["sha_gl"] = rec["sha_gr"] = None
    gu_sha, gu.sha = gu.sha, (lambda a: None)
    try:
        f, gl, gr, ex = gu.lowdim_inputs(rec)
    finally:
        gu.sha = gu_sha
    rec["sha_f"] = sha(f)
    if gl is not None:
        rec["sha_gl"], rec["sha_gr"] = sha(gl), sha(gr)
    u, grid2, secs = ref_solve_low(dim, periodic, p, f=f, gl=gl, gr=gr)
    assert np.array_equal(grid, grid2)
    rec.update({"sha_u": sha(u), "max_abs_u": float(np.abs(u).max()), "ref_seconds": secs})
    # How far rounding alone moves the reference's algorithm on this case:
    # the oracle built with fused multiply-adds against the reference's output.
    from test_oracle_lowdim import as3d, oracle_desc
    d = oracle_desc(rec)
    F, GL, GR = as3d(rec, f, gl, gr)
    uf = (po.solve_periodic(d, F, fma=True) if periodic else po.solve(d, F, GL, GR, fma=True)).reshape(u.shape)
    rec["rounding_spread"] = float(np.abs(uf - u).max() / np.abs(u).max())
    if ex is not None:
        rec["err_vs_exact"] = float(np.abs(u - ex).max())
    if u.size <= 20000:
        rec["u"] = hexs(u)
    else:
        rng = wl.SplitMix64(99)
        idx = [[rng.randint(0, s - 1) for s in u.shape] for _ in range(4096)]
        rec["sample_idx"] = idx
        rec["sample_u"] = hexs([u[i, k] for i, k in idx])
    with open(os.path.join(HERE, f"low_{name}.json"), "w") as fp:
        json.dump(rec, fp, indent=0)
    print(f"{name}: K {K} elements {rec['elements']} ref {secs:.3f}s err_vs_exact {rec.get('err_vs_exact')} "
          f"rounding_spread {rec['rounding_spread']:.2e}")


def lowdim_fixtures():
    two = dict(xl=0.0, xr=1.0, yl=0.0, yr=1.0)
    robin_a = dict(al=1.0, bl=0.0, ar=0.0, br=1.0)
    none = dict(al=0.0, bl=1.0, ar=0.0, br=1.0)
    shipped = {"kind": "shipped"}
    for size in (1, 4):  # TestNonPeriodic2D_MPI.cpp:84-118
        lowdim_case(f"2d_shipped_np{size}", 2, False, dict(Nx=512, Ny=512, N=10, **two, **robin_a, lam=1.0, size=size),
                    shipped)
    lowdim_case("p2d_shipped_np1", 2, True, dict(Nx=512, Ny=512, N=10, **two, **none, lam=1.0, size=1), shipped)
    tau = 2 * math.pi
    lowdim_case("2d_random_64_N17", 2, False,
                dict(Nx=64, Ny=51, N=17, xl=0.0, xr=tau, yl=-1.0, yr=1.0, al=0.5, bl=1.0, ar=0.25, br=2.0, lam=10.0,
                     size=1), {"kind": "random", "seed": 11, "kmax": 16, "lmax": 16, "zl": -1.0})
    lowdim_case("2d_random_128_N33", 2, False,
                dict(Nx=128, Ny=33, N=33, xl=0.0, xr=tau, yl=-1.0, yr=1.0, **none, lam=5.0, size=1),
                {"kind": "random", "seed": 12, "kmax": 32, "lmax": 32, "zl": -1.0})
    lowdim_case("p2d_random_64_N17", 2, True,
                dict(Nx=64, Ny=51, N=17, xl=0.0, xr=tau, yl=-1.0, yr=1.0, **none, lam=3.0, size=1),
                {"kind": "random", "seed": 13, "kmax": 16, "lmax": 16, "zl": -1.0})
    one = dict(xl=0.0, xr=1.0)
    for size in (1, 4):  # TestNonPeriodic1D_MPI.cpp:91-118, first problem
        lowdim_case(f"1d_shipped_a_np{size}", 1, False, dict(Nx=512, N=8, **one, **robin_a, lam=1.0, size=size),
                    shipped)
    lowdim_case("1d_shipped_b_np1", 1, False,  # :136-149, second problem
                dict(Nx=512, N=8, **one, al=1.0, bl=1.0, ar=0.0, br=1.0, lam=5.0, size=1), shipped)
    lowdim_case("p1d_shipped_np1", 1, True, dict(Nx=512, N=8, **one, **none, lam=1.0, size=1), shipped)
    lowdim_case("1d_random_N33", 1, False,
                dict(Nx=66, N=33, xl=-1.0, xr=1.0, al=0.5, bl=1.0, ar=0.25, br=2.0, lam=10.0, size=1),
                {"kind": "random", "seed": 14, "kmax": 0, "lmax": 32, "zl": -1.0})
    lowdim_case("p1d_random_N17", 1, True,
                dict(Nx=68, N=17, xl=-1.0, xr=1.0, **none, lam=2.0, size=1),
                {"kind": "random", "seed": 15, "kmax": 0, "lmax": 16, "zl": -1.0})


def sample_idx(shape, count, seed):
    rng = wl.SplitMix64(seed)
    return [[rng.randint(0, s - 1) for s in shape] for _ in range(count)]


def solve_case(name, desc, f, gl, gr, exact=None, full=False, gen=None):
    u, secs = ref_solve(desc, f, gl, gr)
    rec = {"desc": {k: getattr(desc, k) for k, _ in desc._fields_}, "generator": gen,
           "sha_f": sha(f), "sha_gl": sha(gl), "sha_gr": sha(gr), "sha_u": sha(u),
           "max_abs_u": float(np.abs(u).max()), "ref_seconds": secs}
    if exact is not None:
        rec["err_vs_exact"] = float(np.abs(u - exact).max())
    idx = sample_idx(u.shape, 4096 if not full else 256, 99)
    rec["sample_idx"] = idx
    rec["sample_u"] = hexs([u[i, j, k] for i, j, k in idx])
    np_path = os.path.join(HERE, f"solve_{name}.json")
    with open(np_path, "w") as fp:
        json.dump(rec, fp, indent=0)
    print(f"{name}: ref {secs:.3f}s err_vs_exact {rec.get('err_vs_exact')}")


def periodic_case(name, desc, f, exact=None, gen=None):
    u, secs = ref_solve_periodic(desc, f)
    rec = {"desc": {k: getattr(desc, k) for k, _ in desc._fields_}, "generator": gen, "periodic": True,
           "sha_f": sha(f), "sha_u": sha(u), "max_abs_u": float(np.abs(u).max()), "ref_seconds": secs}
    if exact is not None:
        rec["err_vs_exact"] = float(np.abs(u - exact).max())
    idx = sample_idx(u.shape, 4096, 99)
    rec["sample_idx"] = idx
    rec["sample_u"] = hexs([u[i, j, k] for i, j, k in idx])
    with open(os.path.join(HERE, f"psolve_{name}.json"), "w") as fp:
        json.dump(rec, fp, indent=0)
    print(f"periodic {name}: ref {secs:.3f}s err_vs_exact {rec.get('err_vs_exact')}")


def periodic_shipped_inputs(size):
    """Inputs of the reference's periodic test program (TestPeriodic3D_MPI.cpp:27-41, :96-107)."""
    PI = 3.1415926535897932
    d = po.make_desc(64, 32, 250, 10, size=size, xl=0, xr=1, yl=0, yr=1, zl=0, zr=1, lam=1.0)
    z = po.z_grid(d)
    x = np.arange(64) / 64.0
    y = np.arange(32) / 32.0
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    f = -(64 * PI * PI + 36 * PI * PI + 16 * PI * PI + 1.0) * np.sin(8 * PI * X) * np.cos(6 * PI * Y) * np.cos(4 * PI * Z)
    ex = np.sin(8 * PI * X) * np.cos(6 * PI * Y) * np.cos(4 * PI * Z)
    return d, np.ascontiguousarray(f), ex


def periodic_random_inputs(nx, N, nz, lam, seed):
    d = po.make_desc(nx, nx, nz, N, size=1, xl=0, xr=2 * math.pi, yl=0, yr=2 * math.pi, zl=-1, zr=1, lam=lam)
    x, y, zeta = wl.grid(nx, nx, po.z_grid(d))
    f, _, _ = wl.random_problem(x, y, zeta, seed=seed, nmodes=32, kmax=nx // 4, lmax=min(N - 1, 32))
    return d, np.ascontiguousarray(f)


def shipped_inputs(size):
    """The reference's own test program, TestNonPeriodic3D_MPI.cpp:26-128."""
    PI = 3.1415926535897932
    d = po.make_desc(64, 32, 250, 10, size=size, xl=0, xr=1, yl=0, yr=1, zl=0, zr=1,
                     alphal=1.0, betal=0.0, alphar=0.0, betar=1.0, lam=1.0)
    z = po.z_grid(d)
    x = np.arange(64) / 64.0
    y = np.arange(32) / 32.0
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    f = -(64 * PI * PI + 36 * PI * PI + 16 * PI * PI + 1.0) * np.sin(8 * PI * X) * np.cos(6 * PI * Y) * np.cos(4 * PI * Z)
    XX, YY = np.meshgrid(x, y, indexing="ij")
    gl = (1.0 * 4.0 * PI * np.sin(0.0) / 1.0 + 0.0 * np.cos(0.0)) * np.sin(8 * PI * XX) * np.cos(6 * PI * YY)
    gr = (-0.0 * 4.0 * PI * np.sin(4 * PI) / 1.0 + 1.0 * np.cos(4 * PI)) * np.sin(8 * PI * XX) * np.cos(6 * PI * YY)
    ex = np.sin(8 * PI * X) * np.cos(6 * PI * Y) * np.cos(4 * PI * Z)
    return d, f, gl, gr, ex


def config1_inputs(N, size):
    d = po.make_desc(32, 32, 33, N, size=size, xl=0, xr=2 * math.pi, yl=0, yr=2 * math.pi, zl=-1, zr=1,
                     alphal=0, betal=1, alphar=0, betar=1, lam=1.0)
    x, y, zeta = wl.grid(32, 32, po.z_grid(d))
    f, gl, gr, u = wl.manufactured_config1(x, y, zeta, 1.0)
    return d, f, gl, gr, u


def config2_inputs(N, size, lam=10.0):
    d = po.make_desc(256, 256, 129, N, size=size, xl=0, xr=2 * math.pi, yl=0, yr=2 * math.pi, zl=-1, zr=1,
                     alphal=0, betal=1, alphar=0, betar=1, lam=lam)
    x, y, zeta = wl.grid(256, 256, po.z_grid(d))
    f, gl, gr = wl.random_problem(x, y, zeta, seed=2, nmodes=64, kmax=32, lmax=64)
    return d, f, gl, gr


def config4_inputs():
    """512 x 512 x 257 (one element), lambda = 100, 256 random modes."""
    d = po.make_desc(512, 512, 257, 257, size=1, xl=0, xr=2 * math.pi, yl=0, yr=2 * math.pi, zl=-1, zr=1,
                     alphal=0, betal=1, alphar=0, betar=1, lam=100.0)
    x, y, zeta = wl.grid(512, 512, po.z_grid(d))
    f, gl, gr = wl.random_problem(x, y, zeta, seed=4, nmodes=256, kmax=32, lmax=64)
    return d, f, gl, gr


def main():
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -C oracle ref")
    if "--lowdim" in sys.argv:
        lowdim_fixtures()
        return
    if "--fderiv" in sys.argv:
        with open(os.path.join(HERE, "fderiv.json"), "w") as fp:
            json.dump(fderiv_fixture(), fp, indent=0)
        return
    if "--periodic" in sys.argv:
        for size in (1, 4):
            d, f, ex = periodic_shipped_inputs(size)
            periodic_case(f"shipped_np{size}", d, f, ex, gen="TestPeriodic3D_MPI.cpp inputs")
        for nx, N, nz, lam, seed in ((64, 33, 129, 10.0, 7), (128, 129, 129, 10.0, 8), (32, 9, 33, 0.5, 9)):
            d, f = periodic_random_inputs(nx, N, nz, lam, seed)
            periodic_case(f"random_{nx}_N{N}", d, f,
                          gen=f"random_problem(seed={seed}, nmodes=32, kmax={nx // 4}, lmax={min(N - 1, 32)}) lam={lam}")
        return
    if "--config4" in sys.argv:  # minutes of CPU time: on request only
        d, f, gl, gr = config4_inputs()
        solve_case("config4_N257_np1", d, f, gl, gr,
                   gen="random_problem(seed=4, nmodes=256, kmax=32, lmax=64) lam=100")
        return
    with open(os.path.join(HERE, "gll.json"), "w") as fp:
        json.dump(gll_fixture(), fp, indent=0)
    with open(os.path.join(HERE, "lu.json"), "w") as fp:
        json.dump(lu_fixture(), fp, indent=0)
    with open(os.path.join(HERE, "fft.json"), "w") as fp:
        json.dump(fft_fixture(), fp, indent=0)
    for size in (1, 4):
        d, f, gl, gr, ex = shipped_inputs(size)
        solve_case(f"shipped_np{size}", d, f, gl, gr, ex, gen="TestNonPeriodic3D_MPI.cpp inputs")
    for N, size in ((33, 1), (9, 1), (9, 4)):
        d, f, gl, gr, ex = config1_inputs(N, size)
        solve_case(f"config1_N{N}_np{size}", d, f, gl, gr, ex, gen="manufactured_config1 lam=1")
    for N, size in ((129, 1), (33, 4)):
        d, f, gl, gr = config2_inputs(N, size)
        solve_case(f"config2_N{N}_np{size}", d, f, gl, gr,
                   gen="random_problem(seed=2, nmodes=64, kmax=32, lmax=64) lam=10")


if __name__ == "__main__":
    main()
