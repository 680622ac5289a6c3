"""Seeded synthetic inputs for the benchmark configurations.

Random numbers come from a SplitMix64 stream (Steele, Lea & Flood 2014)
with Box-Muller normals, written here so that the same seed gives the same
inputs on every machine.  Fields are evaluated in numpy on the host in the
solver's layout [nx, ny, nz] (z fastest).

Domain of every configuration: x, y in [0, 2 pi), zeta in [-1, 1].  The
reference's z array runs from 0 (NonPeriodic3D_MPI.cpp:113-119), so fields
are evaluated at zeta = z + zl.
"""
import math

import numpy as np

_MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & _MASK

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self):
        """Uniform in [0, 1) with 53 random bits."""
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def randint(self, lo, hi):
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1))

    def normal(self):
        u1 = self.uniform()
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(1.0 - u1)) * math.cos(2.0 * math.pi * u2)


def uniform_stream(count, seed):
    """The first `count` uniforms of SplitMix64(seed), vectorised: the i-th
    state is seed + (i+1) * gamma (mod 2^64), then the same finaliser."""
    i = np.arange(1, count + 1, dtype=np.uint64)
    z = np.uint64(seed) + i * np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def uniform_field(shape, lo, hi, seed):
    """Seeded field with entries uniform in [lo, hi)."""
    n = int(np.prod(shape))
    return (lo + (hi - lo) * uniform_stream(n, seed)).reshape(shape)


def legendre_table(zeta, lmax):
    """P_l(zeta) for l = 0..lmax, shape [lmax+1, len(zeta)] (three-term recurrence)."""
    P = np.zeros((lmax + 1, len(zeta)))
    P[0] = 1.0
    if lmax >= 1:
        P[1] = zeta
    for l in range(2, lmax + 1):
        P[l] = ((2 * l - 1) * zeta * P[l - 1] - (l - 1) * P[l - 2]) / l
    return P


def random_modes_field(x, y, zeta, rng, nmodes, kmax, lmax, amp=1.0):
    """sum_m c_m cos(a_m x + b_m y + phi_m) P_{l_m}(zeta); a, b ~ U{0..kmax}, l ~ U{0..lmax}."""
    a = np.array([rng.randint(0, kmax) for _ in range(nmodes)], dtype=np.float64)
    b = np.array([rng.randint(0, kmax) for _ in range(nmodes)], dtype=np.float64)
    l = np.array([rng.randint(0, lmax) for _ in range(nmodes)])
    c = np.array([amp * rng.normal() for _ in range(nmodes)])
    ph = np.array([2 * math.pi * rng.uniform() for _ in range(nmodes)])
    P = legendre_table(zeta, lmax)[l]                        # [m, nz]
    arg = a[:, None, None] * x[None, :, None] + b[:, None, None] * y[None, None, :] + ph[:, None, None]
    C = c[:, None, None] * np.cos(arg)                       # [m, nx, ny]
    return np.einsum("mij,mk->ijk", C, P, optimize=True)


def random_modes_face(x, y, rng, nmodes, kmax, amp):
    p = np.array([rng.randint(0, kmax) for _ in range(nmodes)], dtype=np.float64)
    q = np.array([rng.randint(0, kmax) for _ in range(nmodes)], dtype=np.float64)
    d = np.array([amp * rng.normal() for _ in range(nmodes)])
    ps = np.array([2 * math.pi * rng.uniform() for _ in range(nmodes)])
    arg = p[:, None, None] * x[None, :, None] + q[:, None, None] * y[None, None, :] + ps[:, None, None]
    return np.einsum("m,mij->ij", d, np.cos(arg))


def grid(nx, ny, z, xl=0.0, xr=2 * math.pi, yl=0.0, yr=2 * math.pi, zl=-1.0):
    x = xl + np.arange(nx) * ((xr - xl) / nx)
    y = yl + np.arange(ny) * ((yr - yl) / ny)
    return x, y, np.asarray(z) + zl


def manufactured_config1(x, y, zeta, lam=1.0):
    """u = sin x cos 2y (1 - zeta^2) e^zeta and its f, g_l = g_r = 0 (Dirichlet)."""
    X, Y, Z = np.meshgrid(x, y, zeta, indexing="ij")
    u = np.sin(X) * np.cos(2 * Y) * (1 - Z ** 2) * np.exp(Z)
    f = np.sin(X) * np.cos(2 * Y) * np.exp(Z) * ((-1 - 4 * Z - Z ** 2) - (5 + lam) * (1 - Z ** 2))
    g = np.zeros((len(x), len(y)))
    return f, g, g.copy(), u


def random_problem(x, y, zeta, seed, nmodes, kmax, lmax, bc_modes=8, bc_kmax=8, bc_amp=0.1):
    """Config-2 style inputs: f from `nmodes` random modes, Dirichlet faces from
    `bc_modes` smooth random modes of amplitude N(0, bc_amp)."""
    rng = SplitMix64(seed)
    f = random_modes_field(x, y, zeta, rng, nmodes, kmax, lmax)
    gl = random_modes_face(x, y, rng, bc_modes, bc_kmax, bc_amp)
    gr = random_modes_face(x, y, rng, bc_modes, bc_kmax, bc_amp)
    return f, gl, gr


def batch_specs(count=512, seed=3):
    """Sizes and shifts of the many-problem workload: nx, ny uniform in
    {32, 64, 128}, N uniform over the odd values 17..129 (one element),
    lambda log-uniform in [1e-2, 1e3]."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        nx = (32, 64, 128)[rng.randint(0, 2)]
        ny = (32, 64, 128)[rng.randint(0, 2)]
        N = 17 + 2 * rng.randint(0, 56)
        lam = 10.0 ** (-2.0 + 5.0 * rng.uniform())
        out.append({"nx": nx, "ny": ny, "N": N, "lam": lam, "seed": rng.next_u64() & 0xFFFFFFFF})
    return out


def batch_fields(spec, zeta, nmodes=16, bc_modes=8):
    """f = `nmodes` random modes cos(a x + b y + phi) P_l(zeta) with
    a <= nx/2 - 1, b <= ny/2 - 1, l <= N - 1; Dirichlet faces from
    `bc_modes` smooth modes of amplitude N(0, 0.1)."""
    nx, ny, N = spec["nx"], spec["ny"], spec["N"]
    x = np.arange(nx) * (2 * math.pi / nx)
    y = np.arange(ny) * (2 * math.pi / ny)
    rng = SplitMix64(spec["seed"])
    a = np.array([rng.randint(0, nx // 2 - 1) for _ in range(nmodes)], dtype=np.float64)
    b = np.array([rng.randint(0, ny // 2 - 1) for _ in range(nmodes)], dtype=np.float64)
    l = np.array([rng.randint(0, N - 1) for _ in range(nmodes)])
    c = np.array([rng.normal() for _ in range(nmodes)])
    ph = np.array([2 * math.pi * rng.uniform() for _ in range(nmodes)])
    P = legendre_table(np.asarray(zeta), N - 1)[l]
    arg = a[:, None, None] * x[None, :, None] + b[:, None, None] * y[None, None, :] + ph[:, None, None]
    f = np.einsum("mij,mk->ijk", c[:, None, None] * np.cos(arg), P, optimize=True)
    kx, ky = min(8, nx // 2 - 1), min(8, ny // 2 - 1)
    faces = []
    for _ in range(2):
        p = np.array([rng.randint(0, kx) for _ in range(bc_modes)], dtype=np.float64)
        q = np.array([rng.randint(0, ky) for _ in range(bc_modes)], dtype=np.float64)
        d = np.array([0.1 * rng.normal() for _ in range(bc_modes)])
        ps = np.array([2 * math.pi * rng.uniform() for _ in range(bc_modes)])
        arg = p[:, None, None] * x[None, :, None] + q[:, None, None] * y[None, None, :] + ps[:, None, None]
        faces.append(np.einsum("m,mij->ij", d, np.cos(arg)))
    return np.ascontiguousarray(f), faces[0], faces[1]
