"""LU test matrices shared by the golden generator and the tests."""
import numpy as np

from mpi_amd import workloads as wl
from oracle import pyoracle as po


def lu_cases():
    cases = []
    for N, lz in ((10, 1.0), (33, 2.0), (129, 2.0)):
        n = N - 2
        _, _, Dzz = po.plan_operators(N, 1, 1, 0, lz)
        for s in (1e-2, 1.0, 10.0, 1e3, 3.3e4):
            cases.append((f"gll{N}_shift{s:g}", n, Dzz[1:-1, 1:-1] - s * np.eye(n)))
    rng = wl.SplitMix64(20260101)
    for n in (5, 8, 31, 64, 127):
        A = np.array([[rng.normal() for _ in range(n)] for _ in range(n)])
        cases.append((f"random{n}", n, A))
    n = 12
    A = np.array([[rng.normal() * 10.0 ** (3 * (i % 4) - 4) for _ in range(n)] for i in range(n)])
    cases.append(("scaled12", n, A))
    return cases
