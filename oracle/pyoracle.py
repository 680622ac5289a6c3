"""ctypes front end of oracle/libhz_oracle.so (test infrastructure only).

The C library restates gpritam/MPI's DirectPseudoSpectral solver
(HelmholtzSolvers/DirectPseudoSpectral/src/*.cpp) in the reference's own
operation order; this module only marshals numpy arrays.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libhz_oracle.so")
# The same C source built with fused multiply-adds (oracle/Makefile): the
# reference's algorithm under a different rounding, used only to measure how
# far rounding alone moves a result (the parity bar of ill-conditioned cases).
_FMA_PATH = os.path.join(_HERE, "libhz_oracle_fma.so")
_lib = None
_fma = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


class Desc(ctypes.Structure):
    _fields_ = [("Nx", ctypes.c_int), ("Ny", ctypes.c_int), ("Nz", ctypes.c_int),
                ("N", ctypes.c_int), ("size", ctypes.c_int),
                ("xl", ctypes.c_double), ("xr", ctypes.c_double),
                ("yl", ctypes.c_double), ("yr", ctypes.c_double),
                ("zl", ctypes.c_double), ("zr", ctypes.c_double),
                ("alphal", ctypes.c_double), ("betal", ctypes.c_double),
                ("alphar", ctypes.c_double), ("betar", ctypes.c_double),
                ("lam", ctypes.c_double)]


class OracleError(RuntimeError):
    pass


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib(fma=False):
    global _lib, _fma
    if (_fma if fma else _lib) is None:
        path = _FMA_PATH if fma else _LIB_PATH
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.hzo_error_string.restype = ctypes.c_char_p
        L.hzo_legendre.restype = ctypes.c_double
        L.hzo_legendre.argtypes = [ctypes.c_double, ctypes.c_int, _I]
        L.hzo_shift.restype = ctypes.c_double
        L.hzo_shift.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        for name in ("hzo_gll_points",):
            getattr(L, name).argtypes = [_D, ctypes.c_int]
        L.hzo_gll_weights.argtypes = [_D, _D, ctypes.c_int]
        L.hzo_first_derivative_matrix.argtypes = [_D, _D, ctypes.c_int]
        L.hzo_second_derivative_matrix.argtypes = [_D, _D, ctypes.c_int]
        L.hzo_plan_operators.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double, _D, _D, _D]
        L.hzo_lu_decompose.argtypes = [_D, _I, _D, ctypes.c_int]
        L.hzo_lu_solve.argtypes = [_D, _I, _D, ctypes.c_int]
        L.hzo_fft_xy.argtypes = [_D, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.hzo_elements_per_rank.argtypes = [ctypes.POINTER(Desc)]
        L.hzo_total_z_points.argtypes = [ctypes.POINTER(Desc)]
        L.hzo_z_grid.argtypes = [ctypes.POINTER(Desc), _D]
        L.hzo_solve.argtypes = [ctypes.POINTER(Desc), _D, _D, _D, _D]
        L.hzo_derivative_z.argtypes = [ctypes.POINTER(Desc), ctypes.c_int, _D, _D]
        L.hzo_solve_periodic.argtypes = [ctypes.POINTER(Desc), _D, _D]
        L.hzo_solve_periodic_alpha.argtypes = [ctypes.POINTER(Desc), _D, _D, _D]
        L.hzo_fundamental_solutions.argtypes = [ctypes.c_int, ctypes.c_int, _D, _I, _D, _D, _D, _D, _D]
        if fma:
            _fma = L
        else:
            _lib = L
    return _fma if fma else _lib


def _p(a):
    return a.ctypes.data_as(_D)


def _check(code):
    if code:
        raise OracleError(lib().hzo_error_string(code).decode())


def gll(N):
    """Nodes, weights, D and Dzz (unscaled) on [-1, 1]."""
    L = lib()
    x = np.zeros(N)
    w = np.zeros(N)
    D = np.zeros((N, N))
    Dzz = np.zeros((N, N))
    _check(L.hzo_gll_points(_p(x), N))
    _check(L.hzo_gll_weights(_p(w), _p(x), N))
    _check(L.hzo_first_derivative_matrix(_p(D), _p(x), N))
    L.hzo_second_derivative_matrix(_p(Dzz), _p(D), N)
    return x, w, D, Dzz


def plan_operators(N, Ns, size, rank, Lz):
    z = np.zeros(Ns * (N - 1) + 1)
    Dz = np.zeros((N, N))
    Dzz = np.zeros((N, N))
    _check(lib().hzo_plan_operators(N, Ns, size, rank, Lz, _p(z), _p(Dz), _p(Dzz)))
    return z, Dz, Dzz


def lu(A):
    """Reference LU of one matrix: (factors in LINPACK storage, pivots[n-1], det)."""
    A = np.array(A, dtype=np.float64, order="C", copy=True)
    n = A.shape[0]
    piv = np.zeros(n, dtype=np.int32)
    det = ctypes.c_double(0.0)
    _check(lib().hzo_lu_decompose(_p(A), piv.ctypes.data_as(_I), ctypes.byref(det), n))
    return A, piv[: n - 1].copy(), det.value


def lu_solve(LU, piv, b):
    n = LU.shape[0]
    LU = np.ascontiguousarray(LU, dtype=np.float64)
    pv = np.zeros(n, dtype=np.int32)
    pv[: n - 1] = piv
    x = np.array(b, dtype=np.float64, copy=True)
    lib().hzo_lu_solve(_p(LU), pv.ctypes.data_as(_I), _p(x), n)
    return x


def fft_xy(A, isign):
    """In-place reference transform of A[2Nx][2Ny][Nz] (or [2Nx][2Ny])."""
    A = np.array(A, dtype=np.float64, order="C", copy=True)
    nz = 1 if A.ndim == 2 else A.shape[2]
    _check(lib().hzo_fft_xy(_p(A), isign, A.shape[0] // 2, A.shape[1] // 2, nz))
    return A


def shift(lam, iw, jw, Lx, Ly):
    return lib().hzo_shift(lam, iw, jw, Lx, Ly)


def make_desc(Nx, Ny, Nz, N, size=1, xl=0.0, xr=1.0, yl=0.0, yr=1.0, zl=0.0, zr=1.0,
              alphal=0.0, betal=1.0, alphar=0.0, betar=1.0, lam=1.0):
    return Desc(Nx, Ny, Nz, N, size, xl, xr, yl, yr, zl, zr, alphal, betal, alphar, betar, lam)


def total_z_points(desc):
    return lib().hzo_total_z_points(ctypes.byref(desc))


def z_grid(desc):
    z = np.zeros(total_z_points(desc))
    _check(lib().hzo_z_grid(ctypes.byref(desc), _p(z)))
    return z


def solve(desc, f, gl, gr, fma=False):
    """Full reference SolvePDE with desc.size ranks simulated in-process
    (fma: the fused-multiply-add build, for rounding-sensitivity checks)."""
    K = total_z_points(desc)
    f = np.ascontiguousarray(f, dtype=np.float64).reshape(desc.Nx, desc.Ny, K)
    gl = np.ascontiguousarray(gl, dtype=np.float64).reshape(desc.Nx, desc.Ny)
    gr = np.ascontiguousarray(gr, dtype=np.float64).reshape(desc.Nx, desc.Ny)
    u = np.zeros_like(f)
    _check(lib(fma).hzo_solve(ctypes.byref(desc), _p(f), _p(gl), _p(gr), _p(u)))
    return u


def solve_periodic(desc, f, fma=False, alpha=False):
    """Reference periodic-z SolvePDE (Periodic3D_MPI.cpp) at size = 1; with
    alpha=True also its closure coefficients alphaPDE[Nx][2 Ny] (zeros for
    one element)."""
    K = total_z_points(desc)
    f = np.ascontiguousarray(f, dtype=np.float64).reshape(desc.Nx, desc.Ny, K)
    u = np.zeros_like(f)
    if not alpha:
        _check(lib(fma).hzo_solve_periodic(ctypes.byref(desc), _p(f), _p(u)))
        return u
    a = np.zeros((desc.Nx, 2 * desc.Ny))
    _check(lib(fma).hzo_solve_periodic_alpha(ctypes.byref(desc), _p(f), _p(u), _p(a)))
    return u, a


def derivative_z(desc, u, order):
    """FirstDerivativeZ (order 1) / SecondDerivativeZ (order 2) of element
    arrays u[size*Ns][Nx][Ny][N]."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros_like(u)
    _check(lib().hzo_derivative_z(ctypes.byref(desc), order, _p(u), _p(out)))
    return out


def fundamental_solutions(N, LU, piv, Dzz, rhs):
    """ConstructFundamentalSolutions for one column: rhs[Ns][N] (interior
    entries used) -> (U[Ns][N], ul[N], ur[N]) with factors from lu()."""
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    Ns = rhs.shape[0]
    LU = np.ascontiguousarray(LU, dtype=np.float64)
    pv = np.zeros(N - 2, dtype=np.int32)
    pv[: N - 3] = piv
    Dzz = np.ascontiguousarray(Dzz, dtype=np.float64)
    U = np.zeros((Ns, N))
    ul = np.zeros(N)
    ur = np.zeros(N)
    _check(lib().hzo_fundamental_solutions(N, Ns, _p(LU), pv.ctypes.data_as(_I), _p(Dzz), _p(rhs), _p(U), _p(ul),
                                           _p(ur)))
    return U, ul, ur
