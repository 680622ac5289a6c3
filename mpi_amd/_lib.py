"""ctypes binding of libhz.so (the C ABI declared in include/hz.h).

The shared library is built in-tree by ``make -C mpi_amd/csrc`` (see
``__graft_entry__.build``).  Loading fails loudly when it is missing: there
is no fallback implementation.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhz.so")

HZ_OK = 0
HZ_ERR_ARG = 1
HZ_ERR_NOT_POW2 = 2
HZ_ERR_N_TOO_SMALL = 3
HZ_ERR_SINGULAR = 4
HZ_ERR_ALLOC = 5
HZ_ERR_HIP = 6
HZ_ERR_UNSUPPORTED = 7
HZ_ERR_NO_DEVICE = 8

FORM_PLAIN_LU = 0
FORM_DIAG = 1
FORM_MIXED = 2

# Every symbol include/hz.h declares, with its ctypes signature.
_D = ctypes.c_double
_PD = ctypes.POINTER(ctypes.c_double)
_PI = ctypes.POINTER(ctypes.c_int)
_VP = ctypes.c_void_p


class PlanDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("n", ctypes.c_int), ("ne", ctypes.c_int),
                ("xl", _D), ("xr", _D), ("yl", _D), ("yr", _D), ("zl", _D), ("zr", _D),
                ("lam", _D), ("alpha_l", _D), ("beta_l", _D), ("alpha_r", _D), ("beta_r", _D),
                ("form", ctypes.c_int), ("refine_sweeps", ctypes.c_int), ("mixed_fp64_cond", _D),
                ("periodic_z", ctypes.c_int)]


SIGNATURES = {
    "hz_plan_create": (ctypes.c_int, [ctypes.POINTER(PlanDesc), ctypes.POINTER(_VP)]),
    "hz_plan_destroy": (ctypes.c_int, [_VP]),
    "hz_plan_nz_total": (ctypes.c_int, [_VP, _PI]),
    "hz_plan_z_nodes": (ctypes.c_int, [_VP, _PD]),
    "hz_plan_operators": (ctypes.c_int, [_VP, _PD, _PD]),
    "hz_plan_num_modes": (ctypes.c_int, [_VP, _PI]),
    "hz_plan_set_lambda": (ctypes.c_int, [_VP, _D]),
    "hz_plan_set_robin": (ctypes.c_int, [_VP, _D, _D, _D, _D]),
    "hz_solve": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP, _VP]),
    "hz_solve_operators": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "hz_solve_status": (ctypes.c_int, [_VP, _VP]),
    "hz_adjoint": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP, _VP, _PD, _VP]),
    "hz_adjoint_noalloc": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "hz_derivative_z": (ctypes.c_int, [_VP, ctypes.c_int, _VP, _VP, _VP]),
    "hz_derivative_xy": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, _VP, _VP, ctypes.c_int, _VP]),
    "hz_fourier_xy": (ctypes.c_int, [_VP, ctypes.c_int, _VP, _VP, _VP, _VP, ctypes.c_int, _VP]),
    "hz_lu_factor_batched": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, _VP, _VP, _VP, _VP]),
    "hz_lu_solve_batched": (ctypes.c_int, [_VP, _VP, _VP, ctypes.c_int, ctypes.c_int, _VP]),
    "hz_batch_create": (ctypes.c_int, [ctypes.POINTER(_VP), ctypes.c_int, ctypes.POINTER(_VP)]),
    "hz_batch_solve": (ctypes.c_int, [_VP, ctypes.POINTER(_VP), ctypes.POINTER(_VP), ctypes.POINTER(_VP),
                                      ctypes.POINTER(_VP), _VP]),
    "hz_batch_destroy": (ctypes.c_int, [_VP]),
    "hz_solve_batched": (ctypes.c_int, [ctypes.POINTER(_VP), ctypes.c_int, ctypes.POINTER(_VP), ctypes.POINTER(_VP),
                                        ctypes.POINTER(_VP), ctypes.POINTER(_VP), _VP]),
    "hz_batch_status": (ctypes.c_int, [_VP]),
    "hz_batch_tuning": (ctypes.c_int, [_VP, _PI, _PI, _PI, _PI, _PD, ctypes.c_int]),
    "hz_batch_force_variant": (ctypes.c_int, [_VP, ctypes.c_int]),
    "hz_mixed_history": (ctypes.c_int, [_VP, _PD, ctypes.c_int, _PI]),
    "hz_mixed_fp64_modes": (ctypes.c_int, [_VP, _PI]),
    "hz_plan_tuning": (ctypes.c_int, [_VP, _PI, _PD, ctypes.c_int, _PI, _PD]),
    "hz_plan_diag_candidates": (ctypes.c_int, [_VP, _PI, ctypes.c_int, _PI]),
    "hz_plan_set_diag_tile": (ctypes.c_int, [_VP, ctypes.c_int]),
    "hz_plan_retune": (ctypes.c_int, [_VP]),
    "hz_host_plan_operators": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _D, _PD, _PD, _PD, _PD]),
    "hz_host_eigen": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _D, _PD, _PD, _PD]),
    "hz_host_eigen_split": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _D, ctypes.c_int, _PD, _PD, _PD, _PD]),
    "hz_plan_profile": (ctypes.c_int, [_VP, ctypes.c_int]),
    "hz_plan_profile_read": (ctypes.c_int, [_VP, _PD, ctypes.c_int, _PI, _PI, ctypes.c_char_p, ctypes.c_int]),
    "hz_last_error_string": (ctypes.c_char_p, []),
    "hz_version": (ctypes.c_char_p, []),
    "hz_probe_ceilings": (ctypes.c_int, [_PD, _PD, ctypes.c_size_t]),
    "hz_plan_periodic_closure": (ctypes.c_int, [_VP, _PD]),
    "hz_cubic_rhs": (ctypes.c_int, [_VP, _VP, ctypes.c_longlong, _D, _D, _VP]),
    "hz_solve_cubic": (ctypes.c_int, [_VP, _VP, _D, _D, _VP, _VP, _VP, _VP]),
    "hz_cubic_adjoint": (ctypes.c_int, [_VP, _VP, _VP, _VP, _D, _D, _D, _D, _VP, _VP, _VP]),
}

_lib = None


class HzError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"hz error {code}: {msg}")
        self.code = code


def load(path=LIB_PATH):
    """Load libhz.so and bind every declared symbol (raises if any is missing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise ImportError(f"{path} is not built; run __graft_entry__.build() "
                              "(make -C mpi_amd/csrc)")
        try:  # torch ships its own HIP runtime: load it first so that libhz binds to the same one
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(code):
    if code != HZ_OK:
        raise HzError(code, load().hz_last_error_string().decode())
    return code
