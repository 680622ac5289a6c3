"""Host-side plan object over the C ABI (include/hz.h).

``HelmholtzPlan`` owns one ``hz_plan``.  Fields are torch float64 tensors on
the current HIP device, shaped ``[nx, ny, nz_total]`` (z fastest) for volume
data and ``[nx, ny]`` for face data.  Nothing here computes: every call goes
through libhz.so, and a missing library or device raises.

Reference counterpart: CreatePDEPlan / SolvePDE / DestroyPDEPlan,
HelmholtzSolvers/DirectPseudoSpectral/src/NonPeriodic3D_MPI.cpp:40-591.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib
from ._lib import check, load

FORMS = {"plain": _lib.FORM_PLAIN_LU, "diag": _lib.FORM_DIAG, "mixed": _lib.FORM_MIXED}


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream_ptr(stream):
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


class HelmholtzPlan:
    def __init__(self, nx, ny, n, ne=1, xl=0.0, xr=2 * math.pi, yl=0.0, yr=2 * math.pi,
                 zl=-1.0, zr=1.0, lam=1.0, alpha_l=0.0, beta_l=1.0, alpha_r=0.0, beta_r=1.0,
                 form="plain", refine_sweeps=0, mixed_fp64_cond=0.0, periodic_z=False):
        if not torch.cuda.is_available():
            raise RuntimeError("HelmholtzPlan needs a HIP device")
        self.lib = load()
        self.desc = _lib.PlanDesc(nx, ny, n, ne, xl, xr, yl, yr, zl, zr, lam, alpha_l, beta_l,
                                  alpha_r, beta_r, FORMS[form] if isinstance(form, str) else form,
                                  refine_sweeps, mixed_fp64_cond, 1 if periodic_z else 0)
        self.handle = ctypes.c_void_p()
        check(self.lib.hz_plan_create(ctypes.byref(self.desc), ctypes.byref(self.handle)))
        k = ctypes.c_int()
        check(self.lib.hz_plan_nz_total(self.handle, ctypes.byref(k)))
        self.nx, self.ny, self.n, self.ne, self.nz = nx, ny, n, ne, k.value
        m = ctypes.c_int()
        check(self.lib.hz_plan_num_modes(self.handle, ctypes.byref(m)))
        self.num_modes = m.value
        self.form = form

    # ---- geometry ------------------------------------------------------
    def z_nodes(self):
        """The reference's z array (starts at 0, NonPeriodic3D_MPI.cpp:111-120)."""
        z = np.zeros(self.nz)
        check(self.lib.hz_plan_z_nodes(self.handle, z.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return z

    def operators(self):
        Dz = np.zeros((self.n, self.n))
        Dzz = np.zeros((self.n, self.n))
        P = ctypes.POINTER(ctypes.c_double)
        check(self.lib.hz_plan_operators(self.handle, Dz.ctypes.data_as(P), Dzz.ctypes.data_as(P)))
        return Dz, Dzz

    def grid(self):
        """Physical node coordinates x[i], y[j], zeta[k] (zeta = z + zl)."""
        d = self.desc
        x = d.xl + np.arange(self.nx) * ((d.xr - d.xl) / self.nx)
        y = d.yl + np.arange(self.ny) * ((d.yr - d.yl) / self.ny)
        return x, y, self.z_nodes() + d.zl

    def set_lambda(self, lam):
        check(self.lib.hz_plan_set_lambda(self.handle, float(lam)))
        self.desc.lam = float(lam)

    def set_robin(self, alpha_l, beta_l, alpha_r, beta_r):
        check(self.lib.hz_plan_set_robin(self.handle, float(alpha_l), float(beta_l), float(alpha_r), float(beta_r)))
        self.desc.alpha_l, self.desc.beta_l = float(alpha_l), float(beta_l)
        self.desc.alpha_r, self.desc.beta_r = float(alpha_r), float(beta_r)

    # ---- solves --------------------------------------------------------
    def empty_field(self):
        return torch.empty((self.nx, self.ny, self.nz), dtype=torch.float64, device="cuda")

    def _check_field(self, t, shape):
        if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous() or tuple(t.shape) != shape:
            raise ValueError(f"expected contiguous float64 HIP tensor of shape {shape}")

    def solve(self, f, gl=None, gr=None, out=None, stream=None):
        """u for right-hand side f and face data gl, gr (None for a periodic-z plan)."""
        self._check_field(f, (self.nx, self.ny, self.nz))
        if not self.desc.periodic_z:
            self._check_field(gl, (self.nx, self.ny))
            self._check_field(gr, (self.nx, self.ny))
        u = out if out is not None else self.empty_field()
        check(self.lib.hz_solve(self.handle, _ptr(f), _ptr(gl), _ptr(gr), _ptr(u), _stream_ptr(stream)))
        return u

    def solve_cubic(self, u, a3, a1, gl=None, gr=None, out=None, stream=None):
        """solve() with the right-hand side a3 u^3 + a1 u formed inside the
        forward transform (hz_solve_cubic; equal to solve(cubic_rhs(u)) bit for
        bit)."""
        self._check_field(u, (self.nx, self.ny, self.nz))
        if not self.desc.periodic_z:
            self._check_field(gl, (self.nx, self.ny))
            self._check_field(gr, (self.nx, self.ny))
        v = out if out is not None else self.empty_field()
        check(self.lib.hz_solve_cubic(self.handle, _ptr(u), float(a3), float(a1), _ptr(gl), _ptr(gr), _ptr(v),
                                      _stream_ptr(stream)))
        return v

    def cubic_adjoint(self, fbar, u_in, u_out, a3, a1, b1, b3, acc, ubar=None, stream=None):
        """One backward step through solve_cubic's right-hand side
        (hz_cubic_adjoint): returns ubar = fbar (3 a3 u_in^2 + a1) (in fbar's
        storage unless ubar is given) and adds sum fbar (b1 (u_out - u_in) +
        b3 (u_in^3 - u_in)) to the device scalar acc."""
        for t in (fbar, u_in, u_out):
            self._check_field(t, (self.nx, self.ny, self.nz))
        if ubar is None:
            ubar = fbar
        check(self.lib.hz_cubic_adjoint(self.handle, _ptr(fbar), _ptr(u_in), _ptr(u_out), float(a3), float(a1),
                                        float(b1), float(b3), _ptr(ubar), _ptr(acc), _stream_ptr(stream)))
        return ubar

    def solve_operators(self, f, gl, gr, op_diag=None, op_full=None, out=None, stream=None):
        u = out if out is not None else self.empty_field()
        check(self.lib.hz_solve_operators(self.handle, _ptr(f), _ptr(gl), _ptr(gr), _ptr(op_diag),
                                          _ptr(op_full), _ptr(u), _stream_ptr(stream)))
        return u

    def status(self, stream=None):
        check(self.lib.hz_solve_status(self.handle, _stream_ptr(stream)))

    def adjoint(self, ubar, u=None, stream=None):
        """(dL/df, dL/dg_l, dL/dg_r, dL/dlambda or None) for dL/du = ubar; the
        face gradients are None for a periodic-z plan."""
        fbar = self.empty_field()
        if self.desc.periodic_z:
            glbar = grbar = None
        else:
            glbar = torch.empty((self.nx, self.ny), dtype=torch.float64, device="cuda")
            grbar = torch.empty_like(glbar)
        lam_bar = ctypes.c_double(0.0)
        check(self.lib.hz_adjoint(self.handle, _ptr(ubar), _ptr(u), _ptr(fbar), _ptr(glbar), _ptr(grbar),
                                  ctypes.byref(lam_bar) if u is not None else None, _stream_ptr(stream)))
        return fbar, glbar, grbar, (lam_bar.value if u is not None else None)

    def derivative_z(self, u, order=1, out=None, stream=None):
        o = out if out is not None else self.empty_field()
        check(self.lib.hz_derivative_z(self.handle, int(order), _ptr(u), _ptr(o), _stream_ptr(stream)))
        return o

    def derivative_xy(self, u, axis, order=1, out=None, stream=None):
        """Fourier derivative of u[nx, ny, lanes] along x (axis 0) or y (axis 1)."""
        if u.dim() != 3 or u.shape[0] != self.nx or u.shape[1] != self.ny or not u.is_contiguous():
            raise ValueError("expected a contiguous [nx, ny, lanes] tensor")
        o = out if out is not None else torch.empty_like(u)
        check(self.lib.hz_derivative_xy(self.handle, int(axis), int(order), _ptr(u), _ptr(o), u.shape[2],
                                        _stream_ptr(stream)))
        return o

    def fourier_xy(self, x, lanes=None, stream=None):
        """Forward half-spectrum transform of real x[nx, ny, lanes] -> (re, im)[nx, ny//2+1, lanes]."""
        lanes = x.shape[2] if x.dim() == 3 else 1
        M = self.ny // 2 + 1
        re = torch.empty((self.nx, M, lanes), dtype=torch.float64, device="cuda")
        im = torch.empty_like(re)
        check(self.lib.hz_fourier_xy(self.handle, 1, _ptr(x), _ptr(re), _ptr(im), None, lanes,
                                     _stream_ptr(stream)))
        return re, im

    def inverse_fourier_xy(self, re, im, stream=None):
        lanes = re.shape[2]
        out = torch.empty((self.nx, self.ny, lanes), dtype=torch.float64, device="cuda")
        re = re.clone()
        im = im.clone()
        check(self.lib.hz_fourier_xy(self.handle, -1, None, _ptr(re), _ptr(im), _ptr(out), lanes,
                                     _stream_ptr(stream)))
        return out

    def profile(self, enable=True):
        check(self.lib.hz_plan_profile(self.handle, 1 if enable else 0))

    def profile_read(self):
        """Mean milliseconds per kernel stage over the solves recorded since
        profiling was enabled (HIP events on the solve stream)."""
        ms = np.zeros(32)
        ns = ctypes.c_int(0)
        nsol = ctypes.c_int(0)
        buf = ctypes.create_string_buffer(1024)
        check(self.lib.hz_plan_profile_read(self.handle, ms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 32,
                                            ctypes.byref(ns), ctypes.byref(nsol), buf, 1024))
        names = buf.value.decode().split(";") if ns.value else []
        return dict(zip(names, ms[: ns.value].tolist())), nsol.value

    def mixed_history(self, max_sweeps=16):
        h = np.zeros(max_sweeps)
        n = ctypes.c_int(0)
        check(self.lib.hz_mixed_history(self.handle, h.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                        max_sweeps, ctypes.byref(n)))
        return h[: n.value]

    def tuning(self):
        """Plan-time autotuning: diag tile rows (and the ms of each candidate),
        z lanes per work-group of the four transform passes (and their ms)."""
        P = ctypes.POINTER(ctypes.c_double)
        rt = ctypes.c_int(0)
        ms = np.zeros(8)
        lanes = (ctypes.c_int * 4)()
        fms = np.zeros(4)
        check(self.lib.hz_plan_tuning(self.handle, ctypes.byref(rt), ms.ctypes.data_as(P), 8, lanes,
                                      fms.ctypes.data_as(P)))
        cand = self.diag_candidates()
        return {"diag_tile_rows": rt.value, "choice_ms": [v for v in ms.tolist() if v > 0],
                "choices": {c: ms[i] for i, c in enumerate(cand) if ms[i] > 0},
                "fft_lanes": list(lanes), "fft_ms": fms.tolist()}

    def diag_candidates(self):
        """Mode-kernel shapes offered for this plan: DIAG tile rows, or the
        column capacities of the PLAIN_LU kernel's instances."""
        rows = (ctypes.c_int * 8)()
        k = ctypes.c_int(0)
        check(self.lib.hz_plan_diag_candidates(self.handle, rows, 8, ctypes.byref(k)))
        return list(rows)[:k.value]

    def set_diag_tile(self, rows):
        """Pins the mode kernel's shape (DIAG or PLAIN_LU) to one of diag_candidates()."""
        check(self.lib.hz_plan_set_diag_tile(self.handle, int(rows)))

    def retune(self):
        """Runs the plan's autotuning again (synchronises the device); the
        new choices and times show in tuning()."""
        check(self.lib.hz_plan_retune(self.handle))

    def periodic_closure(self):
        """Closure coefficients of the last periodic-z solve, [2][nx][ny//2+1]
        (real, imaginary plane); zeros with one element."""
        a = np.zeros((2, self.nx, self.ny // 2 + 1))
        check(self.lib.hz_plan_periodic_closure(self.handle, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return a

    def mixed_fp64_modes(self):
        k = ctypes.c_int(0)
        check(self.lib.hz_mixed_fp64_modes(self.handle, ctypes.byref(k)))
        return k.value

    def close(self):
        if self.handle:
            self.lib.hz_plan_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchSolver:
    """Many DIAG- or MIXED-form plans solved together (hz_batch_*): one launch
    per stage and task kind, or one launch per stage over every kind (both
    replayed from a HIP graph), chosen by the autotuner on the first solve.
    Variants: 4, 5 staged and 8, 9 merged (mode-solve tiles of 32, 64 rows);
    6, 7 fused (switched off) and 0-3 refused."""

    NUM_VARIANTS = 10

    def __init__(self, plans):
        self.lib = load()
        self.plans = list(plans)
        arr = (ctypes.c_void_p * len(self.plans))(*[p.handle.value for p in self.plans])
        self.handle = ctypes.c_void_p()
        check(self.lib.hz_batch_create(arr, len(self.plans), ctypes.byref(self.handle)))

    def solve(self, fs, gls, grs, us=None, stream=None):
        n = len(self.plans)
        if us is None:
            us = [p.empty_field() for p in self.plans]
        for p, f, gl, gr, u in zip(self.plans, fs, gls, grs, us):
            p._check_field(f, (p.nx, p.ny, p.nz))
            if not p.desc.periodic_z:  # periodic-z problems take None for both faces
                p._check_field(gl, (p.nx, p.ny))
                p._check_field(gr, (p.nx, p.ny))
            p._check_field(u, (p.nx, p.ny, p.nz))
        P = ctypes.c_void_p * n
        check(self.lib.hz_batch_solve(self.handle, P(*[f.data_ptr() for f in fs]), P(*[_ptr(g) for g in gls]),
                                      P(*[_ptr(g) for g in grs]), P(*[u.data_ptr() for u in us]),
                                      _stream_ptr(stream)))
        return us

    def status(self):
        check(self.lib.hz_batch_status(self.handle))

    def force_variant(self, v):
        check(self.lib.hz_batch_force_variant(self.handle, int(v)))

    def tuning(self):
        v, fused, rt, wgs = (ctypes.c_int(-1) for _ in range(4))
        ms = np.zeros(self.NUM_VARIANTS)
        check(self.lib.hz_batch_tuning(self.handle, ctypes.byref(v), ctypes.byref(fused), ctypes.byref(rt),
                                       ctypes.byref(wgs), ms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                       self.NUM_VARIANTS))
        return {"variant": v.value, "launches": ("per stage and kind", "fused", "per stage")[fused.value],
                "tile_rows": rt.value,
                "workgroups_per_cu": wgs.value, "variant_ms": ms.tolist()}

    def close(self):
        if self.handle:
            self.lib.hz_batch_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def lu_factor_batched(A, stream=None):
    """Reference LU (LU_Decomposition.cpp:7-72) of A[count, n, n] in place on the GPU.
    Returns (pivots[count, n-1] int32, det[count], status[count])."""
    lib = load()
    count, n, _ = A.shape
    piv = torch.zeros((count, max(n - 1, 1)), dtype=torch.int32, device=A.device)
    det = torch.zeros(count, dtype=torch.float64, device=A.device)
    st = torch.zeros(count, dtype=torch.int32, device=A.device)
    check(lib.hz_lu_factor_batched(_ptr(A), n, count, _ptr(piv), _ptr(det), _ptr(st), _stream_ptr(stream)))
    return piv, det, st


def lu_solve_batched(LU, piv, b, stream=None):
    lib = load()
    count, n, _ = LU.shape
    check(lib.hz_lu_solve_batched(_ptr(LU), _ptr(piv), _ptr(b), n, count, _stream_ptr(stream)))
    return b
