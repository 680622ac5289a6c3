"""torch.autograd bridge to the adjoint entry point (hz_adjoint).

``helmholtz_solve(f, g_l, g_r, lam, plan)`` returns u = S_lam(f, g_l, g_r)
with a backward that calls the hand-written adjoint kernels of libhz.so:
given dL/du it returns dL/df, dL/dg_l, dL/dg_r and dL/dlam.  It exists so
the adjoint can be checked against finite differences with torch's tools
(torch.autograd.gradcheck); the product path is the C ABI.

The derivative with respect to lam follows from differentiating the
collocation system (A - lam I) u = f in lam: du/dlam = S_lam(u, 0, 0), so
dL/dlam = <S^T dL/du, u>, the node sum of fbar * u (hz_adjoint computes it on
the device).
"""
import torch

from .solver import HelmholtzPlan


class HelmholtzSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f, gl, gr, lam, plan):
        lam_v = float(lam)
        if plan.desc.lam != lam_v:
            plan.set_lambda(lam_v)
        faces = [None if g is None else g.contiguous() for g in (gl, gr)]  # None for periodic z
        u = plan.solve(f.contiguous(), *faces)
        ctx.plan = plan
        ctx.lam = lam_v
        ctx.lam_like = lam if torch.is_tensor(lam) else None
        ctx.save_for_backward(u)
        return u

    @staticmethod
    def backward(ctx, ubar):
        (u,) = ctx.saved_tensors
        plan: HelmholtzPlan = ctx.plan
        if plan.desc.lam != ctx.lam:
            plan.set_lambda(ctx.lam)
        want_lam = ctx.lam_like is not None and ctx.needs_input_grad[3]
        fbar, glbar, grbar, lam_bar = plan.adjoint(ubar.contiguous(), u if want_lam else None)
        lam_grad = None
        if want_lam:
            lam_grad = torch.tensor(lam_bar, dtype=ctx.lam_like.dtype, device=ctx.lam_like.device)
        return fbar, glbar, grbar, lam_grad, None


def helmholtz_solve(f, gl, gr, lam, plan):
    """u = S_lam(f, g_l, g_r) on `plan`, differentiable in f, g_l, g_r and lam."""
    return HelmholtzSolve.apply(f, gl, gr, lam, plan)
