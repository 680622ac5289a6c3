"""Semi-implicit Allen-Cahn stepping on the Helmholtz solver, differentiated
with the adjoint kernels and step checkpointing.

    u_t = eps^2 lap(u) + u - u^3      on [0, 2 pi)^2 x [-1, 1]

Each step is one Helmholtz solve (z faces homogeneous Neumann:
alpha = 1, beta = 0, g = 0):

    lap(u1) - lam u1 = -lam u0 + (u0^3 - u0) / eps^2,   lam = 1 / (dt eps^2)

The loss is the quadrature  L = sum_k W_k (u_T - u_target)_k^2  with
uniform weights in x, y and Gauss-Lobatto-Legendre weights in z.

``loss_and_grad`` returns L, dL/du_0 and dL/deps.  It runs the forward pass
keeping a checkpoint every ``checkpoint_every`` steps, then walks the
segments backwards: each segment is recomputed from its checkpoint and the
adjoint of each step is applied with hz_adjoint (through the plan).  With
ubar = dL/du1 for the step u0 -> u1 and (fbar, lam_bar) = S^T ubar:

    dL/du0  = fbar * (-lam + (3 u0^2 - 1) / eps^2)
    dL/deps += lam_bar dlam/deps + <fbar, -u0 dlam/deps - 2 (u0^3 - u0) / eps^3>
    dlam/deps = -2 / (dt eps^3)

with lam_bar = <fbar, u1> (du1/dlam solves the same system with right-hand
side u1).  Everything stays on the device: dL/deps is accumulated there and
read back once at the end.  ``capture_run`` records the forward steps as one
HIP graph.  The reference has no time stepper; this module is the
differentiation workload built on the solver.
"""
import math

import numpy as np
import torch

from .solver import HelmholtzPlan


def gll_weights(x):
    """Quadrature weights of GLL nodes x[0..N-1] on [-1, 1]:
    w_i = 2 / (N (N-1) P_{N-1}(x_i)^2)."""
    N = len(x)
    p0 = np.ones_like(x)
    p1 = np.array(x, dtype=np.float64)
    for k in range(2, N):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
    return 2.0 / (N * (N - 1) * p1 * p1)


class AllenCahn:
    def __init__(self, nx=128, ny=128, n=65, eps=0.02, dt=1e-3, steps=64, checkpoint_every=8,
                 form="diag", Lx=2 * math.pi, Ly=2 * math.pi, zl=-1.0, zr=1.0):
        self.nx, self.ny, self.n = nx, ny, n
        self.dt, self.steps, self.every = dt, steps, checkpoint_every
        self.eps0 = eps
        self.plan = HelmholtzPlan(nx, ny, n, 1, xl=0.0, xr=Lx, yl=0.0, yr=Ly, zl=zl, zr=zr,
                                  lam=self.lam_of(eps), alpha_l=1.0, beta_l=0.0, alpha_r=1.0, beta_r=0.0,
                                  form=form)
        self.K = self.plan.nz
        z = self.plan.z_nodes()
        xi = 2.0 * (z - z[0]) / (z[-1] - z[0]) - 1.0
        wz = gll_weights(xi) * (zr - zl) / 2.0
        self.W = torch.from_numpy(np.broadcast_to(wz * (Lx / nx) * (Ly / ny), (nx, ny, self.K)).copy()).cuda()
        self.g0 = torch.zeros((nx, ny), dtype=torch.float64, device="cuda")

    def lam_of(self, eps):
        return 1.0 / (self.dt * eps * eps)

    def _set_eps(self, eps):
        lam = self.lam_of(eps)
        if self.plan.desc.lam != lam:
            self.plan.set_lambda(lam)
        return lam

    def step(self, u, eps, lam, out=None):
        # f = (u^3 - u) / eps^2 - lam u = c u^3 - (c + lam) u, c = 1 / eps^2,
        # formed inside the solve's forward transform (hz_solve_cubic)
        c = 1.0 / (eps * eps)
        return self.plan.solve_cubic(u.contiguous(), c, -(c + lam), self.g0, self.g0, out=out)

    def run(self, u0, eps):
        """u_T after `steps` steps (no checkpoints kept)."""
        lam = self._set_eps(eps)
        u = u0.contiguous()
        for _ in range(self.steps):
            u = self.step(u, eps, lam)
        return u

    def capture_run(self, u0, eps):
        """Records the `steps` steps from a static copy of u0 as one HIP graph
        (torch.cuda.CUDAGraph; the solver's kernels go on the capturing
        stream) and returns (replay, u_in, u_out): replay() runs all steps,
        reading u_in and leaving u_T in u_out.  Replays recompute from u_in,
        so refill u_in (copy_) to start from a new state.  The graph holds
        this eps (the right-hand side's coefficients are recorded in it):
        replay() raises once the plan's lambda has changed, e.g. by a run at
        another eps; capture again then."""
        lam = self._set_eps(eps)
        u_in = u0.contiguous().clone()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):  # warm-up on the side stream, as torch's capture wants
            u = u_in
            for _ in range(2):
                u = self.step(u, eps, lam)
        torch.cuda.current_stream().wait_stream(side)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            u = u_in
            for _ in range(self.steps):
                u = self.step(u, eps, lam)

        def replay():
            if self.plan.desc.lam != lam:
                raise RuntimeError(f"graph recorded at eps = {eps!r}; the plan's lambda has changed since "
                                   "(capture again)")
            graph.replay()
        return replay, u_in, u

    def loss(self, uT, target):
        d = uT - target
        return float(torch.sum(self.W * d * d))

    def loss_and_grad(self, u0, eps, target):
        lam = self._set_eps(eps)
        dlam = -2.0 / (self.dt * eps ** 3)
        # forward with checkpoints
        ckpt = {}
        u = u0.contiguous()
        for t in range(self.steps):
            if t % self.every == 0:
                ckpt[t] = u
            u = self.step(u, eps, lam)
        uT = u
        L = self.loss(uT, target)
        ubar = 2.0 * self.W * (uT - target)
        # dL/deps accumulates on the device (one host read at the end):
        # per step dL/dlam = sum(fbar * u_out) (du/dlam = S(u, 0, 0)) times
        # dlam/deps, plus the explicit eps dependence of the right-hand side
        eps_bar = torch.zeros((), dtype=torch.float64, device="cuda")
        inv_e2 = 1.0 / (eps * eps)
        for s in sorted(ckpt, reverse=True):
            e = min(s + self.every, self.steps)
            states = [ckpt[s]]
            for _ in range(s, e):
                states.append(self.step(states[-1], eps, lam))
            for t in range(e - 1, s - 1, -1):
                u_in, u_out = states[t - s], states[t - s + 1]
                fbar, _, _, _ = self.plan.adjoint(ubar)
                # eps_bar += <fbar, (u_out - u_in) dlam/deps - 2 (u_in^3 - u_in) / eps^3>,
                # ubar = fbar (3 u_in^2 / eps^2 - lam - 1 / eps^2), one fused pass
                ubar = self.plan.cubic_adjoint(fbar, u_in, u_out, inv_e2, -(lam + inv_e2), dlam, -2.0 / eps ** 3,
                                               eps_bar)
        return L, ubar, float(eps_bar)

    def close(self):
        self.plan.close()
