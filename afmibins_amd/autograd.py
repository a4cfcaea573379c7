"""torch.autograd wrapper of the IB adjoint (afm_ib_grad).

    J = IBObjective.apply(k, R, setup)

runs the IB simulation described by `setup` (grid, initial velocity and
marker positions, physics, step count) with per-marker stiffness k (m,) and
rest positions R (m, 2), and returns the scalar
J = 1/2 sum |u(T)|^2 h^2 + sum_m |c_m(T) - c_m(0)|^2.  The gradients with
respect to k and R come from the discrete adjoint on the GPU (checkpoint and
recompute); they are computed with the forward pass and scaled by the
incoming gradient in backward.
"""
from dataclasses import dataclass

import torch

from .api import Context, IBSolver, NSSolver


@dataclass
class IBSetup:
    ctx: Context
    u0: torch.Tensor
    v0: torch.Tensor
    X0: torch.Tensor
    ds: torch.Tensor
    mu: float
    rho: float
    dt: float
    tol: float
    steps: int
    msize: int
    checkpoint_every: int = 20
    precision: int = 0


def run(setup, k, R, need_grad):
    u, v = setup.u0.clone(), setup.v0.clone()
    ns = NSSolver(setup.ctx, setup.mu, setup.rho, setup.dt, setup.tol, u, v, tol_mode=1,
                  precision=setup.precision)
    ib = IBSolver(ns, setup.X0.contiguous(), R.contiguous(), k.contiguous(), setup.ds)
    try:
        return ib.grad(setup.steps, setup.msize, setup.checkpoint_every, need_grad)
    finally:
        ib.close()
        ns.close()


class IBObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, k, R, setup):
        need = k.requires_grad or R.requires_grad
        J, gk, gR = run(setup, k.detach(), R.detach(), need)
        ctx.save_for_backward(*( (gk, gR) if need else () ))
        ctx.need = need
        return torch.tensor(J, dtype=torch.float64, device=k.device)

    @staticmethod
    def backward(ctx, gout):
        if not ctx.need:
            return None, None, None
        gk, gR = ctx.saved_tensors
        return gout * gk, gout * gR, None
