"""torch.autograd wrappers of the plan's linear operators.

Every operator here is complex-linear, so its vector-Jacobian product is the
adjoint applied to the incoming gradient (PyTorch's convention for complex
tensors, the same one ``torch.matmul`` follows):

  matvec    y = Z x           grad_x = Z^H grad_y      (mlfma_matvec_adjoint)
  solve     x = Z^-1 b        grad_b = Z^-H grad_x     (adjoint GMRES)
  farfield  E = F I           grad_I = F^H grad_E      (mlfma_farfield_adjoint)

Chained, they give the gradient of a far-field objective with respect to the
excitation, e.g. J = sum_d |E_far(rhat_d)|^2 with E_far = F Z^-1 b.  The
plan's geometry is fixed: gradients with respect to source positions are not
provided.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _capi as A
from .plan import Plan


def _np(t: torch.Tensor) -> np.ndarray:
    return t.detach().to("cpu", torch.complex128).contiguous().numpy()


class _Matvec(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, plan):
        ctx.plan = plan
        if x.is_cuda:
            return plan.matvec(x.to(torch.complex128).contiguous())
        return torch.from_numpy(plan.matvec(_np(x)))

    @staticmethod
    def backward(ctx, g):
        p = ctx.plan
        if g.is_cuda:
            return p.matvec_adjoint(g.to(torch.complex128).contiguous()), None
        return torch.from_numpy(p.matvec_adjoint(_np(g))), None


class _Solve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, b, plan, opts):
        ctx.plan, ctx.opts, ctx.device = plan, opts, b.device
        x, rep = plan.solve(_np(b), **opts)
        ctx.report = rep
        return torch.from_numpy(x).to(b.device)

    @staticmethod
    def backward(ctx, g):
        gb, _ = ctx.plan.solve(_np(g), adjoint=True, **ctx.opts)
        return torch.from_numpy(gb).to(ctx.device), None, None


class _Farfield(torch.autograd.Function):
    @staticmethod
    def forward(ctx, I, plan, dirs):
        ctx.plan, ctx.dirs, ctx.device = plan, dirs, I.device
        return torch.from_numpy(plan.farfield(_np(I), dirs)).to(I.device)

    @staticmethod
    def backward(ctx, g):
        gI = ctx.plan.farfield_adjoint(_np(g), ctx.dirs)
        return torch.from_numpy(gI).to(ctx.device), None, None


def matvec(plan: Plan, x: torch.Tensor) -> torch.Tensor:
    """y = Z x, differentiable in x (CUDA tensors stay on the device)."""
    return _Matvec.apply(x, plan)


def solve(plan: Plan, b: torch.Tensor, method: int = A.MLFMA_GMRES_RESTART, restart: int = 50,
          maxit: int = 500, tol: float = 1e-8) -> torch.Tensor:
    """x = Z^-1 b by GMRES, differentiable in b (the backward pass is an
    adjoint GMRES solve with the same options)."""
    return _Solve.apply(b, plan, dict(method=method, restart=restart, maxit=maxit, tol=tol))


def farfield(plan: Plan, I: torch.Tensor, dirs) -> torch.Tensor:
    """E = F I at the unit directions dirs (ndir x 3), differentiable in I."""
    dirs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
    return _Farfield.apply(I, plan, dirs)
