"""torch.autograd wrappers of the plan's linear operators.

Every operator here is complex-linear, so its vector-Jacobian product is the
adjoint applied to the incoming gradient (PyTorch's convention for complex
tensors, the same one ``torch.matmul`` follows):

  matvec    y = Z x           grad_x = Z^H grad_y      (mlfma_matvec_adjoint)
  solve     x = Z^-1 b        grad_b = Z^-H grad_x     (adjoint GMRES)
  farfield  E = F I           grad_I = F^H grad_E      (mlfma_farfield_adjoint)

Chained, they give the gradient of a far-field objective with respect to the
excitation, e.g. J = sum_d |E_far(rhat_d)|^2 with E_far = F Z^-1 b.

For dipole plans ``objective_gradients`` also gives dJ/dr, the gradient with
respect to the source positions (adjoint method, tree held fixed):
with I = Z^-1 b, E = F I and lambda = Z^-H F^H E,

  dJ = 2 Re(E^H dF I) - 2 Re(lambda^H dZ I) + 2 Re(lambda^H db).

The far-field pattern of dipole i carries e^{-ik rhat.r_i}, so the first term
is 2 Re(-ik I_i conj(u_i^a)) with u^a = F^H (rhat^a E); the second comes from
mlfma_position_gradient; the third is present when b is the plan's own
plane-wave excitation (b_i proportional to e^{i kvec.r_i}).
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


def _require_converged(rep, what):
    """A GMRES solve that stopped above its tolerance returns an iterate, not
    the solution: the values and gradients built on it would be wrong, so
    the wrappers raise instead (MLFMA_ERR_NOT_CONVERGED in the report)."""
    if rep["status"] != "OK":
        raise RuntimeError(f"{what} GMRES solve did not converge: status {rep['status']}, "
                           f"{rep['iterations']} iterations, true residual {rep['resid_true']:.3e}")


def _solve_any(plan, b, adjoint, opts):
    """GMRES on a CUDA tensor (kept on the device) or a host tensor."""
    if b.is_cuda:
        return plan.solve(b.to(torch.complex128).contiguous(), adjoint=adjoint, **opts)
    x, rep = plan.solve(_np(b), adjoint=adjoint, **opts)
    return torch.from_numpy(x), rep


class _Solve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, b, plan, opts):
        ctx.plan, ctx.opts = plan, opts
        x, rep = _solve_any(plan, b, False, opts)
        _require_converged(rep, "forward")
        ctx.report = rep
        return x

    @staticmethod
    def backward(ctx, g):
        gb, rep = _solve_any(ctx.plan, g, True, ctx.opts)
        _require_converged(rep, "adjoint")
        return gb, None, None


class _Farfield(torch.autograd.Function):
    @staticmethod
    def forward(ctx, I, plan, dirs):
        ctx.plan, ctx.dirs = plan, dirs
        if I.is_cuda:
            return plan.farfield(I.to(torch.complex128).contiguous(), dirs)
        return torch.from_numpy(plan.farfield(_np(I), dirs))

    @staticmethod
    def backward(ctx, g):
        if g.is_cuda:
            return ctx.plan.farfield_adjoint(g.to(torch.complex128).contiguous(), ctx.dirs), None, None
        return torch.from_numpy(ctx.plan.farfield_adjoint(_np(g), ctx.dirs)), None, None


def matvec(plan: Plan, x: torch.Tensor) -> torch.Tensor:
    """y = Z x, differentiable in x (CUDA tensors stay on the device)."""
    return _Matvec.apply(x, plan)


def solve(plan: Plan, b: torch.Tensor, method: int = A.MLFMA_GMRES_RESTART, restart: int = 50,
          maxit: int = 500, tol: float = 1e-8, deflate: int = 0) -> torch.Tensor:
    """x = Z^-1 b by GMRES, differentiable in b (the backward pass is an
    adjoint GMRES solve with the same options; deflate = k: deflated
    restarts keeping k harmonic Ritz vectors)."""
    return _Solve.apply(b, plan, dict(method=method, restart=restart, maxit=maxit, tol=tol, deflate=deflate))


def farfield(plan: Plan, I: torch.Tensor, dirs) -> torch.Tensor:
    """E = F I at the unit directions dirs (ndir x 3), differentiable in I."""
    dirs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
    return _Farfield.apply(I, plan, dirs)


def position_gradient(plan: Plan, I, lam, E, dirs, b=None, wavevector=None) -> np.ndarray:
    """dJ/dr (n x 3) of J = sum_d |E(rhat_d)|^2 for a dipole plan from the
    forward solution I = Z^-1 b, the far fields E = F I at dirs and the
    adjoint solution lam = Z^-H F^H E.  With wavevector (3,) given, b is the
    plane-wave excitation b_i = c_i e^{i wavevector.r_i} and its derivative is
    included."""
    dirs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
    E = np.asarray(E, dtype=np.complex128).reshape(len(dirs), 3)
    I = np.asarray(I, dtype=np.complex128)
    lam = np.asarray(lam, dtype=np.complex128)
    k = plan.pw.wavenum
    g = plan.position_gradient(I, lam)
    for a in range(3):
        u = plan.farfield_adjoint(dirs[:, a:a + 1] * E, dirs)
        g[:, a] += 2.0 * np.real(-1j * k * I * np.conj(u))
    if wavevector is not None:
        b = np.asarray(b, dtype=np.complex128)
        w = np.asarray(wavevector, dtype=np.float64)
        g += 2.0 * np.real((np.conj(lam) * 1j * b)[:, None] * w[None, :])
    return g


def objective_gradients(plan: Plan, dirs, b=None, restart: int = 50, maxit: int = 500, tol: float = 1e-10):
    """J = sum_d |E_far(rhat_d)|^2 with E_far = F Z^-1 b (b = None: the plan's
    plane-wave excitation) and its gradients: returns (J, grad_b, grad_r)
    with dJ = 2 Re(grad_b^H db) + grad_r . dr (grad_r: n x 3, dipole plans).
    Two GMRES solves (forward and adjoint) with the given options."""
    dirs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
    opts = dict(method=A.MLFMA_GMRES_RESTART, restart=restart, maxit=maxit, tol=tol)
    wave = None
    if b is None:
        b = -plan.voltage()
        wave = plan.pw.wavenum * np.asarray(plan.pw.khat[:3], dtype=np.float64)
    I, rep = plan.solve(b, **opts)
    _require_converged(rep, "forward")
    E = plan.farfield(I, dirs)
    lam, rep = plan.solve(plan.farfield_adjoint(E, dirs), adjoint=True, **opts)
    _require_converged(rep, "adjoint")
    J = float(np.sum(np.abs(E) ** 2))
    return J, lam, position_gradient(plan, I, lam, E, dirs, b, wave)


class DipoleScatterer:
    """The far field of a set of point dipoles as a differentiable function of
    the excitation b and the positions r: E = F(r) (Z(r) + s I)^-1 b at fixed
    unit directions (s: an optional shift, e.g. a dipole self term, which the
    reference drops, src/sources/dipole.h:65).  A plan is built for each new
    set of positions (the tree follows the positions, so derivatives hold it
    fixed only while no dipole crosses a box face); the moments, the
    configuration and the solve options are fixed.

        sc = DipoleScatterer(mom, dirs, config_kwargs, shift=s, tol=1e-12)
        E = sc(b, pos)              # b: (n,) complex128, pos: (n, 3) float64
        J = (E.abs() ** 2).sum(); J.backward()   # b.grad, pos.grad
    """

    def __init__(self, mom, dirs, config: dict, pw=None, shift=0j, method=A.MLFMA_GMRES_RESTART, restart=100,
                 maxit=5000, tol=1e-12, device=0):
        self.mom = np.ascontiguousarray(mom, dtype=np.float64)
        self.dirs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
        self.config = dict(config)
        self.config.setdefault("flags", A.MLFMA_FLAG_FRESH_LOCALS)
        self.config["flags"] |= A.MLFMA_FLAG_FRESH_LOCALS  # the adjoint needs the linear operator
        self.pw = pw if pw is not None else A.make_planewave()
        self.opts = dict(method=method, restart=restart, maxit=maxit, tol=tol, shift=shift)
        self.device = device
        self._key = None
        self._plan = None

    def plan(self, pos: np.ndarray) -> Plan:
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        key = pos.tobytes()
        if key != self._key:
            geom = A.GeometryArrays(A.MLFMA_SRC_DIPOLE, pos=pos, mom=self.mom)
            self._plan = Plan(geom, self.pw, A.make_config(device=self.device, **self.config))
            self._key = key
        return self._plan

    def __call__(self, b: torch.Tensor, pos: torch.Tensor) -> torch.Tensor:
        return _DipoleFarField.apply(b, pos, self)


class _DipoleFarField(torch.autograd.Function):
    """E = F(r) (Z(r) + s I)^-1 b.  With the incoming gradient G of E
    (dL = Re(G^H dE)) and lambda = (Z + s I)^-H F^H G:
      grad_b = lambda
      grad_r[i, a] = Re(G^H dF/dr_i^a I) - Re(lambda^H dZ/dr_i^a I)
    The first term is Re(-ik I_i conj(u^a_i)) with u^a = F^H (rhat^a G) (the
    pattern of dipole i carries e^{-ik rhat.r_i}); the second is half of
    mlfma_position_gradient(I, lambda).  One forward and one adjoint GMRES
    solve; I is recomputed in the backward pass rather than stored when the
    plan changed in between."""

    @staticmethod
    def forward(ctx, b, pos, sc):
        p = sc.plan(pos.detach().cpu().numpy())
        dev = torch.device("cuda", sc.device)
        bd = b.detach().to(dev, torch.complex128).contiguous()
        I, rep = p.solve(bd, **sc.opts)
        _require_converged(rep, "forward")
        E = p.farfield(I, sc.dirs)
        ctx.sc, ctx.b_device, ctx.pos_device = sc, b.device, pos.device
        ctx.save_for_backward(pos.detach(), b.detach())
        ctx.I = I
        return E.to(b.device)

    @staticmethod
    def backward(ctx, G):
        sc = ctx.sc
        pos, b = ctx.saved_tensors
        p = sc.plan(pos.cpu().numpy())
        dev = torch.device("cuda", sc.device)
        I = ctx.I
        Gd = G.detach().to(dev, torch.complex128).contiguous()
        lam, rep = p.solve(p.farfield_adjoint(Gd, sc.dirs), adjoint=True, **sc.opts)
        _require_converged(rep, "adjoint")
        k = p.pw.wavenum
        In = I.cpu().numpy()
        grad_r = 0.5 * p.position_gradient(In, lam.cpu().numpy())
        for a in range(3):
            u = p.farfield_adjoint(Gd * torch.from_numpy(sc.dirs[:, a:a + 1]).to(dev), sc.dirs).cpu().numpy()
            grad_r[:, a] += np.real(-1j * k * In * np.conj(u))
        return (lam.to(ctx.b_device), torch.from_numpy(grad_r).to(ctx.pos_device, pos.dtype), None)
