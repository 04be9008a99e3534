"""Adjoint solve, far-field adjoint and the torch.autograd wrappers (gradient
of a far-field objective with respect to the excitation) on the GPU.

Pins: <F I, E> = <I, F^H E> to round-off; the adjoint GMRES solution against a
dense solve of Z^H; the autograd gradient of J = sum_d |E_far(rhat_d)|^2,
E_far = F Z^-1 b, against central finite differences of J (step 1e-6 of |b|)
in random directions."""
import numpy as np
import pytest

from mlfma_amd import _capi as A
from mlfma_amd import geometry as G
from mlfma_amd.plan import Plan
from tests import cases
from tests.test_gpu_solver import _small_problem

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def small_plan():
    # 48 dipoles with far-field interactions; restart >= n makes GMRES exact
    # within n iterations
    geom = _small_problem()
    kw = dict(root_leng=2.0, max_node_srcs=2, max_level=-1, interp_order=3, digits=5,
              flags=A.MLFMA_FLAG_FRESH_LOCALS)
    return geom, Plan(geom, cases.planewave_z(), A.make_config(device=0, **kw))


def dirs_ring(n=13):
    t = np.linspace(0.05, np.pi - 0.05, n)
    return np.stack([np.sin(t), 0 * t, np.cos(t)], axis=1)


def test_farfield_adjoint_identity():
    for name in ("dipole_fixed", "rwg_fixed"):
        geom, kw = cases.make(name)
        p = Plan(geom, cases.planewave_z(), A.make_config(device=0, **kw))
        dirs = dirs_ring()
        I = G.complex_gaussian(geom.n, 1)
        E = (G.complex_gaussian(3 * len(dirs), 2)).reshape(-1, 3)
        lhs = np.vdot(E, p.farfield(I, dirs))
        rhs = np.vdot(p.farfield_adjoint(E, dirs), I)
        say(name, "farfield adjoint", abs(lhs - rhs) / abs(lhs))
        assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_adjoint_solve_against_dense():
    geom, p = small_plan()
    n = geom.n
    Z = np.stack([p.matvec(np.eye(n, dtype=complex)[:, j]) for j in range(n)], axis=1)
    b = G.complex_gaussian(n, 5)
    x, rep = p.solve(b, method=A.MLFMA_GMRES_RESTART, restart=n + 2, maxit=3 * n, tol=1e-12, adjoint=True)
    say("adjoint solve", rep)
    xd = np.linalg.solve(Z.conj().T, b)
    assert rep["status"] == "OK"
    assert cases.relerr(x, xd) <= 1e-9
    x2, _ = p.solve(b, method=A.MLFMA_GMRES_RESTART, restart=n + 2, maxit=3 * n, tol=1e-12)
    assert cases.relerr(x2, np.linalg.solve(Z, b)) <= 1e-9


def test_autograd_farfield_objective():
    import torch
    from mlfma_amd import autograd as AG
    geom, p = small_plan()
    n = geom.n
    dirs = dirs_ring(9)
    b0 = torch.from_numpy(G.complex_gaussian(n, 7))

    def J(b):
        x = AG.solve(p, b, tol=1e-13, restart=n + 2, maxit=3 * n)
        E = AG.farfield(p, x, dirs)
        return (E.abs() ** 2).sum()

    b = b0.clone().requires_grad_(True)
    j0 = J(b)
    j0.backward()
    g = b.grad.detach().numpy()
    rng = np.random.default_rng(0)
    h = 1e-6 * np.linalg.norm(b0.numpy())
    for trial in range(3):
        d = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        d /= np.linalg.norm(d)
        dt = torch.from_numpy(d)
        with torch.no_grad():
            fd = (J(b0 + h * dt) - J(b0 - h * dt)).item() / (2 * h)
        ad = float(np.real(np.vdot(g, d)))   # dJ = Re <grad, db> for real J
        say("directional derivative", ad, "finite difference", fd)
        assert abs(ad - fd) <= 1e-5 * max(abs(fd), 1e-30) + 1e-9 * np.linalg.norm(g)


def test_autograd_matvec_device():
    import torch
    from mlfma_amd import autograd as AG
    geom, kw = cases.make("dipole_fixed")
    p = Plan(geom, cases.planewave_z(), A.make_config(device=0, **kw))
    x = torch.from_numpy(G.complex_gaussian(geom.n, 3)).to("cuda:0").requires_grad_(True)
    w = torch.from_numpy(G.complex_gaussian(geom.n, 4)).to("cuda:0")
    y = AG.matvec(p, x)
    loss = (w.conj() * y).real.sum()
    loss.backward()
    # d/dx Re<w, Z x> = Z^H w (torch convention for complex leaves)
    ref = p.matvec_adjoint(w.detach().cpu().numpy())
    assert cases.relerr(x.grad.detach().cpu().numpy(), ref) <= 1e-13
