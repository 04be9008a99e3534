"""The dense gradient references of tests/_dense_grad.py against central
finite differences of their own functions on small random systems (CPU):
the Green's-function gradient, -2 Re(lambda^H dZ I) and 2 Re(E^H dF I)."""
import numpy as np

from tests import _dense_grad as D

K = 2 * np.pi


def dense_Z(pos, phat):
    n = len(pos)
    return np.stack([D.dense_matvec(pos, phat, np.eye(n, dtype=complex)[:, j], K) for j in range(n)], axis=1)


def dense_F(pos, phat, dirs):
    n = len(pos)
    rh = np.asarray(dirs)
    P = np.eye(3)[None] - rh[:, :, None] * rh[:, None, :]
    ph = np.exp(-1j * K * (rh @ pos.T))
    return D.ck(K) * np.einsum("dij,sj,ds->dis", P, phat, ph)      # (d, 3, s)


def test_green_gradient_finite_differences():
    rng = np.random.default_rng(0)
    for _ in range(5):
        X = rng.normal(size=3) * 0.4
        a, b = rng.normal(size=3), rng.normal(size=3)
        g = D.green_grad(X, a, b, K)
        h = 1e-6
        fd = np.array([(D.green(X + h * e, a, b, K) - D.green(X - h * e, a, b, K)) / (2 * h) for e in np.eye(3)])
        assert np.abs(g - fd).max() <= 1e-7 * np.abs(fd).max()
        # the symmetry used for the column terms
        assert abs(D.green(X, a, b, K) - D.green(-X, b, a, K)) <= 1e-14 * abs(D.green(X, a, b, K))


def test_operator_and_farfield_gradients_finite_differences():
    rng = np.random.default_rng(1)
    n = 6
    pos = rng.uniform(-0.6, 0.6, size=(n, 3))
    phat = rng.normal(size=(n, 3))
    phat /= np.linalg.norm(phat, axis=1)[:, None]
    I = rng.normal(size=n) + 1j * rng.normal(size=n)
    lam = rng.normal(size=n) + 1j * rng.normal(size=n)
    t = np.linspace(0.2, 2.9, 5)
    dirs = np.stack([np.sin(t), 0 * t, np.cos(t)], axis=1)
    E = rng.normal(size=(5, 3)) + 1j * rng.normal(size=(5, 3))
    g_op = D.operator_gradient(pos, phat, I, lam, K)
    g_ff = D.farfield_gradient(pos, phat, I, E, dirs, K)
    h = 1e-6
    for i in range(n):
        for a in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[i, a] += h
            pm[i, a] -= h
            f_op = (-2 * np.real(np.vdot(lam, dense_Z(pp, phat) @ I)) + 2 * np.real(np.vdot(lam, dense_Z(pm, phat) @ I))) / (2 * h)
            f_ff = (2 * np.real(np.vdot(E, np.einsum("dis,s->di", dense_F(pp, phat, dirs), I)))
                    - 2 * np.real(np.vdot(E, np.einsum("dis,s->di", dense_F(pm, phat, dirs), I)))) / (2 * h)
            assert abs(g_op[i, a] - f_op) <= 1e-6 * np.abs(g_op).max()
            assert abs(g_ff[i, a] - f_ff) <= 1e-6 * np.abs(g_ff).max()
