"""Host-side dense algebra of the deflated GMRES restarts (mlfma_amd/csrc/dense.hpp).

The header is compiled with g++ into a small shim (tests/dense_shim.cpp) and
checked against numpy: eigenvalues/eigenvectors, full QR and LU solves; then
the restart itself against a numpy restatement of the device's pipelined
GMRES(m) cycle (classical Gram-Schmidt twice, Givens rotations
G = [[conj c, conj s], [-s, c]]): the kept vectors satisfy A Z_k = V' [R_k; 0]
with V' orthonormal, the residual of the cycle's solution is V' g', and a
warm-started second cycle keeps those relations.  Finally GMRES-DR must
converge in fewer iterations than GMRES(m) on a matrix whose few small
eigenvalues stall the plain restarts (the behaviour deflation exists for)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("dense") / "libdense_shim.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", os.path.join(HERE, "dense_shim.cpp"), "-o", out],
                   check=True)
    L = C.CDLL(out)
    P = C.POINTER(C.c_double)
    L.dense_eig.argtypes = [C.c_int, P, P, P]
    L.dense_qr.argtypes = [C.c_int, C.c_int, P, P, P]
    L.dense_lu_solve.argtypes = [C.c_int, C.c_int, P, P]
    L.dense_dr.argtypes = [C.c_int, P, P, P, P, C.c_int, P, C.c_int, P, P, P, P, P, P]
    return L


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _cm(a):
    """complex matrix -> column-major interleaved doubles"""
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex128).T).view(np.float64).copy()


def _from_cm(buf, r, c):
    return buf.view(np.complex128).reshape(c, r).T.copy()


def eig(L, A):
    n = A.shape[0]
    w = np.zeros(2 * n)
    X = np.zeros(2 * n * n)
    assert L.dense_eig(n, _p(_cm(A)), _p(w), _p(X)) == 1
    return w.view(np.complex128).copy(), _from_cm(X, n, n)


def rand_c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("n", [1, 2, 3, 8, 40, 150])
def test_eig_matches_numpy(lib, n):
    rng = np.random.default_rng(n)
    A = rand_c(rng, n, n)
    w, X = eig(lib, A)
    an = np.linalg.norm(A)
    # every eigenpair is one: |A x - w x| small, |x| = 1
    assert np.allclose(np.linalg.norm(X, axis=0), 1.0, atol=1e-12)
    assert np.linalg.norm(A @ X - X * w, axis=0).max() <= 1e-11 * an
    # the same spectrum as LAPACK (greedy nearest matching)
    ref = list(np.linalg.eigvals(A))
    for v in w:
        i = int(np.argmin([abs(v - r) for r in ref]))
        assert abs(v - ref[i]) <= 1e-9 * an
        ref.pop(i)


def test_eig_hessenberg_with_tiny_eigenvalues(lib):
    # the shape the restart sees: eigenvalues spread over decades
    rng = np.random.default_rng(7)
    lam = np.concatenate([1e-5 * np.exp(2j * np.pi * rng.random(4)), 1 + 0.5 * rand_c(rng, 56)])
    Q, _ = np.linalg.qr(rand_c(rng, 60, 60))
    A = Q @ np.diag(lam) @ Q.conj().T
    w, X = eig(lib, A)
    small = np.sort(np.abs(w))[:4]
    assert np.allclose(small, np.sort(np.abs(lam))[:4], rtol=1e-6)
    assert np.linalg.norm(A @ X - X * w, axis=0).max() <= 1e-11 * np.linalg.norm(A)


@pytest.mark.parametrize("r,c", [(5, 5), (9, 4), (21, 20), (2, 1)])
def test_qr_full(lib, r, c):
    rng = np.random.default_rng(r * 31 + c)
    A = rand_c(rng, r, c)
    Q = np.zeros(2 * r * r)
    R = np.zeros(2 * r * c)
    lib.dense_qr(r, c, _p(_cm(A)), _p(Q), _p(R))
    Q, R = _from_cm(Q, r, r), _from_cm(R, r, c)
    assert np.allclose(Q.conj().T @ Q, np.eye(r), atol=1e-13)
    assert np.allclose(np.tril(R, -1), 0, atol=0)
    assert np.linalg.norm(Q @ R - A) <= 1e-13 * np.linalg.norm(A)


def test_lu_solve(lib):
    rng = np.random.default_rng(3)
    A = rand_c(rng, 30, 30)
    B = rand_c(rng, 30, 4)
    buf = _cm(B)
    assert lib.dense_lu_solve(30, 4, _p(_cm(A)), _p(buf)) == 1
    assert np.allclose(_from_cm(buf, 30, 4), np.linalg.solve(A, B), rtol=1e-11, atol=1e-12)
    assert lib.dense_lu_solve(2, 1, _p(_cm(np.zeros((2, 2)))), _p(_cm(np.ones((2, 1))))) == 0


# ---------------------------------------------------------------- restart
def givens_cycle(A, V, Zin, H0, g0, cs0, sn0, j0, m):
    """numpy restatement of the device cycle from step j0 (columns < j0 given:
    V[:, :j0+1], inputs Zin[:j0], rotated columns H0, rotations cs0/sn0, g0).
    Returns V (m+1 columns), the inputs, R (m x m), cs, sn, g (m+1)."""
    n = A.shape[0]
    V = np.concatenate([V[:, :j0 + 1], np.zeros((n, m - j0), complex)], axis=1)
    Z = list(Zin[:j0])
    R = np.zeros((m, m), complex)
    R[:j0, :j0] = H0
    cs = np.ones(m, complex)
    sn = np.zeros(m, complex)
    cs[:j0], sn[:j0] = cs0, sn0
    g = np.zeros(m + 1, complex)
    g[:len(g0)] = g0
    for j in range(j0, m):
        z = V[:, j]
        Z.append(z)
        w = A @ z
        h = np.zeros(j + 2, complex)
        for _ in range(2):
            c = V[:, :j + 1].conj().T @ w
            w = w - V[:, :j + 1] @ c
            h[:j + 1] += c
        h[j + 1] = np.linalg.norm(w)
        V[:, j + 1] = w / h[j + 1]
        for i in range(j):
            a = np.conj(cs[i]) * h[i] + np.conj(sn[i]) * h[i + 1]
            h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1]
            h[i] = a
        rho = np.hypot(abs(h[j]), abs(h[j + 1]))
        cs[j], sn[j] = h[j] / rho, h[j + 1] / rho
        h[j], h[j + 1] = rho, 0
        R[:j + 1, j] = h[:j + 1]
        g[j + 1] = -sn[j] * g[j]
        g[j] = np.conj(cs[j]) * g[j]
    return V, np.stack(Z, axis=1), R, cs, sn, g


def dr(L, R, cs, sn, gm, k0, Qprev, k):
    m = R.shape[0]
    Pk, U, Rk = np.zeros(2 * m * k), np.zeros(2 * (m + 1) * (k + 1)), np.zeros(2 * k * k)
    g, Qh, dfc = np.zeros(2 * (k + 1)), np.zeros(2 * (k + 1) ** 2), np.zeros(1)
    qp = _cm(Qprev) if k0 > 0 else np.zeros(2)
    ok = L.dense_dr(m, _p(_cm(R)), _p(_cm(cs[None, :])), _p(_cm(sn[None, :])), _p(np.array([gm.real, gm.imag])),
                    k0, _p(qp), k, _p(Pk), _p(U), _p(Rk), _p(g), _p(Qh), _p(dfc))
    assert ok == 1
    return (_from_cm(Pk, m, k), _from_cm(U, m + 1, k + 1), _from_cm(Rk, k, k), g.view(np.complex128).copy(),
            _from_cm(Qh, k + 1, k + 1), float(dfc[0]))


def hard_matrix(n, seed, nsmall=5):
    rng = np.random.default_rng(seed)
    lam = np.concatenate([np.linspace(2e-3, 2e-2, nsmall) * np.exp(1j * rng.random(nsmall)),
                          1.5 + 0.8 * rng.random(n - nsmall) * np.exp(2j * np.pi * rng.random(n - nsmall))])
    S = np.eye(n) + 0.2 * rand_c(rng, n, n) / np.sqrt(n)
    return S @ np.diag(lam) @ np.linalg.inv(S), rand_c(rng, n)


def test_restart_relations_two_cycles(lib):
    n, m, k = 300, 30, 6
    A, b = hard_matrix(n, 11)
    beta = np.linalg.norm(b)
    V0 = (b / beta)[:, None]
    V, Z, R, cs, sn, g = givens_cycle(A, V0, [], np.zeros((0, 0)), [beta], [], [], 0, m)
    assert np.allclose(A @ Z, V @ _unrot(R, cs, sn), atol=1e-10 * np.linalg.norm(A))
    y = np.linalg.solve(R, g[:m])
    x1 = Z @ y
    assert np.isclose(np.linalg.norm(b - A @ x1), abs(g[m]), rtol=1e-8)
    Pk, U, Rk, gp, Qh, dfc = dr(lib, R, cs, sn, g[m], 0, None, k)
    assert dfc < 1e-9
    Zk, Vn = V[:, :m] @ Pk, V @ U
    assert np.allclose(Vn.conj().T @ Vn, np.eye(k + 1), atol=1e-12)
    assert np.allclose(np.tril(Rk, -1), 0)
    an = np.linalg.norm(A @ Zk)
    assert np.linalg.norm(A @ Zk - Vn[:, :k] @ Rk) <= 1e-9 * an
    r1 = b - A @ x1
    assert np.linalg.norm(r1 - Vn @ gp) <= 1e-9 * np.linalg.norm(r1)
    # second cycle warm-started from the deflated block
    V2, Z2, R2, cs2, sn2, g2 = givens_cycle(A, Vn, list(Zk.T), Rk, gp, np.ones(k), np.zeros(k), k, m)
    assert np.allclose(A @ Z2, V2 @ _unrot(R2, cs2, sn2), atol=1e-9 * np.linalg.norm(A))
    y2 = np.linalg.solve(R2, g2[:m])
    x2 = x1 + Z2 @ y2
    r2 = b - A @ x2
    assert np.isclose(np.linalg.norm(r2), abs(g2[m]), rtol=1e-6)
    Pk2, U2, Rk2, gp2, Qh2, dfc2 = dr(lib, R2, cs2, sn2, g2[m], k, Qh, k)
    assert dfc2 < 1e-8
    Zk2, Vn2 = V2[:, :m] @ Pk2, V2 @ U2
    assert np.allclose(Vn2.conj().T @ Vn2, np.eye(k + 1), atol=1e-12)
    assert np.linalg.norm(A @ Zk2 - Vn2[:, :k] @ Rk2) <= 1e-8 * np.linalg.norm(A @ Zk2)
    assert np.linalg.norm(r2 - Vn2 @ gp2) <= 1e-8 * np.linalg.norm(r2)


def _unrot(R, cs, sn):
    """Omega^H [R; 0] (the Hessenberg matrix before the rotations)"""
    m = R.shape[0]
    H = np.vstack([R, np.zeros((1, m))]).astype(complex)
    for i in range(m - 1, -1, -1):
        p, q = H[i].copy(), H[i + 1].copy()
        H[i] = cs[i] * p - np.conj(sn[i]) * q
        H[i + 1] = sn[i] * p + np.conj(cs[i]) * q
    return H


def _solve(lib, A, b, m, k, tol, maxit):
    """GMRES(m) (k = 0) or GMRES-DR(m, k) through the numpy cycle and the
    shim's restart; true residual checked at every restart like the device
    solver.  Returns (iterations, relative residual)."""
    n = A.shape[0]
    bn = np.linalg.norm(b)
    x = np.zeros(n, complex)
    it, state = 0, None
    while it < maxit:
        r = b - A @ x
        if np.linalg.norm(r) / bn <= tol:
            break
        if state is None:
            beta = np.linalg.norm(r)
            V, Z, R, cs, sn, g = givens_cycle(A, (r / beta)[:, None], [], np.zeros((0, 0)), [beta], [], [], 0, m)
            it += m
            k0, Qp = 0, None
        else:
            Zk, Vn, Rk, gp, Qp = state
            V, Z, R, cs, sn, g = givens_cycle(A, Vn, list(Zk.T), Rk, gp, np.ones(k), np.zeros(k), k, m)
            it += m - k
            k0 = k
        x = x + Z @ np.linalg.solve(R, g[:m])
        if k:
            Pk, U, Rk, gp, Qh, _ = dr(lib, R, cs, sn, g[m], k0, Qp, k)
            state = (V[:, :m] @ Pk, V @ U, Rk, gp, Qh)
    return it, np.linalg.norm(b - A @ x) / bn


def test_deflation_beats_plain_restarts(lib):
    A, b = hard_matrix(400, 5)
    it0, r0 = _solve(lib, A, b, 25, 0, 1e-8, 1500)
    it1, r1 = _solve(lib, A, b, 25, 6, 1e-8, 1500)
    assert r1 <= 1e-8
    assert it1 < 0.6 * it0, (it0, r0, it1, r1)
