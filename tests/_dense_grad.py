"""Dense O(N^2) references for dipole position gradients (test-only).

Z_ij = C k G(r_i - r_j; p_i, p_j) for i != j (zero diagonal, as the
reference drops the self term, src/sources/dipole.h:62-70), with
G(X; a, b) = e^{ikr}/r [(1 + i/kr - 1/(kr)^2) a.b - (1 + 3i/kr - 3/(kr)^2)
(a.X^)(b.X^)] (src/math.h:132-142) and C k = -i c0 mu0 k / 4 pi
(src/phys.h:10).  The far field is E_d = C k (I - r^r^) sum_s I_s p_s
e^{-ik r^_d . r_s} (src/fmm/node.cpp:147-178 generalised to any direction).

The gradient of G with respect to its first position is written out
independently of the library's kernel: with A(r) = e^{ikr}(1/r - 1/(k^2 r^3)
+ i/(k r^2)) and B(r) = e^{ikr}(1/r - 3/(k^2 r^3) + 3i/(k r^2)),
G = A a.b - B (a.X^)(b.X^) and
grad G = A' X^ a.b - B' X^ (a.X^)(b.X^) - B [(a - (a.X^)X^)(b.X^) + (a.X^)(b - (b.X^)X^)] / r.
Since G(X; a, b) = G(-X; b, a), dZ_ji/dr_i = -C k grad G(r_i - r_j; p_i, p_j).
"""
from __future__ import annotations

import numpy as np

C0 = 299792458.0
MU0 = 1.256637e-6


def ck(k: float) -> complex:
    return -1j * C0 * MU0 / (4 * np.pi) * k


def _parts(X, a, b, k):
    r = np.linalg.norm(X, axis=-1)
    h = X / r[..., None]
    e = np.exp(1j * k * r)
    A = e * (1 / r - 1 / (k ** 2 * r ** 3) + 1j / (k * r ** 2))
    B = e * (1 / r - 3 / (k ** 2 * r ** 3) + 3j / (k * r ** 2))
    ab = np.sum(a * b, axis=-1)
    ah = np.sum(a * h, axis=-1)
    bh = np.sum(b * h, axis=-1)
    return r, h, e, A, B, ab, ah, bh


def green(X, a, b, k):
    _, _, _, A, B, ab, ah, bh = _parts(X, a, b, k)
    return A * ab - B * ah * bh


def green_grad(X, a, b, k):
    r, h, e, A, B, ab, ah, bh = _parts(X, a, b, k)
    Ad = 1j * k * A + e * (-1 / r ** 2 + 3 / (k ** 2 * r ** 4) - 2j / (k * r ** 3))
    Bd = 1j * k * B + e * (-1 / r ** 2 + 9 / (k ** 2 * r ** 4) - 6j / (k * r ** 3))
    t = ((a - ah[..., None] * h) * bh[..., None] + ah[..., None] * (b - bh[..., None] * h)) / r[..., None]
    return (Ad * ab)[..., None] * h - (Bd * ah * bh)[..., None] * h - B[..., None] * t


def _pairs(pos, phat, rows):
    X = pos[rows, None, :] - pos[None, :, :]
    a = np.broadcast_to(phat[rows, None, :], X.shape)
    b = np.broadcast_to(phat[None, :, :], X.shape)
    same = np.all(X == 0, axis=-1)
    X = np.where(same[..., None], 1.0, X)     # placeholder, masked below
    return X, a, b, same


def dense_matvec(pos, phat, x, k, chunk=256):
    n = len(pos)
    y = np.zeros(n, dtype=complex)
    for s in range(0, n, chunk):
        rows = np.arange(s, min(n, s + chunk))
        X, a, b, same = _pairs(pos, phat, rows)
        Z = np.where(same, 0.0, green(X, a, b, k))
        y[rows] = ck(k) * (Z @ x)
    return y


def operator_gradient(pos, phat, I, lam, k, chunk=256):
    """-2 Re(lam^H (dZ/dr_i^a) I), (n, 3)."""
    n = len(pos)
    g = np.zeros((n, 3))
    lc = np.conj(lam)
    for s in range(0, n, chunk):
        rows = np.arange(s, min(n, s + chunk))
        X, a, b, same = _pairs(pos, phat, rows)
        G = np.where(same[..., None], 0.0, green_grad(X, a, b, k)) * ck(k)   # (m, n, 3)
        row = lc[rows, None] * np.einsum("mjc,j->mc", G, I)
        col = I[rows, None] * np.einsum("mjc,j->mc", G, lc)
        g[rows] = -2.0 * np.real(row + col)
    return g


def farfield_gradient(pos, phat, I, E, dirs, k):
    """2 Re(E^H (dF/dr_i^a) I), (n, 3)."""
    rh = np.asarray(dirs, float)
    P = np.eye(3)[None] - rh[:, :, None] * rh[:, None, :]          # (d, 3, 3)
    PE = np.einsum("dij,dj->di", P, np.conj(E))                     # conj(E_d)^T P
    w = np.einsum("di,si->ds", PE, phat)                            # (d, s)
    ph = np.exp(-1j * k * (rh @ pos.T))                             # (d, s)
    core = ck(k) * w * ph * (-1j * k) * I[None, :]                  # (d, s)
    return 2.0 * np.real(np.einsum("ds,dc->sc", core, rh))
