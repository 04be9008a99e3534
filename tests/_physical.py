"""Test-only references for the physical near field (MLFMA_FLAG_PHYSICAL).

* ``mie_pec_rcs``: bistatic RCS of a PEC sphere from the Mie series (Bohren &
  Huffman, "Absorption and Scattering of Light by Small Particles", 1983,
  §4.4 with the perfect-conductor limit a_n = [x j_n]'/[x h_n]', b_n =
  j_n/h_n).  Pinned in tests by the Rayleigh limit sigma_b = 9 pi a^2 (ka)^4
  and the optical limit sigma_b -> pi a^2.
* ``physical_dense``: the dense physical EFIE matrix of an RWG mesh in the
  library's scheme (outer quadrature rule of the mesh triangles; for close
  triangle pairs the smooth part (e^{ikR}-1)/R sampled by the same rule), but
  with the 1/R part of close pairs integrated numerically by a Duffy
  transform around the projected observation point instead of in closed
  form, so that a comparison isolates the closed-form singular integrals.
"""
from __future__ import annotations

import numpy as np
from scipy.special import spherical_jn, spherical_yn

# 7-point triangle rule (barycentric a, b, c; weight), weights summing to 1/2
RULE7 = [(1 / 3, 1 / 3, 1 / 3, 0.1125)]
for a_, b_, w_ in ((0.059715871789770, 0.470142064105115, 0.066197076394253),
                   (0.797426985353087, 0.101286507323456, 0.0629695902724135)):
    RULE7 += [(a_, b_, b_, w_), (b_, a_, b_, w_), (b_, b_, a_, w_)]


def mie_pec_rcs(ka: float, theta: np.ndarray, wavelength: float = 1.0):
    """(sigma_phi0, sigma_phi90) in wavelength^2 units for an x-polarised
    plane wave travelling along +z; theta in radians (0 = forward)."""
    x = ka
    nmax = int(np.ceil(x + 4.05 * x ** (1 / 3) + 2))
    n = np.arange(1, nmax + 1)
    jn, yn = spherical_jn(n, x), spherical_yn(n, x)
    djn, dyn = spherical_jn(n, x, derivative=True), spherical_yn(n, x, derivative=True)
    hn = jn + 1j * yn
    dxjn = jn + x * djn
    dxhn = hn + x * (djn + 1j * dyn)
    a = dxjn / dxhn
    b = jn / hn
    mu = np.cos(np.asarray(theta, dtype=float))
    pi = np.zeros((nmax + 1, mu.size))
    tau = np.zeros((nmax + 1, mu.size))
    pi[1] = 1.0
    if nmax >= 2:
        pi[2] = 3 * mu
    for m in range(3, nmax + 1):
        pi[m] = (2 * m - 1) / (m - 1) * mu * pi[m - 1] - m / (m - 1) * pi[m - 2]
    for m in range(1, nmax + 1):
        tau[m] = m * mu * pi[m] - (m + 1) * (pi[m - 1] if m > 1 else 0.0)
    c = (2 * n + 1) / (n * (n + 1))
    S1 = (c[:, None] * (a[:, None] * pi[1:] + b[:, None] * tau[1:])).sum(0)
    S2 = (c[:, None] * (a[:, None] * tau[1:] + b[:, None] * pi[1:])).sum(0)
    k = 2 * np.pi / wavelength
    return 4 * np.pi / k ** 2 * np.abs(S2) ** 2, 4 * np.pi / k ** 2 * np.abs(S1) ** 2


def _gl01(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1), 0.5 * w


def _graded(n):
    """Gauss-Legendre on [0, 1] graded cubically toward 0."""
    z, wz = _gl01(n)
    return z ** 3, 3 * z ** 2 * wz


def _inv_r_integrals(r, V, nu=96, nt=48):
    """I = int_T 1/|r-r'| dS', Iv = int_T (r'-r)/|r-r'| dS' over the triangle
    with vertices V (3x3), for one point r (3,) or many (m, 3), by a Duffy
    transform around the projection p of r: T = sum of the signed triangles
    (p, v_i, v_i+1), each mapped from the unit square (r' = p + u w(t),
    w = (v_i - p) + t (v_i+1 - v_i)) with Jacobian u * 2 area.  With d the
    height of r, r' - r = u w - d n and |r' - r| = sqrt(u^2 |w|^2 + d^2), so
    Jacobian / R = 2 area / sqrt(|w|^2 + (d/u)^2) is formed without
    cancellation.  The radial coordinate is graded toward the apex (u = s^4)
    and the base coordinate is split at the foot t* of the perpendicular from
    p with both panels graded toward it: |w(t)| is smallest there, and a
    point close to an edge or a vertex makes the integrand peak sharply."""
    r = np.atleast_2d(np.asarray(r, float))
    e1, e2 = V[1] - V[0], V[2] - V[0]
    nrm = np.cross(e1, e2)
    nrm /= np.linalg.norm(nrm)
    d = (r - V[0]) @ nrm                         # (m,)
    p = r - d[:, None] * nrm                     # (m, 3)
    s, ws = _gl01(nu)
    u, wu = s ** 4, 4 * s ** 3 * ws
    g, wg = _graded(nt)
    I = np.zeros(len(r))
    Iv = np.zeros((len(r), 3))
    for i in range(3):
        a, b = V[i], V[(i + 1) % 3]
        area2 = np.cross(a - p, b - p) @ nrm      # (m,) signed, twice the area
        L2 = np.dot(b - a, b - a)
        ts = np.clip(((p - a) @ (b - a)) / L2, 0.0, 1.0)                      # (m,)
        # panels [0, t*] and [t*, 1], nodes graded toward t*
        t = np.concatenate([ts[:, None] * (1 - g[None, :]), ts[:, None] + (1 - ts[:, None]) * g[None, :]], 1)
        wt = np.concatenate([ts[:, None] * wg[None, :], (1 - ts[:, None]) * wg[None, :]], 1)   # (m, 2nt)
        w = (a - p)[:, None, :] + t[:, :, None] * (b - a)[None, None, :]                     # (m, 2nt, 3)
        w2 = np.sum(w * w, axis=-1)
        ratio = area2[:, None, None] / np.sqrt(w2[:, None, :] + (d[:, None, None] / u[None, :, None]) ** 2)
        W = wu[None, :, None] * wt[:, None, :]                                              # (m, nu, 2nt)
        I += np.sum(W * ratio, axis=(1, 2))
        uw = np.einsum("mut,mtc->mc", W * ratio * u[None, :, None], w)
        Iv += uw - d[:, None] * np.sum(W * ratio, axis=(1, 2))[:, None] * nrm
    if I.size == 1:
        return float(I[0]), Iv[0]
    return I, Iv


def physical_dense(verts, faces, rwgs, k, rule=RULE7):
    """Dense physical EFIE matrix (without the constant C k) in the library's
    normalisation: Z_mn = l_m l_n sum s s' sum_q sum_q' w w' (rho.rho' -
    4/k^2) g, g = e^{ikR}/R, with the 1/R part of close triangle pairs
    (centroid distance < 2 x the longer edge) integrated by Duffy quadrature."""
    verts = np.asarray(verts, float)
    faces = np.asarray(faces, np.int64)
    nt = len(faces)
    bary = np.array([r[:3] for r in rule])
    wq = np.array([r[3] for r in rule])
    TV = verts[faces]                      # (nt, 3, 3)
    Q = np.einsum("qj,tjc->tqc", bary, TV)  # (nt, nq, 3)
    cen = TV.mean(axis=1)
    edges = np.stack([np.linalg.norm(TV[:, 1] - TV[:, 0], axis=1), np.linalg.norm(TV[:, 2] - TV[:, 1], axis=1),
                      np.linalg.norm(TV[:, 0] - TV[:, 2], axis=1)], 1).max(1)
    area = 0.5 * np.linalg.norm(np.cross(TV[:, 1] - TV[:, 0], TV[:, 2] - TV[:, 0]), axis=1)
    n = len(rwgs)
    leng = np.linalg.norm(verts[rwgs[:, 0]] - verts[rwgs[:, 1]], axis=1)
    free = np.zeros((n, 2, 3))
    for i, (v0, v1, tp, tm) in enumerate(rwgs):
        for s, t in enumerate((tp, tm)):
            fv = [v for v in faces[t] if v != v0 and v != v1]
            free[i, s] = verts[fv[0]]
    k2 = 4.0 / k ** 2
    # per triangle pair: closeness and, for close pairs, (I, Iv) at each outer point
    sing = {}

    def pair_terms(to, ts):
        key = (to, ts)
        if key not in sing:
            close = np.linalg.norm(cen[to] - cen[ts]) < 2 * max(edges[to], edges[ts])
            vals = None
            if close:
                Iq, Ivq = _inv_r_integrals(Q[to], TV[ts])
                vals = list(zip(Iq, Ivq))
            sing[key] = (close, vals)
        return sing[key]

    Z = np.zeros((n, n), dtype=complex)
    for m in range(n):
        for q in range(n):
            acc = 0j
            for so, to in enumerate(rwgs[m, 2:]):
                for ss, ts in enumerate(rwgs[q, 2:]):
                    sg = -1.0 if so != ss else 1.0
                    close, vals = pair_terms(to, ts)
                    ro = Q[to]                         # (nq, 3)
                    rs = Q[ts]
                    vo = ro - free[m, so]
                    vs = rs - free[q, ss]
                    R = np.linalg.norm(ro[:, None, :] - rs[None, :, :], axis=-1)
                    h = vo @ vs.T - k2
                    if close:
                        g = np.where(R > 0, (np.exp(1j * k * R) - 1) / np.where(R > 0, R, 1), 1j * k)
                    else:
                        g = np.exp(1j * k * R) / R
                    inner = (h * g) @ wq
                    if close:
                        for i in range(len(wq)):
                            I, Iv = vals[i]
                            rx = ro[i] - free[q, ss]
                            inner[i] += (np.dot(vo[i], Iv + rx * I) - k2 * I) / (2 * area[ts])
                    acc += sg * np.dot(wq, inner)
            Z[m, q] = leng[m] * leng[q] * acc
    return Z
