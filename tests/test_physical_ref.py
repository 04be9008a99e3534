"""CPU pins of the test-only physical references (tests/_physical.py)."""
import numpy as np

from tests import _physical as P


def test_mie_limits():
    # Rayleigh: monostatic sigma = 9 pi a^2 (ka)^4 for a perfect conductor
    ka = 0.02
    a = ka / (2 * np.pi)
    se, sh = P.mie_pec_rcs(ka, np.array([np.pi]))
    assert abs(se[0] / (9 * np.pi * a ** 2 * ka ** 4) - 1) < 1e-3
    assert abs(se[0] - sh[0]) < 1e-12 * se[0]          # backscatter: both cuts agree
    # optical limit: monostatic sigma -> pi a^2
    ka = 60.0
    a = ka / (2 * np.pi)
    se, _ = P.mie_pec_rcs(ka, np.array([np.pi]))
    assert abs(se[0] / (np.pi * a ** 2) - 1) < 0.03


def test_duffy_potentials_against_closed_forms():
    # point above the centroid of a right triangle: I against a fine brute-force
    # product rule; a point inside the triangle (d = 0): finite and positive
    V = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    r = np.array([0.3, 0.3, 0.2])
    I, Iv = P._inv_r_integrals(r, V)
    u, w = np.polynomial.legendre.leggauss(200)
    u = 0.5 * (u + 1); w = 0.5 * w
    X, Y = np.meshgrid(u, u, indexing="ij")
    WW = np.outer(w, w)
    px, py = X, Y * (1 - X)                 # collapsed square onto the triangle
    jac = (1 - X)
    R = np.sqrt((px - r[0]) ** 2 + (py - r[1]) ** 2 + r[2] ** 2)
    Ib = np.sum(WW * jac / R)
    assert abs(I - Ib) < 1e-8 * Ib
    Ivx = np.sum(WW * jac * (px - r[0]) / R)
    assert abs(Iv[0] - Ivx) < 1e-8 * abs(Ib)
    I0, _ = P._inv_r_integrals(np.array([0.25, 0.25, 0.0]), V)
    assert np.isfinite(I0) and I0 > 0
