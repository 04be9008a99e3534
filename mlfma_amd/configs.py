"""The five synthetic workloads the benchmark and the parity tests run, built
from the generators of ``geometry`` (seeded numpy ``default_rng``; the
reference's ``random_device`` is not seedable, src/fileio.h:35-36).

Each builder returns ``(geom, planewave, config kwargs, description)``; the
kwargs go to ``_capi.make_config`` (add ``device=...``).

  c1  4,096 z-directed Hertzian dipoles uniform in a cube of side 4
      wavelengths (seed 0), k = 2 pi, leaf 0.5 wavelength (levels 0-3),
      interpolation order 6, 3 digits, plane wave theta = 0 x-polarised
  c2  PEC sphere of radius 2 wavelengths, geodesic icosphere of frequency 32
      (30,720 RWG unknowns), 7-point triangle rule, leaf 0.25 wavelength
      (levels 0-4)
  c3  PEC sphere of radius 4 wavelengths, icosphere frequency 82 (201,720 RWG
      unknowns), leaf 0.25 wavelength (levels 0-5)
  c4  16 scatterers of uneven size (seed 1): spheres and 2:1:1 ellipsoids of
      equivalent radius uniform in 0.5-3 wavelengths, leaf 0.25 wavelength,
      8 incidence angles each (theta_j = j 180/7 degrees, phi = 0)
  c5  200,000 dipoles uniform in a cube of side 8 wavelengths (seed 2), leaf
      0.5 wavelength (levels 0-4), complex-Gaussian excitation
"""
from __future__ import annotations

import numpy as np

from . import _capi as A
from . import geometry as G

K0 = 2 * np.pi


def planewave_theta0():
    """theta = 0 incidence (along +z), x polarisation, unit amplitude."""
    return A.make_planewave(1.0, (1.0, 0.0, 0.0), (0.0, 0.0, 1.0), K0)


def planewave_xz(theta):
    """Incidence at polar angle theta in the x-z plane (phi = 0), polarised
    in that plane (theta-polarised)."""
    return A.make_planewave(1.0, (np.cos(theta), 0.0, -np.sin(theta)), (np.sin(theta), 0.0, np.cos(theta)), K0)


def _sphere(freq, radius, root, levels, name):
    v, f = G.icosphere(freq, radius)
    r = G.rwg_edges(f)
    geom = A.GeometryArrays(A.MLFMA_SRC_RWG, verts=v, faces=f, rwgs=r)
    kw = dict(root_leng=root, max_node_srcs=0, max_level=levels, interp_order=6, digits=3, over_interp=4.0,
              quad_prec=2)
    desc = {"workload": name, "unknowns": int(len(r)), "triangles": int(len(f)),
            "mean_edge_wavelengths": round(G.mesh_stats(v, f, r)["mean_edge"], 4), "leaf_wavelengths": 0.25,
            "levels": f"0-{levels}", "interp_order": 6, "digits": 3, "over_interp": 4.0, "quad_points": 7,
            "excitation": "plane wave theta=0 x-pol"}
    return geom, planewave_theta0(), kw, desc


def c1():
    pos, mom = G.uniform_dipoles(4096, 4.0, 0)
    geom = A.GeometryArrays(A.MLFMA_SRC_DIPOLE, pos=pos, mom=mom)
    kw = dict(root_leng=4.0, max_node_srcs=0, max_level=3, interp_order=6, digits=3, over_interp=4.0)
    desc = {"workload": "dipoles_4096_cube4", "unknowns": 4096, "leaf_wavelengths": 0.5, "levels": "0-3",
            "interp_order": 6, "digits": 3, "over_interp": 4.0, "excitation": "plane wave theta=0 x-pol"}
    return geom, planewave_theta0(), kw, desc


def c2():
    return _sphere(32, 2.0, 4.0, 4, "rwg_sphere_R2_icosphere32")


def c3():
    return _sphere(82, 4.0, 8.0, 5, "rwg_sphere_R4_icosphere82")


def c4_scatterer(kind, r_eq, mesh_per_wavelength=15.0, scale=1.0, mesh_scale=1.0, quad_prec=2):
    """One scatterer of the c4 family: a sphere of radius r_eq (icosphere of
    frequency round(m r_eq)) or a 2:1:1 ellipsoid of equivalent radius r_eq
    (semi-axes 2a, a, a with a = r_eq / 2^(1/3); an icosphere of frequency
    round(m a) mapped onto it), m = mesh_per_wavelength x mesh_scale.  The
    mesh is then scaled by ``scale`` (same triangles, another electrical
    size).  Fixed depth: root = 0.25 x 2^levels >= 2.05 x the largest
    semi-axis.  Returns a dict with kind, r_eq, geom, kw, n and levels."""
    m = mesh_per_wavelength * mesh_scale
    if kind == "ellipsoid":
        a = r_eq / 2 ** (1 / 3)
        v, f = G.ellipsoid(max(4, int(round(m * a))), (2 * a, a, a))
        ext = 2 * a
    else:
        v, f = G.icosphere(max(4, int(round(m * r_eq))), r_eq)
        ext = r_eq
    v = v * scale
    ext *= scale
    r = G.rwg_edges(f)
    L = 2
    while 0.25 * 2 ** L < 2.05 * ext:
        L += 1
    kw = dict(root_leng=0.25 * 2 ** L, max_node_srcs=0, max_level=L, interp_order=6, digits=3,
              over_interp=4.0, quad_prec=quad_prec)
    return dict(kind=kind, r_eq=float(r_eq), n=int(len(r)), levels=L,
                geom=A.GeometryArrays(A.MLFMA_SRC_RWG, verts=v, faces=f, rwgs=r), kw=kw)


def c4(mesh_per_wavelength=15.0):
    """The 16 scatterers (seed 1): each a sphere or a 2:1:1 ellipsoid with
    probability 1/2, equivalent radius uniform in 0.5-3 wavelengths; built by
    c4_scatterer."""
    rng = np.random.default_rng(1)
    out = []
    for _ in range(16):
        r_eq = rng.uniform(0.5, 3.0)
        ell = rng.uniform() < 0.5
        out.append(c4_scatterer("ellipsoid" if ell else "sphere", r_eq, mesh_per_wavelength))
    return out


def c4_angles():
    return [j * np.pi / 7 for j in range(8)]


def c5(n=200_000):
    pos, mom = G.uniform_dipoles(n, 8.0, 2)
    geom = A.GeometryArrays(A.MLFMA_SRC_DIPOLE, pos=pos, mom=mom)
    kw = dict(root_leng=8.0, max_node_srcs=0, max_level=4, interp_order=6, digits=3, over_interp=4.0,
              flags=A.MLFMA_FLAG_FRESH_LOCALS)
    desc = {"workload": "dipoles_200k_cube8_gradient", "unknowns": n, "leaf_wavelengths": 0.5, "levels": "0-4",
            "interp_order": 6, "digits": 3, "over_interp": 4.0}
    return geom, planewave_theta0(), kw, desc


def c5_dirs():
    """360 bistatic angles theta = 0..359 degrees in the phi = 0 plane."""
    th = np.deg2rad(np.arange(360.0))
    return np.stack([np.sin(th), 0 * th, np.cos(th)], axis=1)
