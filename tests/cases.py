"""Small seeded geometries shared by the parity tests (sizes the oracle
finishes in about a second)."""
from __future__ import annotations

import numpy as np

from mlfma_amd import _capi as A
from mlfma_amd import geometry as G


def dipole_cube(n=600, side=1.8, seed=0, clustered=False, moment=(0.0, 0.0, 1e-8)):
    pos, mom = G.uniform_dipoles(n, side, seed, moment=moment)
    if clustered:  # a dense cluster makes the adaptive tree uneven (lists 3 and 4)
        pos[: n // 3] = pos[: n // 3] * 0.25 + np.array([0.3, -0.2, 0.1])
    return A.GeometryArrays(A.MLFMA_SRC_DIPOLE, pos=pos, mom=mom)


def sphere(freq=5, radius=0.9, center=(0.0, 0.0, 0.0)):
    v, f = G.icosphere(freq, radius, center)
    r = G.rwg_edges(f)
    return A.GeometryArrays(A.MLFMA_SRC_RWG, verts=v, faces=f, rwgs=r)


def ellipsoid(freq=5, axes=(0.9, 0.45, 0.45)):
    v, f = G.ellipsoid(freq, axes)
    r = G.rwg_edges(f)
    return A.GeometryArrays(A.MLFMA_SRC_RWG, verts=v, faces=f, rwgs=r)


def planewave_z():
    """theta = 0 incidence, x polarisation."""
    return A.make_planewave(1.0, (1.0, 0.0, 0.0), (0.0, 0.0, 1.0), 2 * np.pi)


def planewave_x():
    """incidence along +x, z polarisation."""
    return A.make_planewave(1.0, (0.0, 0.0, 1.0), (1.0, 0.0, 0.0), 2 * np.pi)


def two_clusters(seed=11, nA=500, nB=150):
    """Cluster A (side 0.45 about (0.5, 0.5, 0.5)) and cluster B (side 0.2
    about (-0.65, -0.65, -0.65)) in a root box of side 2: A's level-3 boxes
    are all neighbours of each other (empty interaction lists, with children),
    and B reaches A only through the level-2 lists."""
    rng = np.random.default_rng(seed)
    pos = np.vstack([rng.uniform(0.275, 0.725, size=(nA, 3)), rng.uniform(-0.75, -0.55, size=(nB, 3))])
    mom = rng.normal(size=(len(pos), 3))
    return A.GeometryArrays(A.MLFMA_SRC_DIPOLE, pos=pos, mom=mom)


TWO_CLUSTERS_KW = dict(root_leng=2.0, max_node_srcs=8, max_level=-1, interp_order=4, digits=4)


CASES = {
    # name: (geometry factory, config kwargs)
    "dipole_fixed": (lambda: dipole_cube(600, 1.8, 0),
                     dict(root_leng=2.0, max_node_srcs=0, max_level=3, interp_order=6, digits=3)),
    "dipole_adaptive": (lambda: dipole_cube(800, 1.8, 1, clustered=True),
                        dict(root_leng=2.0, max_node_srcs=6, max_level=-1, interp_order=3, digits=3)),
    "rwg_fixed": (lambda: sphere(5, 0.9),
                  dict(root_leng=2.0, max_node_srcs=0, max_level=3, interp_order=3, digits=3, quad_prec=2)),
    "rwg_adaptive": (lambda: ellipsoid(6, (0.95, 0.5, 0.4)),
                     dict(root_leng=2.0, max_node_srcs=10, max_level=-1, interp_order=4, digits=4, quad_prec=1)),
}


def make(name, device=0, **over):
    geom_f, kw = CASES[name]
    kw = dict(kw)
    kw.update(over)
    return geom_f(), kw


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    d = np.linalg.norm((a - b).ravel())
    n = np.linalg.norm(b.ravel())
    return d / n if n > 0 else d
