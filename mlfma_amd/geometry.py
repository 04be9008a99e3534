"""Synthetic scatterers and the reference's on-disk input formats.

Generators produce the arrays that ``mlfma_geometry`` takes (``include/mlfma_c.h``):

* point dipoles: positions and moment vectors;
* RWG surfaces: vertices, triangles and RWG edges ``v0 v1 t_plus t_minus``
  (the layout ``importRWG`` reads, src/fileio.h:174-202).

Readers and writers follow the text formats of the reference:
``config/config.txt`` (src/config.h:50-58), ``config/pwave.txt``
(src/fileio.h:8-27), dipole files (src/fileio.h:106-130) and
``vertices.txt`` / ``faces.txt`` / ``rwgs.txt`` (src/fileio.h:132-202).
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass, field

import numpy as np

# --------------------------------------------------------------------------- shapes


def _icosahedron():
    p = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, p, 0], [1, p, 0], [-1, -p, 0], [1, -p, 0],
                  [0, -1, p], [0, 1, p], [0, -1, -p], [0, 1, -p],
                  [p, 0, -1], [p, 0, 1], [-p, 0, -1], [-p, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def icosphere(freq: int, radius: float = 1.0, center=(0.0, 0.0, 0.0)):
    """Geodesic icosphere of frequency ``freq``: 20 freq^2 triangles,
    10 freq^2 + 2 vertices, 30 freq^2 edges, outward-oriented faces."""
    if freq < 1:
        raise ValueError("freq must be >= 1")
    iv, ifc = _icosahedron()
    pts = []
    tris = []
    # barycentric lattice points of one face, then small triangles
    lat = [(i, j) for i in range(freq + 1) for j in range(freq + 1 - i)]
    index = {ij: n for n, ij in enumerate(lat)}
    up = [(index[(i, j)], index[(i + 1, j)], index[(i, j + 1)])
          for i in range(freq) for j in range(freq - i)]
    down = [(index[(i + 1, j)], index[(i + 1, j + 1)], index[(i, j + 1)])
            for i in range(freq) for j in range(freq - 1 - i)]
    local = np.array(up + down, dtype=np.int64)
    ij = np.array(lat, dtype=np.float64)
    base = 0
    for a, b, c in ifc:
        A, B, C = iv[a], iv[b], iv[c]
        P = A[None, :] + np.outer(ij[:, 0] / freq, B - A) + np.outer(ij[:, 1] / freq, C - A)
        pts.append(P)
        tris.append(local + base)
        base += len(lat)
    P = np.concatenate(pts)
    T = np.concatenate(tris)
    P /= np.linalg.norm(P, axis=1, keepdims=True)
    # merge the copies of edge and corner points shared by neighbouring faces
    key = np.round(P * 1e9).astype(np.int64)
    _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    inv = inv.reshape(-1)
    order = np.argsort(first)            # keep vertices in first-seen order
    remap = np.empty_like(order)
    remap[order] = np.arange(len(order))
    verts = P[first[order]] * radius + np.asarray(center, dtype=np.float64)[None, :]
    faces = remap[inv][T]
    # outward orientation
    cen = verts[faces].mean(axis=1) - np.asarray(center)[None, :]
    nrm = np.cross(verts[faces[:, 1]] - verts[faces[:, 0]], verts[faces[:, 2]] - verts[faces[:, 0]])
    flip = (nrm * cen).sum(axis=1) < 0
    faces[flip] = faces[flip][:, [0, 2, 1]]
    return verts, faces.astype(np.int32)


def ellipsoid(freq: int, semi_axes=(2.0, 1.0, 1.0), center=(0.0, 0.0, 0.0)):
    """Icosphere mapped onto an ellipsoid with the given semi-axes."""
    v, f = icosphere(freq, 1.0)
    v = v * np.asarray(semi_axes, dtype=np.float64)[None, :] + np.asarray(center)[None, :]
    return v, f


def rwg_edges(faces: np.ndarray) -> np.ndarray:
    """RWG functions of a closed or open triangle mesh: one per edge shared by
    exactly two triangles, as rows ``v0 v1 t_plus t_minus`` (src/fileio.h:174-202).
    Edges are numbered in order of first appearance."""
    f = np.asarray(faces, dtype=np.int64)
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    t = np.tile(np.arange(len(f)), 3)
    loc = np.concatenate([np.zeros(len(f)), np.ones(len(f)), 2 * np.ones(len(f))]).astype(np.int64)
    # order by (triangle, local edge) so numbering follows the face list
    o = np.lexsort((loc, t))
    e, t = e[o], t[o]
    key = np.sort(e, axis=1)
    _, first, inv, cnt = np.unique(key, axis=0, return_index=True, return_inverse=True,
                                   return_counts=True)
    inv = inv.reshape(-1)
    rows = []
    owners = {}
    for n in range(len(e)):
        owners.setdefault(inv[n], []).append(n)
    for u in np.argsort(first):
        occ = owners[u]
        if len(occ) != 2:
            continue
        a, b = occ
        rows.append((e[a, 0], e[a, 1], t[a], t[b]))
    return np.array(rows, dtype=np.int32)


def mesh_stats(verts, faces, rwgs):
    el = np.linalg.norm(verts[rwgs[:, 0]] - verts[rwgs[:, 1]], axis=1)
    return {"nverts": len(verts), "ntris": len(faces), "nrwg": len(rwgs),
            "mean_edge": float(el.mean()), "max_edge": float(el.max())}


def uniform_dipoles(n: int, side: float, seed: int, moment=(0.0, 0.0, 1e-8)):
    """n dipoles uniform in the cube [-side/2, side/2]^3 (numpy default_rng(seed));
    every dipole carries the same moment vector."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-side / 2.0, side / 2.0, size=(n, 3))
    mom = np.tile(np.asarray(moment, dtype=np.float64), (n, 1))
    return pos, mom


def complex_gaussian(n: int, seed: int, stream: int = 1):
    """CN(0,1) vector: real and imaginary parts each N(0, 1/2)."""
    rng = np.random.default_rng([seed, stream])
    return (rng.normal(0, np.sqrt(0.5), n) + 1j * rng.normal(0, np.sqrt(0.5), n))


# --------------------------------------------------------------------------- files


@dataclass
class RefConfig:
    """Fields of the reference's config/config.txt in file order (src/config.h:54-57)."""
    mode: int = 0
    pdist: int = 0
    qdist: int = 0
    quadPrec: int = 2
    digits: int = 3
    interpOrder: int = 6
    overInterp: float = 4.0
    rootLeng: float = 4.0
    maxNodeSrcs: int = 16
    evalDirect: int = 0
    nsrcs: int = 0
    extra: dict = field(default_factory=dict)


_CFG_ORDER = ["mode", "pdist", "qdist", "quadPrec", "digits", "interpOrder", "overInterp",
              "rootLeng", "maxNodeSrcs", "evalDirect", "nsrcs"]


def write_config(path: str, cfg: RefConfig):
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    with open(path, "w") as f:
        for k in _CFG_ORDER:
            f.write(f"{k}: {getattr(cfg, k)}\n")


_INT_RE = re.compile(r"\d+")
_FLOAT_RE = re.compile(r"\d+(\.\d*)?")
_EXP_RE = re.compile(r"[eE][+-]?\d+")


def _skip_to_digit(s: str, i: int) -> int:
    while i < len(s) and not s[i].isdigit():
        i += 1
    return i


def _last_number(line: str, kind):
    """The value the reference's ``operator>>`` leaves in a field of type
    ``kind`` (int, float or bool) for one config line (src/config.h:16-48):
    skip to the first digit, extract a value, skip to the next digit, extract
    again, and so on; the last value wins.  An extraction starts at a digit, so
    a sign is skipped like any other non-digit and an int field stops at a
    decimal point.  A float whose exponent has no digits ("1e") fails and
    leaves 0; a bool other than 0 or 1 leaves True and stops; a line without a
    number gives 0 (the reference leaves the field unset)."""
    val = kind(0)
    i = _skip_to_digit(line, 0)
    while i < len(line):
        if kind is float:
            m = _FLOAT_RE.match(line, i)
            j = m.end()
            if j < len(line) and line[j] in "eE":
                e = _EXP_RE.match(line, j)
                if e is None:  # "1e" / "1e+": the extraction fails and stores 0
                    return 0.0
                j = e.end()
            val = float(line[i:j])
        else:
            m = _INT_RE.match(line, i)
            j = m.end()
            v = int(line[i:j])
            if kind is bool:
                if v not in (0, 1):
                    return True
                val = bool(v)
            else:
                val = v
        i = _skip_to_digit(line, j)
    return val


def read_config(path: str) -> RefConfig:
    """config/config.txt as the reference's Config reads it: one field per line
    in file order, the last number on a line wins (src/config.h:16-58)."""
    kinds = {"overInterp": float, "rootLeng": float, "evalDirect": bool}
    cfg = RefConfig()
    with open(path) as f:
        lines = f.read().split("\n")
    for k, line in zip(_CFG_ORDER, lines):
        v = _last_number(line, kinds.get(k, int))
        setattr(cfg, k, int(v) if k == "evalDirect" else v)
    return cfg


def write_pwave(path: str, amplitude, pol, khat, wavenum):
    with open(path, "w") as f:
        f.write("%.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n" %
                (amplitude, *pol, *khat, wavenum))


def read_pwave(path: str):
    with open(path) as f:
        a = [float(x) for x in f.readline().split()]
    return {"amplitude": a[0], "pol": a[1:4], "khat": a[4:7], "wavenum": a[7]}


def write_dipoles(path: str, pos, mom):
    np.savetxt(path, np.hstack([pos, mom]), fmt="%.17g")


def read_dipoles(path: str):
    a = np.loadtxt(path, ndmin=2)
    return a[:, :3].copy(), a[:, 3:6].copy()


def write_rwg_dir(path: str, verts, faces, rwgs):
    os.makedirs(path, exist_ok=True)
    np.savetxt(os.path.join(path, "vertices.txt"), verts, fmt="%.17g")
    np.savetxt(os.path.join(path, "faces.txt"), faces, fmt="%d")
    np.savetxt(os.path.join(path, "rwgs.txt"), rwgs, fmt="%d")


def read_rwg_dir(path: str):
    v = np.loadtxt(os.path.join(path, "vertices.txt"), ndmin=2)
    f = np.loadtxt(os.path.join(path, "faces.txt"), ndmin=2, dtype=np.int64).astype(np.int32)
    r = np.loadtxt(os.path.join(path, "rwgs.txt"), ndmin=2, dtype=np.int64).astype(np.int32)
    return v, f, r
