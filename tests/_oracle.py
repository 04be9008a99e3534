"""ctypes binding of the CPU oracle (oracle/, test infrastructure only).

Only tests, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg load
this: the oracle is the checker, never the thing measured or shipped."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from mlfma_amd import _capi as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "_build", "liboracle.so")
# the same restatement compiled with FMA contraction: only used to measure how
# much the reference's own floating-point result moves under another valid
# compilation
LIB_FMA = os.path.join(ORACLE_DIR, "_build", "liboracle_fma.so")

_libs = {}


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib(variant=""):
    path = LIB_FMA if variant == "fma" else LIB
    if path not in _libs:
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        vp = C.c_void_p
        d = A.dptr
        i32 = A.iptr
        L.orc_last_error.restype = C.c_char_p
        L.orc_create.argtypes = [C.POINTER(A.Geometry), C.POINTER(A.PlaneWave), C.POINTER(A.Config),
                                 C.c_int, C.c_int, C.POINTER(vp)]
        L.orc_destroy.argtypes = [vp]
        L.orc_info.argtypes = [vp, C.POINTER(C.c_int64)]
        L.orc_voltage.argtypes = [vp, d]
        L.orc_matvec.argtypes = [vp, d, d]
        L.orc_dense_matvec.argtypes = [vp, d, d]
        L.orc_solve.argtypes = [vp, C.c_int, C.c_double, d, d, C.POINTER(C.c_int)]
        L.orc_farfield.argtypes = [vp, d, C.c_int, d, d]
        L.orc_solve_record.argtypes = [vp, C.c_int, C.c_double, d, d, C.POINTER(C.c_int), d, d, d, d]
        L.orc_angles.argtypes = [vp, C.c_int, C.POINTER(C.c_int), d, d, d]
        L.orc_translation.argtypes = [vp, C.c_int, d, C.POINTER(C.c_int64)]
        L.orc_alpha.argtypes = [vp, C.c_int, d]
        L.orc_interp.argtypes = [vp, C.c_int, C.c_int, d, i32]
        L.orc_apply_interp.argtypes = [vp, C.c_int, C.c_int, d, d]
        L.orc_nodes.argtypes = [vp, i32, C.c_int64]
        L.orc_pairs.argtypes = [vp, C.c_int, i32, C.c_int64]
        L.orc_pairs.restype = C.c_int64
        L.orc_coeffs.argtypes = [vp, C.c_int, C.c_int, d, C.c_int64]
        L.orc_legendre.argtypes = [C.c_double, C.c_int, d]
        L.orc_legendre.restype = C.c_double
        L.orc_hankel.argtypes = [C.c_double, C.c_int, d]
        L.orc_gauss_legendre.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, d, d]
        L.orc_near_gl_idx.argtypes = [C.c_double, C.c_int, C.c_double, C.c_double]
        L.orc_lagrange.argtypes = [C.c_double, d, C.c_int, C.c_int]
        L.orc_lagrange.restype = C.c_double
        L.orc_inode_distances.argtypes = [d]
        L.orc_inode_directions.argtypes = [d]
        L.orc_inode_distvecs.argtypes = [d]
        L.orc_dyadic.argtypes = [d, C.c_double, d]
        L.orc_givens.argtypes = [d, d, d]
        L.orc_rwg_ipw.argtypes = [vp, C.c_int, d, C.c_int, d]
        L.orc_entry.argtypes = [vp, C.c_int, C.c_int, d]
        L.orc_radpat.argtypes = [vp, C.c_int, d, d, d]
        _libs[path] = L
    return _libs[path]


def _c(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _d(a):
    return a.ctypes.data_as(A.dptr)


class Oracle:
    """One restated reference model (FMM tree, or direct=True for evalDirect)."""

    def __init__(self, geom: A.GeometryArrays, pw: A.PlaneWave, cfg: A.Config, direct=False,
                 threads=None, variant=""):
        self.L = lib(variant)
        self.geom = geom
        self.h = C.c_void_p()
        nt = threads or min(8, os.cpu_count() or 1)
        rc = self.L.orc_create(C.byref(geom.struct), C.byref(pw), C.byref(cfg), 1 if direct else 0,
                               nt, C.byref(self.h))
        if rc != 0:
            raise RuntimeError("oracle create failed: " + self.L.orc_last_error().decode())
        self.n = geom.n
        self.root_leng = cfg.root_leng

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            self.L.orc_destroy(self.h)
            self.h = C.c_void_p()

    def info(self):
        o = np.zeros(6, dtype=np.int64)
        self.L.orc_info(self.h, o.ctypes.data_as(C.POINTER(C.c_int64)))
        return dict(zip(["n", "max_level", "nodes", "leaves", "near_pairs", "nonnear_pairs"], o.tolist()))

    def voltage(self):
        v = np.zeros(self.n, dtype=np.complex128)
        self.L.orc_voltage(self.h, _d(v))
        return v

    def matvec(self, x):
        x = _c(x)
        y = np.zeros(self.n, dtype=np.complex128)
        rc = self.L.orc_matvec(self.h, _d(x), _d(y))
        if rc != 0:
            raise RuntimeError(self.L.orc_last_error().decode())
        return y

    def dense_matvec(self, x):
        x = _c(x)
        y = np.zeros(self.n, dtype=np.complex128)
        self.L.orc_dense_matvec(self.h, _d(x), _d(y))
        return y

    def solve(self, maxit=500, tol=1e-6):
        x = np.zeros(self.n, dtype=np.complex128)
        res = np.zeros(maxit + 1)
        it = C.c_int(0)
        rc = self.L.orc_solve(self.h, maxit, tol, _d(x), _d(res), C.byref(it))
        return x, res[: it.value], it.value, rc

    def solve_record(self, maxit, tol=1e-6):
        n = self.n
        x = np.zeros(n, dtype=np.complex128)
        res = np.zeros(maxit + 1)
        it = C.c_int(0)
        H = np.zeros((maxit, maxit + 1), dtype=np.complex128)   # column j = H[j]
        Q = np.zeros((maxit + 1, n), dtype=np.complex128)
        y = np.zeros(maxit, dtype=np.complex128)
        g0 = np.zeros(1)
        rc = self.L.orc_solve_record(self.h, maxit, tol, _d(x), _d(res), C.byref(it), _d(H), _d(Q), _d(y), _d(g0))
        k = it.value
        return {"x": x, "resid": res[:k], "iters": k, "rc": rc, "H": H[:k, : k + 1].T.copy(),
                "Q": Q[: k + 1].T.copy(), "y": y[:k], "g0": g0[0]}

    def farfield(self, I, dirs):
        I = _c(I)
        dirs = np.ascontiguousarray(dirs, dtype=np.float64)
        E = np.zeros((len(dirs), 3), dtype=np.complex128)
        self.L.orc_farfield(self.h, _d(I), len(dirs), _d(dirs), _d(E))
        return E

    def angles(self, level):
        s = (C.c_int * 3)()
        self.L.orc_angles(self.h, level, s, None, None, None)
        th = np.zeros(s[0]); w = np.zeros(s[0]); ph = np.zeros(s[1])
        self.L.orc_angles(self.h, level, s, _d(th), _d(w), _d(ph))
        return {"thetas": th, "weights": w, "phis": ph, "L": s[2]}

    def translation(self, level):
        a = self.angles(level)
        nd = len(a["thetas"]) * len(a["phis"])
        T = np.zeros((316, nd), dtype=np.complex128)
        ex = np.zeros(3, dtype=np.int64)
        rc = self.L.orc_translation(self.h, level, _d(T), ex.ctypes.data_as(C.POINTER(C.c_int64)))
        if rc != 0:
            raise ValueError("no translation table at level %d" % level)
        return T, {"nps": int(ex[0]), "psi_keys": int(ex[1]), "psi_merged": int(ex[2])}

    def alpha(self, level):
        _, ex = self.translation(level)
        A_ = np.zeros((15, ex["nps"]), dtype=np.complex128)
        self.L.orc_alpha(self.h, level, _d(A_))
        return A_

    def interp(self, level, which, order):
        a = self.angles(level)
        cnt = len(a["thetas"]) if which == 0 else len(a["phis"])
        co = np.zeros((cnt, 2 * order))
        ni = np.zeros(cnt, dtype=np.int32)
        self.L.orc_interp(self.h, level, which, _d(co), ni.ctypes.data_as(A.iptr))
        return co, ni

    def apply_interp(self, level, direction, field_in, out=None):
        """direction 0: level+1 -> level (interp), 1: level -> level+1 (anterp);
        fields are (ndir, 2) complex."""
        fi = _c(field_in)
        tgt = level if direction == 0 else level + 1
        a = self.angles(tgt)
        nd = len(a["thetas"]) * len(a["phis"])
        fo = np.zeros((nd, 2), dtype=np.complex128) if out is None else _c(out).copy()
        self.L.orc_apply_interp(self.h, level, direction, _d(fi), _d(fo))
        return fo

    def nodes(self):
        nn = self.info()["nodes"]
        o = np.zeros((nn, 7), dtype=np.int32)
        self.L.orc_nodes(self.h, o.ctypes.data_as(A.iptr), o.size)
        return o

    def pairs(self, kind):
        n = self.L.orc_pairs(self.h, kind, None, 0)
        o = np.zeros((max(n, 1), 2), dtype=np.int32)
        self.L.orc_pairs(self.h, kind, o.ctypes.data_as(A.iptr), o.size)
        return o[:n]

    def coeffs(self, level, which):
        a = self.angles(level)
        nd = len(a["thetas"]) * len(a["phis"])
        nodes = self.nodes()
        nl = int((nodes[:, 0] == level).sum())
        o = np.zeros((nl, nd, 2), dtype=np.complex128)
        self.L.orc_coeffs(self.h, level, which, _d(o), o.size * 2)
        return o

    def entry(self, i, j):
        o = np.zeros(2)
        self.L.orc_entry(self.h, i, j, _d(o))
        return o[0] + 1j * o[1]

    def rwg_ipw(self, i, kvec, numeric=False):
        o = np.zeros(6)
        kv = np.ascontiguousarray(kvec, dtype=np.float64)
        self.L.orc_rwg_ipw(self.h, i, _d(kv), 1 if numeric else 0, _d(o))
        return o[0::2] + 1j * o[1::2]

    def radpat(self, i, X, khat):
        o = np.zeros(4)
        X = np.ascontiguousarray(X, dtype=np.float64)
        kh = np.ascontiguousarray(khat, dtype=np.float64)
        self.L.orc_radpat(self.h, i, _d(X), _d(kh), _d(o))
        return np.array([o[0] + 1j * o[1], o[2] + 1j * o[3]])


# ---- math helpers
def legendre(x, n):
    dp = np.zeros(1)
    p = lib().orc_legendre(x, n, _d(dp))
    return p, dp[0]


def hankel(x, n):
    o = np.zeros(2)
    lib().orc_hankel(x, n, _d(o))
    return o[0] + 1j * o[1]


def gauss_legendre(l, a=-1.0, b=1.0, eps=1e-9):
    nodes = np.zeros(l); w = np.zeros(l)
    lib().orc_gauss_legendre(l, eps, a, b, _d(nodes), _d(w))
    return nodes, w


def near_gl_idx(x, m, a=-1.0, b=1.0):
    return lib().orc_near_gl_idx(x, m, a, b)


def lagrange(x, xs, k):
    xs = np.ascontiguousarray(xs, dtype=np.float64)
    return lib().orc_lagrange(x, _d(xs), len(xs), k)


def inode_distances():
    o = np.zeros(64)
    n = lib().orc_inode_distances(_d(o))
    return o[:n]


def inode_directions():
    o = np.zeros(3 * 400)
    n = lib().orc_inode_directions(_d(o))
    return o[: 3 * n].reshape(n, 3)


def inode_distvecs():
    o = np.zeros(3 * 400)
    n = lib().orc_inode_distvecs(_d(o))
    return o[: 3 * n].reshape(n, 3)


def dyadic(X, k):
    X = np.ascontiguousarray(X, dtype=np.float64)
    o = np.zeros(18)
    lib().orc_dyadic(_d(X), k, _d(o))
    return (o[0::2] + 1j * o[1::2]).reshape(3, 3)


def givens(z, w):
    zz = np.array([z.real, z.imag]); ww = np.array([w.real, w.imag])
    o = np.zeros(4)
    lib().orc_givens(_d(zz), _d(ww), _d(o))
    return o[0] + 1j * o[1], o[2] + 1j * o[3]
