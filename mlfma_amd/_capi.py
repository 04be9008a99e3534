"""ctypes mirror of ``include/mlfma_c.h``.

The structure layouts are the C-ABI's; ``load_library`` finds the in-tree
``libmlfma_amd.so`` built by ``mlfma_amd/build.py`` and raises if it is absent
(there is no fallback implementation)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmlfma_amd.so")

MLFMA_SRC_DIPOLE = 0
MLFMA_SRC_RWG = 1
MLFMA_FP64 = 0
MLFMA_FP32 = 1
MLFMA_FLAG_FRESH_LOCALS = 1
MLFMA_FLAG_STAGE_TIMES = 2
MLFMA_FLAG_M2L_GROUP = 4
MLFMA_FLAG_M2L_TARGET = 8
MLFMA_FLAG_ASYNC = 16
MLFMA_FLAG_PHYSICAL = 32
MLFMA_FLAG_GRAPH = 64
MLFMA_FLAG_ONE_STREAM = 128
MLFMA_PTR_HOST = 0
MLFMA_PTR_DEVICE = 1
MLFMA_GMRES_COMPAT = 0
MLFMA_GMRES_RESTART = 1
MLFMA_FGMRES_IR = 2
MLFMA_GMRES_IR = 3
MLFMA_PRECOND_NONE = 0
MLFMA_PRECOND_LEAF_BLOCKS = 1
MLFMA_SOLVE_ADJOINT = 0x100
MLFMA_SOLVE_DEVICE = 0x200
MLFMA_PARAM_NEAR_WINDOW = 1
MLFMA_PARAM_NEAR_OVERLAP = 2
MLFMA_PARAM_M2L_SPLIT = 3
MLFMA_PARAM_SOLVE_IN_ORDER = 4
MLFMA_PARAM_PROBE_GROUP = 5
MLFMA_PARAM_NEAR_BLOCKS = 6
MLFMA_PARAM_TREE_WALK = 7
MLFMA_PARAM_TREE_STALLS = 8
MLFMA_PARAM_NEAR_SPLIT = 9
GROUPS = {"S2M": 1, "M2M": 2, "M2L_TARGET": 3, "M2L_SIBLING": 4, "M2L_NEAR": 5, "L2L": 6, "NEAR": 7, "L2T": 8,
          "TREE_UP": 9, "M2L_RUNS": 10}

ERRORS = {0: "OK", -1: "ERR_ARG", -2: "ERR_OOM", -3: "ERR_HIP", -4: "ERR_BREAKDOWN",
          -5: "ERR_NONFINITE", -6: "ERR_STATE", -7: "ERR_GEOMETRY", -8: "ERR_NOT_CONVERGED"}
MLFMA_ERR_BREAKDOWN = -4
MLFMA_ERR_NOT_CONVERGED = -8

dptr = C.POINTER(C.c_double)
iptr = C.POINTER(C.c_int32)


class Geometry(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("n", C.c_int64),
                ("dip_pos", dptr), ("dip_mom", dptr),
                ("nverts", C.c_int64), ("verts", dptr),
                ("ntris", C.c_int64), ("tris", iptr), ("rwgs", iptr)]


class PlaneWave(C.Structure):
    _fields_ = [("amplitude", C.c_double), ("pol", C.c_double * 3), ("khat", C.c_double * 3),
                ("wavenum", C.c_double)]


class Config(C.Structure):
    _fields_ = [("digits", C.c_int32), ("interp_order", C.c_int32), ("over_interp", C.c_double),
                ("root_leng", C.c_double), ("max_node_srcs", C.c_int32), ("max_level", C.c_int32),
                ("quad_prec", C.c_int32), ("device", C.c_int32), ("precision", C.c_int32),
                ("flags", C.c_uint32)]


class Info(C.Structure):
    _fields_ = [("n", C.c_int64), ("max_level", C.c_int32), ("num_levels_m2l", C.c_int32),
                ("num_nodes", C.c_int64), ("num_leaves", C.c_int64),
                ("nth", C.c_int64 * 16), ("nph", C.c_int64 * 16), ("boxes", C.c_int64 * 16),
                ("m2l_pairs", C.c_int64 * 16), ("near_nnz", C.c_int64),
                ("radpat_entries", C.c_int64), ("stale_boxes", C.c_int64),
                ("truncated_l2l", C.c_int64), ("psi_merged", C.c_int64),
                ("device_bytes", C.c_int64)]


class SolveOpts(C.Structure):
    _fields_ = [("method", C.c_int32), ("restart", C.c_int32), ("maxit", C.c_int32),
                ("use_voltage", C.c_int32), ("tol", C.c_double), ("outer_restart", C.c_int32),
                ("precond", C.c_int32), ("inner_tol", C.c_double),
                ("shift_re", C.c_double), ("shift_im", C.c_double),
                ("deflate", C.c_int32)]


class SolveReport(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("status", C.c_int32), ("resid_est", C.c_double),
                ("resid_true", C.c_double), ("seconds", C.c_double)]


# exported symbols and their signatures (kept in step with include/mlfma_c.h)
SIGNATURES = {
    "mlfma_plan_create": (C.c_int, [C.POINTER(Geometry), C.POINTER(PlaneWave), C.POINTER(Config),
                                    C.POINTER(C.c_void_p)]),
    "mlfma_plan_destroy": (C.c_int, [C.c_void_p]),
    "mlfma_plan_create_forest": (C.c_int, [C.POINTER(Geometry), C.POINTER(Config), C.c_int32, C.POINTER(PlaneWave),
                                           C.POINTER(C.c_void_p)]),
    "mlfma_plan_segments": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.c_int32, C.POINTER(C.c_int32)]),
    "mlfma_plan_info": (C.c_int, [C.c_void_p, C.POINTER(Info)]),
    "mlfma_voltage": (C.c_int, [C.c_void_p, dptr]),
    "mlfma_excitation": (C.c_int, [C.c_void_p, C.c_void_p, dptr]),
    "mlfma_matvec": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32,
                               C.c_int32, C.c_void_p]),
    "mlfma_matvec_adjoint": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32,
                                       C.c_int32, C.c_void_p]),
    "mlfma_matvec_direct": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "mlfma_solve": (C.c_int, [C.c_void_p, dptr, dptr, C.POINTER(SolveOpts), C.POINTER(SolveReport),
                              dptr]),
    "mlfma_solve_batch": (C.c_int, [C.c_void_p, dptr, dptr, C.c_int32, C.c_void_p, C.c_void_p, dptr]),
    "mlfma_farfield": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int32,
                                 C.c_void_p]),
    "mlfma_farfield_adjoint": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int32,
                                         C.c_void_p]),
    "mlfma_farfield_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "mlfma_position_gradient": (C.c_int, [C.c_void_p, dptr, dptr, dptr]),
    "mlfma_stage_times": (C.c_int, [C.c_void_p, dptr]),
    "mlfma_plan_set_flags": (C.c_int, [C.c_void_p, C.c_uint32]),
    "mlfma_plan_set_param": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "mlfma_plan_get_param": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]),
    "mlfma_probe_read": (C.c_int, [C.c_void_p, dptr, C.POINTER(C.c_int64)]),
    "mlfma_precond_apply": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                      C.c_void_p]),
    "mlfma_autotune_record": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_int32)]),
    "mlfma_autotune_apply": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.c_int32]),
    "mlfma_autotune": (C.c_int, [C.c_void_p, C.c_int32, dptr, dptr]),
    "mlfma_stage_times_detail": (C.c_int, [C.c_void_p, dptr, C.c_int32, C.POINTER(C.c_int32)]),
    "mlfma_level_boxes": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int64), iptr, iptr, iptr,
                                    C.c_int64]),
    "mlfma_level_pairs": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_int64), iptr,
                                    iptr, C.c_int64]),
    "mlfma_get_coeffs": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, dptr, C.c_int64]),
    "mlfma_get_translation": (C.c_int, [C.c_void_p, C.c_int32, dptr, C.c_int64]),
    "mlfma_last_error": (C.c_char_p, []),
}

_lib = None


def _bind_one_hip_runtime():
    """torch ships its own libamdhip64 (same SONAME as /opt/rocm's).  Whichever
    is loaded first is the one the library binds to; loading torch first keeps
    one HIP runtime in the process, so torch tensors and the library can share
    device memory and streams."""
    try:
        import torch  # noqa: F401
    except Exception:
        pass


def load_library(path: str | None = None):
    """Load the HIP library; raise if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    _bind_one_hip_runtime()
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
                           " (or python mlfma_amd/build.py); there is no CPU fallback")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def as_dptr(a: np.ndarray):
    return a.ctypes.data_as(dptr)


def as_iptr(a: np.ndarray):
    return a.ctypes.data_as(iptr)


class GeometryArrays:
    """Owns the numpy arrays behind a ``Geometry`` struct."""

    def __init__(self, kind, *, pos=None, mom=None, verts=None, faces=None, rwgs=None):
        self.kind = kind
        g = Geometry()
        g.kind = kind
        if kind == MLFMA_SRC_DIPOLE:
            self.pos = np.ascontiguousarray(pos, dtype=np.float64)
            self.mom = np.ascontiguousarray(mom, dtype=np.float64)
            g.n = len(self.pos)
            g.dip_pos = as_dptr(self.pos)
            g.dip_mom = as_dptr(self.mom)
        else:
            self.verts = np.ascontiguousarray(verts, dtype=np.float64)
            self.faces = np.ascontiguousarray(faces, dtype=np.int32)
            self.rwgs = np.ascontiguousarray(rwgs, dtype=np.int32)
            g.n = len(self.rwgs)
            g.nverts = len(self.verts)
            g.verts = as_dptr(self.verts)
            g.ntris = len(self.faces)
            g.tris = as_iptr(self.faces)
            g.rwgs = as_iptr(self.rwgs)
        self.struct = g

    @property
    def n(self):
        return self.struct.n


def make_planewave(amplitude=1.0, pol=(1.0, 0.0, 0.0), khat=(0.0, 0.0, 1.0), wavenum=2 * np.pi):
    p = PlaneWave()
    p.amplitude = amplitude
    for i in range(3):
        p.pol[i] = pol[i]
        p.khat[i] = khat[i]
    p.wavenum = wavenum
    return p


def make_config(*, digits=3, interp_order=6, over_interp=4.0, root_leng=4.0, max_node_srcs=16,
                max_level=-1, quad_prec=2, device=0, precision=MLFMA_FP64, flags=0):
    c = Config()
    c.digits = digits
    c.interp_order = interp_order
    c.over_interp = over_interp
    c.root_leng = root_leng
    c.max_node_srcs = max_node_srcs
    c.max_level = max_level
    c.quad_prec = quad_prec
    c.device = device
    c.precision = precision
    c.flags = flags
    return c
