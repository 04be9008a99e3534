"""Python mirror of the C-ABI (``include/mlfma_c.h``).

``Plan`` follows the reference's setup/matvec/solve contract
(src/mlfma.cpp:21-82): building a plan runs the whole setup; ``matvec`` is
``Solver::evalRvec`` (src/solver.cpp:41-57); ``solve`` is ``Solver::solve``
(src/solver.cpp:116-146); ``farfield`` generalises ``Node::printFarSols``
(src/fmm/node.cpp:147-178).  Every call goes to the HIP library; a missing
library raises."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as A


class MlfmaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{A.ERRORS.get(code, code)}: {msg}")
        self.code = code


def _check(lib, rc):
    if rc != 0:
        raise MlfmaError(rc, lib.mlfma_last_error().decode())


def _cvec(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    if n is not None and a.shape[0] != n:
        raise ValueError("vector length does not match the number of unknowns")
    return a


def _is_cuda(t) -> bool:
    try:
        import torch
    except Exception:  # torch absent
        return False
    return isinstance(t, torch.Tensor) and t.is_cuda


def _dev_vec(t, n):
    import torch
    if t.dtype != torch.complex128 or not t.is_contiguous() or t.numel() != n:
        raise ValueError("device vectors must be contiguous complex128 CUDA tensors of n entries")
    return t


def _dirs_tensor(dirs, device):
    import torch
    if isinstance(dirs, torch.Tensor):
        return dirs.to(device=device, dtype=torch.float64).reshape(-1, 3).contiguous()
    return torch.from_numpy(np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)).to(device)


class Plan:
    def __init__(self, geom: A.GeometryArrays, pw: A.PlaneWave | None = None, cfg: A.Config | None = None,
                 **cfg_kw):
        self.lib = A.load_library()
        self.geom = geom
        self.pw = pw or A.make_planewave()
        self.cfg = cfg or A.make_config(**cfg_kw)
        self._h = C.c_void_p()
        _check(self.lib, self.lib.mlfma_plan_create(C.byref(geom.struct), C.byref(self.pw), C.byref(self.cfg),
                                                    C.byref(self._h)))
        self.n = geom.n

    @classmethod
    def forest(cls, geoms, cfgs, pw=None):
        """One plan for several independent scatterers (mlfma_plan_create_forest):
        every matvec stage runs once for all of them; the unknowns of
        scatterer s follow those of s - 1 (segments())."""
        self = cls.__new__(cls)
        self.lib = A.load_library()
        self.geoms = list(geoms)
        self.geom = None
        self.pw = pw or A.make_planewave()
        cfgs = list(cfgs)
        if len(cfgs) != len(self.geoms):
            raise ValueError("one configuration per scatterer")
        self.cfg = cfgs[0]
        garr = (A.Geometry * len(self.geoms))(*[g.struct for g in self.geoms])
        carr = (A.Config * len(cfgs))(*cfgs)
        self._h = C.c_void_p()
        _check(self.lib, self.lib.mlfma_plan_create_forest(garr, carr, len(cfgs), C.byref(self.pw), C.byref(self._h)))
        self.n = sum(g.n for g in self.geoms)
        return self

    def segments(self) -> np.ndarray:
        """Unknown offsets of the plan's scatterers (count + 1 values)."""
        cnt = C.c_int32(0)
        _check(self.lib, self.lib.mlfma_plan_segments(self._h, None, 0, C.byref(cnt)))
        off = np.zeros(cnt.value + 1, dtype=np.int64)
        _check(self.lib, self.lib.mlfma_plan_segments(self._h, off.ctypes.data_as(C.POINTER(C.c_int64)), len(off),
                                                      C.byref(cnt)))
        return off

    def close(self):
        if self._h and self._h.value:
            self.lib.mlfma_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- info
    def info(self) -> dict:
        i = A.Info()
        _check(self.lib, self.lib.mlfma_plan_info(self._h, C.byref(i)))
        ml = i.max_level
        return {
            "n": i.n, "max_level": ml, "num_nodes": i.num_nodes, "num_leaves": i.num_leaves,
            "nth": list(i.nth[: ml + 1]), "nph": list(i.nph[: ml + 1]), "boxes": list(i.boxes[: ml + 1]),
            "m2l_pairs": list(i.m2l_pairs[: ml + 1]), "near_nnz": i.near_nnz,
            "radpat_entries": i.radpat_entries, "stale_boxes": i.stale_boxes,
            "truncated_l2l": i.truncated_l2l, "device_bytes": i.device_bytes,
        }

    def voltage(self) -> np.ndarray:
        v = np.zeros(self.n, dtype=np.complex128)
        _check(self.lib, self.lib.mlfma_voltage(self._h, v.ctypes.data_as(A.dptr)))
        return v

    # ---------------------------------------------------------------- matvec
    def matvec(self, x, y=None, accumulate=False, stream=None):
        """y = Z x (or y += Z x).  numpy arrays (n,) or (nrhs, n) go through host
        buffers; torch CUDA tensors (complex128, contiguous) are used in place."""
        return self._apply(self.lib.mlfma_matvec, x, y, accumulate, stream)

    def matvec_adjoint(self, x, y=None, accumulate=False, stream=None):
        """y = Z^H x (or y += Z^H x), same conventions as matvec."""
        return self._apply(self.lib.mlfma_matvec_adjoint, x, y, accumulate, stream)

    def _apply(self, fn, x, y, accumulate, stream):
        try:
            import torch
            is_t = isinstance(x, torch.Tensor)
        except Exception:  # torch absent
            is_t = False
        if is_t:
            if x.dtype != torch.complex128 or not x.is_cuda or not x.is_contiguous():
                raise ValueError("device matvec needs a contiguous complex128 CUDA tensor")
            nrhs = 1 if x.dim() == 1 else x.shape[0]
            if y is None:
                y = torch.empty_like(x)
            st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
            _check(self.lib, fn(self._h, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), nrhs,
                                1 if accumulate else 0, A.MLFMA_PTR_DEVICE, C.c_void_p(st)))
            return y
        x = _cvec(x)
        nrhs = 1 if x.ndim == 1 else x.shape[0]
        if x.shape[-1] != self.n:
            raise ValueError("vector length does not match the number of unknowns")
        if y is None:
            y = np.zeros_like(x)
        elif accumulate:
            y = _cvec(y)
        _check(self.lib, fn(self._h, C.c_void_p(x.ctypes.data), C.c_void_p(y.ctypes.data), nrhs,
                            1 if accumulate else 0, A.MLFMA_PTR_HOST, None))
        return y

    def precond(self, x, adjoint=False, kind=A.MLFMA_PRECOND_LEAF_BLOCKS):
        """y = P x (or P^H x): the leaf block-Jacobi preconditioner (host
        arrays (n,) or (nrhs, n))."""
        x = _cvec(x)
        nrhs = 1 if x.ndim == 1 else x.shape[0]
        y = np.zeros_like(x)
        _check(self.lib, self.lib.mlfma_precond_apply(self._h, kind, C.c_void_p(x.ctypes.data), C.c_void_p(y.ctypes.data),
                                                      nrhs, 1 if adjoint else 0, A.MLFMA_PTR_HOST, None))
        return y

    def matvec_direct(self, x):
        x = _cvec(x, self.n)
        y = np.zeros_like(x)
        _check(self.lib, self.lib.mlfma_matvec_direct(self._h, C.c_void_p(x.ctypes.data), C.c_void_p(y.ctypes.data),
                                                      A.MLFMA_PTR_HOST, None))
        return y

    def set_param(self, param: int, value: int) -> None:
        """Set one of the launch parameters the autotuner chooses
        (MLFMA_PARAM_*)."""
        _check(self.lib, self.lib.mlfma_plan_set_param(self._h, int(param), int(value)))

    def probe(self, group: str | int | None) -> None:
        """Bracket every launch of one launch group (a name of _capi.GROUPS or
        its number; None: no probe) with HIP events on the stream it runs
        on."""
        g = 0 if group is None else (A.GROUPS[group] if isinstance(group, str) else int(group))
        self.set_param(A.MLFMA_PARAM_PROBE_GROUP, g)

    def probe_read(self):
        """(milliseconds, launches) of the probed group since the last read."""
        ms, cnt = C.c_double(0), C.c_int64(0)
        _check(self.lib, self.lib.mlfma_probe_read(self._h, C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value

    def get_param(self, param: int) -> int:
        v = C.c_int32(0)
        _check(self.lib, self.lib.mlfma_plan_get_param(self._h, int(param), C.byref(v)))
        return v.value

    def set_flags(self, flags: int) -> None:
        """Replace the MLFMA_FLAG_* set used by later matvecs and solves."""
        _check(self.lib, self.lib.mlfma_plan_set_flags(self._h, int(flags)))
        self._flags = int(flags)

    @property
    def flags(self) -> int:
        """The MLFMA_FLAG_* set in effect (creation flags until set_flags)."""
        return getattr(self, "_flags", int(self.cfg.flags))

    def autotune(self, nrhs: int = 1):
        """Time candidate launch parameters for batches of nrhs right-hand
        sides and keep the fastest; returns (ms before, ms after)."""
        a, b = C.c_double(0), C.c_double(0)
        _check(self.lib, self.lib.mlfma_autotune(self._h, nrhs, C.byref(a), C.byref(b)))
        return a.value, b.value

    def tuning(self) -> list:
        """The launch parameters the autotuner chose (a list of ints)."""
        n = C.c_int32(0)
        _check(self.lib, self.lib.mlfma_autotune_record(self._h, None, 0, C.byref(n)))
        buf = (C.c_int32 * n.value)()
        _check(self.lib, self.lib.mlfma_autotune_record(self._h, buf, n.value, C.byref(n)))
        return list(buf)

    def apply_tuning(self, rec) -> None:
        """Apply a record from tuning() of a plan of the same workload."""
        buf = (C.c_int32 * len(rec))(*[int(v) for v in rec])
        _check(self.lib, self.lib.mlfma_autotune_apply(self._h, buf, len(rec)))

    def stage_times(self) -> dict:
        t = np.zeros(7)
        _check(self.lib, self.lib.mlfma_stage_times(self._h, t.ctypes.data_as(A.dptr)))
        return dict(zip(["S2M", "M2M", "M2L", "L2L", "L2T", "near", "total"], t.tolist()))

    def stage_times_detail(self) -> dict:
        """Device ms of each launch group of the last timed matvec."""
        t = np.zeros(7)
        cnt = C.c_int32(0)
        _check(self.lib, self.lib.mlfma_stage_times_detail(self._h, t.ctypes.data_as(A.dptr), 7, C.byref(cnt)))
        return dict(zip(["S2M", "M2M", "M2L", "L2L", "near_pairs", "near_sum", "L2T"], t.tolist()))

    # ---------------------------------------------------------------- solve
    def solve(self, b=None, method=A.MLFMA_GMRES_COMPAT, restart=50, maxit=500, tol=1e-6, history=False,
              adjoint=False, outer_restart=0, inner_tol=0.0, precond=0, shift=0j, deflate=0):
        """GMRES for Z x = b (b = None: the plane-wave excitation -V), or for
        Z^H x = b with adjoint=True.  MLFMA_FGMRES_IR (FP32 plans): flexible
        GMRES(outer_restart) in FP64 around inner FP32-far-field GMRES solves of
        at most `restart` iterations to a relative inner_tol.  deflate = k > 0
        (MLFMA_GMRES_RESTART, MLFMA_GMRES_IR): deflated restarts keeping k
        harmonic Ritz vectors (GMRES-DR).  Returns (x, report);
        report["iterations"] counts matvecs."""
        o = A.SolveOpts()
        o.method = method | (A.MLFMA_SOLVE_ADJOINT if adjoint else 0)
        o.restart = restart
        o.maxit = maxit
        o.use_voltage = 1 if b is None else 0
        o.tol = tol
        o.outer_restart = outer_restart
        o.inner_tol = inner_tol
        o.precond = precond
        o.shift_re, o.shift_im = complex(shift).real, complex(shift).imag
        o.deflate = deflate
        rep = A.SolveReport()
        hist = np.zeros(maxit)
        if _is_cuda(b):
            # device vectors: b and x stay on the device (the solve runs on the
            # plan's stream, so torch's queued work on b is finished first)
            import torch
            b = _dev_vec(b, self.n)
            torch.cuda.current_stream(b.device).synchronize()
            x = torch.empty_like(b)
            o.method = o.method | A.MLFMA_SOLVE_DEVICE
            _check(self.lib, self.lib.mlfma_solve(self._h, C.cast(C.c_void_p(b.data_ptr()), A.dptr),
                                                  C.cast(C.c_void_p(x.data_ptr()), A.dptr), C.byref(o),
                                                  C.byref(rep), hist.ctypes.data_as(A.dptr)))
        else:
            x = np.zeros(self.n, dtype=np.complex128)
            bp = None
            if b is not None:
                b = _cvec(b, self.n)
                bp = b.ctypes.data_as(A.dptr)
            _check(self.lib, self.lib.mlfma_solve(self._h, bp, x.ctypes.data_as(A.dptr), C.byref(o), C.byref(rep),
                                                  hist.ctypes.data_as(A.dptr)))
        report = {"iterations": rep.iterations, "status": A.ERRORS.get(rep.status, rep.status),
                  "resid_est": rep.resid_est, "resid_true": rep.resid_true, "seconds": rep.seconds}
        if history:
            report["history"] = hist[: rep.iterations].copy()
        return x, report

    def excitation(self, pw) -> np.ndarray:
        """V of the sources under another plane wave (same wavenumber)."""
        v = np.zeros(self.n, dtype=np.complex128)
        _check(self.lib, self.lib.mlfma_excitation(self._h, C.byref(pw), v.ctypes.data_as(A.dptr)))
        return v

    def solve_batch(self, B, restart=50, maxit=500, tol=1e-6, history=False, method=A.MLFMA_GMRES_RESTART,
                    outer_restart=0, inner_tol=0.0, precond=0, shift=0j, deflate=0):
        """GMRES(restart) on the rows of B (nrhs, n) at once, one batched
        matvec per step (method MLFMA_FGMRES_IR on an FP32 plan: FP64
        flexible GMRES around batched FP32-far-field GMRES; MLFMA_GMRES_IR:
        FP64 refinement around it).  Returns (X, reports): one report per
        right-hand side and scatterer (lane r * segments + s; one scatterer
        unless the plan is a forest)."""
        B = np.ascontiguousarray(B, dtype=np.complex128)
        if B.ndim != 2 or B.shape[1] != self.n:
            raise ValueError("B must be (nrhs, n)")
        nrhs = B.shape[0]
        nseg = len(self.segments()) - 1
        o = A.SolveOpts()
        o.method = method
        o.restart = restart
        o.maxit = maxit
        o.use_voltage = 0
        o.tol = tol
        o.outer_restart = outer_restart
        o.inner_tol = inner_tol
        o.precond = precond
        o.shift_re, o.shift_im = complex(shift).real, complex(shift).imag
        o.deflate = deflate  # deflated restarts (GMRES-DR) per lane; not with MLFMA_FGMRES_IR
        reps = (A.SolveReport * (nrhs * nseg))()
        X = np.zeros_like(B)
        hist = np.zeros((nrhs * nseg, maxit))
        _check(self.lib, self.lib.mlfma_solve_batch(self._h, B.ctypes.data_as(A.dptr), X.ctypes.data_as(A.dptr),
                                                    nrhs, C.byref(o), reps, hist.ctypes.data_as(A.dptr)))
        out = []
        for r in range(nrhs * nseg):
            rep = reps[r]
            d = {"iterations": rep.iterations, "status": A.ERRORS.get(rep.status, rep.status),
                 "resid_est": rep.resid_est, "resid_true": rep.resid_true, "seconds": rep.seconds}
            if history:
                d["history"] = hist[r, : rep.iterations].copy()
            out.append(d)
        return X, out

    # ---------------------------------------------------------------- far field
    def farfield(self, I, dirs, stream=None):
        """E = F I at the unit directions dirs (ndir x 3).  numpy arrays go
        through host buffers; with I a CUDA complex128 tensor the directions
        and the result are CUDA tensors too and stay on the device."""
        if _is_cuda(I):
            import torch
            d = _dirs_tensor(dirs, I.device)
            E = torch.empty((d.shape[0], 3), dtype=torch.complex128, device=I.device)
            st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
            _check(self.lib, self.lib.mlfma_farfield(self._h, C.c_void_p(_dev_vec(I, self.n).data_ptr()),
                                                     C.c_void_p(d.data_ptr()), d.shape[0], C.c_void_p(E.data_ptr()),
                                                     A.MLFMA_PTR_DEVICE, C.c_void_p(st)))
            return E
        I = _cvec(I, self.n)
        dirs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
        E = np.zeros((len(dirs), 3), dtype=np.complex128)
        _check(self.lib, self.lib.mlfma_farfield(self._h, C.c_void_p(I.ctypes.data), C.c_void_p(dirs.ctypes.data),
                                                 len(dirs), C.c_void_p(E.ctypes.data), A.MLFMA_PTR_HOST, None))
        return E

    def farfield_adjoint(self, E, dirs, stream=None):
        """I = F^H E: the adjoint of farfield for the same directions (CUDA
        tensors stay on the device, as in farfield)."""
        if _is_cuda(E):
            import torch
            d = _dirs_tensor(dirs, E.device)
            Et = E.to(torch.complex128).reshape(d.shape[0], 3).contiguous()
            I = torch.empty(self.n, dtype=torch.complex128, device=E.device)
            st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
            _check(self.lib, self.lib.mlfma_farfield_adjoint(self._h, C.c_void_p(Et.data_ptr()),
                                                             C.c_void_p(d.data_ptr()), d.shape[0],
                                                             C.c_void_p(I.data_ptr()), A.MLFMA_PTR_DEVICE,
                                                             C.c_void_p(st)))
            return I
        dirs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
        E = np.ascontiguousarray(E, dtype=np.complex128).reshape(len(dirs), 3)
        I = np.zeros(self.n, dtype=np.complex128)
        _check(self.lib, self.lib.mlfma_farfield_adjoint(self._h, C.c_void_p(E.ctypes.data),
                                                         C.c_void_p(dirs.ctypes.data), len(dirs),
                                                         C.c_void_p(I.ctypes.data), A.MLFMA_PTR_HOST, None))
        return I

    def position_gradient(self, I, lam):
        """(n, 3) array -2 Re(lam^H (dZ/dr_i^a) I) for a dipole plan: the
        operator's sensitivity to each source position (tree held fixed)."""
        I = _cvec(I, self.n)
        lam = _cvec(lam, self.n)
        g = np.zeros((self.n, 3))
        _check(self.lib, self.lib.mlfma_position_gradient(self._h, I.ctypes.data_as(A.dptr),
                                                          lam.ctypes.data_as(A.dptr), g.ctypes.data_as(A.dptr)))
        return g

    # ---------------------------------------------------------------- introspection
    def level_boxes(self, level):
        nb = C.c_int64()
        _check(self.lib, self.lib.mlfma_level_boxes(self._h, level, C.byref(nb), None, None, None, 0))
        co = np.zeros((nb.value, 3), np.int32)
        lf = np.zeros(nb.value, np.int32)
        ns = np.zeros(nb.value, np.int32)
        _check(self.lib, self.lib.mlfma_level_boxes(self._h, level, C.byref(nb), co.ctypes.data_as(A.iptr),
                                                    lf.ctypes.data_as(A.iptr), ns.ctypes.data_as(A.iptr), nb.value))
        return co, lf, ns

    def level_pairs(self, level, kind):
        npr = C.c_int64()
        _check(self.lib, self.lib.mlfma_level_pairs(self._h, level, kind, C.byref(npr), None, None, 0))
        pr = np.zeros((max(npr.value, 1), 2), np.int32)
        sl = np.zeros(max(npr.value, 1), np.int32)
        _check(self.lib, self.lib.mlfma_level_pairs(self._h, level, kind, C.byref(npr), pr.ctypes.data_as(A.iptr),
                                                    sl.ctypes.data_as(A.iptr), pr.size))
        return pr[: npr.value], sl[: npr.value]

    def coeffs(self, level, which):
        info = self.info()
        nd = info["nth"][level] * info["nph"][level]
        nb = len(self.level_boxes(level)[0])
        out = np.zeros((nb, nd, 2), dtype=np.complex128)
        _check(self.lib, self.lib.mlfma_get_coeffs(self._h, level, which, out.ctypes.data_as(A.dptr), out.size * 2))
        return out

    def translation(self, level):
        info = self.info()
        nd = info["nth"][level] * info["nph"][level]
        T = np.zeros((316, nd), dtype=np.complex128)
        _check(self.lib, self.lib.mlfma_get_translation(self._h, level, T.ctypes.data_as(A.dptr), T.size * 2))
        return T
