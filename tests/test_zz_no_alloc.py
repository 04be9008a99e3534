"""mlfma_farfield_noalloc (the far field on device pointers with its partial
sums in a fixed per-device array, no allocation) against mlfma_farfield (the
same kernels with a library-allocated partial buffer): equal results on two
calls in a row, the second with a smaller input, and the size limit."""
import ctypes as C

import numpy as np
import pytest

from mlfma_amd import _capi as A
from mlfma_amd import geometry as G
from mlfma_amd.plan import Plan
from tests import cases

pytestmark = pytest.mark.gpu

LIMIT = 65536


def dirs(nd, seed):
    v = np.random.default_rng(seed).normal(size=(nd, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def noalloc(plan, I, d, E, stream):
    return plan.lib.mlfma_farfield_noalloc(plan._h, C.c_void_p(I.data_ptr()), C.c_void_p(d.data_ptr()), d.shape[0],
                                           C.c_void_p(E.data_ptr()), C.c_void_p(stream))


def test_noalloc_farfield_matches_farfield():
    import torch
    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream().cuda_stream
    calls = [("rwg_adaptive", 3000), ("dipole_fixed", 361)]   # the second call smaller
    for name, nd in calls:
        geom, kw = cases.make(name)
        p = Plan(geom, cases.planewave_z(), A.make_config(device=0, **kw))
        I = torch.from_numpy(G.complex_gaussian(geom.n, 5)).to(dev)
        d = torch.from_numpy(dirs(nd, nd)).to(dev)
        E_ref = p.farfield(I, d)                       # mlfma_farfield, device pointers
        E = torch.full((nd, 3), float("nan"), dtype=torch.complex128, device=dev)
        rc = noalloc(p, I, d, E, st)
        assert rc == 0, p.lib.mlfma_last_error().decode()
        torch.cuda.synchronize()
        e = cases.relerr(E.cpu().numpy(), E_ref.cpu().numpy())
        print(name, geom.n, "sources", nd, "directions: no-alloc vs mlfma_farfield", e, flush=True)
        assert torch.isfinite(E).all()
        assert e <= 1e-13
        # host-side check of the same far field through host buffers
        Eh = p.farfield(I.cpu().numpy(), d.cpu().numpy())
        assert cases.relerr(E.cpu().numpy(), Eh) <= 1e-13


def test_noalloc_farfield_rejects_too_many_directions():
    import torch
    geom, kw = cases.make("dipole_fixed")
    p = Plan(geom, cases.planewave_z(), A.make_config(device=0, **kw))
    I = torch.zeros(geom.n, dtype=torch.complex128, device="cuda:0")
    d = torch.zeros((LIMIT + 1, 3), dtype=torch.float64, device="cuda:0")
    d[:, 2] = 1.0
    E = torch.zeros((LIMIT + 1, 3), dtype=torch.complex128, device="cuda:0")
    rc = noalloc(p, I, d, E, torch.cuda.current_stream().cuda_stream)
    assert A.ERRORS[rc] == "ERR_ARG"
    assert b"65536" in p.lib.mlfma_last_error()
