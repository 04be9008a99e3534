"""Forest plans (mlfma_plan_create_forest): several scatterers in one plan,
every matvec stage launched once for all of them.  The forest's operator is
the block diagonal of the scatterers' own operators: its matvec, adjoint and
batched solves equal those of one plan per scatterer (bitwise where the
launches read the same data in the same order: the per-target M2L kernel is
forced, since the kernel choice per level depends on the level's totals)."""
import numpy as np
import pytest

from mlfma_amd import _capi as A
from mlfma_amd import configs as CF
from mlfma_amd import geometry as G
from mlfma_amd.plan import Plan
from tests import cases

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def small_forest(flags=0, precision=A.MLFMA_FP64):
    """Four scatterers with roots 1, 2, 2 and 4 wavelengths (leaf 0.25)."""
    out = []
    for kind, size, freq, L in [("sphere", 0.45, 4, 2), ("ellipsoid", 0.45, 5, 3), ("sphere", 0.8, 6, 3),
                                ("ellipsoid", 0.9, 7, 4)]:
        if kind == "sphere":
            v, f = G.icosphere(freq, size)
        else:
            v, f = G.ellipsoid(freq, (2 * size / 2 ** (1 / 3), size / 2 ** (1 / 3), size / 2 ** (1 / 3)))
        r = G.rwg_edges(f)
        geom = A.GeometryArrays(A.MLFMA_SRC_RWG, verts=v, faces=f, rwgs=r)
        kw = dict(root_leng=0.25 * 2 ** L, max_node_srcs=0, max_level=L, interp_order=4, digits=3, quad_prec=2,
                  flags=flags, precision=precision)
        out.append((geom, kw))
    return out


def test_forest_matvec_equals_separate_plans():
    flags = A.MLFMA_FLAG_FRESH_LOCALS | A.MLFMA_FLAG_M2L_TARGET
    sc = small_forest(flags)
    pw = cases.planewave_z()
    cfgs = [A.make_config(device=0, **kw) for _, kw in sc]
    F = Plan.forest([g for g, _ in sc], cfgs, pw)
    off = F.segments()
    assert len(off) == len(sc) + 1 and off[-1] == F.n
    sep = [Plan(g, pw, A.make_config(device=0, **kw)) for g, kw in sc]
    X = np.stack([G.complex_gaussian(F.n, 30 + r) for r in range(3)])
    Y = F.matvec(X)
    Ya = F.matvec_adjoint(X[0])
    for s, p in enumerate(sep):
        seg = slice(off[s], off[s + 1])
        ys = p.matvec(np.ascontiguousarray(X[:, seg]))
        say("scatterer", s, p.n, "forest vs own plan", cases.relerr(Y[:, seg], ys))
        assert np.array_equal(Y[:, seg], ys)
        assert np.array_equal(Ya[seg], p.matvec_adjoint(np.ascontiguousarray(X[0, seg])))
    # the excitation of the forest is the scatterers' excitations side by side
    V = F.voltage()
    for s, p in enumerate(sep):
        assert np.array_equal(V[off[s]:off[s + 1]], p.voltage())
    # without forcing the M2L kernel: equal to round-off
    F2 = Plan.forest([g for g, _ in sc], [A.make_config(device=0, **dict(kw, flags=A.MLFMA_FLAG_FRESH_LOCALS))
                                          for _, kw in sc], pw)
    assert cases.relerr(F2.matvec(X[0]), Y[0]) <= 1e-13


@pytest.mark.parametrize("method", [A.MLFMA_GMRES_RESTART, A.MLFMA_FGMRES_IR])
def test_forest_batched_solve_equals_separate(method):
    flags = A.MLFMA_FLAG_FRESH_LOCALS | A.MLFMA_FLAG_M2L_TARGET | A.MLFMA_FLAG_PHYSICAL
    prec = A.MLFMA_FP32 if method == A.MLFMA_FGMRES_IR else A.MLFMA_FP64
    sc = small_forest(flags, prec)
    pw = cases.planewave_z()
    F = Plan.forest([g for g, _ in sc], [A.make_config(device=0, **kw) for _, kw in sc], pw)
    off = F.segments()
    angles = CF.c4_angles()[:3]
    B = np.stack([-F.excitation(CF.planewave_xz(t)) for t in angles])
    opts = dict(restart=40, maxit=1500, tol=1e-6, method=method)
    Xf, rf = F.solve_batch(B, **opts)
    nseg = len(sc)
    assert len(rf) == len(angles) * nseg
    for s, (g, kw) in enumerate(sc):
        p = Plan(g, pw, A.make_config(device=0, **kw))
        seg = slice(off[s], off[s + 1])
        Xs, rs = p.solve_batch(np.ascontiguousarray(B[:, seg]), **opts)
        for a in range(len(angles)):
            lane = rf[a * nseg + s]
            say("method", method, "scatterer", s, "angle", a, lane["iterations"], rs[a]["iterations"],
                lane["resid_true"])
            assert lane["status"] == rs[a]["status"] and lane["iterations"] == rs[a]["iterations"]
            assert np.array_equal(Xf[a, seg], Xs[a])


def test_forest_rejects_mismatched_scatterers():
    sc = small_forest()
    pw = cases.planewave_z()
    cfgs = [A.make_config(device=-1, **kw) for _, kw in sc]
    cfgs[1].digits = 5
    with pytest.raises(Exception):
        Plan.forest([g for g, _ in sc], cfgs, pw)
    cfgs = [A.make_config(device=-1, **kw) for _, kw in sc]
    cfgs[2].root_leng = 3.0      # not the largest root over a power of two
    with pytest.raises(Exception):
        Plan.forest([g for g, _ in sc], cfgs, pw)
