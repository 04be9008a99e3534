"""Several scatterers solved concurrently (mlfma_amd.batch.solve_many): the
same results as one after another, bitwise (the plans share nothing)."""
import numpy as np
import pytest

from mlfma_amd import _capi as A
from mlfma_amd import geometry as G
from mlfma_amd.batch import solve_many
from mlfma_amd.plan import Plan

pytestmark = pytest.mark.gpu


def scatterers():
    out = []
    rng = np.random.default_rng(1)
    for i in range(4):
        r = float(rng.uniform(0.3, 0.6))
        if i % 2:
            v, f = G.ellipsoid(5, (2 * r, r, r))
        else:
            v, f = G.icosphere(5, r)
        rw = G.rwg_edges(f)
        geom = A.GeometryArrays(A.MLFMA_SRC_RWG, verts=v, faces=f, rwgs=rw)
        kw = dict(root_leng=4.0 * r + 0.5, max_node_srcs=0, max_level=2, interp_order=4, digits=3,
                  flags=A.MLFMA_FLAG_PHYSICAL | A.MLFMA_FLAG_FRESH_LOCALS)
        p = Plan(geom, A.make_planewave(1.0, (1.0, 0.0, 0.0), (0.0, 0.0, 1.0), 2 * np.pi),
                 A.make_config(device=0, **kw))
        ang = np.radians(np.arange(3) * 60.0)
        B = np.stack([-p.excitation(A.make_planewave(1.0, (np.cos(t), 0.0, -np.sin(t)),
                                                     (np.sin(t), 0.0, np.cos(t)), 2 * np.pi)) for t in ang])
        out.append((p, B))
    return out


def test_concurrent_scatterers_match_sequential():
    sc = scatterers()
    plans, rhs = [s[0] for s in sc], [s[1] for s in sc]
    seq = [p.solve_batch(b, restart=40, maxit=1500, tol=1e-7) for p, b in sc]
    par = solve_many(plans, rhs, restart=40, maxit=1500, tol=1e-7)
    for (xs, rs), (xp, rp) in zip(seq, par):
        assert all(r["status"] == "OK" for r in rp), rp
        assert np.array_equal(xs, xp)
        assert [r["iterations"] for r in rs] == [r["iterations"] for r in rp]
