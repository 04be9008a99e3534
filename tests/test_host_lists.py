"""The library's host-side tree and interaction lists (geometric construction
from integer box coordinates) equal the oracle's restatement of the
reference's pointer-walking rules (src/fmm/nbor.cpp, src/fmm/node.cpp:61-99,
src/fmm/leaf.cpp:70-80) on adaptive and fixed-depth trees.  CPU only: the
library plan is built host-only (device = -1)."""
import numpy as np
import pytest

from mlfma_amd import _capi as A
from mlfma_amd import geometry as G
from mlfma_amd.plan import Plan
from tests import _oracle as O
from tests import cases


def product_structure(plan):
    info = plan.info()
    boxes = {lv: plan.level_boxes(lv) for lv in range(info["max_level"] + 1)}
    nodes = set()
    for lv, (co, lf, ns) in boxes.items():
        for i in range(len(co)):
            nodes.add((lv, *co[i], int(lf[i]), int(ns[i])))
    pairs = {}
    for kind in range(3):
        s = set()
        for lv in boxes:
            pr, sl = plan.level_pairs(lv, kind)
            for q in range(len(pr)):
                a = (lv, *boxes[lv][0][pr[q, 0]])
                b = (int(sl[q]), *boxes[sl[q]][0][pr[q, 1]])
                s.add(tuple(sorted([a, b])) if kind == 1 else (a, b))
        pairs[kind] = s
    return nodes, pairs, info


def oracle_structure(o):
    nd = o.nodes()
    key = {int(r[6]): (int(r[0]), int(r[1]), int(r[2]), int(r[3])) for r in nd}
    cnt = {int(r[6]): int(r[5]) for r in nd}
    nodes = set((int(r[0]), int(r[1]), int(r[2]), int(r[3]), int(r[4]), int(r[5])) for r in nd)
    pairs = {0: set((key[a], key[b]) for a, b in o.pairs(0)),
             1: set(tuple(sorted([key[a], key[b]])) for a, b in o.pairs(1) if cnt[a] > 0 and cnt[b] > 0),
             2: set((key[a], key[b]) for a, b in o.pairs(2))}
    return nodes, pairs, o.info()


def _dipoles(seed, n, side, cluster):
    pos, mom = G.uniform_dipoles(n, side, seed)
    pos[: n // 3] *= cluster
    return A.GeometryArrays(A.MLFMA_SRC_DIPOLE, pos=pos, mom=mom)


CASES = [
    ("dipoles_mns4", lambda: _dipoles(0, 300, 1.8, 0.2), 2.0, 4, -1),
    ("dipoles_mns8", lambda: _dipoles(1, 1000, 3.6, 0.2), 4.0, 8, -1),
    ("dipoles_mns3_deep", lambda: _dipoles(2, 2000, 3.6, 0.2), 4.0, 3, -1),
    ("dipoles_mns1", lambda: _dipoles(3, 500, 1.8, 0.2), 2.0, 1, -1),
    ("dipoles_capped", lambda: _dipoles(4, 700, 1.8, 0.3), 2.0, 2, 4),
    ("sphere_mns5", lambda: cases.sphere(6, 0.95), 2.0, 5, -1),
    ("sphere_fixed3", lambda: cases.sphere(6, 0.95), 2.0, 0, 3),
    ("ellipsoid_mns7", lambda: cases.ellipsoid(7, (0.95, 0.5, 0.3)), 2.0, 7, -1),
]


@pytest.mark.parametrize("name,geom_f,root,mns,cap", CASES, ids=[c[0] for c in CASES])
def test_tree_and_lists_match_oracle(name, geom_f, root, mns, cap):
    g = geom_f()
    cfg = A.make_config(root_leng=root, max_node_srcs=mns, max_level=cap, device=-1, interp_order=2)
    pn, pp, pi = product_structure(Plan(g, cases.planewave_z(), cfg))
    on, op, oi = oracle_structure(O.Oracle(g, cases.planewave_z(), cfg))
    assert pi["max_level"] == oi["max_level"]
    assert pn == on
    assert pp[0] == op[0], "interaction lists (list 2)"
    assert pp[1] == op[1], "adjacent leaf pairs (list 1)"
    assert pp[2] == op[2], "list-4 / list-3 pairs"


def test_interaction_list_bounds_and_offsets():
    # at most 6^3 - 3^3 = 189 entries (src/fmm/node.cpp:83); every M2L offset is
    # one of the 316 table vectors
    g = _dipoles(7, 3000, 3.9, 0.5)
    p = Plan(g, cases.planewave_z(), A.make_config(root_leng=4.0, max_node_srcs=0, max_level=4, device=-1))
    for lv in range(2, 5):
        co, _, _ = p.level_boxes(lv)
        pr, sl = p.level_pairs(lv, 0)
        assert (sl == lv).all()
        d = co[pr[:, 0]] - co[pr[:, 1]]
        m = np.abs(d).max(axis=1)
        assert ((m >= 2) & (m <= 3)).all()
        counts = np.bincount(pr[:, 0], minlength=len(co))
        assert counts.max() <= 189
