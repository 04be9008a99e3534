"""The oracle's neighbour search against the reference's own case table and
against the geometry it must produce (CPU only).

The reference spells the search out direction by direction and branch by
branch (src/fmm/nbor.cpp:10-483 getNeighborGeqSize, 490-644
getNeighborsLeqSize); the oracle restates it as one bit rule
(oracle/oracle.cpp neighborRule, neighborChildFaces).  The rule is pinned
by the table read from nbor.cpp's text (tests/golden/nbor_table.json, written
by tests/golden/make_nbor_table.py): for all 26 directions x 8 branch indices,
sibling or parent step, the parent's direction, the child taken and whether a
Leaf answer is kept.  The lists built from it are then checked on adaptive
trees against their geometric meaning, computed from box coordinates alone:
  * a leaf's near neighbours are exactly the other leaves whose closed boxes
    touch it, each once;
  * a node's neighbours of equal or greater size are, over the 26 adjacent
    cells, the deepest node of level <= its own containing the cell (a node
    of its level or a coarser leaf), each once;
  * a node's interaction list is the non-empty boxes of its level that are
    children of its parent's neighbours and do not touch it, and its leaf
    list the non-empty coarser leaves next to its parent that do not touch it
    (src/fmm/node.cpp:61-84)."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from mlfma_amd import _capi as A
from tests import _oracle as O
from tests import cases

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = json.load(open(os.path.join(HERE, "golden", "nbor_table.json")))
DIRNAMES = TABLE["dirs"]
VEC = {"W": (-1, 0, 0), "E": (1, 0, 0), "S": (0, -1, 0), "N": (0, 1, 0), "D": (0, 0, -1), "U": (0, 0, 1)}


def dvec(name):
    return tuple(sum(VEC[c][a] for c in name) for a in range(3))


def test_table_covers_the_26_directions():
    vecs = [dvec(n) for n in DIRNAMES]
    assert len(set(vecs)) == 26 and (0, 0, 0) not in vecs
    assert all(max(abs(v) for v in x) == 1 for x in vecs)


@pytest.mark.parametrize("d", range(26))
def test_neighbor_rule_matches_reference_table(d):
    L = O.lib()
    out = (C.c_int32 * 4)()
    for b in range(8):
        assert L.orc_nbor_rule(d, b, out) == 0
        ref = TABLE["geq"][DIRNAMES[d]][b]
        if ref[0] == "sibling":
            assert (out[0], out[1]) == (0, ref[1]), (DIRNAMES[d], b, list(out), ref)
        else:
            _, exit_dir, child, leaf = ref
            got = (out[0], DIRNAMES[out[2]], out[1], "leaf" if out[3] else "drop")
            assert got == (1, exit_dir, child, leaf), (DIRNAMES[d], b, got, ref)


def test_leq_children_match_reference_table():
    L = O.lib()
    out = (C.c_int32 * 9)()
    for d in range(26):
        assert L.orc_nbor_leq_children(d, out) == 0
        assert list(out[1:1 + out[0]]) == TABLE["leq"][DIRNAMES[d]], DIRNAMES[d]


def _tree(name):
    geom, kw = cases.make(name)
    o = O.Oracle(geom, cases.planewave_z(), A.make_config(device=-1, **kw))
    nodes = o.nodes()   # [level, ix, iy, iz, leaf, nsrcs, id]
    return o, nodes


def _touch(a, b):
    """Closed boxes of nodes a, b (rows of nodes()) share at least a point."""
    L = max(a[0], b[0])
    sa, sb = 1 << (L - a[0]), 1 << (L - b[0])
    for k in range(3):
        lo1, hi1 = a[1 + k] * sa, (a[1 + k] + 1) * sa
        lo2, hi2 = b[1 + k] * sb, (b[1 + k] + 1) * sb
        if lo1 > hi2 or lo2 > hi1:
            return False
    return True


@pytest.fixture(scope="module", params=["dipole_adaptive", "rwg_adaptive", "dipole_fixed"])
def tree(request):
    o, nodes = _tree(request.param)
    byid = {int(n[6]): n for n in nodes}
    at = {(int(n[0]), int(n[1]), int(n[2]), int(n[3])): n for n in nodes}
    return request.param, o, nodes, byid, at


def test_near_neighbors_are_touching_leaves(tree):
    name, o, nodes, byid, at = tree
    leaves = nodes[nodes[:, 4] == 1]
    assert len({int(n[0]) for n in leaves}) > 1 or name == "dipole_fixed"
    got = {}
    for a, b in o.pairs(4):
        got.setdefault(int(a), []).append(int(b))
    # every leaf pair at once, in units of the finest level
    Lm = int(leaves[:, 0].max())
    s = (1 << (Lm - leaves[:, 0]))[:, None]
    lo = leaves[:, 1:4] * s
    hi = lo + s
    for i, leaf in enumerate(leaves):
        t = np.all((lo <= hi[i]) & (lo[i] <= hi), axis=1)
        t[i] = False
        want = sorted(int(x) for x in leaves[t, 6])
        have = got.get(int(leaf[6]), [])
        assert len(have) == len(set(have)), ("duplicate near neighbour", name, leaf)
        assert sorted(have) == want, (name, leaf, sorted(have), want)


def test_geq_neighbors_are_the_deepest_boxes_of_the_adjacent_cells(tree):
    name, o, nodes, byid, at = tree
    got = {}
    for a, b in o.pairs(3):
        got.setdefault(int(a), []).append(int(b))
    for n in nodes:
        lv = int(n[0])
        if lv == 0:
            continue
        want = set()
        for dname in DIRNAMES:
            c = [int(n[1 + k]) + dvec(dname)[k] for k in range(3)]
            if min(c) < 0 or max(c) >= (1 << lv):
                continue
            for j in range(lv, -1, -1):
                m = at.get((j, c[0] >> (lv - j), c[1] >> (lv - j), c[2] >> (lv - j)))
                if m is not None:
                    assert j == lv or m[4] == 1   # a coarser box on the way down is a leaf
                    want.add(int(m[6]))
                    break
        have = got.get(int(n[6]), [])
        assert len(have) == len(set(have)), ("duplicate neighbour", name, n)
        assert set(have) == want, (name, n, sorted(have), sorted(want))


def test_interaction_lists_are_geometric(tree):
    name, o, nodes, byid, at = tree
    il, leafil = {}, {}
    for a, b in o.pairs(0):
        il.setdefault(int(a), []).append(int(b))
    for a, b in o.pairs(2):            # (leaf in the node's leaf list, node)
        leafil.setdefault(int(b), []).append(int(a))
    nb = {}
    for a, b in o.pairs(3):
        nb.setdefault(int(a), set()).add(int(b))
    for n in nodes:
        lv = int(n[0])
        if lv == 0:
            continue
        p = [int(n[1 + k]) >> 1 for k in range(3)]
        parent = at[(lv - 1, *p)]
        want_il, want_leaf = set(), set()
        for q in nb.get(int(parent[6]), ()):
            m = byid[q]
            if m[5] == 0:
                continue
            if m[4] == 1:                                   # a leaf next to the parent
                if not _touch(m, n):
                    want_leaf.add(q)
                continue
            for c in range(8):                              # children of a stem next to it
                cc = [2 * int(m[1 + k]) + ((c >> k) & 1) for k in range(3)]
                ch = at.get((lv, *cc))
                if ch is not None and ch[5] > 0 and not _touch(ch, n):
                    want_il.add(int(ch[6]))
        # the same sets from coordinates alone: boxes of this level whose parent
        # touches (or is) the parent, that do not touch the node
        geo = set()
        for dx in range(-3, 4):
            for dy in range(-3, 4):
                for dz in range(-3, 4):
                    cc = (int(n[1]) + dx, int(n[2]) + dy, int(n[3]) + dz)
                    ch = at.get((lv, *cc))
                    if ch is None or ch[5] == 0 or _touch(ch, n):
                        continue
                    if max(abs((cc[k] >> 1) - p[k]) for k in range(3)) <= 1:
                        geo.add(int(ch[6]))
        assert want_il == geo, (name, n)
        have = il.get(int(n[6]), [])
        assert len(have) == len(set(have)) and set(have) == want_il, (name, n, sorted(have), sorted(want_il))
        hl = leafil.get(int(n[6]), [])
        assert len(hl) == len(set(hl)) and set(hl) == want_leaf, (name, n, sorted(hl), sorted(want_leaf))
