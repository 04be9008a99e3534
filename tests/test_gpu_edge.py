"""Edge cases of the GPU path against the CPU oracle (same inputs, FP64 bar
1e-10): a single source, two sources, every source in one leaf (no far field
at all), coincident sources (the reference drops their entry,
src/sources/dipole.h:62-70), sources exactly on the split planes of the tree
(ties go to the lower octant, src/fmm/stem.cpp:13), a one-level tree whose
leaves are all near each other, and a solve on an operator that is exactly
zero (a breakdown reported as a status, not a crash), and two clusters
coupled only through level-2 lists above stems with empty lists (the
reference drops the L2L below such stems: rows exactly zero)."""
import numpy as np
import pytest

from mlfma_amd import _capi as A
from mlfma_amd import geometry as G
from mlfma_amd.plan import Plan
from tests import _oracle as O
from tests import cases

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def dipoles(pos, mom=None):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    if mom is None:
        mom = np.tile([0.0, 0.0, 1e-8], (len(pos), 1))
    return A.GeometryArrays(A.MLFMA_SRC_DIPOLE, pos=pos, mom=np.ascontiguousarray(mom, dtype=np.float64))


def check(geom, kw, tag, tol=1e-10):
    pw = cases.planewave_z()
    p = Plan(geom, pw, A.make_config(device=0, **kw))
    ref = O.Oracle(geom, pw, A.make_config(device=-1, **kw))
    x = G.complex_gaussian(geom.n, 11)
    y, yr = p.matvec(x), ref.matvec(x)
    e = cases.relerr(y, yr) if np.linalg.norm(yr) > 0 else float(np.abs(y).max())
    say(tag, "n", geom.n, "info", {k: p.info()[k] for k in ("max_level", "near_nnz", "m2l_pairs")}, "err", e)
    assert e <= tol
    yd = p.matvec_direct(x)
    return p, x, y, yd


def test_single_source():
    geom = dipoles([[0.1, -0.2, 0.3]])
    p, x, y, yd = check(geom, dict(root_leng=2.0, max_node_srcs=4), "single")
    assert np.all(y == 0) and np.all(yd == 0)       # the self term is dropped


def test_two_sources_and_one_leaf():
    rng = np.random.default_rng(3)
    geom = dipoles(rng.uniform(-0.2, 0.2, size=(2, 3)))
    p, x, y, yd = check(geom, dict(root_leng=2.0, max_node_srcs=4), "two")
    assert cases.relerr(y, yd) <= 1e-13
    # every source in the root leaf: no far field, the matvec is the direct sum
    geom = dipoles(rng.uniform(-0.3, 0.3, size=(40, 3)), rng.normal(size=(40, 3)))
    p, x, y, yd = check(geom, dict(root_leng=2.0, max_node_srcs=64), "one leaf")
    assert p.info()["max_level"] == 0
    assert cases.relerr(y, yd) <= 1e-13


def test_coincident_sources():
    rng = np.random.default_rng(4)
    pos = rng.uniform(-0.9, 0.9, size=(60, 3))
    pos[10] = pos[3]
    pos[40] = pos[3]
    pos[55] = pos[20]
    geom = dipoles(pos, rng.normal(size=(60, 3)))
    check(geom, dict(root_leng=2.0, max_node_srcs=3), "coincident")


def test_sources_on_split_planes():
    # x, y or z equal to a box boundary at several levels (0, +-0.5, +-0.25)
    rng = np.random.default_rng(5)
    pos = rng.uniform(-0.95, 0.95, size=(120, 3))
    pos[:40, 0] = rng.choice([0.0, 0.5, -0.5, 0.25, -0.25], size=40)
    pos[40:80, 1] = rng.choice([0.0, 0.5, -0.5], size=40)
    pos[80:100, :] = rng.choice([0.0, 0.5, -0.25], size=(20, 3))
    geom = dipoles(pos, rng.normal(size=(120, 3)))
    check(geom, dict(root_leng=2.0, max_node_srcs=2), "split planes")
    check(geom, dict(root_leng=2.0, max_node_srcs=0, max_level=3, interp_order=4, digits=4), "split planes fixed")


def test_one_level_tree_is_all_near():
    # leaves at level 1 (2x2x2 boxes) are all neighbours: no M2L pairs
    rng = np.random.default_rng(6)
    geom = dipoles(rng.uniform(-0.9, 0.9, size=(80, 3)), rng.normal(size=(80, 3)))
    p, x, y, yd = check(geom, dict(root_leng=2.0, max_node_srcs=0, max_level=1), "level 1")
    assert sum(p.info()["m2l_pairs"]) == 0
    assert cases.relerr(y, yd) <= 1e-13


def test_zero_operator_solve_reports_breakdown():
    from mlfma_amd.plan import MlfmaError
    geom = dipoles([[0.1, -0.2, 0.3]])
    p = Plan(geom, cases.planewave_z(), A.make_config(device=0, root_leng=2.0, max_node_srcs=4))
    with pytest.raises(MlfmaError) as e:
        p.solve(np.ones(1, complex), method=A.MLFMA_GMRES_RESTART, restart=5, maxit=10)
    say("zero operator", e.value)
    assert e.value.code == A.MLFMA_ERR_BREAKDOWN and "breakdown" in str(e.value)
    # the plan stays usable
    assert np.all(p.matvec(np.ones(1, complex)) == 0)


def test_two_clusters_no_l2l_below_empty_list_stems():
    geom = cases.two_clusters()
    p, x, y, yd = check(geom, dict(cases.TWO_CLUSTERS_KW), "two clusters")
    # driven from cluster B alone, cluster A's rows are exactly zero
    xb = np.zeros(geom.n, complex)
    xb[500:] = x[500:]
    yb = p.matvec(xb)
    say("two clusters: max |y| on A", np.abs(yb[:500]).max(), "on B", np.abs(yb[500:]).max())
    assert np.all(yb[:500] == 0) and np.abs(yb[500:]).max() > 0
