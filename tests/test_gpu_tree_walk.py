"""The persistent tree walk (MLFMA_PARAM_TREE_WALK): S2M and every M2M level
of one right-hand side as one launch over a packed worklist, each unit
waiting on the completion counters of the units that write its inputs.
Every unit does the arithmetic of the staged launches (same kernels' bodies,
same row runs), so the matvec is the staged one bit for bit; a wait that hit
its poll bound would show in MLFMA_PARAM_TREE_STALLS.  Run on the small
parity cases (fixed and adaptive trees, leaves at several levels), the
two-cluster tree (boxes with empty interaction lists and children), the
30,720-RWG sphere and a forest plan of several scatterers, in FP64 and with
the FP32 far field, under every stream schedule and in graph replays."""
import numpy as np
import pytest

from mlfma_amd import _capi as A
from mlfma_amd import geometry as G
from mlfma_amd.plan import Plan
from tests import cases

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _walk_vs_staged(p, x, reps=3):
    """Matvecs with the walk off, then on (reps times, so the counters are
    reused after the last block's reset): the same bits, no stalls."""
    p.set_param(A.MLFMA_PARAM_TREE_WALK, 0)
    y0 = p.matvec(x).cpu().numpy()
    p.set_param(A.MLFMA_PARAM_TREE_WALK, 1)
    assert p.get_param(A.MLFMA_PARAM_TREE_WALK) == 1
    for rep in range(reps):
        y1 = p.matvec(x).cpu().numpy()
        e = cases.relerr(y1, y0)
        say("rep", rep, "walk vs staged", e)
        assert e == 0.0, rep
    assert p.get_param(A.MLFMA_PARAM_TREE_STALLS) == 0
    return y0


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("name", list(cases.CASES) + ["two_clusters"])
def test_walk_matches_staged(name, fp32):
    import torch
    say("case", name, "fp32", fp32)
    if name == "two_clusters":
        geom, kw = cases.two_clusters(), dict(cases.TWO_CLUSTERS_KW)
    else:
        geom, kw = cases.make(name)
    prec = A.MLFMA_FP32 if fp32 else A.MLFMA_FP64
    # fresh locals: with the reference's stale locals (boxes with empty
    # interaction lists keep the previous matvec's) consecutive matvecs differ
    p = Plan(geom, cases.planewave_z(),
             A.make_config(device=0, precision=prec, flags=A.MLFMA_FLAG_FRESH_LOCALS, **kw))
    x = torch.from_numpy(G.complex_gaussian(geom.n, 41)).to("cuda:0")
    _walk_vs_staged(p, x)


@pytest.mark.parametrize("fp32", [False, True])
def test_walk_every_schedule_sphere30k(fp32):
    """The 30,720-RWG sphere (levels 2-4 of M2M units, about 1,400 leaves):
    the walk gives the staged bits under every stream schedule (near field in
    order, forked before or after the upward pass; finest M2L beside or in
    line) and in graph replays with eager work in between."""
    import torch
    from mlfma_amd import configs as CF
    geom, pw, kw, _ = CF.c2()
    prec = A.MLFMA_FP32 if fp32 else A.MLFMA_FP64
    flags = A.MLFMA_FLAG_FRESH_LOCALS
    p = Plan(geom, pw, A.make_config(device=0, precision=prec, flags=flags, **kw))
    x = torch.from_numpy(G.complex_gaussian(geom.n, 43)).to("cuda:0")
    y0 = _walk_vs_staged(p, x)
    for overlap in (0, 1, 2, 3):
        for split in (0, 1):
            p.set_param(A.MLFMA_PARAM_NEAR_OVERLAP, overlap)
            p.set_param(A.MLFMA_PARAM_M2L_SPLIT, split)
            p.set_param(A.MLFMA_PARAM_NEAR_BLOCKS, 0)
            y1 = p.matvec(x).cpu().numpy()
            say("overlap", overlap, "split", split, cases.relerr(y1, y0))
            assert cases.relerr(y1, y0) == 0.0
    p.set_flags(flags | A.MLFMA_FLAG_GRAPH)
    y = torch.empty_like(x)
    for rep in range(4):
        p.matvec(x, y)
        assert cases.relerr(y.cpu().numpy(), y0) == 0.0, rep
        (x * 2).sum()  # eager torch work between replays
        torch.cuda.synchronize()
    p.set_flags(flags)
    assert p.get_param(A.MLFMA_PARAM_TREE_STALLS) == 0
    # the record carries the choice and applies to a second plan
    rec = p.tuning()
    assert rec[0] == 4 and rec[-2] == 1
    q = Plan(geom, pw, A.make_config(device=0, precision=prec, flags=flags, **kw))
    q.apply_tuning(rec)
    assert q.get_param(A.MLFMA_PARAM_TREE_WALK) == 1
    assert cases.relerr(q.matvec(x).cpu().numpy(), y0) == 0.0


def test_walk_forest():
    """A forest plan (four scatterers with roots of 1-4 wavelengths in one
    plan, levels aligned by box size): one walk covers every scatterer's
    leaves and stems."""
    import torch
    from tests.test_gpu_forest import small_forest
    sc = small_forest(A.MLFMA_FLAG_FRESH_LOCALS)
    p = Plan.forest([g for g, _ in sc], [A.make_config(device=0, **kw) for _, kw in sc], cases.planewave_z())
    x = torch.from_numpy(G.complex_gaussian(p.n, 47)).to("cuda:0")
    _walk_vs_staged(p, x)
