"""The HIP path on the benchmark configurations themselves, at full size,
against the CPU oracle (the restatement of the reference's FP64 matvec and
GMRES, src/solver.cpp:41-146) and through size-independent properties.

  C1  4,096 dipoles (src/mlfma.cpp:12-110 on the reference's own case):
      matvec and per-level fields vs the oracle, compat GMRES iterates
  C2  30,720-RWG sphere: matvec vs the oracle
  C3  201,720-RWG sphere: matvec vs the oracle; FP32 far field vs FP64,
      8-right-hand-side batch vs single, adjoint identity
  C4  16 scatterers: concurrent solves = sequential (bitwise), batched =
      single, oracle matvec on the smallest and largest

FP64 tolerance policy (as tests/test_gpu_parity.py): 1e-10, widened only to
ten times the spread between two valid compilations of the reference's
arithmetic (the oracle without and with FMA contraction), measured here on
the same input.  The spread comes from the cancellation next to the
|alpha| < 1e-6 branches of the RWG plane-wave integral
(src/sources/rwg.cpp:56-70)."""
import os

import numpy as np
import pytest

from mlfma_amd import _capi as A
from mlfma_amd import configs as CF
from mlfma_amd import geometry as G
from mlfma_amd.plan import Plan
from tests import _oracle as O
from tests import cases

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def say(*a):
    print(*a, flush=True)


def oracle(geom, pw, kw, variant=""):
    return O.Oracle(geom, pw, A.make_config(device=-1, **kw), threads=THREADS, variant=variant)


def fp64_tolerance(geom, pw, kw, x, y_ref):
    """max(1e-10, 10 x the FMA / no-FMA spread of the oracle on this input)."""
    b = oracle(geom, pw, kw, "fma")
    spread = cases.relerr(b.matvec(x), y_ref)
    del b
    return max(1e-10, 10 * spread), spread


# ---------------------------------------------------------------- C1
@pytest.fixture(scope="module")
def c1():
    geom, pw, kw, desc = CF.c1()
    say("C1", desc)
    return geom, pw, kw, Plan(geom, pw, A.make_config(device=0, **kw)), oracle(geom, pw, kw)


def test_c1_matvec_and_level_fields(c1):
    geom, pw, kw, plan, ref = c1
    x = G.complex_gaussian(geom.n, 0)
    yp = plan.matvec(x)
    yo = ref.matvec(x)
    tol, spread = fp64_tolerance(geom, pw, kw, x, yo)
    err = cases.relerr(yp, yo)
    say("C1 matvec vs oracle", err, "spread", spread, "tolerance", tol)
    assert err <= tol
    nodes = ref.nodes()
    for lv in range(2, plan.info()["max_level"] + 1):
        co, lf, ns = plan.level_boxes(lv)
        idx = {tuple(c): i for i, c in enumerate(co)}
        on = nodes[nodes[:, 0] == lv]
        order = [idx[(r[1], r[2], r[3])] for r in on]
        keep = on[:, 5] > 0
        for which in (0, 1):
            e = cases.relerr(plan.coeffs(lv, which)[order][keep], ref.coeffs(lv, which)[keep])
            say("C1 level", lv, "local" if which else "multipole", e)
            assert e <= 100 * tol


def test_c1_compat_gmres_iterates(c1):
    geom, pw, kw, plan, ref = c1
    x, rep = plan.solve(method=A.MLFMA_GMRES_COMPAT, maxit=40, tol=1e-6, history=True)
    xo, ro, it, rc = ref.solve(maxit=40, tol=1e-6)
    say("C1 compat GMRES", rep["iterations"], it, "last residual", ro[-1], "solution diff", cases.relerr(x, xo))
    assert rep["iterations"] == it
    assert np.abs(rep["history"] - ro).max() <= 1e-10 * max(1.0, np.abs(ro).max())
    assert cases.relerr(x, xo) <= 1e-10


# ---------------------------------------------------------------- C2
def test_c2_matvec_vs_oracle():
    geom, pw, kw, desc = CF.c2()
    say("C2", desc)
    plan = Plan(geom, pw, A.make_config(device=0, **kw))
    x = G.complex_gaussian(geom.n, 0)
    yp = plan.matvec(x)
    yo = oracle(geom, pw, kw).matvec(x)
    tol, spread = fp64_tolerance(geom, pw, kw, x, yo)
    err = cases.relerr(yp, yo)
    say("C2 matvec vs oracle", err, "spread", spread, "tolerance", tol)
    assert err <= tol


# ---------------------------------------------------------------- C3
@pytest.fixture(scope="module")
def c3():
    geom, pw, kw, desc = CF.c3()
    say("C3", desc)
    # linear operator (boxes with an empty interaction list get zero locals
    # every matvec): a first matvec equals the reference's, whose stale locals
    # start at zero, and batches and the adjoint are defined
    kw = dict(kw, flags=A.MLFMA_FLAG_FRESH_LOCALS)
    return geom, pw, kw, Plan(geom, pw, A.make_config(device=0, **kw))


@pytest.mark.timeout(900)
def test_c3_matvec_vs_oracle(c3):
    geom, pw, kw, plan = c3
    x = G.complex_gaussian(geom.n, 0)
    yp = plan.matvec(x)
    ref = oracle(geom, pw, kw)
    yo = ref.matvec(x)
    del ref
    tol, spread = fp64_tolerance(geom, pw, kw, x, yo)
    err = cases.relerr(yp, yo)
    say("C3 matvec vs oracle", err, "spread", spread, "tolerance", tol)
    assert err <= tol


def test_c3_fp32_far_field(c3):
    geom, pw, kw, plan = c3
    p32 = Plan(geom, pw, A.make_config(device=0, precision=A.MLFMA_FP32, **kw))
    x = G.complex_gaussian(geom.n, 1)
    e = cases.relerr(p32.matvec(x), plan.matvec(x))
    say("C3 FP32 far field vs FP64", e)
    assert 1e-12 < e <= 1e-6


def test_c3_batch_and_adjoint(c3):
    import torch
    geom, pw, kw, plan = c3
    n = geom.n
    X = torch.from_numpy(np.stack([G.complex_gaussian(n, 20 + r) for r in range(8)])).to("cuda:0")
    Y = plan.matvec(X)
    for r in (0, 3, 7):
        e = cases.relerr(Y[r].cpu().numpy(), plan.matvec(X[r].contiguous()).cpu().numpy())
        say("C3 batch lane", r, "vs single", e)
        assert e <= 1e-13
    x, w = X[0].contiguous(), X[1].contiguous()
    zx = plan.matvec(x)
    zhw = plan.matvec_adjoint(w)
    lhs = torch.vdot(w, zx).item()
    rhs = torch.vdot(zhw, x).item()
    rel = abs(lhs - rhs) / (torch.linalg.norm(zx).item() * torch.linalg.norm(w).item())
    say("C3 adjoint identity", rel)
    assert rel <= 1e-13


# ---------------------------------------------------------------- C4
@pytest.fixture(scope="module")
def c4():
    sc = CF.c4()
    say("C4", [(s["kind"], round(s["r_eq"], 3), s["n"], s["levels"]) for s in sc])
    return sc


@pytest.mark.timeout(900)
def test_c4_oracle_smallest_and_largest(c4):
    order = sorted(range(len(c4)), key=lambda i: c4[i]["n"])
    pw = CF.planewave_theta0()
    for i in (order[0], order[-1]):
        s = c4[i]
        plan = Plan(s["geom"], pw, A.make_config(device=0, **s["kw"]))
        x = G.complex_gaussian(s["n"], i)
        yp = plan.matvec(x)
        yo = oracle(s["geom"], pw, s["kw"]).matvec(x)
        tol, spread = fp64_tolerance(s["geom"], pw, s["kw"], x, yo)
        err = cases.relerr(yp, yo)
        say("C4 scatterer", i, s["kind"], s["n"], "matvec vs oracle", err, "spread", spread, "tolerance", tol)
        assert err <= tol


def test_c4_concurrent_batched_single(c4):
    from mlfma_amd.batch import solve_many
    flags = A.MLFMA_FLAG_PHYSICAL | A.MLFMA_FLAG_FRESH_LOCALS
    plans, rhs = [], []
    for s in c4:
        p = Plan(s["geom"], CF.planewave_theta0(), A.make_config(device=0, precision=A.MLFMA_FP32, flags=flags,
                                                                  **s["kw"]))
        rhs.append(np.stack([-p.excitation(CF.planewave_xz(t)) for t in CF.c4_angles()]))
        plans.append(p)
    # batched 8-right-hand-side matvecs equal one at a time: FP32 far field
    # (the batched kernels sum in another order), so to FP32 round-off
    for i in (0, 5, 12):
        Y = plans[i].matvec(rhs[i])
        for r in (0, 4, 7):
            e = cases.relerr(Y[r], plans[i].matvec(rhs[i][r]))
            say("C4 scatterer", i, "batch lane", r, "vs single (FP32 far field)", e)
            assert e <= 1e-6, (i, r, e)
    # a bounded number of iterations of the mixed-precision batched solve,
    # all scatterers concurrently and one after another: bitwise equal
    opts = dict(restart=20, maxit=40, tol=1e-6, method=A.MLFMA_FGMRES_IR)
    par = solve_many(plans, rhs, **opts)
    seq = [p.solve_batch(b, **opts) for p, b in zip(plans, rhs)]
    for i, ((xs, rs), (xp, rp)) in enumerate(zip(seq, par)):
        assert np.array_equal(xs, xp), i
        assert [r["iterations"] for r in rs] == [r["iterations"] for r in rp]
    say("C4 concurrent = sequential over", len(plans), "scatterers,", sum(s["n"] for s in c4), "unknowns")
