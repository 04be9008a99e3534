"""GMRES on the GPU: the reference-compatible unrestarted solver against the
oracle's iterates, and the restarted / mixed-precision solvers against a dense
LU solve of the same operator."""
import numpy as np
import pytest

from mlfma_amd import _capi as A
from mlfma_amd import geometry as G
from mlfma_amd.plan import Plan
from tests import _oracle as O
from tests import cases

pytestmark = pytest.mark.gpu


def test_compat_gmres_matches_oracle_iterates():
    geom, kw = cases.make("dipole_fixed")
    pw = cases.planewave_z()
    p = Plan(geom, pw, A.make_config(device=0, **kw))
    ref = O.Oracle(geom, pw, A.make_config(device=-1, **kw))
    x, rep = p.solve(method=A.MLFMA_GMRES_COMPAT, maxit=40, tol=1e-6, history=True)
    xo, ro, it, rc = ref.solve(maxit=40, tol=1e-6)
    print("compat", rep["iterations"], it, flush=True)
    assert rep["iterations"] == it
    assert np.abs(rep["history"] - ro).max() <= 1e-10 * max(1.0, np.abs(ro).max())
    assert cases.relerr(x, xo) <= 1e-10


def _small_problem(n=48, seed=9):
    pos, mom = G.uniform_dipoles(n, 1.6, seed)
    mom = np.random.default_rng(seed).normal(size=(n, 3))
    return A.GeometryArrays(A.MLFMA_SRC_DIPOLE, pos=pos, mom=mom)


def test_restarted_and_mixed_gmres_vs_dense_solve():
    geom = _small_problem()
    kw = dict(root_leng=2.0, max_node_srcs=2, max_level=-1, interp_order=3, digits=5,
              flags=A.MLFMA_FLAG_FRESH_LOCALS)
    pw = A.make_planewave(1.0, (0.0, 0.0, 1.0), (1.0, 0.0, 0.0), 2 * np.pi)
    p = Plan(geom, pw, A.make_config(device=0, **kw))
    n = geom.n
    Z = np.stack([p.matvec(np.eye(n)[j].astype(complex)) for j in range(n)], 1)
    b = G.complex_gaussian(n, 4)
    xd = np.linalg.solve(Z, b)
    # restart >= n: GMRES reaches the exact solution within n iterations
    x, rep = p.solve(b, method=A.MLFMA_GMRES_RESTART, restart=n + 2, maxit=3 * n, tol=1e-12)
    assert rep["status"] == "OK" and rep["resid_true"] <= 1e-11
    assert cases.relerr(x, xd) <= 1e-8
    p32 = Plan(geom, pw, A.make_config(device=0, precision=A.MLFMA_FP32, **kw))
    x32, r32 = p32.solve(b, method=A.MLFMA_FGMRES_IR, restart=n + 2, maxit=6 * n, tol=1e-9)
    print("IR", r32, flush=True)
    assert r32["status"] == "OK" and r32["resid_true"] <= 1e-9
    assert cases.relerr(x32, xd) <= 1e-6


def test_zero_rhs_and_errors():
    geom = _small_problem(20, 3)
    p = Plan(geom, cases.planewave_z(), A.make_config(device=0, root_leng=2.0, max_node_srcs=4))
    x, rep = p.solve(np.zeros(geom.n, complex), method=A.MLFMA_GMRES_RESTART, restart=10, maxit=20)
    assert np.all(x == 0) and rep["iterations"] == 0
    with pytest.raises(Exception):
        p.solve(np.full(geom.n, np.nan + 0j), method=A.MLFMA_GMRES_RESTART)
    with pytest.raises(Exception):
        p.solve(method=A.MLFMA_FGMRES_IR)   # needs an FP32 plan


def test_batched_solve_matches_single_and_dense():
    geom = _small_problem()
    kw = dict(root_leng=2.0, max_node_srcs=2, max_level=-1, interp_order=3, digits=5,
              flags=A.MLFMA_FLAG_FRESH_LOCALS)
    pw = A.make_planewave(1.0, (0.0, 0.0, 1.0), (1.0, 0.0, 0.0), 2 * np.pi)
    p = Plan(geom, pw, A.make_config(device=0, **kw))
    n = geom.n
    Z = np.stack([p.matvec(np.eye(n)[j].astype(complex)) for j in range(n)], 1)
    B = np.stack([G.complex_gaussian(n, 40 + r) for r in range(5)])
    B[3] = 0.0                                           # a zero right-hand side in the batch
    X, reps = p.solve_batch(B, restart=n + 2, maxit=3 * n, tol=1e-12)
    for r in range(5):
        if r == 3:
            assert np.all(X[r] == 0)
            continue
        assert reps[r]["status"] == "OK" and reps[r]["resid_true"] <= 1e-11, (r, reps[r])
        assert cases.relerr(X[r], np.linalg.solve(Z, B[r])) <= 1e-8
        xs, rs = p.solve(B[r], method=A.MLFMA_GMRES_RESTART, restart=n + 2, maxit=3 * n, tol=1e-12)
        assert cases.relerr(X[r], xs) <= 1e-9
    # batched FP64 refinement around FP32-far-field inner solves
    p32 = Plan(geom, pw, A.make_config(device=0, precision=A.MLFMA_FP32, **kw))
    X32, reps32 = p32.solve_batch(B, restart=n + 2, maxit=6 * n, tol=1e-9, method=A.MLFMA_FGMRES_IR)
    for r in (0, 1, 2, 4):
        assert reps32[r]["status"] == "OK" and reps32[r]["resid_true"] <= 1e-9, (r, reps32[r])
        assert cases.relerr(X32[r], np.linalg.solve(Z, B[r])) <= 1e-6
    assert np.all(X32[3] == 0)
    # restarted (small m) batch on a larger problem: every column converges
    geom2, kw2 = cases.make("rwg_fixed")
    q = Plan(geom2, cases.planewave_z(), A.make_config(device=0, flags=A.MLFMA_FLAG_PHYSICAL, **kw2))
    B2 = np.stack([G.complex_gaussian(geom2.n, 60 + r) for r in range(8)])
    X2, reps2 = q.solve_batch(B2, restart=30, maxit=2000, tol=1e-8)
    for r in range(8):
        y = q.matvec(X2[r])
        assert cases.relerr(y, B2[r]) <= 2e-8, (r, reps2[r])
