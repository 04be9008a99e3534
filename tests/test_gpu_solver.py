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
    # the FP64 outer of the mixed solvers reaches an FP64 tolerance although
    # every inner step runs the FP32 far field: flexible GMRES (several inner
    # budgets and outer restarts) and iterative refinement.  (Short inner
    # budgets with short outer cycles, e.g. GMRES(5) inside FGMRES(8),
    # stagnate on this indefinite operator; a NumPy FGMRES on its dense matrix
    # does the same.)
    for method, kw_s in [(A.MLFMA_FGMRES_IR, dict(restart=n + 2)),
                         (A.MLFMA_FGMRES_IR, dict(restart=12, outer_restart=200, inner_tol=1e-3)),
                         (A.MLFMA_GMRES_IR, dict(restart=n + 2))]:
        x32, r32 = p32.solve(b, method=method, maxit=20 * n, tol=1e-10, **kw_s)
        print("mixed", method, kw_s, r32, cases.relerr(x32, xd), flush=True)
        assert r32["status"] == "OK" and r32["resid_true"] <= 1e-10
        # the outer estimate is the true FP64 residual of the returned x
        assert abs(r32["resid_est"] - r32["resid_true"]) <= 1e-3 * r32["resid_true"] + 1e-15
        assert cases.relerr(x32, xd) <= 1e-7


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
    # batched FP64 flexible GMRES / refinement around FP32-far-field inner solves
    p32 = Plan(geom, pw, A.make_config(device=0, precision=A.MLFMA_FP32, **kw))
    for method, kw_s in [(A.MLFMA_FGMRES_IR, dict(restart=n + 2)),
                         (A.MLFMA_FGMRES_IR, dict(restart=12, outer_restart=200, inner_tol=1e-3)),
                         (A.MLFMA_GMRES_IR, dict(restart=n + 2))]:
        X32, reps32 = p32.solve_batch(B, maxit=20 * n, tol=1e-9, method=method, **kw_s)
        print("batched mixed", method, kw_s, [r["iterations"] for r in reps32], flush=True)
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


def _gmres_restart_numpy(Z, b, m, maxit, tol):
    """GMRES(m) from zero on a dense matrix (CGS2, unitary Givens rotations,
    restarts with the true residual), returning x, the iteration count and
    the residual estimate of every iteration: the synchronous algorithm the
    device solver pipelines."""
    n = len(b)
    x = np.zeros(n, complex)
    bn = np.linalg.norm(b)
    hist, total = [], 0
    while True:
        r = b - Z @ x if total else b.copy()
        beta = np.linalg.norm(r)
        if beta / bn <= tol or total >= maxit:
            return x, total, np.array(hist)
        Q = [r / beta]
        H = np.zeros((m + 1, m), complex)
        cs, sn = np.zeros(m, complex), np.zeros(m, complex)
        g = np.zeros(m + 1, complex)
        g[0] = beta
        j = 0
        stop = False
        while j < m and total < maxit:
            w = Z @ Q[j]
            h = np.zeros(j + 2, complex)
            for _ in range(2):
                c = np.array([np.vdot(q, w) for q in Q])
                w = w - sum(ci * q for ci, q in zip(c, Q))
                h[:j + 1] += c
            hn = np.linalg.norm(w)
            h[j + 1] = hn
            for i in range(j):
                a = np.conj(cs[i]) * h[i] + np.conj(sn[i]) * h[i + 1]
                h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1]
                h[i] = a
            rho = np.sqrt(abs(h[j]) ** 2 + abs(h[j + 1]) ** 2)
            cs[j], sn[j] = h[j] / rho, h[j + 1] / rho
            h[j], h[j + 1] = rho, 0
            H[:j + 2, j] = h
            g[j + 1] = -sn[j] * g[j]
            g[j] = np.conj(cs[j]) * g[j]
            total += 1
            est = abs(g[j + 1]) / bn
            hist.append(est)
            j += 1
            if est <= tol or hn == 0:
                stop = True
                break
            Q.append(w / hn)
        y = np.linalg.solve(np.triu(H[:j, :j]), g[:j])
        x = x + sum(yi * q for yi, q in zip(y, Q[:j]))


def test_pipelined_gmres_matches_synchronous_algorithm():
    """The device GMRES(m) queues steps ahead of its stopping test; its
    iteration count and residual history must be those of the synchronous
    algorithm on the same operator (physical RWG sphere, restart 30: several
    cycles and stops in the middle of one), single and batched."""
    geom, kw = cases.make("rwg_fixed")
    p = Plan(geom, cases.planewave_z(),
             A.make_config(device=0, **dict(kw, flags=A.MLFMA_FLAG_PHYSICAL | A.MLFMA_FLAG_FRESH_LOCALS)))
    n = geom.n
    Z = np.stack([p.matvec(np.eye(n)[j].astype(complex)) for j in range(n)], 1)
    m, maxit = 30, 2000
    mid = False
    for seed, tol in ((4, 1e-4), (5, 1e-6), (6, 1e-7)):
        b = G.complex_gaussian(n, seed)
        xr, itr, hr = _gmres_restart_numpy(Z, b, m, maxit, tol)
        x, rep = p.solve(b, method=A.MLFMA_GMRES_RESTART, restart=m, maxit=maxit, tol=tol, history=True)
        print("pipelined", rep["iterations"], "synchronous", itr, flush=True)
        assert rep["iterations"] == itr and itr < maxit
        mid = mid or itr % m != 0
        assert np.allclose(rep["history"], hr, rtol=1e-6, atol=0)
        assert cases.relerr(x, xr) <= 1e-9
    assert mid  # at least one stop inside a cycle
    B = np.stack([G.complex_gaussian(n, 60 + r) for r in range(3)])
    X, reps = p.solve_batch(B, restart=m, maxit=maxit, tol=1e-6, history=True)
    for r in range(3):
        xr, itr, hr = _gmres_restart_numpy(Z, B[r], m, maxit, 1e-6)
        assert reps[r]["iterations"] == itr, (r, reps[r]["iterations"], itr)
        assert np.allclose(reps[r]["history"][:itr], hr, rtol=1e-6, atol=0)
        assert cases.relerr(X[r], xr) <= 1e-9
