"""Deflated restarts of the device GMRES(m) (GMRES-DR, opts.deflate = k).

The device solver against an independent NumPy restatement of GMRES-DR in
Morgan's form (R. B. Morgan, SIAM J. Sci. Comput. 24 (2002) 20-37) on the
dense matrix of the same operator: Arnoldi with CGS2, least-squares residual
estimates, harmonic Ritz vectors from numpy.linalg.eig of
H_m + |h_{m+1,m}|^2 H_m^{-H} e_m e_m^T, the restart basis V_{m+1} P with the
residual direction appended.  The device keeps the same subspaces in another
basis (a triangular leading block, so its Givens pipeline runs unchanged):
iteration counts, residual histories and solutions must agree.  Then the
FP32-basis path (iterative refinement), right preconditioning, the adjoint,
argument errors, the batched solver (lanes in lockstep) against single
solves, and a resonant scatterer where plain restarts stall."""
import numpy as np
import pytest

from mlfma_amd import _capi as A
from mlfma_amd import configs as CF
from mlfma_amd import geometry as G
from mlfma_amd.plan import Plan
from tests import cases

pytestmark = pytest.mark.gpu


def _cgs2(V, w):
    h = np.zeros(len(V) + 1, complex)
    for _ in range(2):
        c = np.array([np.vdot(q, w) for q in V])
        w = w - sum(ci * q for ci, q in zip(c, V))
        h[:len(V)] += c
    h[len(V)] = np.linalg.norm(w)
    return h, w


def gmres_dr_numpy(Z, b, m, k, maxit, tol):
    """GMRES-DR(m, k) from zero on a dense Z; restarts check the true
    residual (like the device solver); a cycle that stops early on its
    estimate is followed, when the true residual is still above tol, by a
    plain restart.  Returns x, iterations, per-iteration estimates |r|/|b|."""
    n = len(b)
    bn = np.linalg.norm(b)
    x = np.zeros(n, complex)
    hist, total = [], 0
    keep = None  # (V_{k+1} list, Hbar (k+1) x k, c (k+1))
    while True:
        r = b - Z @ x if total else b.copy()
        beta = np.linalg.norm(r)
        if beta / bn <= tol or total >= maxit:
            return x, total, np.array(hist)
        H = np.zeros((m + 1, m), complex)
        if keep is None:
            V, c, j0 = [r / beta], np.zeros(m + 1, complex), 0
            c[0] = beta
        else:
            V, Hk, ck = keep
            j0 = len(V) - 1
            H[:j0 + 1, :j0] = Hk
            c = np.zeros(m + 1, complex)
            c[:j0 + 1] = ck
        j, stop = j0, False
        while j < m and total < maxit:
            h, w = _cgs2(V, Z @ V[j])
            H[:j + 2, j] = h
            total += 1
            y, *_ = np.linalg.lstsq(H[:j + 2, :j + 1], c[:j + 2], rcond=None)
            est = np.linalg.norm(c[:j + 2] - H[:j + 2, :j + 1] @ y) / bn
            hist.append(est)
            j += 1
            if est <= tol or h[-1] == 0:
                stop = True
                break
            V.append(w / h[-1])
        y, *_ = np.linalg.lstsq(H[:j + 1, :j], c[:j + 1], rcond=None)
        x = x + np.stack(V[:j], 1) @ y
        if stop or j < m:
            keep = None
            continue
        # harmonic Ritz vectors of the cycle, residual direction appended
        Hm = H[:m, :m]
        f = np.linalg.solve(Hm.conj().T, np.eye(m)[:, m - 1])
        F = Hm + abs(H[m, m - 1]) ** 2 * np.outer(f, np.eye(m)[m - 1])
        th, Gv = np.linalg.eig(F)
        idx = np.argsort(np.abs(th), kind="stable")[:k]
        res = c - H @ y
        P = np.zeros((m + 1, k + 1), complex)
        P[:m, :k] = Gv[:, idx]
        P[:, k] = res
        P, _ = np.linalg.qr(P)
        Vm1 = np.stack(V[:m + 1], 1)
        Vn = Vm1 @ P
        Hk = P.conj().T @ H @ P[:m, :k]
        keep = ([Vn[:, i] for i in range(k + 1)], Hk, P.conj().T @ res)


def _sphere_plan():
    geom, kw = cases.make("rwg_fixed")
    p = Plan(geom, cases.planewave_z(),
             A.make_config(device=0, **dict(kw, flags=A.MLFMA_FLAG_PHYSICAL | A.MLFMA_FLAG_FRESH_LOCALS)))
    Z = np.stack([p.matvec(np.eye(geom.n)[j].astype(complex)) for j in range(geom.n)], 1)
    return geom, kw, p, Z


def test_deflated_gmres_matches_numpy_gmres_dr():
    geom, kw, p, Z = _sphere_plan()
    n = geom.n
    for seed, m, k, tol in ((4, 20, 5, 1e-7), (5, 30, 8, 1e-9)):
        b = G.complex_gaussian(n, seed)
        xr, itr, hr = gmres_dr_numpy(Z, b, m, k, 4000, tol)
        x, rep = p.solve(b, method=A.MLFMA_GMRES_RESTART, restart=m, maxit=4000, tol=tol, history=True, deflate=k)
        x0, rep0 = p.solve(b, method=A.MLFMA_GMRES_RESTART, restart=m, maxit=4000, tol=tol)
        print("GMRES-DR(%d,%d)" % (m, k), "device", rep["iterations"], "numpy", itr, "plain", rep0["iterations"],
              rep["resid_true"], flush=True)
        assert rep["status"] == "OK" and rep["resid_true"] <= tol
        assert abs(rep["iterations"] - itr) <= 2
        c = min(len(hr), len(rep["history"]))
        assert np.allclose(rep["history"][:c], hr[:c], rtol=1e-4, atol=0)
        assert cases.relerr(x, xr) <= 1e-6
        assert rep["iterations"] < rep0["iterations"]


def test_deflated_fp32_basis_preconditioned_and_adjoint():
    geom, kw, p, Z = _sphere_plan()
    n = geom.n
    b = G.complex_gaussian(n, 7)
    xd = np.linalg.solve(Z, b)
    # right preconditioning (leaf blocks) and the adjoint, FP64
    xp, rp = p.solve(b, method=A.MLFMA_GMRES_RESTART, restart=20, maxit=4000, tol=1e-9, precond=1, deflate=6)
    assert rp["status"] == "OK" and rp["resid_true"] <= 1e-9
    assert cases.relerr(xp, xd) <= 1e-6
    xa, ra = p.solve(b, method=A.MLFMA_GMRES_RESTART, restart=20, maxit=4000, tol=1e-9, adjoint=True, deflate=6)
    assert ra["status"] == "OK" and ra["resid_true"] <= 1e-9
    assert cases.relerr(xa, np.linalg.solve(Z.conj().T, b)) <= 1e-6
    # FP32 far field, FP32 Krylov basis inside FP64 refinement
    p32 = Plan(geom, cases.planewave_z(), A.make_config(
        device=0, precision=A.MLFMA_FP32, **dict(kw, flags=A.MLFMA_FLAG_PHYSICAL | A.MLFMA_FLAG_FRESH_LOCALS)))
    x32, r32 = p32.solve(b, method=A.MLFMA_GMRES_IR, restart=20, maxit=6000, tol=1e-9, deflate=6)
    print("IR + DR", r32, flush=True)
    assert r32["status"] == "OK" and r32["resid_true"] <= 1e-9
    assert cases.relerr(x32, xd) <= 1e-6


def test_batched_deflation_matches_single_solves():
    geom, kw, p, Z = _sphere_plan()
    n = geom.n
    B = np.stack([G.complex_gaussian(n, 20 + r) for r in range(3)])
    X, reps = p.solve_batch(B, restart=20, maxit=4000, tol=1e-7, deflate=5)
    for r in range(3):
        x, rep = p.solve(B[r], method=A.MLFMA_GMRES_RESTART, restart=20, maxit=4000, tol=1e-7, deflate=5)
        print("batched lane", r, reps[r]["iterations"], "single", rep["iterations"], flush=True)
        assert reps[r]["status"] == "OK" and reps[r]["resid_true"] <= 1e-7
        assert abs(reps[r]["iterations"] - rep["iterations"]) <= 2
        assert cases.relerr(X[r], x) <= 1e-6
        assert cases.relerr(X[r], np.linalg.solve(Z, B[r])) <= 1e-5
    p32 = Plan(geom, cases.planewave_z(), A.make_config(
        device=0, precision=A.MLFMA_FP32, **dict(kw, flags=A.MLFMA_FLAG_PHYSICAL | A.MLFMA_FLAG_FRESH_LOCALS)))
    X32, r32 = p32.solve_batch(B, method=A.MLFMA_GMRES_IR, restart=20, maxit=6000, tol=1e-9, deflate=6)
    print("batched IR + DR", [r["iterations"] for r in r32], flush=True)
    for r in range(3):
        assert r32[r]["status"] == "OK" and r32[r]["resid_true"] <= 1e-9
        assert cases.relerr(X32[r], np.linalg.solve(Z, B[r])) <= 1e-6


def test_deflation_argument_errors():
    geom, kw = cases.make("rwg_fixed")
    p = Plan(geom, cases.planewave_z(), A.make_config(device=0, **dict(kw, flags=A.MLFMA_FLAG_FRESH_LOCALS)))
    b = G.complex_gaussian(geom.n, 1)
    for bad in (dict(restart=10, deflate=9), dict(restart=10, deflate=-1), dict(restart=200, deflate=65)):
        with pytest.raises(Exception):
            p.solve(b, method=A.MLFMA_GMRES_RESTART, maxit=50, **bad)
    with pytest.raises(Exception):
        p.solve(b, method=A.MLFMA_GMRES_COMPAT, maxit=50, deflate=4)
    with pytest.raises(Exception):  # the flexible outer keeps plain restarts
        p.solve_batch(np.stack([b, b]), restart=20, maxit=50, deflate=4, method=A.MLFMA_FGMRES_IR)
    with pytest.raises(Exception):
        p.solve_batch(np.stack([b, b]), restart=10, maxit=50, deflate=9)
    # still usable after the refusals
    x, rep = p.solve(b, method=A.MLFMA_GMRES_RESTART, restart=20, maxit=20, tol=1e-3, deflate=4)
    assert np.isfinite(rep["resid_true"])


def test_deflation_converges_on_a_resonant_ellipsoid():
    """The smallest seed-1 ellipsoid scaled to k a = 8.16 (its long
    semi-axis), one of the EFIE interior resonances where FP64 GMRES(100)
    stalls: deflating 20 harmonic Ritz vectors per restart converges."""
    shapes = CF.c4()
    ells = [sh for sh in shapes if sh["kind"] == "ellipsoid"]
    small = min(ells, key=lambda sh: sh["n"])
    extent = 2 * small["r_eq"] / 2 ** (1 / 3)
    sh = CF.c4_scatterer("ellipsoid", small["r_eq"], scale=8.16 / (2 * np.pi * extent))
    pw0 = CF.planewave_xz(0.0)
    p = Plan(sh["geom"], pw0, A.make_config(device=0, flags=A.MLFMA_FLAG_PHYSICAL | A.MLFMA_FLAG_FRESH_LOCALS,
                                            **sh["kw"]))
    b = -p.excitation(pw0)
    _, r0 = p.solve(b, method=A.MLFMA_GMRES_RESTART, restart=100, maxit=3000, tol=1e-6)
    _, r1 = p.solve(b, method=A.MLFMA_GMRES_RESTART, restart=100, maxit=3000, tol=1e-6, deflate=20)
    print("resonant ellipsoid n=%d: GMRES(100)" % sh["n"], r0, "GMRES-DR(100,20)", r1, flush=True)
    assert r1["status"] == "OK" and r1["resid_true"] <= 1e-6
    assert r1["iterations"] < r0["iterations"]


def test_forest_deflation_converges_every_lane():
    """Three scatterers in one forest plan, 3 right-hand sides: 9 lanes in
    lockstep whose refinement steps stop at different inner tolerances, so
    lanes without a deflated start restart plainly inside deflated cycles.
    Every lane converges, and each scatterer's solutions agree with its own
    plan's deflated solves.  (The fourth scatterer of the small forest
    stalls at restart 16 with or without deflation, and in a batched
    refinement its lanes would spend the shared iteration budget.)"""
    from tests.test_gpu_forest import small_forest
    fl = A.MLFMA_FLAG_FRESH_LOCALS | A.MLFMA_FLAG_PHYSICAL
    sc = small_forest(fl, A.MLFMA_FP32)[:3]
    pw = cases.planewave_z()
    cfgs = [A.make_config(device=0, **kw) for _, kw in sc]
    F = Plan.forest([g for g, _ in sc], cfgs, pw)
    off = F.segments()
    B = np.stack([G.complex_gaussian(F.n, 40 + r) for r in range(3)])
    for method in (A.MLFMA_GMRES_IR, A.MLFMA_GMRES_RESTART):
        X, reps = F.solve_batch(B, method=method, restart=16, maxit=4000, tol=1e-8, deflate=4)
        print("forest", method, [r["iterations"] for r in reps], flush=True)
        assert all(r["status"] == "OK" and r["resid_true"] <= 1e-8 for r in reps)
        for s, (g, kw) in enumerate(sc):
            p = Plan(g, pw, A.make_config(device=0, **kw))
            seg = slice(off[s], off[s + 1])
            Xs, rs = p.solve_batch(np.ascontiguousarray(B[:, seg]), method=method, restart=16, maxit=4000, tol=1e-8,
                                   deflate=4)
            assert all(r["status"] == "OK" for r in rs)
            assert cases.relerr(X[:, seg], Xs) <= 1e-6
