"""The leaf block-Jacobi preconditioner (MLFMA_PRECOND_LEAF_BLOCKS; no
reference counterpart): with a single root leaf the self block is the whole
operator, so P Z x = x; with many leaves, right-preconditioned GMRES(m)
(single, batched, and inside FGMRES) reaches the same solution as the plain
solve; the adjoint applies P^H."""
import numpy as np
import pytest

from mlfma_amd import _capi as A
from mlfma_amd import geometry as G
from mlfma_amd.plan import Plan
from tests import cases

pytestmark = pytest.mark.gpu
PB = A.MLFMA_PRECOND_LEAF_BLOCKS


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("name,flags", [("rwg_fixed", 0), ("rwg_fixed", A.MLFMA_FLAG_PHYSICAL),
                                        ("dipole_fixed", 0)])
def test_single_leaf_block_is_the_inverse(name, flags):
    geom, kw = cases.make(name)
    n = geom.n
    # one root leaf: every interaction is near field and the leaf's self
    # block is Z itself (rows split into pieces of 128, columns into items)
    kw = dict(kw, max_node_srcs=n, max_level=0, flags=flags | A.MLFMA_FLAG_FRESH_LOCALS)
    p = Plan(geom, cases.planewave_z(), A.make_config(device=0, **kw))
    assert p.info()["num_leaves"] == 1 and n > 256
    X = np.stack([G.complex_gaussian(n, 70 + r) for r in range(3)])
    Y = p.precond(p.matvec(X))
    e = cases.relerr(Y, X)
    Ya = p.precond(p.matvec_adjoint(X), adjoint=True)
    ea = cases.relerr(Ya, X)
    say(name, flags, "P Z x vs x", e, "P^H Z^H x vs x", ea)
    assert e <= 1e-8 and ea <= 1e-8
    # preconditioned GMRES on that plan: P = Z^-1, one step
    b = G.complex_gaussian(n, 5)
    x, rep = p.solve(b, method=A.MLFMA_GMRES_RESTART, restart=10, maxit=10, tol=1e-9, precond=PB)
    say("one-leaf preconditioned GMRES", rep)
    assert rep["status"] == "OK" and rep["iterations"] <= 2


def test_preconditioned_solves_match_plain():
    geom, kw = cases.make("rwg_fixed")
    flags = A.MLFMA_FLAG_PHYSICAL | A.MLFMA_FLAG_FRESH_LOCALS
    p = Plan(geom, cases.planewave_z(), A.make_config(device=0, flags=flags, **kw))
    B = np.stack([G.complex_gaussian(geom.n, 90 + r) for r in range(3)])
    x0, r0 = p.solve(B[0], method=A.MLFMA_GMRES_RESTART, restart=30, maxit=3000, tol=1e-9)
    x1, r1 = p.solve(B[0], method=A.MLFMA_GMRES_RESTART, restart=30, maxit=3000, tol=1e-9, precond=PB)
    say("plain", r0, "preconditioned", r1, cases.relerr(x1, x0))
    assert r0["status"] == "OK" and r1["status"] == "OK" and r1["resid_true"] <= 1e-9
    assert cases.relerr(x1, x0) <= 1e-7
    # adjoint solve with P^H
    z0, q0 = p.solve(B[1], method=A.MLFMA_GMRES_RESTART, restart=30, maxit=3000, tol=1e-9, adjoint=True)
    z1, q1 = p.solve(B[1], method=A.MLFMA_GMRES_RESTART, restart=30, maxit=3000, tol=1e-9, adjoint=True, precond=PB)
    say("adjoint plain", q0, "preconditioned", q1)
    assert q1["status"] == "OK" and cases.relerr(z1, z0) <= 1e-7
    # batched lanes equal single preconditioned solves
    Xb, rb = p.solve_batch(B, restart=30, maxit=3000, tol=1e-9, precond=PB)
    for r in range(3):
        xs, rs = p.solve(B[r], method=A.MLFMA_GMRES_RESTART, restart=30, maxit=3000, tol=1e-9, precond=PB)
        assert rb[r]["status"] == "OK" and cases.relerr(Xb[r], xs) <= 1e-8, (r, rb[r])
    # inside the mixed-precision flexible GMRES
    p32 = Plan(geom, cases.planewave_z(), A.make_config(device=0, precision=A.MLFMA_FP32, flags=flags, **kw))
    xf, rf = p32.solve(B[0], method=A.MLFMA_FGMRES_IR, restart=30, maxit=3000, tol=1e-9, precond=PB)
    say("FGMRES preconditioned", rf)
    assert rf["status"] == "OK" and cases.relerr(xf, x0) <= 1e-7


def test_zero_diagonal_blocks_are_handled():
    """The reference's dipole kernel has a zero diagonal, so a one-source leaf
    has a zero self block: it gets a scaled identity (finite, nonzero
    output).  Block Jacobi does not help this indefinite operator, so no
    convergence is asserted."""
    geom, kw = cases.make("dipole_adaptive")
    p = Plan(geom, cases.planewave_z(), A.make_config(device=0, flags=A.MLFMA_FLAG_FRESH_LOCALS, **kw))
    b = G.complex_gaussian(geom.n, 3)
    y = p.precond(b)
    assert np.all(np.isfinite(y)) and np.linalg.norm(y) > 0
    x1, r1 = p.solve(b, method=A.MLFMA_GMRES_RESTART, restart=60, maxit=300, tol=1e-9, precond=PB)
    assert np.all(np.isfinite(x1)) and np.isfinite(r1["resid_true"])
