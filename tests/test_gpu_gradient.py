"""Position gradients of a far-field objective for dipole sources on the GPU.

J(r) = sum_d |E(rhat_d)|^2 with E = F(r) Z(r)^-1 b.  Pins:
* the adjoint-method gradient (mlfma_position_gradient for the operator,
  the far-field adjoint for the pattern) against central finite differences
  of J along random directions of the positions, with Z(r) assembled column
  by column from the plan's own matvec and solved densely, so that the
  differences see exactly the operator the gradient differentiates;
* the same with b the plan's plane-wave excitation (its position term);
* the GMRES-based objective_gradients against the dense-solve gradient;
* a 300-dipole case whose interactions are mostly far field;
* the error paths (RWG plan, operator with stale locals).
Every perturbed plan is checked to have the same tree (box counts per level),
since the gradient holds the tree fixed."""
import numpy as np
import pytest

from mlfma_amd import _capi as A
from mlfma_amd import autograd as AG
from mlfma_amd import geometry as G
from mlfma_amd.plan import Plan, MlfmaError
from tests import cases
from tests.test_gpu_solver import _small_problem

pytestmark = pytest.mark.gpu

KW = dict(root_leng=2.0, max_node_srcs=2, max_level=-1, interp_order=3, digits=5, flags=A.MLFMA_FLAG_FRESH_LOCALS)


def say(*a):
    print(*a, flush=True)


def dirs_ring(n=11):
    t = np.linspace(0.1, np.pi - 0.1, n)
    return np.stack([np.sin(t) * np.cos(0.3), np.sin(t) * np.sin(0.3), np.cos(t)], axis=1)


def plan_at(geom, pos, kw=KW):
    g = A.GeometryArrays(A.MLFMA_SRC_DIPOLE, pos=np.ascontiguousarray(pos), mom=geom.mom)
    return g, Plan(g, cases.planewave_z(), A.make_config(device=0, **kw))


def tree_signature(p):
    info = p.info()
    return [tuple(np.concatenate([a.ravel() for a in p.level_boxes(l)]).tolist())
            for l in range(info["max_level"] + 1)]


def dense_Z(p):
    n = p.n
    return np.stack([p.matvec(np.eye(n, dtype=complex)[:, j]) for j in range(n)], axis=1)


def dense_J(p, b, dirs):
    I = np.linalg.solve(dense_Z(p), b)
    E = p.farfield(I, dirs)
    return float(np.sum(np.abs(E) ** 2))


def dense_gradient(p, b, dirs, wave=None):
    Z = dense_Z(p)
    I = np.linalg.solve(Z, b)
    E = p.farfield(I, dirs)
    lam = np.linalg.solve(Z.conj().T, p.farfield_adjoint(E, dirs))
    return AG.position_gradient(p, I, lam, E, dirs, b, wave)


def check_fd(geom, p0, g, Jfun, rng, h, trials, rtol):
    sig = tree_signature(p0)
    for t in range(trials):
        D = rng.standard_normal((geom.n, 3))
        D /= np.linalg.norm(D)
        _, pp = plan_at(geom, geom.pos + h * D)
        _, pm = plan_at(geom, geom.pos - h * D)
        assert tree_signature(pp) == sig and tree_signature(pm) == sig, "perturbation changed the tree"
        fd = (Jfun(pp) - Jfun(pm)) / (2 * h)
        ad = float(np.sum(g * D))
        say("trial", t, "adjoint", ad, "finite difference", fd, "rel", abs(ad - fd) / abs(fd))
        assert abs(ad - fd) <= rtol * abs(fd) + 1e-9 * np.linalg.norm(g)


def test_position_gradient_fixed_b_vs_finite_differences():
    geom = _small_problem()
    _, p = plan_at(geom, geom.pos)
    assert p.info()["max_level"] >= 3
    dirs = dirs_ring()
    b = G.complex_gaussian(geom.n, 21)
    g = dense_gradient(p, b, dirs)
    say("gradient norm", np.linalg.norm(g))
    check_fd(geom, p, g, lambda q: dense_J(q, b, dirs), np.random.default_rng(1), 1e-6, 3, 1e-6)


def test_position_gradient_single_components():
    # components one at a time: a wrong sign or a swapped row/column term in
    # one source's gradient cannot hide in a random combination
    geom = _small_problem()
    _, p = plan_at(geom, geom.pos)
    dirs = dirs_ring(7)
    b = G.complex_gaussian(geom.n, 22)
    g = dense_gradient(p, b, dirs)
    sig = tree_signature(p)
    h = 1e-6
    for i, a in ((0, 0), (5, 1), (17, 2), (40, 0)):
        pos = geom.pos.copy()
        pos[i, a] += h
        _, pp = plan_at(geom, pos)
        pos[i, a] -= 2 * h
        _, pm = plan_at(geom, pos)
        assert tree_signature(pp) == sig and tree_signature(pm) == sig
        fd = (dense_J(pp, b, dirs) - dense_J(pm, b, dirs)) / (2 * h)
        say("source", i, "axis", a, "adjoint", g[i, a], "finite difference", fd)
        assert abs(g[i, a] - fd) <= 1e-5 * abs(fd) + 1e-8 * np.abs(g).max()


def test_position_gradient_planewave_excitation():
    geom = _small_problem()
    _, p = plan_at(geom, geom.pos)
    dirs = dirs_ring()
    wave = p.pw.wavenum * np.asarray(p.pw.khat[:3])
    g = dense_gradient(p, -p.voltage(), dirs, wave)
    check_fd(geom, p, g, lambda q: dense_J(q, -q.voltage(), dirs), np.random.default_rng(2), 1e-6, 2, 1e-6)
    # the GMRES-based helper agrees with the dense solves
    J, lam, g2 = AG.objective_gradients(p, dirs, restart=geom.n + 2, maxit=3 * geom.n, tol=1e-13)
    say("objective", J, "gmres vs dense gradient", cases.relerr(g2, g))
    assert abs(J - dense_J(p, -p.voltage(), dirs)) <= 1e-10 * J
    assert cases.relerr(g2, g) <= 1e-8


def test_position_gradient_more_levels():
    # 300 dipoles, up to 4 per leaf: most interactions go through the far field
    geom = _small_problem(300, 4)
    kw = dict(KW, max_node_srcs=4, interp_order=4, digits=4)
    _, p = plan_at(geom, geom.pos, kw)
    info = p.info()
    say("levels", info["max_level"], "m2l pairs", info["m2l_pairs"], "near nnz", info["near_nnz"])
    assert info["near_nnz"] < 0.5 * geom.n ** 2
    dirs = dirs_ring(9)
    b = G.complex_gaussian(geom.n, 23)
    g = dense_gradient(p, b, dirs)
    sig = tree_signature(p)
    rng = np.random.default_rng(4)
    D = rng.standard_normal((geom.n, 3))
    D /= np.linalg.norm(D)
    h = 1e-6
    _, pp = plan_at(geom, geom.pos + h * D, kw)
    _, pm = plan_at(geom, geom.pos - h * D, kw)
    assert tree_signature(pp) == sig and tree_signature(pm) == sig
    fd = (dense_J(pp, b, dirs) - dense_J(pm, b, dirs)) / (2 * h)
    ad = float(np.sum(g * D))
    say("adjoint", ad, "finite difference", fd)
    assert abs(ad - fd) <= 1e-6 * abs(fd)


def test_position_gradient_errors():
    geom, kw = cases.make("rwg_fixed")
    p = Plan(geom, cases.planewave_z(), A.make_config(device=0, **kw))
    z = np.zeros(geom.n, complex)
    with pytest.raises(MlfmaError):
        p.position_gradient(z, z)
    geom = _small_problem()
    q = Plan(geom, cases.planewave_z(), A.make_config(device=0, **dict(KW, flags=0)))
    if q.info()["stale_boxes"] > 0:
        with pytest.raises(MlfmaError):
            q.position_gradient(np.zeros(geom.n, complex), np.zeros(geom.n, complex))


def test_operator_gradient_against_dense_4096():
    """A 4,096-dipole subset of the 200k-dipole set (seed 2, cube of side 8
    wavelengths): the library's -2 Re(lambda^H dZ I) and far-field term
    against the dense O(N^2) references of tests/_dense_grad.py for random I
    and lambda.  The MLFMA operator differs from the dense one by its
    truncation error, measured here as the matvec error against the dense
    product; the gradient must agree to a small multiple of it.  The far-field
    term involves no approximation (1e-10)."""
    from tests import _dense_grad as DG
    pos, mom = G.uniform_dipoles(200_000, 8.0, 2)
    pos, mom = np.ascontiguousarray(pos[:4096]), np.ascontiguousarray(mom[:4096])
    geom = A.GeometryArrays(A.MLFMA_SRC_DIPOLE, pos=pos, mom=mom)
    # digits 6: the level-2 grid of the adjoint of M2M exceeds one block's LDS,
    # so this also runs its two-pass path
    kw = dict(root_leng=8.0, max_node_srcs=0, max_level=4, interp_order=6, digits=6,
              flags=A.MLFMA_FLAG_FRESH_LOCALS)
    p = Plan(geom, cases.planewave_z(), A.make_config(device=0, **kw))
    n = geom.n
    phat = mom / np.linalg.norm(mom, axis=1)[:, None]
    k = p.pw.wavenum
    I = G.complex_gaussian(n, 41)
    lam = G.complex_gaussian(n, 42)
    y = p.matvec(I)
    e_mv = cases.relerr(y, DG.dense_matvec(pos, phat, I, k))
    zl = p.matvec_adjoint(lam)
    adj = abs(np.vdot(lam, y) - np.vdot(zl, I)) / (np.linalg.norm(y) * np.linalg.norm(lam))
    say("adjoint identity", adj)
    assert adj <= 1e-13
    g = p.position_gradient(I, lam)
    gd = DG.operator_gradient(pos, phat, I, lam, k)
    e_g = cases.relerr(g, gd)
    say("matvec vs dense", e_mv, "operator gradient vs dense", e_g)
    # the reference's operator (its psi-interpolated translations) is accurate
    # to a few 1e-3 here (5.6e-3 measured at 4 digits)
    assert e_mv < 2e-2
    assert e_g <= max(20 * e_mv, 1e-8)
    dirs = dirs_ring(13)
    E = p.farfield(I, dirs)
    gf = AG.position_gradient(p, I, lam, E, dirs) - g
    gfd = DG.farfield_gradient(pos, phat, I, E, dirs, k)
    say("far-field term vs dense", cases.relerr(gf, gfd))
    assert cases.relerr(gf, gfd) <= 1e-10
