"""torch.autograd through the dipole positions (mlfma_amd.autograd.DipoleScatterer):
E = F(r) (Z(r) + s I)^-1 b, J = sum |E|^2; pos.grad and b.grad against central
finite differences of J (positions moved along random directions that keep
every dipole inside its leaf box, so the tree is the same; h = 1e-6
wavelength), with the forward and adjoint solves on the device."""
import numpy as np
import pytest
import torch

from mlfma_amd import _capi as A
from mlfma_amd import autograd as AG
from mlfma_amd import configs as CF
from mlfma_amd import geometry as G

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def inside_leaf_mask(pos, root, leaf, margin=1e-5):
    """1 for dipoles farther than margin from every face of their leaf box."""
    f = (pos + root / 2) / leaf
    d = np.abs(f - np.round(f)) * leaf
    return (d.min(axis=1) > margin).astype(float)


def self_shift():
    """A real diagonal of 30 |C k| k^3 (a dipole self term; the reference has none)."""
    ck = 299792458.0 * 1.256637e-6 / (4 * np.pi) * 2 * np.pi
    return 30.0 * ck * (2 * np.pi) ** 3


def fd_check(sc, b0, pos0, root, leaf, trials, h=1e-6, rtol=1e-6, seed=0):
    b = torch.from_numpy(b0).requires_grad_(True)
    pos = torch.from_numpy(pos0.copy()).requires_grad_(True)
    E = sc(b, pos)
    J = (E.abs() ** 2).sum()
    J.backward()
    gp, gb = pos.grad.numpy(), b.grad.numpy()
    rng = np.random.default_rng(seed)
    mask = inside_leaf_mask(pos0, root, leaf)[:, None]

    def Jat(bb, pp):
        with torch.no_grad():
            return float((sc(torch.from_numpy(bb), torch.from_numpy(pp)).abs() ** 2).sum())
    # directions with a large component along the gradient (a random unit
    # direction in 3n dimensions sees only ~1/sqrt(3n) of it, and the
    # difference would drown in the solve tolerance): the gradient itself,
    # then the gradient plus a random direction of the same length
    gm = gp * mask
    gm /= np.linalg.norm(gm)
    gbn = gb / np.linalg.norm(gb)
    for t in range(trials):
        R = rng.standard_normal(pos0.shape) * mask
        D = gm + (t > 0) * R / np.linalg.norm(R)
        D /= np.linalg.norm(D)
        fd = (Jat(b0, pos0 + h * D) - Jat(b0, pos0 - h * D)) / (2 * h)
        ad = float(np.sum(gp * D))
        say("positions trial", t, "autograd", ad, "finite difference", fd, "rel", abs(ad - fd) / abs(fd))
        assert abs(ad - fd) <= rtol * abs(fd)
        rb = G.complex_gaussian(len(b0), 100 + t)
        db = gbn + (t > 0) * rb / np.linalg.norm(rb)
        db /= np.linalg.norm(db)
        fdb = (Jat(b0 + h * db, pos0) - Jat(b0 - h * db, pos0)) / (2 * h)
        adb = float(np.real(np.vdot(gb, db)))
        say("excitation trial", t, "autograd", adb, "finite difference", fdb, "rel", abs(adb - fdb) / abs(fdb))
        assert abs(adb - fdb) <= rtol * abs(fdb)
    return float(J.detach())


def test_autograd_positions_small():
    n = 500
    pos, mom = G.uniform_dipoles(n, 1.8, 7)
    mom = np.random.default_rng(7).normal(size=(n, 3))
    kw = dict(root_leng=2.0, max_node_srcs=0, max_level=3, interp_order=6, digits=6)
    dirs = CF.c5_dirs()[::10]
    sc = AG.DipoleScatterer(mom, dirs, kw, pw=CF.planewave_theta0(), shift=self_shift(), tol=1e-13)
    fd_check(sc, G.complex_gaussian(n, 2), pos, 2.0, 0.25, trials=2, rtol=1e-6)


@pytest.mark.timeout(900)
def test_c5_gradient_full_size():
    """C5 at its own size: 200,000 dipoles in a cube of side 8 wavelengths
    (seed 2), 360 bistatic angles; the forward and adjoint solves of the
    shifted operator converge, and the gradient matches central differences
    along random directions."""
    geom, pw, kw, desc = CF.c5()
    kw = {k: v for k, v in kw.items() if k != "flags"}
    sc = AG.DipoleScatterer(geom.mom, CF.c5_dirs(), kw, pw=pw, shift=self_shift(), tol=1e-12)
    J = fd_check(sc, G.complex_gaussian(geom.n, 2), geom.pos, 8.0, 0.5, trials=3, rtol=1e-6)
    say("C5 J", J)
