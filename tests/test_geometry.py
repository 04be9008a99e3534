"""Mesh generators and the reference's on-disk formats (src/config.h:50-58,
src/fileio.h:8-202).  CPU only."""
import os

import numpy as np
import pytest

from mlfma_amd import geometry as G


@pytest.mark.parametrize("freq", [1, 2, 5, 8])
def test_icosphere_counts_and_closure(freq):
    v, f = G.icosphere(freq, 2.0)
    assert len(v) == 10 * freq ** 2 + 2
    assert len(f) == 20 * freq ** 2
    assert np.allclose(np.linalg.norm(v, axis=1), 2.0)
    r = G.rwg_edges(f)
    assert len(r) == 30 * freq ** 2                  # closed surface: every edge is an RWG
    # each RWG's edge is shared by its two triangles, which differ
    for v0, v1, tp, tm in r[:: max(1, len(r) // 50)]:
        assert tp != tm
        assert {v0, v1} <= set(f[tp]) and {v0, v1} <= set(f[tm])
    # outward orientation
    c = v[f].mean(axis=1)
    nrm = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    assert ((nrm * c).sum(axis=1) > 0).all()


def test_sphere_workload_sizes():
    # R = 2 wavelengths, geodesic frequency 32: 30,720 RWG unknowns
    v, f = G.icosphere(32, 2.0)
    r = G.rwg_edges(f)
    s = G.mesh_stats(v, f, r)
    assert s["nrwg"] == 30720 and s["ntris"] == 20480
    assert 0.06 < s["mean_edge"] < 0.08


def test_ellipsoid_axes():
    v, f = G.ellipsoid(4, (2.0, 1.0, 0.5))
    assert np.allclose(np.abs(v).max(axis=0), [2.0, 1.0, 0.5], atol=1e-12)


def test_file_round_trips(tmp_path):
    cfg = G.RefConfig(quadPrec=2, digits=3, interpOrder=6, overInterp=4.0, rootLeng=4.0, maxNodeSrcs=20,
                      evalDirect=0, nsrcs=30720)
    p = os.path.join(tmp_path, "config", "config.txt")
    G.write_config(p, cfg)
    back = G.read_config(p)
    for k in ["quadPrec", "digits", "interpOrder", "overInterp", "rootLeng", "maxNodeSrcs", "nsrcs"]:
        assert getattr(back, k) == getattr(cfg, k)
    G.write_pwave(os.path.join(tmp_path, "pwave.txt"), 1.0, (1, 0, 0), (0, 0, 1), 2 * np.pi)
    pw = G.read_pwave(os.path.join(tmp_path, "pwave.txt"))
    assert pw["wavenum"] == 2 * np.pi and pw["pol"] == [1, 0, 0]
    pos, mom = G.uniform_dipoles(7, 2.0, 1)
    G.write_dipoles(os.path.join(tmp_path, "d.txt"), pos, mom)
    p2, m2 = G.read_dipoles(os.path.join(tmp_path, "d.txt"))
    assert np.array_equal(p2, pos) and np.array_equal(m2, mom)
    v, f = G.icosphere(2, 1.0)
    r = G.rwg_edges(f)
    G.write_rwg_dir(os.path.join(tmp_path, "rwg"), v, f, r)
    v2, f2, r2 = G.read_rwg_dir(os.path.join(tmp_path, "rwg"))
    assert np.array_equal(v2, v) and np.array_equal(f2, f) and np.array_equal(r2, r)


def test_seeded_inputs_are_reproducible():
    a, _ = G.uniform_dipoles(5, 4.0, 0)
    b, _ = G.uniform_dipoles(5, 4.0, 0)
    assert np.array_equal(a, b) and a.min() >= -2 and a.max() <= 2
    x = G.complex_gaussian(20000, 0)
    assert abs(np.mean(np.abs(x) ** 2) - 1) < 0.05
