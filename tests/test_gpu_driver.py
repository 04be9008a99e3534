"""The `mlfma` executable (the reference's main() flow, src/mlfma.cpp:12-110,
over the C-ABI) on files in the reference's formats (config/config.txt,
config/pwave.txt, config/rwg/n<N>/{vertices,faces,rwgs}.txt or a dipole file):
its written solutions must equal the Python API's solve of the same inputs
through the same library, for the FMM plan and for the direct plan of
evalDirect (a single root leaf, src/mlfma.cpp:84-107)."""
import os
import subprocess

import numpy as np
import pytest

from mlfma_amd import _capi as A
from mlfma_amd import build as B
from mlfma_amd import geometry as G
from mlfma_amd.plan import Plan
from tests import cases

pytestmark = pytest.mark.gpu


def read_sol(path):
    a = np.loadtxt(path, ndmin=2)
    return a[:, 0] + 1j * a[:, 1]


def run_driver(root, extra):
    cmd = [B.DRIVER, "--root", str(root), "--device", "0"] + extra
    r = subprocess.run(cmd, cwd=str(root), capture_output=True, text=True, timeout=240)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0, r.stderr


def write_common(root, kw, n, eval_direct):
    cfg = G.RefConfig(quadPrec=kw.get("quad_prec", 2), digits=kw["digits"], interpOrder=kw["interp_order"],
                      overInterp=4.0, rootLeng=kw["root_leng"], maxNodeSrcs=kw["max_node_srcs"],
                      evalDirect=eval_direct, nsrcs=n)
    G.write_config(os.path.join(root, "config", "config.txt"), cfg)
    G.write_pwave(os.path.join(root, "config", "pwave.txt"), 1.0, (1.0, 0.0, 0.0), (0.0, 0.0, 1.0), 2 * np.pi)


def test_driver_rwg_matches_api(tmp_path):
    geom, kw = cases.make("rwg_fixed")
    n = geom.n
    write_common(tmp_path, kw, n, 1)
    G.write_rwg_dir(os.path.join(tmp_path, "config", "rwg", f"n{n}"), geom.verts, geom.faces, geom.rwgs)
    run_driver(tmp_path, ["--method", "gmres", "--restart", "30", "--maxit", "120", "--tol", "1e-8",
                          "--max-level", str(kw["max_level"])])
    x_drv = read_sol(os.path.join(tmp_path, "out", "sol", "curr_nq7.txt"))
    pw = cases.planewave_z()
    p = Plan(geom, pw, A.make_config(device=0, **kw))
    x, rep = p.solve(method=A.MLFMA_GMRES_RESTART, restart=30, maxit=120, tol=1e-8)
    print("fmm", rep, flush=True)
    assert len(x_drv) == n and cases.relerr(x_drv, x) <= 1e-12
    # evalDirect: the same solve on a single root leaf
    x_dir = read_sol(os.path.join(tmp_path, "out", "sol", "currDir_nq7.txt"))
    kd = dict(kw, max_node_srcs=n, max_level=0)
    pd = Plan(geom, pw, A.make_config(device=0, **kd))
    xd, repd = pd.solve(method=A.MLFMA_GMRES_RESTART, restart=30, maxit=120, tol=1e-8)
    assert cases.relerr(x_dir, xd) <= 1e-12


def test_driver_dipoles_compat_matches_api(tmp_path):
    geom, kw = cases.make("dipole_fixed")
    n = geom.n
    write_common(tmp_path, kw, n, 0)
    # x-directed moments so that the x-polarised plane wave excites them
    mom = np.tile([1e-8, 0.0, 0.0], (n, 1))
    G.write_dipoles(os.path.join(tmp_path, "dip.txt"), geom.pos, mom)
    run_driver(tmp_path, ["--method", "compat", "--maxit", "60", "--tol", "1e-6", "--max-level", str(kw["max_level"]),
                          "--dipoles", os.path.join(tmp_path, "dip.txt")])
    x_drv = read_sol(os.path.join(tmp_path, "out", "sol", "curr_nq7.txt"))
    g2 = A.GeometryArrays(A.MLFMA_SRC_DIPOLE, pos=geom.pos, mom=mom)
    p = Plan(g2, cases.planewave_z(), A.make_config(device=0, **kw))
    x, rep = p.solve(method=A.MLFMA_GMRES_COMPAT, maxit=60, tol=1e-6)
    print("compat", rep, flush=True)
    assert np.any(x != 0) and cases.relerr(x_drv, x) <= 1e-12
