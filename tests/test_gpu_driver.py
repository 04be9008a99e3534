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


def write_config_lines(root, kw, n, eval_direct):
    """config.txt with several numbers on some lines (the reference keeps the
    last one, src/config.h:37-44)."""
    lines = ["mode (0 read, 1 write): 0", "pdist: 0", "qdist: 0", f"quadPrec (0-3): {kw.get('quad_prec', 2)}",
             f"digits (1-15): {kw['digits']}", f"interpOrder: {kw['interp_order']}", "overInterp (x nth-1): 4.0",
             f"rootLeng 1.0 or {kw['root_leng']}", f"maxNodeSrcs: {kw['max_node_srcs']}",
             f"evalDirect (0/1): {eval_direct}", f"nsrcs: {n}"]
    os.makedirs(os.path.join(root, "config"), exist_ok=True)
    with open(os.path.join(root, "config", "config.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")
    G.write_pwave(os.path.join(root, "config", "pwave.txt"), 1.0, (1.0, 0.0, 0.0), (0.0, 0.0, 1.0), 2 * np.pi)


@pytest.mark.parametrize("name", ["rwg_fixed", "dipole_fixed"])
def test_driver_compat_matches_oracle(tmp_path, name):
    """The executable's compat GMRES solution on reference-format files against
    the oracle's restatement of the reference's solve (src/solver.cpp:116-146)
    on the same inputs, read back from the same files."""
    from tests import _oracle as O
    geom, kw = cases.make(name)
    n = geom.n
    write_config_lines(tmp_path, kw, n, 0)
    extra = []
    if geom.kind == A.MLFMA_SRC_RWG:
        G.write_rwg_dir(os.path.join(tmp_path, "config", "rwg", f"n{n}"), geom.verts, geom.faces, geom.rwgs)
    else:
        G.write_dipoles(os.path.join(tmp_path, "dip.txt"), geom.pos, geom.mom)
        extra = ["--dipoles", os.path.join(tmp_path, "dip.txt")]
    run_driver(tmp_path, ["--method", "compat", "--maxit", "40", "--tol", "1e-6", "--max-level",
                          str(kw["max_level"])] + extra)
    x_drv = read_sol(os.path.join(tmp_path, "out", "sol", "curr_nq7.txt"))
    rc = G.read_config(os.path.join(tmp_path, "config", "config.txt"))
    pwf = G.read_pwave(os.path.join(tmp_path, "config", "pwave.txt"))
    assert (rc.digits, rc.interpOrder, rc.rootLeng) == (kw["digits"], kw["interp_order"], kw["root_leng"])
    if geom.kind == A.MLFMA_SRC_RWG:
        v, f, r = G.read_rwg_dir(os.path.join(tmp_path, "config", "rwg", f"n{rc.nsrcs}"))
        g2 = A.GeometryArrays(A.MLFMA_SRC_RWG, verts=v, faces=f, rwgs=r)
    else:
        pos, mom = G.read_dipoles(os.path.join(tmp_path, "dip.txt"))
        g2 = A.GeometryArrays(A.MLFMA_SRC_DIPOLE, pos=pos, mom=mom)
    cfg = A.make_config(device=-1, root_leng=rc.rootLeng, max_node_srcs=rc.maxNodeSrcs, max_level=kw["max_level"],
                        interp_order=rc.interpOrder, digits=rc.digits, over_interp=rc.overInterp,
                        quad_prec=rc.quadPrec)
    pw = A.make_planewave(pwf["amplitude"], pwf["pol"], pwf["khat"], pwf["wavenum"])
    o = O.Oracle(g2, pw, cfg)
    xo, res, it, orc = o.solve(maxit=40, tol=1e-6)
    err = cases.relerr(x_drv, xo)
    print(name, "driver vs oracle", err, "iterations", it, flush=True)
    # the files hold 16 significant digits; the GPU and the oracle agree to
    # 1e-10 on these cases (RWG: the tolerance policy of test_gpu_parity)
    assert len(x_drv) == n and np.any(xo != 0) and err <= 1e-9
