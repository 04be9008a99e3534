"""Benchmark: MLFMA matvec throughput (and GMRES solve time) on one MI355X.

Workload (default): plain FP64 RWG EFIE on a PEC sphere of radius 2 wavelengths,
geodesic icosphere of frequency 32 (20,480 triangles, 30,720 RWG unknowns),
7-point triangle quadrature, leaf boxes of 0.25 wavelength (root 4, levels
0-4), interpolation order 6, 3 digits, plane wave theta = 0 x-polarised.

A step is one full matvec y = Z x (S2M, M2M, M2L, L2L, near field, L2T) with
x and y resident in HBM.  W untimed warmup steps, then K timed steps bracketed
by torch.cuda.synchronize() give `value`; one JSON line.

The line also carries:
  stage_ms      per-stage device times from HIP events on the library's stream,
                recorded over a second pass of the same K steps;
  roofline      the near-field pair kernel's achieved algorithmic GB/s (its
                launch time from the events of that second pass) against 8 TB/s;
  batched       8 right-hand sides per call (each operator read shared by 8);
  cpu_baseline  the CPU oracle (FP64 restatement of the reference, one core)
                timed on a bounded number of matvecs of the same workload;
  solve         GMRES(50) to 1e-6 in FP64 and the FP32-far-field refinement
                solve (the reference's kernel), with iteration counts and true
                residuals; and GMRES(50) with the physical near field plus the
                bistatic RCS at 361 angles against the Mie series.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name):
    from mlfma_amd import _capi as A
    from mlfma_amd import geometry as G
    if name == "sphere30k":
        v, f = G.icosphere(32, 2.0)
        r = G.rwg_edges(f)
        geom = A.GeometryArrays(A.MLFMA_SRC_RWG, verts=v, faces=f, rwgs=r)
        kw = dict(root_leng=4.0, max_node_srcs=0, max_level=4, interp_order=6, digits=3, over_interp=4.0,
                  quad_prec=2)
        pw = A.make_planewave(1.0, (1.0, 0.0, 0.0), (0.0, 0.0, 1.0), 2 * np.pi)
        desc = {"workload": "rwg_sphere_R2_icosphere32", "unknowns": int(len(r)), "triangles": int(len(f)),
                "mean_edge_wavelengths": round(G.mesh_stats(v, f, r)["mean_edge"], 4), "leaf_wavelengths": 0.25,
                "levels": "0-4", "interp_order": 6, "digits": 3, "over_interp": 4.0, "quad_points": 7,
                "excitation": "plane wave theta=0 x-pol"}
        return geom, pw, kw, desc
    if name == "sphere200k":
        v, f = G.icosphere(82, 4.0)
        r = G.rwg_edges(f)
        geom = A.GeometryArrays(A.MLFMA_SRC_RWG, verts=v, faces=f, rwgs=r)
        kw = dict(root_leng=8.0, max_node_srcs=0, max_level=5, interp_order=6, digits=3, over_interp=4.0,
                  quad_prec=2)
        pw = A.make_planewave(1.0, (1.0, 0.0, 0.0), (0.0, 0.0, 1.0), 2 * np.pi)
        desc = {"workload": "rwg_sphere_R4_icosphere82", "unknowns": int(len(r)), "triangles": int(len(f)),
                "mean_edge_wavelengths": round(G.mesh_stats(v, f, r)["mean_edge"], 4), "leaf_wavelengths": 0.25,
                "levels": "0-5", "interp_order": 6, "digits": 3, "over_interp": 4.0, "quad_points": 7,
                "excitation": "plane wave theta=0 x-pol"}
        return geom, pw, kw, desc
    if name == "dipoles4k":
        pos, mom = G.uniform_dipoles(4096, 4.0, 0)
        geom = A.GeometryArrays(A.MLFMA_SRC_DIPOLE, pos=pos, mom=mom)
        kw = dict(root_leng=4.0, max_node_srcs=0, max_level=3, interp_order=6, digits=3, over_interp=4.0)
        pw = A.make_planewave(1.0, (1.0, 0.0, 0.0), (0.0, 0.0, 1.0), 2 * np.pi)
        desc = {"workload": "dipoles_4096_cube4", "unknowns": 4096, "leaf_wavelengths": 0.5, "levels": "0-3",
                "interp_order": 6, "digits": 3}
        return geom, pw, kw, desc
    raise SystemExit(f"unknown workload {name}")


def measured_traffic(kernel, tag):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary (calibrated against a plain copy), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{tag}_traffic.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if line.startswith("#") or not line.strip():
            continue
        k, v = line.split()[:2]
        vals[k] = v
    return float(vals["traffic_bytes"]) if "traffic_bytes" in vals else None


def near_bytes(info):
    # algorithmic bytes of one kNearSym launch: every stored near-field value
    # (FP64 complex, 16 B) is read once; the column indices, x gathers and
    # partial sums are a few percent on top and are not counted
    return info["near_nnz"] * 16


def byte_model(info, n):
    """Algorithmic bytes of one FP64 matvec per stage (c = 16 B, two
    components per direction): S2M = L2T = (patterns + leaf fields), M2L = per
    level pairs x directions x 2c (source read per pair) + 316 operators x
    directions x c (once) + boxes x directions x 2c (locals), M2M = L2L =
    child and parent fields once per level, near = stored values x c.  Also
    the compulsory bytes: every array touched once (patterns, near values,
    operators, multipole and local fields, x and y)."""
    c = 16
    ml = info["max_level"]
    nd = [info["nth"][l] * info["nph"][l] for l in range(ml + 1)]
    B = info["boxes"]
    P = info["m2l_pairs"]
    leaf_fields = info["num_leaves"] * nd[ml] * 2 * c
    s2m = info["radpat_entries"] * 2 * c + leaf_fields
    m2l = sum(P[l] * nd[l] * 2 * c + 316 * nd[l] * c + B[l] * nd[l] * 2 * c for l in range(2, ml + 1) if P[l] > 0)
    m2m = sum((B[l + 1] * nd[l + 1] + B[l] * nd[l]) * 2 * c for l in range(2, ml))
    near = info["near_nnz"] * c
    stages = {"S2M": s2m, "M2M": m2m, "M2L": m2l, "L2L": m2m, "L2T": s2m, "near": near}
    fields = sum(B[l] * nd[l] * 2 * c for l in range(2, ml + 1))
    compulsory = (info["radpat_entries"] * 2 * c + near + sum(316 * nd[l] * c for l in range(2, ml + 1) if P[l] > 0)
                  + 2 * fields + 2 * n * c)
    return stages, compulsory


def physical_solve(geom, pw, kw, radius=2.0, restart=50, maxit=3000):
    """GMRES(restart) to 1e-6 with the physical near field (closed-form
    singular integrals) on the sphere, and its bistatic RCS at 361 angles (theta 0-180
    degrees in 0.5-degree steps, phi = 0 and 90) against the Mie series."""
    from mlfma_amd import _capi as A
    from mlfma_amd.plan import Plan
    from tests import _physical as P
    t0 = time.time()
    p = Plan(geom, pw, A.make_config(device=0, flags=A.MLFMA_FLAG_PHYSICAL, **kw))
    setup = time.time() - t0
    p.autotune(1)  # launch parameters for single right-hand sides (not timed)
    I, rep = p.solve(method=A.MLFMA_GMRES_RESTART, restart=restart, maxit=maxit, tol=1e-6)
    log("physical GMRES", rep["iterations"], rep["status"], round(rep["seconds"], 3), "s")
    # the same with the FP32 far field inside FP64 iterative refinement
    p32 = Plan(geom, pw, A.make_config(device=0, precision=A.MLFMA_FP32, flags=A.MLFMA_FLAG_PHYSICAL, **kw))
    p32.autotune(1)
    I32, rep32 = p32.solve(method=A.MLFMA_FGMRES_IR, restart=restart, maxit=maxit, tol=1e-6)
    log("physical FP32-IR", rep32["iterations"], rep32["status"], round(rep32["seconds"], 3), "s")
    mixed = {"method": "FP64 refinement, FP32 far field, MLFMA_FLAG_PHYSICAL", "iterations": rep32["iterations"],
             "status": rep32["status"], "resid_true": rep32["resid_true"], "seconds": rep32["seconds"],
             "solution_vs_fp64_rel": float(np.linalg.norm(I32 - I) / np.linalg.norm(I))}
    # eight incidence angles theta_j = j 180/7 degrees (phi = 0, theta-polarised):
    # one batched solve against eight single solves
    angles = np.radians(np.arange(8) * 180.0 / 7.0)
    B = np.stack([-p.excitation(A.make_planewave(1.0, (np.cos(t), 0.0, -np.sin(t)), (np.sin(t), 0.0, np.cos(t)),
                                                 2 * np.pi)) for t in angles])
    ts = time.perf_counter()
    singles = []
    for j in range(8):
        singles.append(p.solve(B[j], method=A.MLFMA_GMRES_RESTART, restart=restart, maxit=maxit, tol=1e-6))
        log("physical single", j, singles[-1][1]["iterations"])
    t_single = time.perf_counter() - ts
    p.autotune(8)  # launch parameters for batches of 8 (not timed)
    tb = time.perf_counter()
    Xb, repb = p.solve_batch(B, restart=restart, maxit=maxit, tol=1e-6)
    t_batch = time.perf_counter() - tb
    log("physical batched sweep", round(t_batch, 3), "s")
    # batched and mixed precision together: FP64 refinement around batched
    # FP32-far-field GMRES, against the plain FP64 single solves above
    p32.autotune(8)
    tm = time.perf_counter()
    Xm, repm = p32.solve_batch(B, restart=restart, maxit=maxit, tol=1e-6, method=A.MLFMA_FGMRES_IR)
    t_mixed = time.perf_counter() - tm
    log("physical batched mixed sweep", round(t_mixed, 3), "s")
    dirs = []
    for phi in (0.0, np.pi / 2):
        t_ = np.radians(np.linspace(0.0, 180.0, 361))
        dirs.append(np.stack([np.sin(t_) * np.cos(phi), np.sin(t_) * np.sin(phi), np.cos(t_)], axis=1))
    dirs = np.concatenate(dirs)
    rcs_diff = 0.0
    for j in range(8):
        s64 = np.sum(np.abs(p.farfield(singles[j][0], dirs)) ** 2, axis=1)
        smx = np.sum(np.abs(p.farfield(Xm[j], dirs)) ** 2, axis=1)
        rcs_diff = max(rcs_diff, float(np.abs(smx - s64).max() / s64.max()))
    sweep_mixed = {"seconds": round(t_mixed, 4), "speedup_vs_sequential_fp64": round(t_single / t_mixed, 3),
                   "all_converged": all(r["status"] == "OK" for r in repm),
                   "max_resid_true": max(r["resid_true"] for r in repm),
                   "rcs_max_rel_diff_vs_fp64": rcs_diff}
    sweep = {"angles": 8, "batched_seconds": round(t_batch, 4), "sequential_seconds": round(t_single, 4),
             "batched_mixed": sweep_mixed,
             "iterations": [r["iterations"] for r in repb], "all_converged": all(r["status"] == "OK" for r in repb),
             "max_resid_true": max(r["resid_true"] for r in repb),
             "batched_vs_sequential_solution_rel": max(float(np.linalg.norm(Xb[j] - singles[j][0]) /
                                                             np.linalg.norm(singles[j][0])) for j in range(8))}
    th = np.radians(np.linspace(0.0, 180.0, 361))
    se, sh = P.mie_pec_rcs(2 * np.pi * radius, th)
    errs = []
    for phi, ref in ((0.0, se), (np.pi / 2, sh)):
        d = np.stack([np.sin(th) * np.cos(phi), np.sin(th) * np.sin(phi), np.cos(th)], axis=1)
        E = p.farfield(I, d)
        errs.append(10 * np.log10(4 * np.pi * np.sum(np.abs(E) ** 2, axis=1)) - 10 * np.log10(ref))
    errs = np.concatenate(errs)
    return {"method": f"GMRES({restart}), FP64, MLFMA_FLAG_PHYSICAL", "maxit": maxit, "iterations": rep["iterations"],
            "status": rep["status"], "resid_true": rep["resid_true"], "seconds": rep["seconds"],
            "plan_setup_s": round(setup, 3),
            "rcs_vs_mie_db": {"angles": 2 * len(th), "rms": round(float(np.sqrt(np.mean(errs ** 2))), 4),
                              "max": round(float(np.abs(errs).max()), 4)},
            "mixed": mixed, "incidence_sweep": sweep}


def cpu_baseline(geom, pw, kw, budget_s=20.0, y_gpu=None):
    from mlfma_amd import _capi as A
    from tests import _oracle as O
    from mlfma_amd import geometry as G
    t0 = time.time()
    nt = max(1, min(16, os.cpu_count() or 1))
    o = O.Oracle(geom, pw, A.make_config(device=-1, **kw), threads=nt)
    setup = time.time() - t0
    x = G.complex_gaussian(geom.n, 0)
    y_cpu = o.matvec(x)  # warm; also the parity check of the timed GPU matvec
    parity = None
    if y_gpu is not None:
        parity = float(np.linalg.norm(y_gpu - y_cpu) / np.linalg.norm(y_cpu))
    times = []
    t_start = time.time()
    while time.time() - t_start < budget_s and len(times) < 20:
        t = time.time()
        o.matvec(x)
        times.append(time.time() - t)
    med = float(np.median(times))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": round(1.0 / med, 4), "unit": "matvec/s", "cores": 1, "kind": "port",
            "host": {"cpu_model": model, "logical_cpus": os.cpu_count()},
            "gpu_vs_oracle_rel_diff": parity,
            "sample": f"{len(times)} FP64 matvecs of the same workload by the CPU oracle (single-threaded "
                      f"restatement of the reference); median {med:.3f} s; oracle setup {setup:.1f} s on {nt} threads "
                      f"(not timed)"}


def gradient_job():
    """The gradient pipeline at 200,000 dipoles (uniform-random positions,
    seed 2, cube of side 8 wavelengths, complex-Gaussian right-hand side):
    J = sum over 360 bistatic angles (theta = 0..359 degrees in the phi = 0
    plane) of |E_far|^2; forward GMRES, adjoint GMRES and the position-gradient
    pass, timed.  The dipole operator has a zero diagonal (the reference drops
    the self term), and at this density GMRES(50) does not converge; the solves
    run a fixed 500 iterations and their residuals are reported, so this
    measures the cost of the pipeline, not a converged gradient (that is
    checked on smaller systems by the tests)."""
    import torch
    from mlfma_amd import _capi as A
    from mlfma_amd import geometry as G
    from mlfma_amd import autograd as AG
    from mlfma_amd.plan import Plan
    n = 200_000
    pos, mom = G.uniform_dipoles(n, 8.0, 2)
    geom = A.GeometryArrays(A.MLFMA_SRC_DIPOLE, pos=pos, mom=mom)
    kw = dict(root_leng=8.0, max_node_srcs=0, max_level=4, interp_order=6, digits=3, over_interp=4.0,
              flags=A.MLFMA_FLAG_FRESH_LOCALS)
    pw = A.make_planewave(1.0, (1.0, 0.0, 0.0), (0.0, 0.0, 1.0), 2 * np.pi)
    th = np.deg2rad(np.arange(360.0))
    dirs = np.stack([np.sin(th), 0 * th, np.cos(th)], axis=1)
    b = G.complex_gaussian(n, 2)
    t0 = time.time()
    plan = Plan(geom, pw, A.make_config(device=0, **kw))
    setup_s = time.time() - t0
    info = plan.info()
    log("plan", {k: info[k] for k in ["max_level", "boxes", "m2l_pairs", "near_nnz", "device_bytes"]},
        "setup s", round(setup_s, 2))
    plan.autotune(1)
    opts = dict(method=A.MLFMA_GMRES_RESTART, restart=50, maxit=500, tol=1e-10)
    torch.cuda.synchronize()
    t0 = time.time()
    I, rep_f = plan.solve(b, **opts)
    t_fwd = time.time() - t0
    t0 = time.time()
    E = plan.farfield(I, dirs)
    g_E = plan.farfield_adjoint(E, dirs)
    t_ff = time.time() - t0
    t0 = time.time()
    lam, rep_a = plan.solve(g_E, adjoint=True, **opts)
    t_adj = time.time() - t0
    t0 = time.time()
    grad = AG.position_gradient(plan, I, lam, E, dirs)
    t_grad = time.time() - t0
    log("forward", rep_f, "adjoint", rep_a, "gradient s", round(t_grad, 3))
    keys = ["iterations", "status", "resid_true"]
    out = {
        "metric": "adjoint gradient pipeline of a far-field objective, seconds (FP64, one MI355X)",
        "value": round(t_fwd + t_ff + t_adj + t_grad, 3), "unit": "s", "n_gpus": 1, "higher_is_better": False,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "dipoles_200k_cube8_gradient", "unknowns": n, "angles": 360, "leaf_wavelengths": 0.5,
                   "levels": "0-4", "interp_order": 6, "digits": 3, "gmres": "GMRES(50), 500 iterations"},
        "setup_s": round(setup_s, 2),
        "forward": {"s": round(t_fwd, 3), "ms_per_iteration": round(1e3 * t_fwd / max(rep_f["iterations"], 1), 3),
                    **{k: rep_f[k] for k in keys}},
        "adjoint": {"s": round(t_adj, 3), "ms_per_iteration": round(1e3 * t_adj / max(rep_a["iterations"], 1), 3),
                    **{k: rep_a[k] for k in keys}},
        "farfield_and_adjoint_s": round(t_ff, 4), "position_gradient_s": round(t_grad, 4),
        "gradient_norm": float(np.linalg.norm(grad)),
        "note": "zero-diagonal dipole operator: GMRES(50) stagnates at this density, so the solves stop at 500 "
                "iterations; gradient correctness is pinned by the finite-difference and dense-gradient tests",
    }
    print(json.dumps(out), flush=True)


def multi_scatterer_job():
    """16 scatterers of uneven size (seed 1: sphere or 2:1:1 ellipsoid, equivalent
    radius uniform in 0.5-3 wavelengths; icosphere frequency 12 per wavelength
    of radius, of the short semi-axis for ellipsoids), each under 8 plane waves (theta_j = j 180/7 degrees, phi = 0,
    x-z polarisation in the plane of incidence): 128 right-hand sides.  Each
    scatterer's 8 right-hand sides go through one batched FP64-refined solve
    with an FP32 far field (physical near field, leaf 0.25 wavelength,
    GMRES(50) to 1e-6); the scatterers run concurrently (one plan and stream
    set each) and, for comparison, one after another."""
    import torch
    from mlfma_amd import _capi as A
    from mlfma_amd import geometry as G
    from mlfma_amd.plan import Plan
    from mlfma_amd.batch import solve_many
    rng = np.random.default_rng(1)
    shapes = []
    for i in range(16):
        r_eq = rng.uniform(0.5, 3.0)
        ell = rng.uniform() < 0.5
        if ell:
            a = r_eq / 2 ** (1 / 3)
            axes = (2 * a, a, a)
            freq = max(4, int(round(12 * a)))
            v, f = G.ellipsoid(freq, axes)
            ext = 2 * a
        else:
            freq = max(4, int(round(12 * r_eq)))
            v, f = G.icosphere(freq, r_eq)
            ext = r_eq
        r = G.rwg_edges(f)
        L = 2
        while 0.25 * 2 ** L < 2.05 * ext:
            L += 1
        kw = dict(root_leng=0.25 * 2 ** L, max_node_srcs=0, max_level=L, interp_order=6, digits=3, over_interp=4.0,
                  quad_prec=2)
        shapes.append(dict(kind="ellipsoid" if ell else "sphere", r_eq=round(float(r_eq), 3), n=int(len(r)),
                           levels=L, geom=A.GeometryArrays(A.MLFMA_SRC_RWG, verts=v, faces=f, rwgs=r), kw=kw))
    k = 2 * np.pi
    angles = [j * np.pi / 7 for j in range(8)]
    t0 = time.time()
    plans, rhs = [], []
    for sh in shapes:
        p = Plan(sh["geom"], A.make_planewave(1.0, (1.0, 0.0, 0.0), (0.0, 0.0, 1.0), k),
                 A.make_config(device=0, precision=A.MLFMA_FP32,
                               flags=A.MLFMA_FLAG_PHYSICAL | A.MLFMA_FLAG_FRESH_LOCALS, **sh["kw"]))
        B = np.stack([-p.excitation(A.make_planewave(1.0, (np.cos(t), 0.0, -np.sin(t)), (np.sin(t), 0.0, np.cos(t)), k))
                      for t in angles])
        plans.append(p)
        rhs.append(B)
    setup_s = time.time() - t0
    log("scatterers", [(sh["kind"], sh["r_eq"], sh["n"]) for sh in shapes], "setup s", round(setup_s, 2))
    opts = dict(restart=50, maxit=3000, tol=1e-6, method=A.MLFMA_FGMRES_IR)
    torch.cuda.synchronize()
    t0 = time.time()
    res_c = solve_many(plans, rhs, **opts)
    t_conc = time.time() - t0
    t0 = time.time()
    res_s = [p.solve_batch(B, **opts) for p, B in zip(plans, rhs)]
    t_seq = time.time() - t0
    diff = max(float(np.linalg.norm(a[0] - b[0]) / np.linalg.norm(b[0])) for a, b in zip(res_c, res_s))
    per = []
    for sh, (X, reps) in zip(shapes, res_c):
        per.append({"kind": sh["kind"], "r_eq": sh["r_eq"], "unknowns": sh["n"], "levels": sh["levels"],
                    "iterations": [r["iterations"] for r in reps],
                    "converged": all(r["status"] == "OK" for r in reps),
                    "max_resid_true": max(r["resid_true"] for r in reps)})
    out = {
        "metric": "multi-scatterer batched solve seconds (FP32 far field, FP64 refinement, one MI355X)",
        "value": round(t_conc, 3), "unit": "s", "n_gpus": 1, "higher_is_better": False, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "16_scatterers_x_8_angles", "rhs": 128, "unknowns_total": sum(sh["n"] for sh in shapes),
                   "leaf_wavelengths": 0.25, "interp_order": 6, "digits": 3, "gmres": "FGMRES-IR(50), tol 1e-6",
                   "near_field": "physical"},
        "setup_s": round(setup_s, 2), "sequential_scatterers_s": round(t_seq, 3),
        "concurrent_vs_sequential_solution_rel": diff, "scatterers": per,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="sphere30k")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-solve", action="store_true", help="skip the GMRES solves")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--mode", default="matvec", choices=["matvec", "copy"],
                    help="copy: only run a plain 256 MiB device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: the library runs the matvec on one GPU and has no multi-GPU "
              "path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from mlfma_amd import _capi as A
    from mlfma_amd import geometry as G
    from mlfma_amd.plan import Plan

    if args.mode == "copy":
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda:0")   # 256 MiB
        dst = torch.empty_like(src)
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"mode": "copy", "bytes_read_per_launch": src.numel() * 8,
                          "bytes_written_per_launch": src.numel() * 8, "launches": 10}))
        return

    if args.workload == "dipoles200k_grad":
        gradient_job()
        return
    if args.workload == "multi16":
        multi_scatterer_job()
        return
    geom, pw, kw, desc = workload(args.workload)
    log("workload", desc)
    t0 = time.time()
    plan = Plan(geom, pw, A.make_config(device=0, **kw))
    setup_s = time.time() - t0
    info = plan.info()
    log("plan", {k: info[k] for k in ["max_level", "boxes", "m2l_pairs", "near_nnz", "radpat_entries",
                                      "stale_boxes", "truncated_l2l", "device_bytes"]}, "setup s", round(setup_s, 2))
    n = geom.n
    dev = torch.device("cuda:0")
    x = torch.from_numpy(G.complex_gaussian(n, 0)).to(dev)
    y = torch.empty_like(x)
    stages = {k: 0.0 for k in ["S2M", "M2M", "M2L", "L2L", "L2T", "near", "total"]}
    near_pairs_ms = 0.0
    # the library's autotuner picks M2M / M2L launch shapes for this plan
    tune_before, tune_after = plan.autotune(1)
    log("autotune ms", round(tune_before, 4), "->", round(tune_after, 4))
    for _ in range(args.warmup):
        plan.matvec(x, y)
    # timed region: K matvecs, uninstrumented, queued on torch's stream (the
    # library returns once a matvec is queued; the synchronize closes it)
    plan.set_flags(A.MLFMA_FLAG_ASYNC)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.matvec(x, y)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1000.0 * elapsed / args.steps
    value = args.steps / elapsed
    # the same with every stage in order on one stream (the default runs the
    # near-field pairs and the finest M2L beside the far-field chain)
    plan.set_flags(A.MLFMA_FLAG_ASYNC | A.MLFMA_FLAG_ONE_STREAM)
    torch.cuda.synchronize()
    t0n = time.perf_counter()
    for _ in range(args.steps):
        plan.matvec(x, y)
    torch.cuda.synchronize()
    value_serial = args.steps / (time.perf_counter() - t0n)
    # the same K matvecs replayed from a captured hipGraph
    plan.set_flags(A.MLFMA_FLAG_ASYNC | A.MLFMA_FLAG_GRAPH)
    plan.matvec(x, y)
    torch.cuda.synchronize()
    tg = time.perf_counter()
    for _ in range(args.steps):
        plan.matvec(x, y)
    torch.cuda.synchronize()
    value_graph = args.steps / (time.perf_counter() - tg)
    # the same with a completed result per call (the synchronous entry point)
    plan.set_flags(0)
    torch.cuda.synchronize()
    t0s = time.perf_counter()
    for _ in range(args.steps):
        plan.matvec(x, y)
    torch.cuda.synchronize()
    value_sync = args.steps / (time.perf_counter() - t0s)
    # the same K matvecs again with the per-stage HIP events recorded on the
    # library's stream: stage times and the roofline kernel's launch time
    plan.set_flags(A.MLFMA_FLAG_STAGE_TIMES)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        plan.matvec(x, y)
        st = plan.stage_times()
        for k in stages:
            stages[k] += st[k]
        near_pairs_ms += plan.stage_times_detail()["near_pairs"]
    torch.cuda.synchronize()
    with_events = args.steps / (time.perf_counter() - t1)
    plan.set_flags(0)
    stage_ms = {k: round(v / args.steps, 4) for k, v in stages.items()}
    log("stage ms", stage_ms)
    log("matvec/s", round(value, 2), "one stream", round(value_serial, 2), "graph replay", round(value_graph, 2),
        "synchronous calls", round(value_sync, 2),
        "with stage events",
        round(with_events, 2))
    # batched right-hand sides: 8 vectors per call share every operator read
    nb = 8
    Xb = torch.from_numpy(np.stack([G.complex_gaussian(n, 50 + r) for r in range(nb)])).to(dev)
    Yb = torch.empty_like(Xb)
    tb_before, tb_after = plan.autotune(nb)
    plan.set_flags(A.MLFMA_FLAG_ASYNC)
    for _ in range(2):
        plan.matvec(Xb, Yb)
    torch.cuda.synchronize()
    calls = max(1, args.steps // 2)
    t2 = time.perf_counter()
    for _ in range(calls):
        plan.matvec(Xb, Yb)
    torch.cuda.synchronize()
    el_b = time.perf_counter() - t2
    plan.set_flags(0)
    batched = {"nrhs": nb, "matvec_per_s": round(nb * calls / el_b, 3), "ms_per_call": round(1000 * el_b / calls, 4),
               "autotune_ms": {"before": round(tb_before, 4), "after": round(tb_after, 4)}}
    log("batched", batched)

    # roofline of the near-field pair kernel (one launch per matvec; HIP events
    # on the library's stream right before and after that launch)
    near_ms = near_pairs_ms / args.steps
    achieved = near_bytes(info) / (near_ms * 1e-3) / 1e9
    traffic = measured_traffic("kNearSym", "sphere30k_fp64") if args.workload == "sphere30k" else None
    roofline = {"bound": "hbm", "kernel": "kNearSym", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": round(traffic) if traffic else None,
                "bytes_per_launch": near_bytes(info), "avg_launch_ms": round(near_ms, 5)}

    # streamed-byte model per stage against the stage times (in-order,
    # event-timed matvecs) and for the whole matvec against ms_per_step
    sb, comp = byte_model(info, n)
    total_b = sum(sb.values())
    bytes_model = {
        "stages": {k: {"bytes": int(v), "ms": stage_ms[k], "GBps": round(v / (stage_ms[k] * 1e-3) / 1e9, 1),
                       "frac_hbm_peak": round(v / (stage_ms[k] * 1e-3) / 1e9 / HBM_PEAK_GBS, 3)}
                   for k, v in sb.items() if stage_ms[k] > 0},
        "matvec_streamed_bytes": int(total_b),
        "matvec_effective_GBps": round(total_b / (ms_per_step * 1e-3) / 1e9, 1),
        "matvec_frac_hbm_peak": round(total_b / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 3),
        "compulsory_bytes": int(comp),
        "compulsory_GBps": round(comp / (ms_per_step * 1e-3) / 1e9, 1),
        "note": "M2L re-reads each source field and operator from L2/MALL, so its streamed-byte rate exceeds "
                "the HBM peak; its HBM traffic is a fraction of the streamed bytes",
    }
    log("byte model", bytes_model)

    result = {
        "metric": "MLFMA matvec/s (FP64, one MI355X)", "value": round(value, 3), "unit": "matvec/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": desc, "roofline": roofline, "stage_ms": stage_ms, "setup_s": round(setup_s, 2),
        "autotune_ms": {"before": round(tune_before, 4), "after": round(tune_after, 4)},
        "value_one_stream": round(value_serial, 3), "value_graph_replay": round(value_graph, 3),
        "value_synchronous_calls": round(value_sync, 3),
        "value_with_stage_events": round(with_events, 3),
        "batched": batched,
        "byte_model": bytes_model,
    }

    if not args.no_solve:
        sol = {}
        p64 = Plan(geom, pw, A.make_config(device=0, **kw))
        p64.autotune(1)  # launch parameters and stage order (not timed)
        x64, r64 = p64.solve(method=A.MLFMA_GMRES_RESTART, restart=50, maxit=500, tol=1e-6)
        sol["fp64_gmres50"] = {k: r64[k] for k in ["iterations", "status", "resid_true", "seconds"]}
        p32 = Plan(geom, pw, A.make_config(device=0, precision=A.MLFMA_FP32, **kw))
        p32.autotune(1)
        x32, r32 = p32.solve(method=A.MLFMA_FGMRES_IR, restart=50, maxit=500, tol=1e-6)
        sol["mixed_fp32_ir"] = {k: r32[k] for k in ["iterations", "status", "resid_true", "seconds"]}
        sol["mixed_vs_fp64_solution_rel"] = float(np.linalg.norm(x32 - x64) / max(np.linalg.norm(x64), 1e-300))
        # FP32 far-field matvec throughput (queued on torch's stream, as the
        # FP64 figure)
        xx = torch.from_numpy(G.complex_gaussian(n, 0)).to(dev)
        yy = torch.empty_like(xx)
        p32.set_flags(A.MLFMA_FLAG_ASYNC)
        for _ in range(3):
            p32.matvec(xx, yy)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            p32.matvec(xx, yy)
        torch.cuda.synchronize()
        p32.set_flags(0)
        sol["fp32_farfield_matvec_per_s"] = round(args.steps / (time.perf_counter() - t1), 3)
        if args.workload in ("sphere30k", "sphere200k"):
            if args.workload == "sphere30k":
                sol["physical"] = physical_solve(geom, pw, kw, radius=2.0)
            else:  # ka = 8 pi lies 0.03 from an EFIE interior resonance: longer restarts
                sol["physical"] = physical_solve(geom, pw, kw, radius=4.0, restart=100, maxit=6000)
        result["solve"] = sol
        log("solve", sol)

    if not args.no_cpu:
        log("cpu baseline ...")
        y_gpu = plan.matvec(G.complex_gaussian(n, 0))
        result["cpu_baseline"] = cpu_baseline(geom, pw, kw, args.cpu_budget, y_gpu)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
