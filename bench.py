"""Benchmark: MLFMA matvec throughput (and GMRES solve time) on one MI355X.

Workload (default, --workload sphere200k): plain FP64 RWG EFIE on a PEC
sphere of radius 4 wavelengths, geodesic icosphere of frequency 82 (134,480
triangles, 201,720 RWG unknowns), 7-point triangle quadrature, leaf boxes of
0.25 wavelength (root 8, levels 0-5), interpolation order 6, 3 digits, plane
wave theta = 0 x-polarised.  Also: sphere30k (30,720 RWG, radius 2),
dipoles4k (the reference's 4,096-dipole case), multi16 (16 scatterers) and
dipoles200k_grad (position gradient of 200,000 dipoles).

A step is one full matvec y = Z x (S2M, M2M, M2L, L2L, near field, L2T) with
x and y resident in HBM, in the schedule the library's autotuner picks.  W
untimed warmup steps, then K timed steps bracketed by torch.cuda.synchronize()
give `value`; one JSON line.

The line also carries:
  kernels       per launch group (kernel), the device time of its launches from
                HIP events recorded on the stream each launch runs on, over a
                second pass of the same K steps in the same schedule;
  roofline      the kernel with the most device time per matvec: algorithmic
                bytes per launch / its average launch time against 8 TB/s, and
                the HBM bytes per launch from the committed rocprofv3 counter
                summary (profiles/<kernel>_<workload>_traffic.txt);
  stage_ms      per-stage device times with every stage in order (the
                reference's ClockTimes);
  batched       8 right-hand sides per call (each operator read shared by 8);
  cpu_baseline  the CPU oracle (FP64 restatement of the reference, one core)
                timed on a bounded number of matvecs of the same workload;
  solve         GMRES to 1e-6: the reference's kernel (stalls), then the
                physical near field in FP64 GMRES and mixed-precision FGMRES,
                8 incidence angles one at a time and batched, RCS against the
                Mie series.
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


WORKLOADS = {"dipoles4k": "c1", "sphere30k": "c2", "sphere200k": "c3"}


def workload(name):
    from mlfma_amd import configs as CF
    if name not in WORKLOADS:
        raise SystemExit(f"unknown workload {name}")
    return getattr(CF, WORKLOADS[name])()


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


def solves(geom, pw, kw, desc):
    """GMRES to 1e-6 on the workload, plain FP64 against mixed precision and
    batched right-hand sides.

    The reference's kernel drops the singular self and coincident-triangle
    terms (src/sources/rwg.cpp:97,113-116); GMRES stalls on that operator
    (reported: 500 iterations of GMRES(50), FP64 and mixed).  The solve-time
    comparison therefore runs on the physical near field (closed-form
    singular integrals; same far field): one incidence (theta = 0) in FP64
    GMRES(m) and in mixed precision (FP64 FGMRES outer, FP32-far-field GMRES
    inner), then 8 incidence angles (theta_j = j 180/7 degrees, phi = 0) one
    FP64 solve at a time against one batched mixed-precision solve.  RCS at
    722 bistatic angles (theta 0-180 in 0.5-degree steps, phi = 0 and 90) of
    the mixed solutions against the FP64 ones, and of the FP64 theta = 0
    solution against the Mie series."""
    from mlfma_amd import _capi as A
    from mlfma_amd import configs as CF
    from mlfma_amd.plan import Plan
    from tests import _physical as P
    radius = {"rwg_sphere_R2_icosphere32": 2.0, "rwg_sphere_R4_icosphere82": 4.0}.get(desc["workload"])
    restart = 100 if radius == 4.0 else 50   # ka = 8 pi lies 0.03 from an EFIE interior resonance
    out = {}
    # reference kernel
    p64 = Plan(geom, pw, A.make_config(device=0, **kw))
    p64.autotune(1)
    x64, r64 = p64.solve(method=A.MLFMA_GMRES_RESTART, restart=50, maxit=500, tol=1e-6)
    p32 = Plan(geom, pw, A.make_config(device=0, precision=A.MLFMA_FP32, **kw))
    p32.autotune(1)
    x32, r32 = p32.solve(method=A.MLFMA_FGMRES_IR, restart=30, outer_restart=30, inner_tol=0.1, maxit=500,
                         tol=1e-6)
    keys = ["iterations", "status", "resid_true", "seconds"]
    out["reference_kernel"] = {"fp64_gmres50": {k: r64[k] for k in keys}, "mixed_fgmres": {k: r32[k] for k in keys},
                               "note": "the reference's operator (self and coincident-triangle terms dropped): "
                                       "GMRES stalls; 500 iterations"}
    del p64, p32
    if radius is None:
        return out
    fl = A.MLFMA_FLAG_PHYSICAL | A.MLFMA_FLAG_FRESH_LOCALS
    t0 = time.time()
    q64 = Plan(geom, pw, A.make_config(device=0, flags=fl, **kw))
    setup = time.time() - t0
    q64.autotune(1)
    q32 = Plan(geom, pw, A.make_config(device=0, precision=A.MLFMA_FP32, flags=fl, **kw))
    q32.autotune(1)
    maxit = 6000
    I64, s64 = q64.solve(method=A.MLFMA_GMRES_RESTART, restart=restart, maxit=maxit, tol=1e-6)
    log("physical FP64", s64)
    # deflated restarts: 20 harmonic Ritz vectors kept per restart
    Idr, sdr = q64.solve(method=A.MLFMA_GMRES_RESTART, restart=restart, maxit=maxit, tol=1e-6, deflate=20)
    log("physical FP64 GMRES-DR", sdr)
    mixed = dict(method=A.MLFMA_FGMRES_IR, restart=30, outer_restart=30, inner_tol=0.1)
    I32, s32 = q32.solve(maxit=maxit, tol=1e-6, **mixed)
    log("physical mixed", s32)
    dirs = []
    th = np.radians(np.linspace(0.0, 180.0, 361))
    for phi in (0.0, np.pi / 2):
        dirs.append(np.stack([np.sin(th) * np.cos(phi), np.sin(th) * np.sin(phi), np.cos(th)], axis=1))
    dirs = np.concatenate(dirs)

    def rcs(p, I):
        return 4 * np.pi * np.sum(np.abs(p.farfield(I, dirs)) ** 2, axis=1)

    def rcs_diff(a, b):
        return float(np.abs(a - b).max() / np.abs(b).max())

    se, sh = P.mie_pec_rcs(2 * np.pi * radius, th)
    mie = np.concatenate([se, sh])
    db = 10 * np.log10(rcs(q64, I64)) - 10 * np.log10(mie)
    single = {"fp64": {k: s64[k] for k in keys}, "mixed": {k: s32[k] for k in keys},
              "fp64_deflated": dict({k: sdr[k] for k in keys}, deflate=20,
                                    speedup_vs_fp64=round(s64["seconds"] / sdr["seconds"], 3),
                                    solution_rel_diff_vs_fp64=float(np.linalg.norm(Idr - I64) / np.linalg.norm(I64))),
              "mixed_speedup": round(s64["seconds"] / s32["seconds"], 3),
              "mixed_rcs_rel_diff_vs_fp64": rcs_diff(rcs(q32, I32), rcs(q64, I64)),
              "mixed_solution_rel_diff_vs_fp64": float(np.linalg.norm(I32 - I64) / np.linalg.norm(I64)),
              "fp64_rcs_vs_mie_db": {"rms": round(float(np.sqrt(np.mean(db ** 2))), 4),
                                     "max": round(float(np.abs(db).max()), 4)}}
    B = np.stack([-q64.excitation(CF.planewave_xz(t)) for t in CF.c4_angles()])
    ts = time.perf_counter()
    seq = [q64.solve(B[j], method=A.MLFMA_GMRES_RESTART, restart=restart, maxit=maxit, tol=1e-6) for j in range(8)]
    t_seq = time.perf_counter() - ts
    log("physical 8 angles FP64 one at a time", round(t_seq, 3), [r["iterations"] for _, r in seq])
    q32.autotune(8)
    tm = time.perf_counter()
    Xm, rm = q32.solve_batch(B, maxit=maxit, tol=1e-6, **mixed)
    t_mix = time.perf_counter() - tm
    log("physical 8 angles batched mixed", round(t_mix, 3), [r["iterations"] for r in rm])
    q64.autotune(8)
    tb = time.perf_counter()
    Xb, rb = q64.solve_batch(B, restart=restart, maxit=maxit, tol=1e-6)
    t_b64 = time.perf_counter() - tb
    tb = time.perf_counter()
    _, rbd = q64.solve_batch(B, restart=restart, maxit=maxit, tol=1e-6, deflate=20)
    t_b64d = time.perf_counter() - tb
    log("physical 8 angles batched FP64 GMRES-DR", round(t_b64d, 3), [r["iterations"] for r in rbd])
    rd = max(rcs_diff(rcs(q32, Xm[j]), rcs(q64, seq[j][0])) for j in range(8))
    sweep = {"angles": 8, "fp64_one_at_a_time_s": round(t_seq, 4), "batched_fp64_s": round(t_b64, 4),
             "batched_fp64_deflated_s": round(t_b64d, 4),
             "batched_fp64_deflated_converged": all(r["status"] == "OK" for r in rbd),
             "batched_mixed_s": round(t_mix, 4), "speedup_batched_mixed_vs_fp64": round(t_seq / t_mix, 3),
             "all_converged": all(r["status"] == "OK" for r in rm) and all(r["status"] == "OK" for _, r in seq),
             "max_resid_true_mixed": max(r["resid_true"] for r in rm),
             "mixed_rcs_max_rel_diff_vs_fp64": rd,
             "iterations_fp64": [r["iterations"] for _, r in seq], "iterations_mixed": [r["iterations"] for r in rm]}
    out["physical"] = {"near_field": "MLFMA_FLAG_PHYSICAL (closed-form singular integrals)",
                       "gmres": f"FP64 GMRES({restart}); mixed: FP64 FGMRES(30) around FP32-far-field GMRES(30) "
                                f"to 1e-1 per step; deflated: FP64 GMRES-DR({restart}, 20); tol 1e-6",
                       "plan_setup_s": round(setup, 3),
                       "single_incidence": single, "incidence_sweep": sweep}
    return out


def cpu_baseline(geom, pw, kw, budget_s, y_gpu, x):
    """The CPU oracle (FP64 restatement of the reference, one thread for the
    matvec as the reference) on a bounded number of matvecs of the same
    workload; its first matvec also checks the GPU's timed matvec of x."""
    from mlfma_amd import _capi as A
    from tests import _oracle as O
    t0 = time.time()
    nt = max(1, min(16, os.cpu_count() or 1))
    o = O.Oracle(geom, pw, A.make_config(device=-1, **kw), threads=nt)
    setup = time.time() - t0
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
    seed 2, cube of side 8 wavelengths, complex-Gaussian excitation):
    J = sum over 360 bistatic angles (theta = 0..359 degrees, phi = 0) of
    |E_far|^2 with E_far = F (Z + s I)^-1 b, and dJ/db, dJ/dr through
    torch.autograd (forward GMRES, far field, adjoint GMRES, position-gradient
    pass), timed, then checked against central finite differences (h = 1e-6
    wavelength) along the gradient direction.  The reference's operator has a
    zero diagonal (src/sources/dipole.h:65) and GMRES does not converge on it
    at this density (its residual after 500 iterations is reported); the
    pipeline solves the operator with a real dipole self term s = 30 |C k| k^3
    on the diagonal."""
    import torch
    from mlfma_amd import _capi as A
    from mlfma_amd import autograd as AG
    from mlfma_amd import configs as CF
    from mlfma_amd import geometry as G
    from mlfma_amd.plan import Plan
    geom, pw, kw, desc = CF.c5()
    n = geom.n
    dirs = CF.c5_dirs()
    b0 = G.complex_gaussian(n, 2)
    ck = 299792458.0 * 1.256637e-6 / (4 * np.pi) * 2 * np.pi
    shift = 30.0 * ck * (2 * np.pi) ** 3
    t0 = time.time()
    ref = Plan(geom, pw, A.make_config(device=0, **kw))
    setup_s = time.time() - t0
    ref.autotune(1)
    _, rep_ref = ref.solve(b0, method=A.MLFMA_GMRES_RESTART, restart=50, maxit=500, tol=1e-6)
    log("reference operator GMRES(50)", rep_ref)
    del ref
    cfg = {k: v for k, v in kw.items() if k != "flags"}
    sc = AG.DipoleScatterer(geom.mom, dirs, cfg, pw=pw, shift=shift, tol=1e-10)
    sc.plan(geom.pos).autotune(1)
    b = torch.from_numpy(b0).requires_grad_(True)
    pos = torch.from_numpy(geom.pos.copy()).requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.time()
    E = sc(b, pos)
    J = (E.abs() ** 2).sum()
    t_fwd = time.time() - t0
    t0 = time.time()
    J.backward()
    torch.cuda.synchronize()
    t_bwd = time.time() - t0
    gp = pos.grad.numpy()
    # finite difference along the (box-interior) gradient direction
    f = (geom.pos + 4.0) / 0.5
    mask = (np.abs(f - np.round(f)).min(axis=1) * 0.5 > 1e-5).astype(float)[:, None]
    D = gp * mask
    D /= np.linalg.norm(D)
    h = 1e-6

    def Jat(p_):
        with torch.no_grad():
            return float((sc(torch.from_numpy(b0), torch.from_numpy(p_)).abs() ** 2).sum())
    fd = (Jat(geom.pos + h * D) - Jat(geom.pos - h * D)) / (2 * h)
    ad = float(np.sum(gp * D))
    out = {
        "metric": "adjoint gradient pipeline of a far-field objective, seconds (FP64, one MI355X)",
        "value": round(t_fwd + t_bwd, 3), "unit": "s", "n_gpus": 1, "higher_is_better": False,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": desc["workload"], "unknowns": n, "angles": 360, "leaf_wavelengths": 0.5,
                   "levels": "0-4", "interp_order": 6, "digits": 3,
                   "gmres": "GMRES(100) to 1e-10, shift 30 |C k| k^3"},
        "setup_s": round(setup_s, 2), "forward_s": round(t_fwd, 3), "backward_s": round(t_bwd, 3),
        "objective": float(J.detach()),
        "fd_check": {"direction": "gradient (dipoles within 1e-5 of a leaf face held)", "h": h,
                     "autograd": ad, "finite_difference": fd, "rel_diff": abs(ad - fd) / abs(fd)},
        "reference_operator_gmres50_500": {k: rep_ref[k] for k in ["iterations", "status", "resid_true"]},
    }
    print(json.dumps(out), flush=True)


def multi_scatterer_job():
    """16 scatterers of uneven size (seed 1: sphere or 2:1:1 ellipsoid, equivalent
    radius uniform in 0.5-3 wavelengths; icosphere frequency 15 per wavelength
    of radius, of the short semi-axis for ellipsoids), each under 8 plane
    waves (theta_j = j 180/7 degrees, phi = 0, polarised in the plane of
    incidence): 128 right-hand sides, physical near field, leaf 0.25
    wavelength, mixed precision (FP64 FGMRES(30) around FP32-far-field
    GMRES(30)) to 1e-6.  Timed three ways: one forest plan holding all 16
    scatterers (every matvec stage one launch per level for all of them, one
    batched solve over 128 lanes), one plan per scatterer solved concurrently
    from host threads, and one after another."""
    import torch
    from mlfma_amd import _capi as A
    from mlfma_amd import configs as CF
    from mlfma_amd.plan import Plan
    from mlfma_amd.batch import solve_many
    shapes = CF.c4()
    flags = A.MLFMA_FLAG_PHYSICAL | A.MLFMA_FLAG_FRESH_LOCALS
    pw = CF.planewave_theta0()
    angles = CF.c4_angles()
    t0 = time.time()
    cfgs = [A.make_config(device=0, precision=A.MLFMA_FP32, flags=flags, **sh["kw"]) for sh in shapes]
    F = Plan.forest([sh["geom"] for sh in shapes], cfgs, pw)
    forest_setup = time.time() - t0
    off = F.segments()
    B = np.stack([-F.excitation(CF.planewave_xz(t)) for t in angles])
    t0 = time.time()
    plans = [Plan(sh["geom"], pw, c) for sh, c in zip(shapes, cfgs)]
    sep_setup = time.time() - t0
    rhs = [np.ascontiguousarray(B[:, off[s]:off[s + 1]]) for s in range(len(shapes))]
    log("scatterers", [(sh["kind"], round(sh["r_eq"], 3), sh["n"]) for sh in shapes], "setup s", round(forest_setup, 2),
        round(sep_setup, 2))
    # matvec throughput: one forest matvec of the 8 right-hand sides against
    # the 16 plans' own 8-right-hand-side matvecs one after another
    F.autotune(8)
    for p in plans:
        p.autotune(8)
    Bd = torch.from_numpy(B).to("cuda:0")
    Yd = torch.empty_like(Bd)
    parts = [torch.from_numpy(r).to("cuda:0") for r in rhs]
    outs = [torch.empty_like(x) for x in parts]
    for p in [F] + plans:
        p.set_flags(flags | A.MLFMA_FLAG_ASYNC)
    reps = 20
    for _ in range(2):
        F.matvec(Bd, Yd)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        F.matvec(Bd, Yd)
    torch.cuda.synchronize()
    mv_forest = (time.perf_counter() - t0) / reps
    for p, x, y in zip(plans, parts, outs):
        p.matvec(x, y)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        for p, x, y in zip(plans, parts, outs):
            p.matvec(x, y)
    torch.cuda.synchronize()
    mv_sep = (time.perf_counter() - t0) / reps
    for p in [F] + plans:
        p.set_flags(flags)
    log("8-RHS matvec ms: forest", round(1e3 * mv_forest, 3), "16 plans", round(1e3 * mv_sep, 3))
    opts = dict(restart=30, outer_restart=30, inner_tol=0.1, maxit=6000, tol=1e-6, method=A.MLFMA_FGMRES_IR)
    torch.cuda.synchronize()
    t0 = time.time()
    Xf, rf = F.solve_batch(B, **opts)
    t_forest = time.time() - t0
    log("forest", round(t_forest, 3))
    t0 = time.time()
    res_c = solve_many(plans, rhs, **opts)
    t_conc = time.time() - t0
    log("concurrent plans", round(t_conc, 3))
    t0 = time.time()
    res_s = [p.solve_batch(b, **opts) for p, b in zip(plans, rhs)]
    t_seq = time.time() - t0
    log("sequential plans", round(t_seq, 3))
    nseg = len(shapes)
    diff = max(float(np.linalg.norm(Xf[:, off[s]:off[s + 1]] - res_s[s][0]) / np.linalg.norm(res_s[s][0]))
               for s in range(nseg))
    per = []
    for s, sh in enumerate(shapes):
        lanes = [rf[a * nseg + s] for a in range(len(angles))]
        per.append({"kind": sh["kind"], "r_eq": round(sh["r_eq"], 3), "unknowns": sh["n"], "levels": sh["levels"],
                    "iterations": [r["iterations"] for r in lanes],
                    "converged": all(r["status"] == "OK" for r in lanes),
                    "max_resid_true": max(r["resid_true"] for r in lanes)})
    out = {
        "metric": "multi-scatterer batched solve seconds (FP32 far field, FP64 FGMRES, one MI355X)",
        "value": round(t_forest, 3), "unit": "s", "n_gpus": 1, "higher_is_better": False, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "16_scatterers_x_8_angles", "rhs": 128, "unknowns_total": int(F.n),
                   "leaf_wavelengths": 0.25, "interp_order": 6, "digits": 3,
                   "gmres": "FGMRES(30) around FP32-far-field GMRES(30), tol 1e-6, maxit 6000",
                   "near_field": "physical"},
        "matvec_8rhs_ms": {"forest": round(1e3 * mv_forest, 4), "sixteen_plans_queued": round(1e3 * mv_sep, 4)},
        "forest_setup_s": round(forest_setup, 2), "concurrent_plans_s": round(t_conc, 3),
        "sequential_plans_s": round(t_seq, 3), "forest_vs_sequential_solution_rel": diff,
        "converged_scatterers": sum(p["converged"] for p in per), "scatterers": per,
    }
    print(json.dumps(out), flush=True)


def c4_stalls_job():
    """Why some c4 ellipsoids stall: an interior resonance of the EFIE or the
    mesh.  (1) Every c4 ellipsoid under theta = 0 incidence with the
    multi-scatterer solver (FP64 FGMRES(30) around FP32-far-field GMRES(30)
    to 0.1, tol 1e-6, at most 6,000 iterations).  (2) Every one that stalls
    again with (a) the 13-point triangle rule, (b) a mesh 1.5x finer (the same
    2:1 mapping of an icosphere, smaller triangles) and (c) the same mesh
    scaled by 0.98, 0.99, 1.01 and 1.02 (the same discretisation at another
    electrical size).  A stall that survives the finer rule and mesh and
    goes away with a 1-2% size change is a resonance; one that follows the
    mesh is the discretisation.  (3) Dense: the smallest singular value of the
    physical EFIE matrix (one-leaf plan: every pair in the near field) and the
    FP64 GMRES(100) iterations over a size sweep, for a sphere of radius 0.6
    across its TM_2 interior resonance (kr = 3.870, first zero of
    (x j_2(x))') as the calibration, and for the smallest ellipsoid."""
    from mlfma_amd import _capi as A
    from mlfma_amd import configs as CF
    from mlfma_amd.plan import Plan
    flags = A.MLFMA_FLAG_PHYSICAL | A.MLFMA_FLAG_FRESH_LOCALS
    pw0 = CF.planewave_xz(0.0)
    opts = dict(restart=30, outer_restart=30, inner_tol=0.1, maxit=6000, tol=1e-6, method=A.MLFMA_FGMRES_IR)

    def solve(sh):
        p = Plan(sh["geom"], pw0, A.make_config(device=0, precision=A.MLFMA_FP32, flags=flags, **sh["kw"]))
        _, rep = p.solve(-p.excitation(pw0), **opts)
        p.close()
        return {"unknowns": sh["n"], "iterations": rep["iterations"], "status": rep["status"],
                "resid_true": float("%.3g" % rep["resid_true"])}

    shapes = CF.c4()
    ells = [(i, sh) for i, sh in enumerate(shapes) if sh["kind"] == "ellipsoid"]
    first = {}
    for i, sh in ells:
        first[i] = solve(sh)
        log("c4 scatterer", i, "r_eq", round(sh["r_eq"], 3), first[i])
    variants = {}
    for i, sh in ells:
        if first[i]["status"] == "OK":
            continue
        v = {"c4": first[i]}
        v["rule13"] = solve(CF.c4_scatterer("ellipsoid", sh["r_eq"], quad_prec=3))
        v["mesh1.5x"] = solve(CF.c4_scatterer("ellipsoid", sh["r_eq"], mesh_scale=1.5))
        for s in (0.98, 0.99, 1.01, 1.02):
            v["size%.2f" % s] = solve(CF.c4_scatterer("ellipsoid", sh["r_eq"], scale=s))
        log("stalled scatterer", i, v)
        variants[str(i)] = {"r_eq": round(sh["r_eq"], 4), **v}

    def dense_sweep(kind, r_eq, scales):
        out = []
        for s in scales:
            sh = CF.c4_scatterer(kind, r_eq, scale=s)
            n = sh["n"]
            kw = dict(sh["kw"], max_node_srcs=n, max_level=0)
            p = Plan(sh["geom"], pw0, A.make_config(device=0, flags=flags, **kw))
            Z = np.empty((n, n), dtype=np.complex128)
            for j0 in range(0, n, 8):
                j1 = min(n, j0 + 8)
                E = np.zeros((j1 - j0, n), dtype=np.complex128)
                E[np.arange(j1 - j0), np.arange(j0, j1)] = 1.0
                Z[:, j0:j1] = p.matvec(E).T
            sv = np.linalg.svd(Z, compute_uv=False)
            _, rep = p.solve(-p.excitation(pw0), method=A.MLFMA_GMRES_RESTART, restart=100, maxit=3000, tol=1e-6)
            p.close()
            extent = r_eq if kind == "sphere" else 2 * r_eq / 2 ** (1 / 3)
            row = {"scale": round(s, 4), "k_times_semi_axis": round(2 * np.pi * extent * s, 4), "unknowns": n,
                   "sigma_min_over_max": float("%.4g" % (sv[-1] / sv[0])), "gmres100_iterations": rep["iterations"],
                   "status": rep["status"]}
            log("dense", kind, r_eq, row)
            out.append(row)
        return out

    sphere = dense_sweep("sphere", 0.6, [0.96 + 0.01 * j for j in range(14)])
    small = min((sh["n"], i) for i, sh in ells)[1]
    ellipsoid = dense_sweep("ellipsoid", shapes[small]["r_eq"], [0.96 + 0.01 * j for j in range(9)])
    print(json.dumps({"mode": "c4-stalls", "solver": "FGMRES(30) around FP32-far-field GMRES(30) to 0.1, tol 1e-6, "
                      "maxit 6000, theta=0 incidence, physical near field",
                      "ellipsoids": {str(i): {"r_eq": round(sh["r_eq"], 4), **first[i]} for i, sh in ells},
                      "stalled_variants": variants,
                      "dense_sphere_r0.6": sphere,
                      "dense_smallest_ellipsoid": {"scatterer": small, "r_eq": round(shapes[small]["r_eq"], 4),
                                                   "sweep": ellipsoid}}), flush=True)


def deflation_job(args):
    """Deflated restarts (GMRES-DR) on the physical near field: the workload
    sphere under theta = 0 incidence, FP64 GMRES(m) against GMRES-DR(m, k)
    for a few (m, k), the mixed solvers (FP64 FGMRES around FP32-far-field
    GMRES; FP64 refinement around FP32-far-field GMRES-DR), then the seed-1
    eight incidence angles batched (FP64 GMRES(100) and GMRES-DR(100, 20);
    mixed FGMRES and refinement around GMRES-DR), the sixteen c4 scatterers
    x 8 angles with refinement around FP32-far-field GMRES-DR (one forest
    plan; 16 concurrent plans), then the seed-1 c4 ellipsoids under
    theta = 0 in FP64 GMRES(100) and GMRES-DR(100, 20), tol 1e-6.
    Iterations count matvecs of the Krylov loops; seconds are the solver's
    wall time (the host's small dense problem of every restart included).
    --parts picks the parts (single, batched, c4, ellipsoids)."""
    from mlfma_amd import _capi as A
    from mlfma_amd import configs as CF
    from mlfma_amd.plan import Plan
    geom, pw, kw, desc = workload(args.workload)
    fl = A.MLFMA_FLAG_PHYSICAL | A.MLFMA_FLAG_FRESH_LOCALS
    keys = ["iterations", "status", "resid_true", "seconds"]
    parts = set(args.parts.split(","))
    out = {"mode": "deflation", "workload": desc["workload"], "unknowns": int(geom.n), "tol": 1e-6, "solves": {}}
    if parts & {"single", "batched", "dr"}:
        deflation_sphere(A, CF, Plan, geom, pw, kw, fl, keys, out, parts)
    if "c4" in parts:
        deflation_c4(A, CF, Plan, fl, out)
    if "ellipsoids" in parts:
        deflation_ellipsoids(A, CF, Plan, fl, keys, out)
    print(json.dumps(out), flush=True)


def deflation_sphere(A, CF, Plan, geom, pw, kw, fl, keys, out, parts):
    q64 = Plan(geom, pw, A.make_config(device=0, flags=fl, **kw))
    q64.autotune(1)
    q32 = Plan(geom, pw, A.make_config(device=0, precision=A.MLFMA_FP32, flags=fl, **kw))
    q32.autotune(1)
    runs = [("fp64_gmres100", q64, dict(method=A.MLFMA_GMRES_RESTART, restart=100)),
            ("fp64_gmres_dr100_20", q64, dict(method=A.MLFMA_GMRES_RESTART, restart=100, deflate=20)),
            ("fp64_gmres_dr60_15", q64, dict(method=A.MLFMA_GMRES_RESTART, restart=60, deflate=15)),
            ("fp64_gmres_dr100_30", q64, dict(method=A.MLFMA_GMRES_RESTART, restart=100, deflate=30)),
            ("fp64_gmres_dr40_10", q64, dict(method=A.MLFMA_GMRES_RESTART, restart=40, deflate=10)),
            ("mixed_fgmres30_30", q32, dict(method=A.MLFMA_FGMRES_IR, restart=30, outer_restart=30, inner_tol=0.1)),
            ("mixed_ir_dr100_20", q32, dict(method=A.MLFMA_GMRES_IR, restart=100, deflate=20)),
            ("mixed_ir_dr60_15", q32, dict(method=A.MLFMA_GMRES_IR, restart=60, deflate=15))]
    if "single" not in parts:  # "dr": the deflated FP64 solve alone (for kernel profiles)
        runs = [r for r in runs if "dr" in parts and r[0] == "fp64_gmres_dr100_20"]
    ref = None
    for name, p, o in runs:
        x, r = p.solve(maxit=6000, tol=1e-6, **o)
        if ref is None:
            ref = x
        e = {k: r[k] for k in keys}
        e["solution_rel_diff_vs_fp64_gmres"] = float(np.linalg.norm(x - ref) / np.linalg.norm(ref))
        out["solves"][name] = e
        log(name, e)
    if "batched" not in parts:
        return
    # eight incidence angles batched (one multi-RHS matvec per step)
    B = np.stack([-q64.excitation(CF.planewave_xz(t)) for t in CF.c4_angles()])
    q64.autotune(8)
    q32.autotune(8)
    out["batched_8_angles"] = {}
    for name, p, o in [("fp64_gmres100", q64, dict(method=A.MLFMA_GMRES_RESTART, restart=100)),
                       ("fp64_gmres_dr100_20", q64, dict(method=A.MLFMA_GMRES_RESTART, restart=100, deflate=20)),
                       ("mixed_fgmres30_30", q32, dict(method=A.MLFMA_FGMRES_IR, restart=30, outer_restart=30,
                                                       inner_tol=0.1)),
                       ("mixed_ir_dr60_15", q32, dict(method=A.MLFMA_GMRES_IR, restart=60, deflate=15)),
                       ("mixed_ir_dr100_20", q32, dict(method=A.MLFMA_GMRES_IR, restart=100, deflate=20))]:
        t = time.perf_counter()
        _, reps = p.solve_batch(B, maxit=6000, tol=1e-6, **o)
        e = {"seconds": round(time.perf_counter() - t, 4), "iterations": [r["iterations"] for r in reps],
             "converged": sum(r["status"] == "OK" for r in reps), "max_resid_true": max(r["resid_true"] for r in reps)}
        out["batched_8_angles"][name] = e
        log("8 angles", name, e)


def deflation_c4(A, CF, Plan, fl, out):
    # sixteen c4 scatterers x 8 angles: one forest plan and 16 concurrent
    # plans, FP64 refinement around FP32-far-field GMRES-DR
    from mlfma_amd.batch import solve_many
    shapes = CF.c4()
    cfgs = [A.make_config(device=0, precision=A.MLFMA_FP32, flags=fl, **sh["kw"]) for sh in shapes]
    pwc = CF.planewave_theta0()
    F = Plan.forest([sh["geom"] for sh in shapes], cfgs, pwc)
    off = F.segments()
    Bc = np.stack([-F.excitation(CF.planewave_xz(t)) for t in CF.c4_angles()])
    plans = [Plan(sh["geom"], pwc, c) for sh, c in zip(shapes, cfgs)]
    rhs = [np.ascontiguousarray(Bc[:, off[i]:off[i + 1]]) for i in range(len(shapes))]
    F.autotune(8)
    for p in plans:
        p.autotune(8)
    nseg = len(shapes)
    out["c4_16_scatterers_x_8_angles"] = {}
    for name, o in [("ir_dr60_15", dict(method=A.MLFMA_GMRES_IR, restart=60, deflate=15)),
                    ("ir_dr100_20", dict(method=A.MLFMA_GMRES_IR, restart=100, deflate=20)),
                    ("fp64_dr100_20", dict(method=A.MLFMA_GMRES_RESTART, restart=100, deflate=20))]:
        t = time.perf_counter()
        _, rf = F.solve_batch(Bc, maxit=6000, tol=1e-6, **o)
        tf = time.perf_counter() - t
        t = time.perf_counter()
        rc = solve_many(plans, rhs, maxit=6000, tol=1e-6, **o)
        tc = time.perf_counter() - t
        conv_f = sum(all(rf[a * nseg + i]["status"] == "OK" for a in range(8)) for i in range(nseg))
        conv_c = sum(all(r["status"] == "OK" for r in reps) for _, reps in rc)
        e = {"forest_s": round(tf, 3), "concurrent_plans_s": round(tc, 3), "forest_converged_scatterers": conv_f,
             "concurrent_converged_scatterers": conv_c,
             "max_iterations": max(r["iterations"] for r in rf)}
        out["c4_16_scatterers_x_8_angles"][name] = e
        log("c4 x 8", name, e)
    for p in plans:
        p.close()
    F.close()


def deflation_ellipsoids(A, CF, Plan, fl, keys, out):
    pw0 = CF.planewave_xz(0.0)
    ells = {}
    for i, sh in enumerate(CF.c4()):
        if sh["kind"] != "ellipsoid":
            continue
        p = Plan(sh["geom"], pw0, A.make_config(device=0, flags=fl, **sh["kw"]))
        b = -p.excitation(pw0)
        e = {"r_eq": round(sh["r_eq"], 4), "unknowns": sh["n"]}
        for name, o in [("gmres100", dict(restart=100)), ("gmres_dr100_20", dict(restart=100, deflate=20))]:
            _, r = p.solve(b, method=A.MLFMA_GMRES_RESTART, maxit=6000, tol=1e-6, **o)
            e[name] = {k: r[k] for k in keys}
        p.close()
        log("c4 ellipsoid", i, e)
        ells[str(i)] = e
    out["c4_ellipsoids_theta0_fp64"] = ells


def m2l_ceiling(args):
    """Cache-resident gather ceiling for M2L: per level, each target sums the
    fields of its interaction-list sources (32 B per direction) and the
    operators of the pair offsets (16 B per direction), read in the lists'
    order, through torch.nn.functional.embedding_bag (gather and sum, no
    product), timed with CUDA events.  The bytes are the M2L's streamed bytes
    (source field and operator read per pair); the rate is what this access
    pattern reaches from the caches, the reference for kM2Lw / kM2LNear."""
    import torch
    import torch.nn.functional as Fn
    from mlfma_amd import _capi as A
    from mlfma_amd.plan import Plan
    geom, pw, kw, desc = workload(args.workload)
    plan = Plan(geom, pw, A.make_config(device=0, **kw))
    info = plan.info()
    dev = torch.device("cuda:0")
    out = {"mode": "m2l-ceiling", "workload": desc["workload"], "levels": {}}
    tot_b, tot_ms = 0, 0.0
    for l in range(2, info["max_level"] + 1):
        pairs, _ = plan.level_pairs(l, 0)
        co, _, ns = plan.level_boxes(l)
        nbox = len(co)
        # pairs between non-empty boxes: the M2L worklists hold only those
        pairs = pairs[(ns[pairs[:, 0]] > 0) & (ns[pairs[:, 1]] > 0)] if len(pairs) else pairs
        if len(pairs) == 0:
            continue
        nd = info["nth"][l] * info["nph"][l]
        order = np.argsort(pairs[:, 0], kind="stable")
        tgt, src = pairs[order, 0], pairs[order, 1]
        ut, counts = np.unique(tgt, return_counts=True)
        offsets = torch.from_numpy(np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64)).to(dev)
        M = torch.randn(nbox, 4 * nd, dtype=torch.float64, device=dev)     # 32 B per direction
        T = torch.randn(316, 2 * nd, dtype=torch.float64, device=dev)      # 16 B per direction
        si = torch.from_numpy(src.astype(np.int64)).to(dev)
        oi = torch.from_numpy((np.arange(len(src)) % 316).astype(np.int64)).to(dev)
        for _ in range(3):
            Fn.embedding_bag(si, M, offsets, mode="sum")
            Fn.embedding_bag(oi, T, offsets, mode="sum")
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 20
        e0.record()
        for _ in range(reps):
            Fn.embedding_bag(si, M, offsets, mode="sum")
            Fn.embedding_bag(oi, T, offsets, mode="sum")
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        b = len(src) * nd * (32 + 16)
        tot_b += b
        tot_ms += ms
        out["levels"][l] = {"pairs": int(len(src)), "directions": nd, "streamed_bytes": int(b), "ms": round(ms, 4),
                            "GBps": round(b / (ms * 1e-3) / 1e9, 1)}
    out["total"] = {"streamed_bytes": int(tot_b), "ms": round(tot_ms, 4),
                    "GBps": round(tot_b / (tot_ms * 1e-3) / 1e9, 1)}
    print(json.dumps(out), flush=True)


def hbm_rates():
    """What torch's own streaming kernels reach on this device: a 1 GiB
    device copy (read + write) and a sum over 2 GiB (read only), timed with
    CUDA events over 20 launches each."""
    import torch
    dev = torch.device("cuda:0")
    a = torch.ones(128 * 1024 * 1024, dtype=torch.float64, device=dev)      # 1 GiB
    b = torch.empty_like(a)
    c = torch.ones(256 * 1024 * 1024, dtype=torch.float64, device=dev)      # 2 GiB
    out = {"mode": "hbm"}
    for name, fn, nbytes in [("copy_1GiB", lambda: b.copy_(a), 2 * a.numel() * 8),
                             ("sum_2GiB", lambda: c.sum(), c.numel() * 8)]:
        for _ in range(3):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(20):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 20
        out[name] = {"bytes": nbytes, "ms": round(ms, 4), "GBps": round(nbytes / (ms * 1e-3) / 1e9, 1)}
    print(json.dumps(out), flush=True)


def kernel_bytes(group, info, n):
    """Algorithmic bytes per matvec of the launches of a launch group (FP64,
    c = 16 B): the per-unit figures times the units the group's launches
    process in one matvec (one launch for S2M, L2T, NEAR, TREE_UP; one per
    level for M2M and L2L; the M2L levels in one or two launches).
      S2M, L2T  32 B per (source, leaf direction) of radiation patterns, plus
                the leaf fields (32 B per leaf direction) once
      NEAR      16 B per stored near-field value
      M2L       compulsory: operators (16 B per offset and direction), source
                multipoles and target locals (32 B per box direction) once;
                streamed: 32 B per (pair, direction) of source reads instead
      M2L_NEAR  the M2L bytes of the per-target levels plus NEAR
      M2M       per parent level: the child multipoles read and the parent
                multipoles written once (32 B per box direction)
      L2L       per child level: the parent locals read, the child locals
                read and written once (32 B per box direction)
      TREE_UP   S2M plus M2M
    Returns (bytes, streamed bytes)."""
    c = 16
    ml = info["max_level"]
    nd = [info["nth"][l] * info["nph"][l] for l in range(ml + 1)]
    B, P = info["boxes"], info["m2l_pairs"]
    leaf = info["radpat_entries"] * 2 * c + info["boxes"][ml] * nd[ml] * 2 * c
    near = info["near_nnz"] * c
    m2l_comp = sum(316 * nd[l] * c + 2 * B[l] * nd[l] * 2 * c for l in range(2, ml + 1) if P[l] > 0)
    m2l_str = sum(P[l] * nd[l] * 2 * c + 316 * nd[l] * c + B[l] * nd[l] * 2 * c for l in range(2, ml + 1) if P[l] > 0)
    if group in ("S2M", "L2T"):
        return leaf, leaf
    if group == "NEAR":
        return near, near
    if group in ("M2L_TARGET", "M2L_RUNS"):
        return m2l_comp, m2l_str
    if group == "M2L_NEAR":
        return m2l_comp + near, m2l_str + near
    m2m = sum((B[l + 1] * nd[l + 1] + B[l] * nd[l]) * 2 * c for l in range(2, ml))
    if group == "M2M":
        return m2m, m2m
    if group == "L2L":
        l2l = sum((B[l - 1] * nd[l - 1] + 2 * B[l] * nd[l]) * 2 * c for l in range(3, ml + 1))
        return l2l, l2l
    if group == "TREE_UP":
        return leaf + m2m, leaf + m2m
    return None, None


KERNEL_NAMES = {"S2M": "kS2Mx", "L2T": "kL2TS", "NEAR": "kNearSym", "M2L_TARGET": "kM2Lw", "M2L_NEAR": "kM2LNear",
                "M2M": "kM2MFused", "L2L": "kL2LFused", "M2L_SIBLING": "kM2LGroup", "TREE_UP": "kTreeUp",
                "M2L_RUNS": "kM2LT"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="sphere200k",
                    help="sphere200k (default: the 201,720-RWG sphere), sphere30k, dipoles4k, multi16, "
                         "dipoles200k_grad")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-solve", action="store_true", help="skip the GMRES solves")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--tuning", default=None,
                    help="JSON file of launch parameters: written after autotuning when absent, applied instead "
                         "of autotuning when present (profiled runs then launch only the timed schedule)")
    ap.add_argument("--timed-only", action="store_true",
                    help="only the timed matvecs and their probe passes (one schedule; for kernel profiles)")
    ap.add_argument("--nrhs", type=int, default=1,
                    help="with --timed-only: right-hand sides per matvec call (1, 2, 4 or 8)")
    ap.add_argument("--fp32", action="store_true",
                    help="with --timed-only: the plan's far field in FP32 (the inner matvec of the "
                         "mixed-precision solves) instead of FP64")
    ap.add_argument("--parts", default="single,batched,c4,ellipsoids",
                    help="with --mode deflation: which parts to run (single, batched, c4, ellipsoids; dr: "
                         "only the FP64 GMRES-DR(100, 20) solve of the workload)")
    ap.add_argument("--mode", default="matvec", choices=["matvec", "copy", "m2l-ceiling", "hbm", "c4-stalls",
                                                              "deflation"],
                    help="copy: only run a plain 256 MiB device copy (HBM counter calibration); m2l-ceiling: the "
                         "M2L access pattern (per target, the sum of its sources' fields and operators) through "
                         "torch's gather-sum (embedding_bag) on the workload's interaction lists; hbm: torch's "
                         "copy and sum rates on GiB arrays; c4-stalls: the stalled c4 ellipsoids against finer "
                         "rules and meshes and 1-2% size changes, and dense singular values over size sweeps; "
                         "deflation: GMRES(m) against deflated restarts (GMRES-DR) on the physical near field of "
                         "the workload and of the c4 ellipsoids")
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

    if args.mode == "m2l-ceiling":
        m2l_ceiling(args)
        return

    if args.mode == "hbm":
        hbm_rates()
        return

    if args.mode == "c4-stalls":
        c4_stalls_job()
        return

    if args.mode == "deflation":
        deflation_job(args)
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
    if args.fp32 and not args.timed_only:
        raise SystemExit("--fp32 goes with --timed-only")
    prec = A.MLFMA_FP32 if args.fp32 else A.MLFMA_FP64
    plan = Plan(geom, pw, A.make_config(device=0, precision=prec, **kw))
    setup_s = time.time() - t0
    info = plan.info()
    log("plan", {k: info[k] for k in ["max_level", "boxes", "m2l_pairs", "near_nnz", "radpat_entries",
                                      "stale_boxes", "truncated_l2l", "device_bytes"]}, "setup s", round(setup_s, 2))
    n = geom.n
    dev = torch.device("cuda:0")
    if args.nrhs != 1 and not args.timed_only:
        raise SystemExit("--nrhs goes with --timed-only")
    if args.nrhs == 1:
        x = torch.from_numpy(G.complex_gaussian(n, 0)).to(dev)
    else:
        x = torch.from_numpy(np.stack([G.complex_gaussian(n, r) for r in range(args.nrhs)])).to(dev)
        plan.set_flags(A.MLFMA_FLAG_FRESH_LOCALS)
    y = torch.empty_like(x)
    # the library's autotuner picks launch shapes and the stage schedule
    if args.tuning and os.path.exists(args.tuning):
        with open(args.tuning) as f:
            plan.apply_tuning(json.load(f)["record"])
        tune_before = tune_after = None
    else:
        tune_before, tune_after = plan.autotune(args.nrhs)
        if args.tuning:
            with open(args.tuning, "w") as f:
                json.dump({"workload": desc["workload"], "record": plan.tuning()}, f)
    schedule = {"near_overlap": plan.get_param(A.MLFMA_PARAM_NEAR_OVERLAP),
                "m2l_split": plan.get_param(A.MLFMA_PARAM_M2L_SPLIT),
                "near_window": plan.get_param(A.MLFMA_PARAM_NEAR_WINDOW),
                "near_blocks_per_cu": plan.get_param(A.MLFMA_PARAM_NEAR_BLOCKS),
                "tree_walk": plan.get_param(A.MLFMA_PARAM_TREE_WALK),
                "near_split_pct": plan.get_param(A.MLFMA_PARAM_NEAR_SPLIT)}
    log("autotune ms", tune_before, "->", tune_after, schedule)

    def kname(g):
        # eight right-hand sides: the matrix-core near field and L2T (with the
        # near-field fold); one: beside the chain with resident blocks the
        # near field runs kNearSymP
        if args.nrhs == 8 and g in ("NEAR", "L2T"):
            return {"NEAR": "kNearSymW", "L2T": "kNearSum8+kL2TM"}[g]
        if args.nrhs > 1 and g == "L2T":
            return "kL2Tx"
        if g == "NEAR" and args.nrhs == 1 and schedule["near_overlap"] and schedule["near_blocks_per_cu"]:
            return "kNearSymP"
        return KERNEL_NAMES[g]

    fresh = A.MLFMA_FLAG_FRESH_LOCALS if args.nrhs > 1 else 0
    plan.set_flags(A.MLFMA_FLAG_ASYNC | fresh)
    for _ in range(args.warmup):
        plan.matvec(x, y)
    # timed region: K matvecs, uninstrumented, queued on torch's stream (the
    # library returns once a matvec is queued; the synchronize closes it)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.matvec(x, y)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1000.0 * elapsed / args.steps
    if plan.get_param(A.MLFMA_PARAM_TREE_STALLS):
        raise SystemExit("a tree-walk wait hit its poll bound: the timed matvecs are wrong")
    value = args.steps * args.nrhs / elapsed
    y_timed = y.cpu().numpy()
    # the same K matvecs once per launch group with that group's launches
    # bracketed by HIP events on the stream each runs on (same schedule)
    groups = {}
    for g in ["TREE_UP", "S2M", "M2M", "M2L_RUNS", "M2L_TARGET", "M2L_SIBLING", "M2L_NEAR", "L2L", "NEAR", "L2T"]:
        plan.probe(g)
        for _ in range(args.steps):
            plan.matvec(x, y)
        torch.cuda.synchronize()
        ms, cnt = plan.probe_read()
        if cnt:
            groups[g] = {"kernel": kname(g), "launches_per_matvec": round(cnt / args.steps, 3),
                         "avg_launch_ms": round(ms / cnt, 5), "ms_per_matvec": round(ms / args.steps, 5)}
    plan.probe(None)
    log("launch groups", groups)
    if args.timed_only:
        print(json.dumps({"metric": f"MLFMA matvec/s ({'FP32 far field' if args.fp32 else 'FP64'}, one MI355X)",
                          "value": round(value, 3), "nrhs": args.nrhs,
                          "unit": "matvec/s", "ms_per_step": round(ms_per_step, 4), "steps": args.steps,
                          "config": desc, "schedule": schedule, "kernels": groups}), flush=True)
        return
    # other schedules of the same matvec
    rates = {}
    for name, fl in [("one_stream", A.MLFMA_FLAG_ASYNC | A.MLFMA_FLAG_ONE_STREAM),
                     ("graph_replay", A.MLFMA_FLAG_ASYNC | A.MLFMA_FLAG_GRAPH), ("synchronous_calls", 0)]:
        plan.set_flags(fl)
        plan.matvec(x, y)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            plan.matvec(x, y)
        torch.cuda.synchronize()
        rates[name] = round(args.steps / (time.perf_counter() - t1), 3)
    # per-stage device times, every stage in order with events (the
    # reference's ClockTimes)
    plan.set_flags(A.MLFMA_FLAG_STAGE_TIMES)
    stages = {k: 0.0 for k in ["S2M", "M2M", "M2L", "L2L", "L2T", "near", "total"]}
    near_pairs = 0.0
    for _ in range(args.steps):
        plan.matvec(x, y)
        for k, v in plan.stage_times().items():
            stages[k] += v
        near_pairs += plan.stage_times_detail()["near_pairs"]
    plan.set_flags(0)
    stage_ms = {k: round(v / args.steps, 4) for k, v in stages.items()}
    near_pairs_ms = near_pairs / args.steps
    log("stage ms", stage_ms, "rates", rates)

    # roofline of the kernel that takes the most device time per matvec in
    # the timed schedule (among kernels with a byte model)
    cand = {g: v for g, v in groups.items() if kernel_bytes(g, info, n)[0]}
    dom = max(cand, key=lambda g: cand[g]["ms_per_matvec"])
    kb, kstream = kernel_bytes(dom, info, n)
    avg = groups[dom]["avg_launch_ms"]
    per = groups[dom]["ms_per_matvec"]  # = avg for a group of one launch per matvec
    achieved = kb / (per * 1e-3) / 1e9
    tag = desc["workload"]
    traffic = measured_traffic(kname(dom), tag)
    # (a group of several launches per matvec: the per-launch bytes are the
    # mean over its launches, so bytes / average launch time = bytes per
    # matvec / its launches' time per matvec)
    nl = groups[dom]["launches_per_matvec"]
    roofline = {"bound": "hbm", "kernel": kname(dom), "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": round(traffic) if traffic else None, "bytes_per_launch": int(kb / nl),
                "avg_launch_ms": avg, "launches_per_matvec": nl, "bytes_per_matvec": int(kb), "launch_group": dom}
    if kstream != kb:
        roofline["streamed_bytes_per_launch"] = int(kstream)
        roofline["streamed_frac"] = round(kstream / (per * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    if dom == "NEAR" and near_pairs_ms > 0:
        # the same pair products alone, every stage in order (one block per
        # work item, kNearSym): the kernel's own rate, against the rate it
        # keeps beside the far-field chain in the timed schedule
        roofline["alone_in_order"] = {"kernel": "kNearSym", "launch_ms": round(near_pairs_ms, 5),
                                      "achieved": round(kb / (near_pairs_ms * 1e-3) / 1e9, 1),
                                      "frac": round(kb / (near_pairs_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    log("roofline", roofline)
    kernels = {}
    for g, v in groups.items():
        b, bs = kernel_bytes(g, info, n)
        e = dict(v)
        if b:
            e["bytes_per_matvec"] = int(b)
            e["frac_hbm_peak"] = round(b / (v["ms_per_matvec"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        kernels[g] = e

    # batched right-hand sides: 8 vectors per call share every operator read
    nb = 8
    Xb = torch.from_numpy(np.stack([G.complex_gaussian(n, 50 + r) for r in range(nb)])).to(dev)
    Yb = torch.empty_like(Xb)
    plan.set_flags(A.MLFMA_FLAG_FRESH_LOCALS)
    tb_before, tb_after = plan.autotune(nb)
    plan.set_flags(A.MLFMA_FLAG_ASYNC | A.MLFMA_FLAG_FRESH_LOCALS)
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
    batched = {"nrhs": nb, "matvec_per_s": round(nb * calls / el_b, 3), "ms_per_call": round(1000 * el_b / calls, 4)}
    log("batched", batched)

    sb, comp = byte_model(info, n)
    total_b = sum(sb.values())
    bytes_model = {
        "matvec_streamed_bytes": int(total_b),
        "matvec_streamed_frac_hbm_peak": round(total_b / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 3),
        "compulsory_bytes": int(comp),
        "compulsory_frac_hbm_peak": round(comp / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 3),
    }

    result = {
        "metric": "MLFMA matvec/s (FP64, one MI355X); GMRES solve s to 1e-6, mixed/batched vs plain FP64",
        "value": round(value, 3), "unit": "matvec/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": desc, "roofline": roofline, "kernels": kernels, "schedule": schedule, "stage_ms": stage_ms,
        "setup_s": round(setup_s, 2), "rates_other_schedules": rates, "batched": batched,
        "byte_model": bytes_model,
    }

    if not args.no_solve:
        result["solve"] = solves(geom, pw, kw, desc)

    if not args.no_cpu:
        log("cpu baseline ...")
        x0 = G.complex_gaussian(n, 0)
        if info["stale_boxes"]:
            # the reference's stale locals start at zero: compare a fresh plan's first matvec
            y_timed = Plan(geom, pw, A.make_config(device=0, **kw)).matvec(x0)
        result["cpu_baseline"] = cpu_baseline(geom, pw, kw, args.cpu_budget, y_timed, x0)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
