/*
 * mlfma_c.h — C-ABI of the MI355X (gfx950) Helmholtz MLFMA matvec and GMRES
 * Method-of-Moments solve.
 *
 * The reference (ellilu11/MLFMA) has no FFI: its boundary is a C++ class
 * contract (src/mlfma.cpp:21-82).  Each entry point below replaces one step of
 * that contract; the citation names the reference code it stands in for.
 *
 * Conventions
 *   - Complex numbers are interleaved FP64 (re, im) pairs.
 *   - Unknowns are numbered as in the input files (the reference's Source::idx,
 *     src/sources/source.h:36-38); the library reorders internally.
 *   - Return codes: 0 on success, a negative MLFMA_ERR_* code on failure; the
 *     message of the last failure on this thread is given by mlfma_last_error().
 *     No C++ exception crosses this boundary.
 *   - One plan = one HIP device.  Plans are independent of each other; calls on
 *     one plan must not overlap (the plan owns its scratch buffers and, in the
 *     reference-compatible mode, the persistent local coefficients).
 */
#ifndef MLFMA_C_H
#define MLFMA_C_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status codes */
#define MLFMA_OK 0
#define MLFMA_ERR_ARG (-1)          /* bad argument                               */
#define MLFMA_ERR_OOM (-2)          /* device or host allocation failed           */
#define MLFMA_ERR_HIP (-3)          /* HIP runtime error                          */
#define MLFMA_ERR_BREAKDOWN (-4)    /* GMRES breakdown (zero or non-finite norm)  */
#define MLFMA_ERR_NONFINITE (-5)    /* non-finite value in the input              */
#define MLFMA_ERR_STATE (-6)        /* call not valid for this plan (host-only)   */
#define MLFMA_ERR_GEOMETRY (-7)     /* geometry outside the root box, bad mesh    */
#define MLFMA_ERR_NOT_CONVERGED (-8)/* GMRES stopped at maxit above tolerance     */

/* ------------------------------------------------------------ inputs */
#define MLFMA_SRC_DIPOLE 0   /* src/sources/dipole.h */
#define MLFMA_SRC_RWG 1      /* src/sources/rwg.{h,cpp} */

/* Discretisation.  Replaces importFromConfig / importDipoles / importRWG
 * (src/fileio.h:106-202, 220-263): the caller passes the parsed arrays. */
typedef struct mlfma_geometry {
    int32_t kind;            /* MLFMA_SRC_DIPOLE or MLFMA_SRC_RWG                   */
    int32_t reserved;
    int64_t n;               /* number of unknowns                                  */
    const double* dip_pos;   /* dipoles: n x 3 positions                            */
    const double* dip_mom;   /* dipoles: n x 3 moment vectors (Dipole pol)          */
    int64_t nverts;          /* RWG: vertices                                       */
    const double* verts;     /*      nverts x 3                                     */
    int64_t ntris;           /* RWG: triangles                                      */
    const int32_t* tris;     /*      ntris x 3 vertex indices (0-based)             */
    const int32_t* rwgs;     /* RWG: n x 4 = v0 v1 t_plus t_minus                   */
} mlfma_geometry;

/* Incident plane wave (src/excite.h:6-28, src/fileio.h:8-27). */
typedef struct mlfma_planewave {
    double amplitude;
    double pol[3];
    double khat[3];
    double wavenum;
} mlfma_planewave;

#define MLFMA_FP64 0
#define MLFMA_FP32 1         /* far-field stages in FP32 complex, FP32 accumulate;
                                the FP32 path also reads an FP32 copy of the
                                near-field values (FP64 products) */

/* flags */
#define MLFMA_FLAG_FRESH_LOCALS 1u  /* zero the locals of boxes with an empty
                                       interaction list every matvec instead of
                                       keeping the previous matvec's values
                                       (the reference keeps them, src/fmm/node.cpp:106) */
#define MLFMA_FLAG_STAGE_TIMES 2u   /* record per-stage device events on the first
                                       right-hand side of each matvec (the
                                       reference's ClockTimes, src/clock.h:8-17);
                                       each event costs a few microseconds of
                                       device idle time */
#define MLFMA_FLAG_M2L_GROUP 4u     /* M2L through the sibling-group kernel on
                                       every level (default: the target-run
                                       kernel, or the autotuned choice per
                                       level) */
#define MLFMA_FLAG_M2L_TARGET 8u    /* M2L through the per-target kernel on
                                       every level */
#define MLFMA_FLAG_PHYSICAL 32u     /* RWG only, fixed at creation: near-field
                                       entries in the physical mixed-potential
                                       form with the singular 1/R integrals done
                                       in closed form, coincident triangles and
                                       the diagonal included (the reference drops
                                       them, src/sources/rwg.cpp:97,113-116) */
#define MLFMA_FLAG_GRAPH 64u        /* capture a matvec's launch sequence into a
                                       hipGraph once per (x, y, nrhs,
                                       accumulate, precision) and replay it
                                       (up to 4 cached; cleared when the flags
                                       change; not with stage times or a
                                       probe group) */
#define MLFMA_FLAG_ONE_STREAM 128u /* run every stage in order on the
                                       matvec's stream, instead of the
                                       near-field pair products and the
                                       finest level's M2L beside the
                                       far-field chain on two more streams of
                                       the plan (stage-timed matvecs always
                                       run in order) */
#define MLFMA_FLAG_ASYNC 16u        /* matvecs on device pointers return once
                                       queued on the stream instead of after
                                       completion (order by the stream; host
                                       pointers always complete) */

/* Numerical configuration (src/config.h:50-73) plus device selection. */
typedef struct mlfma_config {
    int32_t digits;          /* excess-bandwidth digits (src/fmm/angles.cpp:9-10)  */
    int32_t interp_order;    /* stencil half-width p; 2p taps (src/fmm/tables.cpp)  */
    double over_interp;      /* psi oversampling (src/fmm/tables.cpp:97)            */
    double root_leng;        /* root box side, centred at the origin                */
    int32_t max_node_srcs;   /* split a box while it holds more sources            */
    int32_t max_level;       /* -1: no cap (reference rule); else also stop at this level */
    int32_t quad_prec;       /* 0,1,2,3 = 1,3,7,13-point triangle rule              */
    int32_t device;          /* HIP device ordinal; -1 = host-only plan             */
    int32_t precision;       /* MLFMA_FP64 or MLFMA_FP32                            */
    uint32_t flags;          /* MLFMA_FLAG_*                                        */
} mlfma_config;

typedef struct mlfma_plan mlfma_plan;

/* Plan summary (counts printed by src/mlfma.cpp:34-36 and
 * src/fmm/angles.cpp:30, plus sizes of the device-resident data). */
typedef struct mlfma_info {
    int64_t n;
    int32_t max_level;
    int32_t num_levels_m2l;        /* levels with M2L work                          */
    int64_t num_nodes;
    int64_t num_leaves;
    int64_t nth[16], nph[16];      /* angular grid per level                        */
    int64_t boxes[16];             /* non-empty boxes per level                     */
    int64_t m2l_pairs[16];         /* interaction-list pairs per level              */
    int64_t near_nnz;              /* stored near-field entries: each unordered pair 
                                      once, self blocks as full squares             */
    int64_t radpat_entries;        /* sum over leaves of srcs x directions          */
    int64_t stale_boxes;           /* non-empty boxes (level >= 2) with empty iList */
    int64_t truncated_l2l;         /* boxes whose parent's iList is empty (level>=3)*/
    int64_t psi_merged;            /* psi keys merged by the translation-table dedup*/
    int64_t device_bytes;          /* device memory held by the plan                */
} mlfma_info;

/* --------------------------------------------------------------- plan */

/* Build everything the matvec needs, in the order of src/mlfma.cpp:21-64:
 * sources and RHS (Node::initStatic / Source::buildVoltage), tree
 * (Stem/Leaf ctors, src/fmm/stem.cpp:3-26), lists (buildLists,
 * src/fmm/node.cpp:61-99, src/fmm/nbor.cpp), near field (Leaf::buildNearRads,
 * src/fmm/leaf.cpp:108-180), angular grids and tables (Node::buildTables,
 * src/fmm/node.cpp:46-56), coefficient storage (resizeCoeffs) and radiation
 * patterns (Leaf::buildRadPats, src/fmm/leaf.cpp:185-207).  With
 * config->device = -1 only the host part (tree, lists, grids, interpolation
 * tables) is built; no HIP call is made. */
int mlfma_plan_create(const mlfma_geometry* geom, const mlfma_planewave* pw,
                      const mlfma_config* cfg, mlfma_plan** out);

int mlfma_plan_destroy(mlfma_plan* plan);

/* One plan for `count` independent scatterers (no reference counterpart: the
 * reference holds one tree per process in static state,
 * src/fmm/node.h:93-105).  Scatterer s has geometry geoms[s] and configuration
 * cfgs[s]; all share the plane wave, the source kind, digits, interpolation
 * order, oversampling, quadrature, precision, flags and device, and every
 * root side must be the largest one divided by a power of two (equal leaf
 * boxes at equal depth below it).  Levels are aligned by box size, so every
 * matvec stage runs once for all scatterers over the concatenated worklists
 * of each level (one launch per stage and level) and the far-field tables
 * are shared.  The unknowns of scatterer s follow those of scatterer s - 1
 * (mlfma_plan_segments); the operator is block diagonal.  mlfma_matvec,
 * mlfma_matvec_adjoint, mlfma_solve_batch (GMRES per scatterer and
 * right-hand side), mlfma_excitation and mlfma_farfield (over all sources)
 * take forest plans; mlfma_solve and the level introspection calls do not. */
int mlfma_plan_create_forest(const mlfma_geometry* geoms, const mlfma_config* cfgs, int32_t count,
                             const mlfma_planewave* pw, mlfma_plan** out);

/* Unknown offsets of the plan's scatterers: offsets[0..count] (count = 1 for
 * a single-scatterer plan); at most cap values written. */
int mlfma_plan_segments(const mlfma_plan* plan, int64_t* offsets, int32_t cap, int32_t* count);

int mlfma_plan_info(const mlfma_plan* plan, mlfma_info* out);

/* Excitation vector V (Source::getVoltage, src/sources/source.h:20), host
 * buffer of n complex values.  The reference solves Z I = -V
 * (src/solver.cpp:31-32). */
int mlfma_voltage(const mlfma_plan* plan, double* v_host);

/* The excitation V of the same sources under another plane wave of the
 * plan's wavenumber (buildVoltage, src/sources/dipole.h:29-32 and
 * src/sources/rwg.h:18-21, for the given incidence): n complex (host, input
 * order).  For sweeps of incidence angles; the solve's right-hand side is -V. */
int mlfma_excitation(mlfma_plan* plan, const mlfma_planewave* pw, double* v_host);

#define MLFMA_PTR_HOST 0
#define MLFMA_PTR_DEVICE 1

/* y = Z x, or y += Z x when accumulate != 0 — the drop-in for
 * Solver::evalRvec (src/solver.cpp:41-57), which reads Node::lvec and adds into
 * Node::rvec.  x and y hold nrhs column vectors of n complex values each
 * (column stride n).  ptr_mode says whether x and y are host or device
 * pointers; stream is a hipStream_t (NULL = the plan's own stream).  The call
 * returns after the result is complete in y (device pointers with
 * MLFMA_FLAG_ASYNC: once the work is queued on the stream). */
int mlfma_matvec(mlfma_plan* plan, const double* x, double* y, int32_t nrhs,
                 int32_t accumulate, int32_t ptr_mode, void* stream);

/* y = Z^H x, or y += Z^H x: the exact adjoint of the matvec above (the
 * reference has none; it is the transposed stage chain: S2M <-> L2T swapped,
 * L2L and M2M transposed with conjugated phases, M2L with conjugated operators
 * of the reversed pairs, conjugated near field).  Arguments as mlfma_matvec.
 * MLFMA_ERR_STATE when the plan keeps stale locals (info.stale_boxes > 0 and
 * MLFMA_FLAG_FRESH_LOCALS not set): that operator is not linear. */
int mlfma_matvec_adjoint(mlfma_plan* plan, const double* x, double* y, int32_t nrhs,
                         int32_t accumulate, int32_t ptr_mode, void* stream);

/* Dense reference matvec over the same near-field kernel, all pairs
 * (the reference's evalDirect path, src/mlfma.cpp:84-107); for small n. */
int mlfma_matvec_direct(mlfma_plan* plan, const double* x, double* y, int32_t ptr_mode,
                        void* stream);

/* --------------------------------------------------------------- solve */
#define MLFMA_GMRES_COMPAT 0   /* src/solver.cpp: unrestarted, complex Givens without
                                  conjugates, stops on |g_k+1|/g0                  */
#define MLFMA_GMRES_RESTART 1  /* GMRES(m), unitary Givens, true residual at restart */
#define MLFMA_FGMRES_IR 2      /* mixed precision: FP64 flexible GMRES(outer_restart)
                                  whose variable preconditioner is an inner GMRES
                                  with the FP32 far field (at most `restart`
                                  iterations or a relative inner_tol per step);
                                  FP64 true residual at every outer restart     */
#define MLFMA_GMRES_IR 3       /* mixed precision by iterative refinement: FP64
                                  residual, restarted FP32-far-field GMRES(restart)
                                  corrections (no outer Krylov space), at most 20
                                  (30 batched, each inner solve at most 30 x
                                  restart steps) corrections                    */
#define MLFMA_SOLVE_ADJOINT 0x100  /* or-ed into method: solve Z^H x = b instead
                                      (the adjoint solve of a gradient; no
                                      reference counterpart)                       */
#define MLFMA_SOLVE_DEVICE 0x200   /* or-ed into method (mlfma_solve): b and x are
                                      device pointers on the plan's device; b
                                      must be complete when the call is made
                                      (the solve runs on the plan's stream)        */

#define MLFMA_PRECOND_NONE 0
#define MLFMA_PRECOND_LEAF_BLOCKS 1  /* block Jacobi: the inverse of each leaf's
                                        self block (Z restricted to the leaf's
                                        sources), built on first use; no
                                        reference counterpart */

typedef struct mlfma_solve_opts {
    int32_t method;       /* MLFMA_GMRES_*                                          */
    int32_t restart;      /* m for GMRES(m) / inner iterations for FGMRES; <= 2048 */
    int32_t maxit;        /* total iteration (matvec) cap (reference: 500)          */
    int32_t use_voltage;  /* 1: b = -V (reference), 0: b from the argument          */
    double tol;           /* reference: 1e-6                                        */
    int32_t outer_restart;/* MLFMA_FGMRES_IR: outer steps per cycle (0: 30)         */
    int32_t precond;      /* MLFMA_PRECOND_*: right preconditioner of the GMRES(m)
                             steps (inner steps for FGMRES_IR / GMRES_IR); not
                             with MLFMA_GMRES_COMPAT                               */
    double inner_tol;     /* MLFMA_FGMRES_IR: relative tolerance of each inner
                             solve (0: 1e-1)                                        */
    double shift_re;      /* solve (Z + s I) x = b with s = shift_re + i shift_im   */
    double shift_im;      /* (0: Z x = b, the reference); the adjoint solves
                             (Z + s I)^H x = b.  E.g. a dipole self term, which
                             the reference drops (src/sources/dipole.h:65)          */
    int32_t deflate;      /* k > 0: deflated restarts (GMRES-DR, Morgan 2002) for
                             MLFMA_GMRES_RESTART and the inner solves of
                             MLFMA_GMRES_IR, single and batched: each restart
                             keeps the k harmonic Ritz vectors of smallest
                             modulus (k + 1 < restart, k <= 64; not with
                             MLFMA_FGMRES_IR).  0: plain restarts.  No reference
                             counterpart (the reference's GMRES never restarts) */
} mlfma_solve_opts;

typedef struct mlfma_solve_report {
    int32_t iterations;   /* matvecs spent in the Krylov loop                       */
    int32_t status;       /* MLFMA_OK, MLFMA_ERR_BREAKDOWN, MLFMA_ERR_NOT_CONVERGED */
    double resid_est;     /* final |g|/g0 (compat) or implicit residual             */
    double resid_true;    /* |b - Z x| / |b| evaluated with one extra FP64 matvec    */
    double seconds;       /* wall time of the solve                                 */
} mlfma_solve_report;

/* Solve Z x = b — the drop-in for Solver::solve (src/solver.cpp:116-146)
 * followed by the currents of Solver::getCurrents.  b and x are host
 * buffers of n complex values (device buffers with MLFMA_SOLVE_DEVICE in
 * opts->method); b may be NULL when opts->use_voltage = 1.
 * resid_hist (optional, host, maxit doubles) receives the per-iteration
 * residual estimate. */
int mlfma_solve(mlfma_plan* plan, const double* b, double* x, const mlfma_solve_opts* opts,
                mlfma_solve_report* report, double* resid_hist);

/* GMRES(m) on nrhs right-hand sides at once (e.g. one scatterer, several
 * incidence angles): one batched matvec per step over all of them (no
 * reference counterpart).  b, x: nrhs x n complex (host, column stride n);
 * opts->method MLFMA_GMRES_RESTART, MLFMA_FGMRES_IR or MLFMA_GMRES_IR;
 * opts->use_voltage is ignored.  Every (right-hand side, scatterer) pair is
 * a lane with its own Krylov space and stopping test (one scatterer per plan
 * except for forest plans, mlfma_plan_segments): reports holds
 * nrhs x segments entries, lane r * segments + s (or NULL); resid_hist the
 * same lanes x maxit (or NULL).  Non-convergence is reported per lane in
 * reports[k].status. */
int mlfma_solve_batch(mlfma_plan* plan, const double* b, double* x, int32_t nrhs,
                      const mlfma_solve_opts* opts, mlfma_solve_report* reports, double* resid_hist);

/* y = P x (adjoint != 0: P^H x) for the preconditioner of kind
 * MLFMA_PRECOND_* (built on first use); x, y: nrhs vectors of n complex
 * (column stride n, input order), ptr_mode and stream as in mlfma_matvec.
 * No reference counterpart. */
int mlfma_precond_apply(mlfma_plan* plan, int32_t kind, const double* x, double* y, int32_t nrhs, int32_t adjoint,
                        int32_t ptr_mode, void* stream);

/* --------------------------------------------------------------- far field */

/* E(rhat_d) = C k (I - rhat rhat^T) sum_n I_n f_n(k rhat_d) for ndir unit
 * directions (generalises Node::printFarSols, src/fmm/node.cpp:147-178, to any
 * direction list).  I: n complex (input order); dirs: ndir x 3;
 * E: ndir x 3 complex.  ptr_mode: all three on the host or all three on the
 * plan's device; stream as in mlfma_matvec (device pointers with
 * MLFMA_FLAG_ASYNC: returns once queued).  The source sums are split into
 * ranges whose partial sums are added in a fixed order (results do not
 * depend on scheduling). */
int mlfma_farfield(mlfma_plan* plan, const double* I, const double* dirs, int64_t ndir, double* E,
                   int32_t ptr_mode, void* stream);

/* mlfma_farfield on device pointers that allocates nothing: no reference
 * counterpart; for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  I (n
 * complex, input order), dirs (ndir x 3) and E (ndir x 3 complex) are on the
 * plan's device; the work is queued on stream (NULL: the plan's stream) and
 * the call returns without synchronising.  The source-range partial sums go
 * to one fixed array per device (3 MiB, set up by mlfma_plan_create, held
 * until the process ends), so ndir is limited to 65536 (MLFMA_ERR_ARG
 * beyond; mlfma_farfield has no limit).  Every call on a device shares that
 * device's array: calls that overlap on one device (other streams, other
 * threads, other plans) race; order them on one stream.  Same kernels and
 * results as mlfma_farfield. */
int mlfma_farfield_noalloc(mlfma_plan* plan, const double* I, const double* dirs, int64_t ndir, double* E,
                           void* stream);

/* The adjoint of mlfma_farfield: I = F^H E for ndir x 3 complex fields E at
 * the unit directions dirs; I: n complex (input order); ptr_mode and stream
 * as in mlfma_farfield.  No reference counterpart: the gradient of a
 * far-field objective. */
int mlfma_farfield_adjoint(mlfma_plan* plan, const double* E, const double* dirs, int64_t ndir, double* I,
                           int32_t ptr_mode, void* stream);

/* Position sensitivity of the operator for dipole sources:
 * grad[i][a] = -2 Re(lambda^H (dZ/dr_i^a) I) for complex n-vectors I and
 * lambda (host, input order); grad: n x 3 (host).  With I = Z^-1 b and
 * lambda = Z^-H F^H E this is the part of the gradient of
 * J = sum_d |E(rhat_d)|^2 that flows through the solve; the far-field term
 * 2 Re(E^H dF I) comes from mlfma_farfield_adjoint.  The tree is held
 * fixed (a derivative, not a re-partition).  No reference counterpart: the
 * reference computes no derivatives.  Requires a linear operator
 * (MLFMA_FLAG_FRESH_LOCALS when boxes keep stale locals); not defined for
 * RWG plans. */
int mlfma_position_gradient(mlfma_plan* plan, const double* I, const double* lambda, double* grad);

/* --------------------------------------------------------------- introspection */

/* Per-stage device time of the last matvec run with MLFMA_FLAG_STAGE_TIMES in
 * milliseconds (the reference's ClockTimes, src/clock.h:8-17): [0] gather+S2M,
 * [1] M2M, [2] M2L, [3] L2L, [4] L2T (with the near-field row sums folded
 * in), [5] near field (pair products), [6] total.
 * MLFMA_ERR_STATE if no such matvec has run. */
int mlfma_stage_times(const mlfma_plan* plan, double* ms7);

/* The same events one launch group at a time: [0] S2M, [1] M2M, [2] M2L,
 * [3] L2L, [4] near-field pair products (kNearSym), [5] near-field sums of
 * rows outside every far-field leaf (kNearSum), [6] L2T with the near-field
 * sums of the other rows folded in (kL2Tx).  Writes min(cap, *count) values;
 * *count = 7.  No reference counterpart (finer than ClockTimes). */
int mlfma_stage_times_detail(const mlfma_plan* plan, double* ms, int32_t cap, int32_t* count);

/* Choose the plan's launch parameters for batches of nrhs (1, 2, 4 or 8)
 * right-hand sides by timing candidates with device events: the block
 * target of the M2M (and adjoint L2L) row split, per level the sibling-group or per-target
 * M2L kernel, whether the near-field pair products run beside the far-field
 * chain (on a second stream of the plan, forked before or after S2M or after
 * the M2M chain; one block per work item or 1-4 resident blocks per CU) or in order (one right-hand side: in the launch of the
 * per-target M2L levels, over an early share of its blocks), and whether the finest level's M2L runs beside the M2M chain (on a
 * third stream) or in line; with nrhs = 1 also whether mlfma_solve runs its
 * matvecs in order (whichever of the two makes short GMRES runs on the
 * plan's excitation faster).  Mean matvec device
 * time before and after (ms, may be NULL).  Results do not depend on the
 * choice beyond round-off.  No reference counterpart. */
int mlfma_autotune(mlfma_plan* plan, int32_t nrhs, double* ms_before, double* ms_after);

/* Launch parameters that mlfma_autotune chooses, set by hand (no reference
 * counterpart).  Results do not depend on them beyond round-off.
 *   MLFMA_PARAM_NEAR_WINDOW    1..100: % of the shared in-order M2L /
 *                              near-field launch over which the near-field
 *                              pair blocks are spread (one right-hand side)
 *   MLFMA_PARAM_NEAR_OVERLAP   0 in order (in the M2L launch), 1 beside the
 *                              chain forked before S2M, 2 after S2M, 3 after M2M
 *   MLFMA_PARAM_M2L_SPLIT      1: the finest level's M2L on its own stream
 *   MLFMA_PARAM_SOLVE_IN_ORDER 1: one-right-hand-side solves run every stage
 *                              in order on one stream
 *   MLFMA_PARAM_NEAR_BLOCKS    0..8: near-field pairs beside the chain run
 *                              as that many resident blocks per CU looping
 *                              over the work items (0: one block per item)
 *   MLFMA_PARAM_TREE_WALK      bit 0: the upward pass of one right-hand side
 *                              (S2M and the M2M of every level) runs as one
 *                              persistent launch walking a packed worklist
 *                              of (leaf) and (level, parent box, row run)
 *                              units, each waiting on the completion
 *                              counters of the units that write its inputs
 *   MLFMA_PARAM_TREE_STALLS    (get only) waits of the tree walks that hit
 *                              their poll bound since the plan was made (a
 *                              non-zero count means a wrong result; reading
 *                              it waits for the plan's stream)
 *   MLFMA_PARAM_NEAR_SPLIT     0..100: beside the chain, % of the near-field
 *                              work items forked at the NEAR_OVERLAP point;
 *                              the rest fork after the M2L, beside the L2L
 *                              launches (100: one launch) */
#define MLFMA_PARAM_NEAR_WINDOW 1
#define MLFMA_PARAM_NEAR_OVERLAP 2
#define MLFMA_PARAM_M2L_SPLIT 3
#define MLFMA_PARAM_SOLVE_IN_ORDER 4
#define MLFMA_PARAM_PROBE_GROUP 5  /* MLFMA_GROUP_* whose launches carry probe
                                      events (0: none); see mlfma_probe_read */
#define MLFMA_PARAM_NEAR_BLOCKS 6
#define MLFMA_PARAM_TREE_WALK 7
#define MLFMA_PARAM_TREE_STALLS 8
#define MLFMA_PARAM_NEAR_SPLIT 9
int mlfma_plan_set_param(mlfma_plan* plan, int32_t param, int32_t value);
int mlfma_plan_get_param(const mlfma_plan* plan, int32_t param, int32_t* value);

/* Launch groups of a matvec.  With MLFMA_PARAM_PROBE_GROUP = g, every launch
 * of group g is bracketed by a pair of HIP events recorded on the stream the
 * launch runs on (whatever the schedule: side streams or in order), so its
 * device time is measured inside the schedule being timed; each record costs
 * that stream a few microseconds.  mlfma_probe_read returns the summed
 * milliseconds and the number of launches since the group was set or last
 * read, and resets both.  No reference counterpart (the reference's stage
 * timers, src/clock.h:8-17, time whole stages). */
#define MLFMA_GROUP_S2M 1          /* kS2Mx                                      */
#define MLFMA_GROUP_M2M 2          /* one M2M level (kM2MFused)                  */
#define MLFMA_GROUP_M2L_TARGET 3   /* per-target M2L levels (kM2Lw / kM2Lx)      */
#define MLFMA_GROUP_M2L_SIBLING 4  /* sibling-group M2L levels (kM2LGroup)       */
#define MLFMA_GROUP_M2L_NEAR 5     /* per-target M2L with the near-field pairs in
                                      the same launch (kM2LNear)                */
#define MLFMA_GROUP_L2L 6          /* one L2L level (kL2LFused)                  */
#define MLFMA_GROUP_NEAR 7         /* near-field pair products (kNearSym)        */
#define MLFMA_GROUP_L2T 8          /* kL2Tx                                      */
#define MLFMA_GROUP_TREE_UP 9      /* S2M + every M2M level as one walk (kTreeUp)*/
#define MLFMA_GROUP_M2L_RUNS 10    /* M2L levels over target runs (kM2LT)        */
#define MLFMA_GROUP_LAST 10
int mlfma_probe_read(mlfma_plan* plan, double* ms, int64_t* launches);

/* The launch parameters mlfma_autotune chose, as a record of int32 values
 * (*len of them; at most cap written), and their application to a plan of
 * the same geometry and configuration (another process, no timing runs).  No
 * reference counterpart. */
int mlfma_autotune_record(const mlfma_plan* plan, int32_t* rec, int32_t cap, int32_t* len);
int mlfma_autotune_apply(mlfma_plan* plan, const int32_t* rec, int32_t len);

/* Replace the plan's MLFMA_FLAG_* set for later matvecs and solves (the flags
 * act at apply time; the plan itself does not change).  No reference
 * counterpart: the reference fixes them at compile time. */
int mlfma_plan_set_flags(mlfma_plan* plan, uint32_t flags);

/* Boxes of one level in the plan's order: coords (nbox x 3 int32 integer box
 * coordinates at that level), node kind (1 leaf, 0 stem) and source count. */
int mlfma_level_boxes(const mlfma_plan* plan, int32_t level, int64_t* nbox, int32_t* coords,
                      int32_t* is_leaf, int32_t* nsrc, int64_t cap);

/* Interaction lists of one level as (target, source) box-index pairs into the
 * order of mlfma_level_boxes: kind 0 = M2L (list 2), 1 = near leaf pairs
 * (list 1, unordered, both leaves), 2 = list-4/list-3 pairs (leaf, node). */
int mlfma_level_pairs(const mlfma_plan* plan, int32_t level, int32_t kind, int64_t* npairs,
                      int32_t* pairs, int32_t* src_level, int64_t cap);

/* Multipole (which = 0) or local (which = 1) coefficients of one level after
 * the last matvec, nbox x ndir x 2 complex (theta, phi components), boxes in
 * the order of mlfma_level_boxes (empty boxes hold zeros).  FP64 plans only. */
int mlfma_get_coeffs(const mlfma_plan* plan, int32_t level, int32_t which, double* out,
                     int64_t cap);

/* Translation operators of one level, 316 x ndir complex, offsets in the
 * order of Math::getINodeDistVecs (src/math.cpp:109-124), without the M2L
 * quadrature weights. */
int mlfma_get_translation(const mlfma_plan* plan, int32_t level, double* out, int64_t cap);

const char* mlfma_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* MLFMA_C_H */
