// plan.hpp — device-resident plan of the MLFMA matvec (one plan = one device,
// one stream).  Layout in HBM:
//   * sources and vectors in tree order (perm maps tree position -> unknown);
//   * radiation patterns: per far-field leaf, [source][direction][theta|phi];
//   * multipole / local fields: per level >= 2, [box slot][direction][theta|phi];
//   * translation operators: per level, [316 offsets][direction], weights folded;
//   * near field: each unordered pair stored once (self blocks full), cut into
//     work items of <= 128 rows; column-major values (C k folded).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <complex>
#include <memory>
#include <string>
#include <vector>

#include "host.hpp"
#include "kernels.hpp"

namespace mlfma {

extern thread_local std::string g_lastError;
void hipCheck(hipError_t e, const char* what);

struct LevelBufs {
    int nth = 0, nph = 0, ndir = 0, nps = 0;
    std::vector<double> khat, dirW;  // host copies
    int64_t Moff = 0;                // first complex of this level in the field arrays
    cplx<double>* T64 = nullptr;
    cplx<float>* T32 = nullptr;
    cplx<double>*phUp64 = nullptr, *phDn64 = nullptr;
    cplx<float>*phUp32 = nullptr, *phDn32 = nullptr;
    // interpolation tables (this level as the coarse one)
    int32_t *thNear = nullptr, *phNear = nullptr;
    double *thW = nullptr, *phW = nullptr;
    int32_t *atPhPtr = nullptr, *atPhSrc = nullptr, *atThPtr = nullptr, *atThSrc = nullptr, *atThFlip = nullptr;
    double *atPhW = nullptr, *atThW = nullptr;
    int nAtPh = 0, nAtTh = 0;  // taps of the gather-form anterpolation tables
    // M2M into this level
    int nstem = 0;
    int32_t *stemSlot = nullptr, *mPtr = nullptr, *mChild = nullptr, *mOct = nullptr;
    // M2L targets of this level
    int ntgt = 0;
    int32_t *tgtSlot = nullptr, *lPtr = nullptr, *lSrc = nullptr, *lOff = nullptr;
    int32_t* lOffNeg = nullptr;  // offsets of the reversed pairs (adjoint)
    // M2L target runs (kM2LT): per run its first target and its source list
    int nrun = 0, runMaxSrc = 0;
    int32_t *runTgt = nullptr, *runSrc = nullptr, *runSlot = nullptr;
    int32_t *lPk = nullptr, *lPkNeg = nullptr;  // per pair: (source index in its run) << 9 | offset
    // M2L sibling groups: 8 target slots per group (-1 none), per group a CSR
    // of distinct sources, 8 offsets per source entry (-1: not in that list)
    int ngrp = 0;
    int64_t nGrpEntries = 0, nPairs = 0;
    int32_t *gTgt = nullptr, *gPtr = nullptr, *gSrc = nullptr, *gOff = nullptr, *gOffNeg = nullptr;
    int32_t* m2mAdjParent = nullptr;  // per M2M child entry: its parent slot (adjoint)
    int nm2mChild = 0;                // M2M child entries into this level
    int32_t* mChildCtr = nullptr;     // per M2M child entry: the tree-walk counter of its writer (-1: none)
    int nstale = 0;
    int32_t* staleSlots = nullptr;
    // L2L into this level
    int nl2l = 0;
    int32_t *l2lChild = nullptr, *l2lParent = nullptr, *l2lOct = nullptr;
    // L2L into this level grouped by parent (adjoint)
    int nl2lPar = 0;
    int32_t *l2lParSlot = nullptr, *l2lParPtr = nullptr, *l2lParChild = nullptr, *l2lParOct = nullptr;

    template <typename R>
    const cplx<R>* T() const;
    template <typename R>
    const cplx<R>* phUp() const;
    template <typename R>
    const cplx<R>* phDn() const;
};
template <>
inline const cplx<double>* LevelBufs::T<double>() const { return T64; }
template <>
inline const cplx<float>* LevelBufs::T<float>() const { return T32; }
template <>
inline const cplx<double>* LevelBufs::phUp<double>() const { return phUp64; }
template <>
inline const cplx<float>* LevelBufs::phUp<float>() const { return phUp32; }
template <>
inline const cplx<double>* LevelBufs::phDn<double>() const { return phDn64; }
template <>
inline const cplx<float>* LevelBufs::phDn<float>() const { return phDn32; }

struct DevicePlan {
    HostPlan hp;
    bool hostOnly = true;
    int device = -1;
    bool useFp32 = false;
    hipStream_t stream = nullptr;
    std::array<hipEvent_t, 9> ev{};
    hipStream_t nearStream = nullptr;  // near-field pairs beside the far-field chain
    hipEvent_t evFork = nullptr, evJoin = nullptr;
    int nearOverlap = 2;               // 0 in order (in the M2L launch), fork 1 before S2M, 2 after S2M,
                                       // 3 after M2M; autotuned
    hipStream_t m2lStream = nullptr;   // finest-level M2L beside the M2M chain
    hipEvent_t evM2L = nullptr, evNear = nullptr;
    bool m2lSplit = true;              // chosen by mlfma_autotune
    int nearWindow = 100;              // % of the shared M2L/near launch the pair blocks spread over; autotuned
    int nearBlocks = 0;                // near field on a side stream: blocks per CU (0: one per item); autotuned
    // near field on a side stream: % of its work items in the launch forked
    // at the nearOverlap point; the rest fork after the M2L (beside the L2L
    // chain) and join before the near-field sums.  100: one launch; autotuned
    int nearSplit = 100;
    hipEvent_t evNearB = nullptr, evJoinB = nullptr;
    int numCU = 0;                     // compute units of the device
    // persistent tree walks (MLFMA_PARAM_TREE_WALK bits; autotuned): 1 the
    // upward pass of one right-hand side (S2M and every M2M level) as one
    // launch over a packed worklist with per-unit completion counters
    int treeWalk = 0;
    int32_t* treeCtr = nullptr;        // counters of the walk, then ticket, exit count, stall count
    int treeCtrLen = 0;
    static constexpr int kTreeS2MPer = 4;  // leaf pieces per S2M unit of the walk
    // M2L target runs: at most this many distinct sources (64 LDS bytes each
    // per block) and targets per run
    static constexpr int kM2LRunSources = 256;
    static constexpr int kM2LRunTargets = 128;
    std::vector<int> treeCtr0;         // per M2M parent level: counter of its stem 0
    std::vector<std::array<int, 4>> treeOcc;  // (float?, block size, LDS bytes, blocks per CU) seen
    int treeGrid(bool fp32, int bs, size_t lds);
    template <typename R>
    bool launchTreeUpWalk(const cplx<double>* x, hipStream_t st);
    int nearGrid() const { return nearBlocks > 0 ? nearBlocks * numCU : 0; }
    bool timesValid = false;
    uint32_t flags = 0;  // MLFMA_FLAG_* in effect for matvecs (settable after creation)
    std::vector<void*> allocs;
    int64_t bytesHeld = 0;

    SrcDev src{};
    int64_t* perm = nullptr;
    std::vector<LevelBufs> levels;
    LevelDev* levelDev = nullptr;
    int64_t fieldLen = 0;
    cplx<double>*M64 = nullptr, *L64 = nullptr, *inner64 = nullptr, *rp64 = nullptr;
    cplx<float>*M32 = nullptr, *L32 = nullptr, *inner32 = nullptr, *rp32 = nullptr;
    LeafItem* leafItems = nullptr;
    int32_t* s2mOrder = nullptr;       // leaf items by source count, largest first (S2M launch order)
    int nLeafItems = 0, maxLeafNdir = 0;
    NearItem* nearItems = nullptr;     // longest first
    int nNearItems = 0;
    int32_t* nearColsIn = nullptr;     // input-order indices of the item columns
    int64_t* nearSumPtr = nullptr;     // per tree position: its partial sums
    int32_t* nearSumIdx = nullptr;
    cplx<double>* nearPart = nullptr;  // partial sums of one near-field apply
    int64_t* farRowPos = nullptr;      // tree positions of the sources of far-field leaves
    int32_t* farRowItem = nullptr;     // their leaf item
    int64_t nFarRows = 0;
    int64_t* nearOnlyRows = nullptr;   // tree positions outside every far-field leaf
    int64_t nNearOnlyRows = 0;
    size_t ldsMax = 65536;
    cplx<double>* nearVal = nullptr;
    cplx<float>* nearVal32 = nullptr;  // FP32 copy (FP32 plans: the FP32 far-field path)
    cplx<double> Ck{0, 0};
    std::vector<std::complex<double>> Vin;  // excitation, input order
    cplx<double>*xt = nullptr, *yt = nullptr, *xs = nullptr, *ys = nullptr;
    int rhsCap = 0;
    cplx<double>* farPart = nullptr;  // far-field partial sums (grown on demand)
    size_t farPartLen = 0;
    // leaf block-Jacobi preconditioner, built on first use: per non-empty leaf
    // the inverse of its self block (Z restricted to the leaf's sources)
    PrecBlock* precBlocks = nullptr;
    int nPrecBlocks = 0, precMaxCnt = 0, precRegularised = 0;
    cplx<double>* precVal = nullptr;
    void buildPrecond();
    // y = P x (P^H x when adjoint) for nb vectors of column stride xs
    void applyPrecond(const cplx<double>* x, cplx<double>* y, int nb, int64_t xs, bool adjoint, hipStream_t st);

    // launch-group probe (MLFMA_PARAM_PROBE_GROUP): an event pair around every
    // launch of one group, recorded on the stream that launch goes to, so
    // that a group's device time is measured in the schedule being timed
    int probeGroup = 0;
    struct ProbePair {
        hipEvent_t a = nullptr, b = nullptr;
    };
    std::vector<ProbePair> probeRing;
    size_t probeHead = 0, probePending = 0;
    double probeMs = 0;
    int64_t probeLaunches = 0;
    void probeBegin(int group, hipStream_t s);
    void probeEnd(int group, hipStream_t s);
    void probeDrain(size_t keep);

    ~DevicePlan();
    void* alloc(size_t bytes);
    void release(void* p);
    template <typename T>
    T* upload(const std::vector<T>& v);
    void build();
    void ensureRhs(int nrhs);
    // y (=|+=) Z x on device vectors in input order
    template <typename R>
    void apply(const cplx<double>* x, cplx<double>* y, int accumulate, hipStream_t st, bool timed, int nb = 1,
               int64_t xs = 0);
    int batchCap = 1;  // right-hand sides the field arrays and partials hold
    cplx<double>* xInner = nullptr;  // eight right-hand sides interleaved (n x 8), for the 8-RHS near field
    // launch parameters chosen by mlfma_autotune (0: compile-time defaults /
    // the per-level rule); tuneMinBlocks[l]: block target of the M2M (and
    // adjoint L2L) row split into parent level l; tuneM2L[l]: 1 sibling-group,
    // 2 per-target, 3 target-run M2L
    std::vector<int> tuneMinBlocks;
    std::vector<int> tuneM2L;
    int minBlocksAt(int l) const { return l < (int)tuneMinBlocks.size() ? tuneMinBlocks[l] : 0; }
    // captured matvecs (MLFMA_FLAG_GRAPH), least recently used last
    struct GraphEntry {
        const void* x;
        void* y;
        int nrhs, accumulate;
        bool fp32;
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs;
    void clearGraphs();
    // nearX: also the near-field pairs of x (one right-hand side), in the
    // launch of the per-target levels; returns whether they were launched
    template <typename R>
    bool launchM2L(const cplx<R>* in, cplx<R>* out, bool adjoint, hipStream_t st, int nb, int lmin, int lmax,
                   const cplx<double>* nearX = nullptr);
    void ensureBatch(int nb);
    // y (=|+=) Z^H x (the adjoint of the linear operator)
    template <typename R>
    void applyAdjoint(const cplx<double>* x, cplx<double>* y, int accumulate, hipStream_t st);
    void applyOpAdjoint(const cplx<double>* x, cplx<double>* y, int accumulate, hipStream_t st, bool fp32);
    void applyOp(const cplx<double>* x, cplx<double>* y, int accumulate, hipStream_t st, bool fp32, bool timed,
                 int nb = 1, int64_t xs = 0);
    // the operator of the running solve: Z, or Z^H for an adjoint solve
    bool solveAdjoint = false;
    // the running solve applies the block preconditioner on the right
    bool solvePrec = false;
    // a one-right-hand-side solve is running with solveSerial set: every
    // stage on the solve's stream; solveSerial is set by mlfma_autotune(1)
    // when short GMRES runs are faster that way
    bool solveInOrder = false;
    bool solveSerial = false;
    // the shift s of a shifted solve, (Z + s I) x = b (zero: none)
    cplx<double> solveShift{0, 0};
    bool shifted() const { return solveShift.re != 0 || solveShift.im != 0; }
    void applySolveOp(const cplx<double>* x, cplx<double>* y, int accumulate, hipStream_t st, bool fp32, bool timed) {
        if (solveAdjoint)
            applyOpAdjoint(x, y, accumulate, st, fp32);
        else
            applyOp(x, y, accumulate, st, fp32, timed);
        if (shifted())  // (Z + s I)^H = Z^H + conj(s) I
            launchAxpy(x, cplx<double>{solveShift.re, solveAdjoint ? -solveShift.im : solveShift.im}, y, hp.src.n, st);
    }
    void matvec(const cplx<double>* x, cplx<double>* y, int nrhs, int accumulate, hipStream_t st, bool fp32);

    template <typename R>
    cplx<R>* fieldM();
    template <typename R>
    cplx<R>* fieldL();
    template <typename R>
    cplx<R>* scratch();
    template <typename R>
    const cplx<R>* radpat();
};
template <>
inline cplx<double>* DevicePlan::fieldM<double>() { return M64; }
template <>
inline cplx<float>* DevicePlan::fieldM<float>() { return M32; }
template <>
inline cplx<double>* DevicePlan::fieldL<double>() { return L64; }
template <>
inline cplx<float>* DevicePlan::fieldL<float>() { return L32; }
template <>
inline cplx<double>* DevicePlan::scratch<double>() { return inner64; }
template <>
inline cplx<float>* DevicePlan::scratch<float>() { return inner32; }
template <>
inline const cplx<double>* DevicePlan::radpat<double>() { return rp64; }
template <>
inline const cplx<float>* DevicePlan::radpat<float>() { return rp32; }

// seconds of iters GMRES iterations on the plan's excitation (solver.hip)
double timeSolveSteps(DevicePlan& d, int iters);

}  // namespace mlfma

// Opaque handle of the C-ABI.
struct mlfma_plan {
    std::unique_ptr<mlfma::DevicePlan> d;
};
