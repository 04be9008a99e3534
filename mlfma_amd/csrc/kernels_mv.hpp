// kernels_mv.hpp — launchers of the per-matvec kernels (kernels_mv.hip).
#pragma once

#include "kernels.hpp"

namespace mlfma {

// Batched launchers take nb right-hand sides (1, 2, 4 or 8): x and y with
// column stride xs (input order), field arrays at stride fs, near-field
// partial sums at stride ps.
template <typename R>
void launchS2Mx(const LeafItem* items, const int32_t* order, int nitems, int maxndir, const cplx<R>* rp,
                const cplx<double>* x, const int64_t* perm, cplx<R>* M, hipStream_t st, int nb = 1, int64_t xs = 0,
                int64_t fs = 0);

// returns false when the level's fields do not fit the LDS limit (caller
// then uses the two-pass kernels); minBlocks: the block count to split parent
// rows for (0: the compile-time default)
template <typename R>
bool launchM2MFused(int nstem, const int32_t* stemSlot, const int32_t* ptr, const int32_t* child, const int32_t* oct,
                    int nthP, int nphP, int mthC, int mphC, int p, const int32_t* thNear, const double* thW,
                    const int32_t* phNear, const double* phW, const cplx<R>* phaseUp, const cplx<R>* Mc, cplx<R>* Mp,
                    size_t ldsMax, hipStream_t st, int nb = 1, int64_t fs = 0, int minBlocks = 0);

// parent theta rows per fused M2M block: split until nstem x row runs reaches
// minBlocks (0: the compile-time default), then until the LDS (*lds bytes)
// fits ldsMax; 0 when one row does not fit
template <typename R>
int m2mRows(int nstem, int nthP, int nphP, int mthC, int mphC, int p, size_t ldsMax, int minBlocks, size_t* lds);

// The upward pass (S2M then M2M of every level) of one right-hand side as one
// persistent launch over a packed worklist (kTreeUp); passed by value.
constexpr int kTreeMaxLevels = 12;
constexpr int kTreeQueues = 8;  // ticket counters; queue q hands out units q, q + 8, q + 16, ...
template <typename R>
struct TreeUpLevel {  // M2M into one parent level
    int nstem = 0, rows = 0, nchunk = 0, unit0 = 0;  // units [unit0, unit0 + nstem nchunk): (stem, row run)
    int nthP = 0, nphP = 0, mthC = 0, mphC = 0;
    int ctr0 = 0;                                    // completion counter of stem 0
    const int32_t *stemSlot, *ptr, *child, *oct;
    const int32_t* childCtr;                         // per child entry: its producer's counter (-1: none)
    const int32_t *thNear, *phNear;
    const double *thW, *phW;
    const cplx<R>* phaseUp;                          // of the child level
    int64_t MoffC = 0, MoffP = 0;                    // child / parent level fields
};
template <typename R>
struct TreeUp {
    int nlev = 0;            // M2M levels, finest parent level first
    int nLeaf = 0;           // leaf items; counters [0, nLeaf) are theirs
    int s2mChunks = 1;       // S2M pieces per leaf item (direction chunks of the block size)
    int s2mPer = 1;          // S2M pieces per unit (consecutive in the S2M order)
    int nS2MPieces = 0;      // nLeaf x s2mChunks
    int nS2M = 0, total = 0; // S2M units (the first nS2M) and all units
    int p = 0;
    int nctr = 0;
    int32_t* ctr = nullptr;  // counters, then kTreeQueues tickets, exit count, stall count
    const LeafItem* items = nullptr;
    const int32_t* order = nullptr;
    const cplx<R>* rp = nullptr;
    const cplx<double>* x = nullptr;
    const int64_t* perm = nullptr;
    cplx<R>* M = nullptr;
    TreeUpLevel<R> lv[kTreeMaxLevels];
};
// bs: 192 or 256 threads; lds: the largest M2M level's; grid: blocks
template <typename R>
void launchTreeUp(const TreeUp<R>& tu, int bs, size_t lds, int grid, hipStream_t st);
// resident blocks per CU of the tree walk at that block size and LDS
template <typename R>
int treeUpOccupancy(int bs, size_t lds);

// M2L work of every level (sibling groups), passed by value to one launch
constexpr int kMaxLevels = 16;
template <typename R>
struct M2LGroups {
    int nlev = 0;
    int ndir[kMaxLevels], ngrp[kMaxLevels], nch[kMaxLevels], per[kMaxLevels], blk[kMaxLevels];
    const int32_t* tgt[kMaxLevels];  // 8 target slots per group (-1: none)
    const int32_t* ptr[kMaxLevels];  // per group: its source entries
    const int32_t* src[kMaxLevels];  // source slot per entry
    const int32_t* off[kMaxLevels];  // 8 offsets per entry (-1: not in that target's list)
    const cplx<R>* T[kMaxLevels];
    const cplx<R>* M[kMaxLevels];
    cplx<R>* L[kMaxLevels];
    int64_t fs = 0;  // field stride between right-hand sides
};

// conjT: conjugated operators (the adjoint pass; offsets then name the
// reversed pairs); nb right-hand sides over grid z
template <typename R>
void launchM2LGroup(M2LGroups<R> lv, bool conjT, hipStream_t st, int nb = 1);

// M2L per target box (one block per target and direction chunk), for levels
// whose sibling lists overlap little; nb right-hand sides share the operator
template <typename R>
struct M2LLevels {
    int nlev = 0;
    int ndir[kMaxLevels], ntgt[kMaxLevels], nch[kMaxLevels], per[kMaxLevels], blk[kMaxLevels];
    const int32_t* tgt[kMaxLevels];
    const int32_t* ptr[kMaxLevels];
    const int32_t* src[kMaxLevels];
    const int32_t* off[kMaxLevels];
    const cplx<R>* T[kMaxLevels];
    const cplx<R>* M[kMaxLevels];
    cplx<R>* L[kMaxLevels];
    int64_t fs = 0;  // field stride between right-hand sides
};

template <typename R>
void launchM2Lx(M2LLevels<R> lv, bool conjT, hipStream_t st, int nb = 1);

// M2L over target runs (kM2LT): per level, runs of consecutive targets and,
// per run, the distinct source boxes of their lists (slots, in one CSR);
// each pair is packed as (index of its source in the run's list) << 9 |
// offset.  A block stages the run's sources and the 316 operators for a few
// directions in LDS and forms every target of the run from there.
constexpr int kM2LOffsets = 316;
template <typename R>
struct M2LTiles {
    int nlev = 0;
    int ndir[kMaxLevels], nrun[kMaxLevels], nch[kMaxLevels], per[kMaxLevels], blk[kMaxLevels];
    const int32_t* tgt[kMaxLevels];   // target slots
    const int32_t* ptr[kMaxLevels];   // per target: its pairs
    const int32_t* pk[kMaxLevels];    // packed pairs
    const int32_t* run[kMaxLevels];   // per run: first target (nrun + 1)
    const int32_t* rsrc[kMaxLevels];  // per run: first entry of its source list (nrun + 1)
    const int32_t* slot[kMaxLevels];  // source slots of every run's list
    const cplx<R>* T[kMaxLevels];
    const cplx<R>* M[kMaxLevels];
    cplx<R>* L[kMaxLevels];
    int64_t fs = 0;
    int maxSrc = 0;                   // longest run source list (LDS size)
};
// directions per block: 2 in FP64, 4 in FP32 (64 bytes of each source)
template <typename R>
constexpr int m2ltDirs() { return sizeof(R) == 8 ? 2 : 4; }
template <typename R>
size_t m2ltLds(int maxSrc) {
    return sizeof(cplx<R>) * (size_t)m2ltDirs<R>() * (2 * (size_t)maxSrc + kM2LOffsets);
}
template <typename R>
void launchM2LT(M2LTiles<R> lv, bool conjT, hipStream_t st, int nb = 1);

template <typename R>
bool launchL2LFused(bool adjoint, int nitems, const int32_t* parent, const int32_t* child, const int32_t* oct, int mthP, int mphP,
                    int nthC, int nphC, const int32_t* atPhPtr, const int32_t* atPhSrc, const double* atPhW, int nAtPh,
                    const int32_t* atThPtr, const int32_t* atThSrc, const double* atThW, const int32_t* atThFlip,
                    int nAtTh, const cplx<R>* phaseDn, const cplx<R>* Lp, cplx<R>* Lc, size_t ldsMax,
                    hipStream_t st, int nb = 1, int64_t fs = 0);

// near field: kNearSym writes partial sums; kL2Tx folds them into the rows of
// far-field leaves and kNearSum into the listed remaining rows
template <typename V>
void launchNearSym(bool conjVal, const NearItem* items, int nitems, const int32_t* colsIn, const int64_t* perm,
                   const cplx<V>* val, const cplx<double>* x, cplx<double>* part, hipStream_t st, int nb = 1,
                   int64_t xs = 0, int64_t ps = 0, int grid = 0);
// eight right-hand sides on the matrix cores; xt: x interleaved by
// launchInterleave8 (unknown-major, eight values per unknown); the partial
// sums come out interleaved the same way (part[j 8 + k]), as the eight-RHS
// folds (launchL2Tx, launchNearSum with nb = 8) read them
template <typename V>
void launchNearSymM8(bool conjVal, const NearItem* items, int nitems, const int32_t* colsIn, const int64_t* perm,
                     const cplx<V>* val, const cplx<double>* xt, cplx<double>* part, hipStream_t st);
void launchInterleave8(const cplx<double>* x, int64_t xs, int64_t n, cplx<double>* xt, hipStream_t st);
// the per-target M2L of one right-hand side and the near-field pairs in one
// launch; conj: conjugated operators (lv holds the reversed pairs' offsets)
// and conjugated values (the adjoint); windowPct: the pair blocks spread over
// that percentage of the launch's blocks, from the first
template <typename R, typename V>
void launchM2LNear(M2LLevels<R> lv, const NearItem* items, int nitems, const int32_t* colsIn, const int64_t* perm,
                   const cplx<V>* val, const cplx<double>* x, cplx<double>* part, hipStream_t st, bool conj,
                   int windowPct);
// out = (float) in, n complex values
void launchToFloat(const cplx<double>* in, cplx<float>* out, int64_t n, hipStream_t st);
// y = P x (adjoint: P^H x) for nb vectors (column stride xs, input order),
// P = block diagonal over the preconditioner blocks; gathers the self-block
// entries of the near-field values into packed column-major blocks
void launchBlockPrec(const PrecBlock* blocks, int nblocks, int maxcnt, const cplx<double>* P, const int64_t* perm,
                     const cplx<double>* x, cplx<double>* y, int nb, int64_t xs, bool adjoint, hipStream_t st);
void launchGatherValues(const cplx<double>* val, const int64_t* idx, int64_t n, cplx<double>* out, hipStream_t st);
// eight right-hand sides: the near-field sums of the far-field rows into y
// (interleaved partials), then y += C k conj(P) L per leaf on the matrix
// cores; false (nothing launched) when a leaf's locals of eight right-hand
// sides exceed the LDS limit
template <typename R>
bool launchL2TM8(const LeafItem* items, int nitems, int maxndir, const cplx<R>* rp, const cplx<R>* Lc, int64_t fs,
                 cplx<double> Ck, const int64_t* rowPos, int64_t nrows, const int64_t* sumPtr, const int32_t* sumIdx,
                 const cplx<double>* part, const int64_t* perm, cplx<double>* y, int64_t xs, int accumulate,
                 size_t ldsMax, hipStream_t st);
void launchNearSum(const int64_t* rows, int64_t nrows, const int64_t* sumPtr, const int32_t* sumIdx,
                   const cplx<double>* part, const int64_t* perm, cplx<double>* y, int accumulate, hipStream_t st,
                   int nb = 1, int64_t xs = 0, int64_t ps = 0);

// y (=|+=) near + L2T for the rows of far-field leaves
template <typename R>
void launchL2Tx(const int64_t* rowPos, const int32_t* rowItem, int64_t nrows, const LeafItem* items, const cplx<R>* rp,
                const cplx<R>* Lc, cplx<double> Ck, const int64_t* sumPtr, const int32_t* sumIdx,
                const cplx<double>* part, const int64_t* perm, cplx<double>* y, int accumulate, hipStream_t st,
                int nb = 1, int64_t fs = 0, int64_t xs = 0, int64_t ps = 0);

// adjoint of L2L: parents gain conj(phase_down) * Interp(child), per parent CSR
template <typename R>
bool launchL2LAdj(int npar, const int32_t* parSlot, const int32_t* ptr, const int32_t* child, const int32_t* oct,
                  int nthP, int nphP, int mthC, int mphC, int p, const int32_t* thNear, const double* thW,
                  const int32_t* phNear, const double* phW, const cplx<R>* phaseDn, const cplx<R>* Lc, cplx<R>* Lp,
                  size_t ldsMax, hipStream_t st, int minBlocks = 0);

template <typename R>
void launchZeroField(cplx<R>* f, int64_t n, hipStream_t st);

// out (n x 3 complex, input order) = sum_d khat_d conj(rp_s[d]) . F[d] per far-field source
void launchL2TGrad(const int64_t* rowPos, const int32_t* rowItem, int64_t nrows, const LeafItem* items,
                   const LevelDev* lv, const cplx<double>* rp, const cplx<double>* F, const int64_t* perm,
                   cplx<double>* out, hipStream_t st);

}  // namespace mlfma
