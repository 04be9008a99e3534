// kernels.hpp — device data layouts and kernel launchers of the MI355X MLFMA
// library.  All arrays live in HBM and are indexed in tree order of the
// unknowns (sources of one leaf contiguous, leaves in depth-first order).
//
// Complex numbers are interleaved (re, im).  A "field" of a box is
// ndir x 2 complex values: for each direction (theta-major, phi-minor, as
// src/fmm/angles.cpp:41-54) the theta and phi components.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mlfma {

template <typename R>
struct alignas(2 * sizeof(R)) cplx {
    R re, im;
};

// ---- per-leaf far-field work item (S2M, L2T, radiation patterns)
struct LeafItem {
    int64_t sbeg;    // first tree position of the leaf's sources
    int32_t cnt;     // number of sources
    int32_t ndir;    // directions of the leaf's level
    int64_t rpOff;   // first complex pair of the leaf's radiation patterns
    int64_t fOff;    // first complex pair of the leaf's field (multipole/local)
    int32_t level;
    int32_t pad;
    double cx, cy, cz;  // leaf centre
};

// ---- near-field work item (one per row block piece; see HostPlan::NearItem)
struct NearItem {
    int64_t rbeg;   // first row (tree position)
    int32_t nr;     // rows, <= 128
    int32_t nc;     // columns
    int32_t nself;  // leading columns of the self block (applied to the rows only)
    int32_t pad;
    int64_t cptr;   // first column entry
    int64_t voff;   // first value (column-major nr x nc)
    int64_t rpo;    // row partial sums [rpo, rpo + nr)
    int64_t cpo;    // column partial sums [cpo, cpo + nc - nself)
};

// ---- one diagonal block of the leaf block-Jacobi preconditioner: the inverse
// of a leaf's self block (rows and columns = its sources, tree positions
// [sbeg, sbeg + cnt)), column-major cnt x cnt at off
struct PrecBlock {
    int64_t sbeg;
    int32_t cnt;
    int32_t pad;
    int64_t off;
};

// ---- source descriptions in tree order
struct DipoleDev {
    double pos[3];
    double phat[3];
};
struct RwgDev {
    double ctr[3];
    double leng;
    int32_t tri[2];
    double xpm[2][3];
};
struct TriDev {
    double X0[3], D0[3], D2[3];
    int32_t nq;
    int32_t pad;
    double q[13][4];
};

struct SrcDev {
    int kind;
    int physical;  // RWG near entries in the physical (singular-integral) form
    const DipoleDev* dip;
    const RwgDev* rwg;
    const TriDev* tri;
};

// ---- one level's angular data
struct LevelDev {
    int32_t nth, nph, ndir, nps;
    const double* khat;   // ndir x 3
    const double* thph;   // ndir x 6 (2x3 projector rows)
};

// ----------------------------------------------------------------- launchers
// setup
void launchRadpat(const SrcDev& s, const LeafItem* items, int nitems, const LevelDev* lv, double k,
                  cplx<double>* rp, int maxndir, hipStream_t st);
void launchNearBuild(const SrcDev& s, const NearItem* items, int nitems, const int64_t* cols, int64_t nnz,
                     double k, cplx<double> Ck, cplx<double>* val, hipStream_t st);
void launchAlpha(int nps, int L, double k, double nodeLeng, const double* dists, cplx<double>* alpha,
                 hipStream_t st);
void launchTransl(int ndir, int nps, int p, const double* psi, const int32_t* nearIdx, const int32_t* distIdx,
                  const cplx<double>* alpha, const double* dirWeight, cplx<double>* T, hipStream_t st);
void launchVoltage(const SrcDev& s, int64_t n, const double* wavevec, const double* pol, double amp,
                   cplx<double>* V, hipStream_t st);
// far field at ndir directions of I (input order): the source sum in
// farfieldChunks(n, ndir) ranges per direction, whose partial sums (3 complex
// each) go to part (ndir x farfieldChunks(n, ndir) x 3 complex), then one
// fixed-order reduction per direction
constexpr int64_t kFarPartCap = 65536;  // partial triples the far-field chunking uses at most
int farfieldChunks(int64_t n, int64_t ndir);
void launchFarfield(const SrcDev& s, int64_t n, const cplx<double>* I, const int64_t* perm, const double* dirs,
                    int64_t ndir, double k, cplx<double> Ck, cplx<double>* part, cplx<double>* E, hipStream_t st);
// the per-device address of the fixed far-field partial-sum array (kFarPartCap
// x 3 complex): initFarfieldScratch(device) once with that device current
// (plan creation), then farfieldScratch(device) (nullptr before)
hipError_t initFarfieldScratch(int device);
cplx<double>* farfieldScratch(int device);
void launchFarfieldAdj(const SrcDev& s, int64_t n, const cplx<double>* E, const double* dirs, int ndir, double k,
                       cplx<double> Ckc, const int64_t* perm, cplx<double>* I, hipStream_t st);
// dipole position gradient, near part: grad (n x 3, input order) += -2 Re(lambda^H dZ_near I)
void launchNearGradDip(const SrcDev& s, const NearItem* items, int nitems, const int32_t* colsIn, const int64_t* perm,
                       const int64_t* iperm, const cplx<double>* I, const cplx<double>* lam, double k, cplx<double> Ck,
                       double* grad, hipStream_t st);

// y += s x (n complex)
void launchAxpy(const cplx<double>* x, cplx<double> s, cplx<double>* y, int64_t n, hipStream_t st);

// matvec stages, templated on the real type of the far field
void launchGather(const cplx<double>* x, const int64_t* perm, cplx<double>* xt, int64_t n, int nrhs,
                  hipStream_t st);
void launchScatter(const cplx<double>* yt, const int64_t* perm, cplx<double>* y, int64_t n, int nrhs,
                   int accumulate, hipStream_t st);

template <typename R>
void launchM2MTheta(int nitems, const int32_t* stemSlot, const int32_t* ptr, const int32_t* child,
                    const int32_t* oct, int nthP, int mthC, int mphC, int p, const int32_t* thNear,
                    const double* thW, const cplx<R>* phaseUp, const cplx<R>* Mc, cplx<R>* inner,
                    hipStream_t st);
template <typename R>
void launchM2MPhi(int nitems, const int32_t* stemSlot, int nthP, int nphP, int mphC, int p,
                  const int32_t* phNear, const double* phW, const cplx<R>* inner, cplx<R>* Mp,
                  hipStream_t st);
template <typename R>
void launchZeroSlots(int n, const int32_t* slots, int ndir, cplx<R>* Lc, hipStream_t st);
template <typename R>
void launchL2LPhi(int nitems, const int32_t* parent, const int32_t* oct, int mthP, int mphP, int nphC,
                  const int32_t* atPtr, const int32_t* atSrc, const double* atW, const cplx<R>* phaseDn,
                  const cplx<R>* Lp, cplx<R>* inner, hipStream_t st);
template <typename R>
void launchL2LTheta(int nitems, const int32_t* child, int mthP, int nthC, int nphC, const int32_t* atPtr,
                    const int32_t* atSrc, const double* atW, const int32_t* atFlip, const cplx<R>* inner,
                    cplx<R>* Lc, hipStream_t st, const cplx<R>* postPhase = nullptr,
                    const int32_t* oct = nullptr);  // postPhase: conj(postPhase[oct]) applied before the sum
void launchDirect(const SrcDev& s, const cplx<double>* xt, int64_t n, double k, cplx<double> Ck, cplx<double>* yt,
                  hipStream_t st);

}  // namespace mlfma
