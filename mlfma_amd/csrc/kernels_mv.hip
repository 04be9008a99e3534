// kernels_mv.hip — the per-matvec kernels (S2M, fused M2M, M2L, fused L2L,
// near field, L2T) for gfx950.  Vectors x and y are in input order; the tree
// permutation is applied inside the kernels that touch them (S2M reads, near
// field and L2T writes), so a matvec needs no separate gather/scatter pass.
#include <algorithm>

#include "cplx_ops.hpp"
#include "kernels_mv.hpp"
#include "host.hpp"  // Error

namespace mlfma {

constexpr int TB = 256;

// launch-shape parameters (compile-time; defaults measured on MI355X)
#ifndef MLFMA_M2M_MIN_BLOCKS
#define MLFMA_M2M_MIN_BLOCKS 1024    // split parent theta rows until this many blocks
#endif
#ifndef MLFMA_M2L_PAIRS
#define MLFMA_M2L_PAIRS 4            // pairs in flight per wave in the one-RHS M2L
#endif
#ifndef MLFMA_NEAR_LOADS
#define MLFMA_NEAR_LOADS 6           // near-field values in flight per lane (5, 6 or 8)
#endif

// ------------------------------------------------------------------ S2M
// M_b[d] = sum_s x_s rp_b[s][d]  (src/fmm/leaf.cpp:212-230).  Block = one leaf
// x blockDim directions; lanes stream consecutive 32-byte (theta, phi) pairs
// of one source row.  Blocks take the leaves in the order `order` (most
// sources first, so that the longest leaves do not form the launch's tail).
// One right-hand side: each wave gathers the x values of 64 rows at a time
// (lane i: row i) and reads them back with v_readlane, so the row loop has no
// dependent loads; four rows' pattern loads are in flight per step (even rows
// add into a, odd rows into b).  Lanes past the leaf's direction count read a
// clamped direction and store nothing.  NB right-hand sides share every
// pattern read: x has column stride xs (input order), the fields of
// right-hand side r start at M + r fs.
//
// One right-hand side, directions [d0, d0 + blockDim.x) of leaf it (d0 below
// the leaf's direction count; uniform per block); COH: device-coherent
// stores of the multipole.
template <typename R, bool COH = false>
__device__ __forceinline__ void s2mOne(const LeafItem& it, int d0, const cplx<R>* __restrict__ rp,
                                       const cplx<double>* __restrict__ x, const int64_t* __restrict__ perm,
                                       cplx<R>* __restrict__ M) {
    const int d = d0 + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int dc = d < it.ndir ? d : it.ndir - 1;
    const int64_t stride = (int64_t)it.ndir * 2;
    cplx<R> a0 = mk<R>(0, 0), a1 = mk<R>(0, 0), b0 = mk<R>(0, 0), b1 = mk<R>(0, 0);
    for (int q0 = 0; q0 < it.cnt; q0 += 64) {
        const int nq = min(64, it.cnt - q0);
        const cplx<double> xl = lane < nq ? x[perm[it.sbeg + q0 + lane]] : mk(0.0, 0.0);
        const cplx<R>* p = rp + it.rpOff + 2 * dc + q0 * stride;
        auto xq = [&](int q) { return cast<R>(mk(readlaneD(xl.re, q), readlaneD(xl.im, q))); };
        int q = 0;
        for (; q + 3 < nq; q += 4) {
            const cplx<R> r00 = p[q * stride], r01 = p[q * stride + 1];
            const cplx<R> r10 = p[(q + 1) * stride], r11 = p[(q + 1) * stride + 1];
            const cplx<R> r20 = p[(q + 2) * stride], r21 = p[(q + 2) * stride + 1];
            const cplx<R> r30 = p[(q + 3) * stride], r31 = p[(q + 3) * stride + 1];
            const cplx<R> x0 = xq(q), x1 = xq(q + 1), x2 = xq(q + 2), x3 = xq(q + 3);
            a0 = a0 + x0 * r00;
            a1 = a1 + x0 * r01;
            b0 = b0 + x1 * r10;
            b1 = b1 + x1 * r11;
            a0 = a0 + x2 * r20;
            a1 = a1 + x2 * r21;
            b0 = b0 + x3 * r30;
            b1 = b1 + x3 * r31;
        }
        for (; q + 1 < nq; q += 2) {
            const cplx<R> r00 = p[q * stride], r01 = p[q * stride + 1];
            const cplx<R> r10 = p[(q + 1) * stride], r11 = p[(q + 1) * stride + 1];
            const cplx<R> x0 = xq(q), x1 = xq(q + 1);
            a0 = a0 + x0 * r00;
            a1 = a1 + x0 * r01;
            b0 = b0 + x1 * r10;
            b1 = b1 + x1 * r11;
        }
        if (q < nq) {  // only in the last run of rows (the others hold 64)
            const cplx<R> x0 = xq(q);
            a0 = a0 + x0 * p[q * stride];
            a1 = a1 + x0 * p[q * stride + 1];
        }
    }
    if (d < it.ndir) {
        stc<R, COH>(M + it.fOff + 2 * d, a0 + b0);
        stc<R, COH>(M + it.fOff + 2 * d + 1, a1 + b1);
    }
}

constexpr int S2M_XROWS = 32;  // rows per staged run of x (NB > 1)
template <typename R, int NB>
__global__ void kS2Mx(const LeafItem* __restrict__ items, const int32_t* __restrict__ order,
                      const cplx<R>* __restrict__ rp, const cplx<double>* __restrict__ x, int64_t xs,
                      const int64_t* __restrict__ perm, cplx<R>* __restrict__ M, int64_t fs) {
    const LeafItem it = items[order[blockIdx.x]];
    const int d = blockIdx.y * blockDim.x + threadIdx.x;
    if constexpr (NB == 1) {
        if (blockIdx.y * blockDim.x >= it.ndir) return;  // uniform per block
        s2mOne<R>(it, blockIdx.y * blockDim.x, rp, x, perm, M);
    } else {
        // the x values of a run of S2M_XROWS rows for the NB right-hand sides
        // are staged in LDS (one coalesced gather per run) and read back as
        // broadcasts; every thread takes part in the staging (lanes past the
        // direction count read a clamped direction and store nothing)
        __shared__ cplx<double> sx[S2M_XROWS * NB];
        if (blockIdx.y * blockDim.x >= it.ndir) return;  // uniform per block
        const int dc = d < it.ndir ? d : it.ndir - 1;
        const cplx<R>* p = rp + it.rpOff + 2 * dc;
        const int64_t stride = (int64_t)it.ndir * 2;
        cplx<R> a[NB][2];
#pragma unroll
        for (int r = 0; r < NB; ++r) a[r][0] = a[r][1] = mk<R>(0, 0);
        for (int q0 = 0; q0 < it.cnt; q0 += S2M_XROWS) {
            const int nq = min(S2M_XROWS, it.cnt - q0);
            if (q0) __syncthreads();  // the previous run's reads are done
            for (int e = threadIdx.x; e < nq * NB; e += blockDim.x) {
                const int q = e / NB, r = e - q * NB;
                sx[e] = x[r * xs + perm[it.sbeg + q0 + q]];
            }
            __syncthreads();
            const cplx<R>* pq = p + q0 * stride;
            int q = 0;
            for (; q + 1 < nq; q += 2) {  // two source rows' pattern loads in flight
                const cplx<R> r00 = pq[q * stride], r01 = pq[q * stride + 1];
                const cplx<R> r10 = pq[(q + 1) * stride], r11 = pq[(q + 1) * stride + 1];
#pragma unroll
                for (int r = 0; r < NB; ++r) {
                    const cplx<R> x0 = cast<R>(sx[q * NB + r]), x1 = cast<R>(sx[(q + 1) * NB + r]);
                    a[r][0] = a[r][0] + x0 * r00 + x1 * r10;
                    a[r][1] = a[r][1] + x0 * r01 + x1 * r11;
                }
            }
            if (q < nq) {
                const cplx<R> r0 = pq[q * stride], r1 = pq[q * stride + 1];
#pragma unroll
                for (int r = 0; r < NB; ++r) {
                    const cplx<R> xr = cast<R>(sx[q * NB + r]);
                    a[r][0] = a[r][0] + xr * r0;
                    a[r][1] = a[r][1] + xr * r1;
                }
            }
        }
        if (d < it.ndir)
#pragma unroll
            for (int r = 0; r < NB; ++r) {
                M[r * fs + it.fOff + 2 * d] = a[r][0];
                M[r * fs + it.fOff + 2 * d + 1] = a[r][1];
            }
    }
}

template <typename R>
void launchS2Mx(const LeafItem* items, const int32_t* order, int nitems, int maxndir, const cplx<R>* rp,
                const cplx<double>* x, const int64_t* perm, cplx<R>* M, hipStream_t st, int nb, int64_t xs,
                int64_t fs) {
    if (nitems <= 0) return;
    const int bs = std::min(TB, ((maxndir + 63) / 64) * 64);
    dim3 g(nitems, (maxndir + bs - 1) / bs);
#define S2M_CASE(B) hipLaunchKernelGGL((kS2Mx<R, B>), g, dim3(bs), 0, st, items, order, rp, x, xs, perm, M, fs)
    if (nb == 1) S2M_CASE(1);
    else if (nb == 2) S2M_CASE(2);
    else if (nb == 4) S2M_CASE(4);
    else if (nb == 8) S2M_CASE(8);
    else throw Error(MLFMA_ERR_ARG, "right-hand-side batch must be 1, 2, 4 or 8");
#undef S2M_CASE
}

// ------------------------------------------------------------------ fused M2M
// One block per (parent box, run of parent theta rows) (src/fmm/stem.cpp:
// 72-103,133-184).  The theta and phi interpolations do not depend on the
// child, so the children are phase-shifted and summed on the child grid first
// (one LDS array, one global read of every child direction) and the sum is
// interpolated once: the same linear map as interpolating every child and
// adding, with a different rounding order.  The interpolation tables of the
// rows are staged in LDS; the theta interpolation (pole rule: reflected node,
// phi + pi, both components negated) goes to LDS, and the periodic phi
// interpolation writes the parent.
//
// m2mFused: the work of one block of BS threads, parent stem `item`, parent
// theta rows [j0, j0 + rows) (clipped to the grid); smem: the LDS bytes of
// m2mRows; COH: device-coherent loads of the children and stores of the
// parent.
template <typename R, int BS, bool COH = false>
__device__ __forceinline__ void m2mFused(int item, int j0, int rows, const int32_t* __restrict__ stemSlot,
                                         const int32_t* __restrict__ ptr, const int32_t* __restrict__ child,
                                         const int32_t* __restrict__ oct, int nthP, int nphP, int mthC, int mphC,
                                         int p, const int32_t* __restrict__ thNear, const double* __restrict__ thW,
                                         const int32_t* __restrict__ phNear, const double* __restrict__ phW,
                                         const cplx<R>* __restrict__ phaseUp, const cplx<R>* __restrict__ Mc,
                                         cplx<R>* __restrict__ Mp, unsigned char* smem) {
    const int nr = min(rows, nthP - j0);
    const int ndirC = mthC * mphC;
    const int nInner = nr * mphC;
    const int tp = 2 * p;
    cplx<R>* sS = reinterpret_cast<cplx<R>*>(smem);  // ndirC x 2: shifted children, summed
    cplx<R>* sI = sS + (size_t)2 * ndirC;            // rows x mphC x 2
    R* tThW = reinterpret_cast<R*>(sI + (size_t)2 * rows * mphC);
    R* tPhW = tThW + rows * tp;
    int* tThNear = reinterpret_cast<int*>(tPhW + nphP * tp);
    int* tPhNear = tThNear + rows;
    // the table loads are issued first and stored to LDS after the children's
    // sum, so that their latency overlaps the children's loads (a loop with
    // one store per iteration waits for each load in turn)
    constexpr int U = 2;
    const int t = threadIdx.x;
    double wTh[U] = {}, wPh[U] = {};
    const int nTh = nr * tp, nPh = nphP * tp;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int i = t + u * BS;
        if (i < nTh) wTh[u] = thW[j0 * tp + i];
        if (i < nPh) wPh[u] = phW[i];
    }
    const int nearTh = t < nr ? thNear[j0 + t] : 0;
    const int nearPh = t < nphP ? phNear[t] : 0;
    const int c0 = ptr[item], nc = ptr[item + 1] - c0;  // at most 8 children
    for (int d = t; d < ndirC; d += BS) {
        // the loads of CG children are issued before their first product:
        // all eight in the staged launch; two in the tree walk, whose
        // registers also hold the S2M stream's (same sums, same order)
        constexpr int CG = COH ? 2 : 8;
        cplx<R> a0 = mk<R>(0, 0), a1 = mk<R>(0, 0);
#pragma unroll
        for (int g = 0; g < 8; g += CG) {
            cplx<R> s[CG], m0[CG], m1[CG];
#pragma unroll
            for (int k = 0; k < CG; ++k)
                if (g + k < nc) {
                    const cplx<R>* Mb = Mc + (int64_t)child[c0 + g + k] * ndirC * 2;
                    s[k] = phaseUp[(int64_t)oct[c0 + g + k] * ndirC + d];
                    m0[k] = ldc<R, COH>(Mb + 2 * d);
                    m1[k] = ldc<R, COH>(Mb + 2 * d + 1);
                }
#pragma unroll
            for (int k = 0; k < CG; ++k)
                if (g + k < nc) {
                    a0 = a0 + s[k] * m0[k];
                    a1 = a1 + s[k] * m1[k];
                }
        }
        sS[2 * d] = a0;
        sS[2 * d + 1] = a1;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int i = t + u * BS;
        if (i < nTh) tThW[i] = (R)wTh[u];
        if (i < nPh) tPhW[i] = (R)wPh[u];
    }
    if (t < nr) tThNear[t] = nearTh;
    if (t < nphP) tPhNear[t] = nearPh;
    for (int i = t + U * BS; i < nTh; i += BS) tThW[i] = (R)thW[j0 * tp + i];
    for (int i = t + U * BS; i < nPh; i += BS) tPhW[i] = (R)phW[i];
    for (int i = t + BS; i < nr; i += BS) tThNear[i] = thNear[j0 + i];
    for (int i = t + BS; i < nphP; i += BS) tPhNear[i] = phNear[i];
    __syncthreads();
    for (int e = threadIdx.x; e < nInner; e += BS) {
        const int jr = e / mphC, iph = e - jr * mphC;
        const int near = tThNear[jr];
        const int iphf = iph + (iph < mphC / 2 ? mphC / 2 : -mphC / 2);
        cplx<R> a0 = mk<R>(0, 0), a1 = mk<R>(0, 0);
#pragma unroll 4
        for (int q = 0; q < tp; ++q) {
            const int i = near + 1 - p + q;
            const int f = i < 0 ? -i - 1 : (i >= mthC ? 2 * mthC - i - 1 : i);
            const bool flip = (i != f);
            const int dd = 2 * (f * mphC + (flip ? iphf : iph));
            const R w = flip ? -tThW[jr * tp + q] : tThW[jr * tp + q];
            a0 = a0 + w * sS[dd];
            a1 = a1 + w * sS[dd + 1];
        }
        sI[2 * e] = a0;
        sI[2 * e + 1] = a1;
    }
    __syncthreads();
    cplx<R>* out = Mp + ((int64_t)stemSlot[item] * nthP + j0) * nphP * 2;
    for (int e = threadIdx.x; e < nr * nphP; e += BS) {
        const int jr = e / nphP, jph = e - jr * nphP;
        const int near = tPhNear[jph];
        const cplx<R>* in = sI + 2 * jr * mphC;
        cplx<R> a0 = mk<R>(0, 0), a1 = mk<R>(0, 0);
#pragma unroll 4
        for (int q = 0; q < tp; ++q) {
            int i = near + 1 - p + q;
            i = i < 0 ? i + mphC : (i >= mphC ? i - mphC : i);
            const R w = tPhW[jph * tp + q];
            a0 = a0 + w * in[2 * i];
            a1 = a1 + w * in[2 * i + 1];
        }
        stc<R, COH>(out + 2 * e, a0);
        stc<R, COH>(out + 2 * e + 1, a1);
    }
}

template <typename R>
__global__ __launch_bounds__(TB) void kM2MFused(const int32_t* __restrict__ stemSlot, const int32_t* __restrict__ ptr,
                                                const int32_t* __restrict__ child, const int32_t* __restrict__ oct,
                                                int nstem, int per, int rows, int nthP, int nphP, int mthC, int mphC,
                                                int p, const int32_t* __restrict__ thNear,
                                                const double* __restrict__ thW, const int32_t* __restrict__ phNear,
                                                const double* __restrict__ phW, const cplx<R>* __restrict__ phaseUp,
                                                const cplx<R>* __restrict__ Mc, cplx<R>* __restrict__ Mp, int64_t fs) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int item = xcdItem(blockIdx.x, per);
    if (item >= nstem) return;
    // right-hand side blockIdx.z; parent theta rows from blockIdx.y * rows
    m2mFused<R, TB>(item, blockIdx.y * rows, rows, stemSlot, ptr, child, oct, nthP, nphP, mthC, mphC, p, thNear, thW,
                    phNear, phW, phaseUp, Mc + blockIdx.z * fs, Mp + blockIdx.z * fs, smem);
}

template <typename R>
bool launchM2MFused(int nstem, const int32_t* stemSlot, const int32_t* ptr, const int32_t* child, const int32_t* oct,
                    int nthP, int nphP, int mthC, int mphC, int p, const int32_t* thNear, const double* thW,
                    const int32_t* phNear, const double* phW, const cplx<R>* phaseUp, const cplx<R>* Mc, cplx<R>* Mp,
                    size_t ldsMax, hipStream_t st, int nb, int64_t fs, int minBlocks) {
    if (nstem <= 0) return true;
    if (minBlocks <= 0) minBlocks = MLFMA_M2M_MIN_BLOCKS;
    // split the parent's theta rows so that there are enough blocks for the
    // device (each block sums the children on the child grid again)
    size_t lds = 0;
    const int rows = m2mRows<R>(nstem, nthP, nphP, mthC, mphC, p, ldsMax, minBlocks, &lds);
    if (rows == 0) return false;
    const int per = (nstem + 7) / 8;
    dim3 g(8 * per, (nthP + rows - 1) / rows, nb);
    hipLaunchKernelGGL((kM2MFused<R>), g, dim3(TB), lds, st, stemSlot, ptr, child, oct, nstem, per, rows, nthP, nphP,
                       mthC, mphC, p, thNear, thW, phNear, phW, phaseUp, Mc, Mp, fs);
    return true;
}

template <typename R>
int m2mRows(int nstem, int nthP, int nphP, int mthC, int mphC, int p, size_t ldsMax, int minBlocks, size_t* lds) {
    if (minBlocks <= 0) minBlocks = MLFMA_M2M_MIN_BLOCKS;
    int rows = nthP;
    while (rows > 2 && (int64_t)nstem * ((nthP + rows - 1) / rows) < minBlocks) rows = (rows + 1) / 2;
    const int tp = 2 * p;
    auto ldsFor = [&](int r) {
        return sizeof(cplx<R>) * 2 * ((size_t)mthC * mphC + (size_t)r * mphC) +
               sizeof(R) * ((size_t)r + nphP) * tp + sizeof(int) * ((size_t)r + nphP);
    };
    while (rows > 1 && ldsFor(rows) > ldsMax) rows = (rows + 1) / 2;
    *lds = ldsFor(rows);
    return *lds > ldsMax ? 0 : rows;
}

// ------------------------------------------------------------------ tree walks
// Persistent kernels over a packed worklist of units of several stages and
// levels.  Units are handed out by kTreeQueues ticket counters: queue q gives
// units q, q + 8, q + 16, ... in turn, and block b draws from queue b mod 8
// (the blocks of one XCD, under the round-robin dispatch), so no single
// counter serialises the walk (one address takes about one atomic per 15-20
// ns).  Every unit waits only for units of lower index.  The lowest
// unfinished unit therefore never waits: it is running, or it is the next
// unit of its queue, whose blocks are all free once every lower unit is done
// (blocks 0-7, one per queue, are dispatched first), so the walk progresses
// whatever the grid size and residency.  A unit waits by polling the
// completion counters of the units it reads from, and counts its own
// completion once its stores are complete.  Values handed between units are
// written and read device-coherently (ldc / stc: past the XCD's L2), and the
// counters are relaxed agent-scope atomics; the fences are workgroup-scope
// (order only), since an agent-scope fence writes back and invalidates the
// whole L2 of the XCD on gfx950.  Waits are bounded: after kTreeSpinLimit
// polls (about a tenth of a second), or once any wait of the plan has
// stalled, a wait counts a stall and goes on, so a wrong count gives wrong
// values, never a hang.  The last block out clears every counter but the
// stall count, so the next launch (also inside a captured graph) starts from
// zero.  Counters: ctr[0, nctr - 10) per producing unit group, then the
// tickets, the exit count and the stall count.
constexpr int kTreeSpinLimit = 1 << 16;

__device__ __forceinline__ void treeWait(const int32_t* c, int need, int32_t* stalls) {
    if (__hip_atomic_load(stalls, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
    for (int i = 0; __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need; ++i) {
        if (i >= kTreeSpinLimit) {
            __hip_atomic_fetch_add(stalls, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
        __builtin_amdgcn_s_sleep(2);
    }
}

// s_waitcnt vmcnt(0): every memory access of the wave has completed
__device__ __forceinline__ void waitVmem() { __builtin_amdgcn_s_waitcnt(0 | (0x7 << 4) | (0xf << 8)); }

// One step of a block's walk: count the block's previous unit, once every
// wave's stores of it have completed (thread i < nsig adds one to counter
// sig(i)), and take the next unit from the block's queue.  The atomics are
// one region between two barriers, and the unit is read into a scalar
// register (readfirstlane), so the loop around the steps stays uniform.
// (With one thread-0 region at the end of the loop body and another at its
// start, and the unit read as a vector value, the structurizer turned the
// loop into two nested ones with thread 0 alone in the outer one: the wave's
// other lanes went on to the next step's barriers with the old unit and
// thread 0 never ran again.)
template <typename SIG>
__device__ __forceinline__ int treeStep(int32_t* ctr, int nctr, int nsig, SIG sig, int* sU) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    waitVmem();
    __syncthreads();
    if ((int)threadIdx.x < nsig) __hip_atomic_fetch_add(ctr + sig((int)threadIdx.x), 1, __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_AGENT);
    if (threadIdx.x == 0) {
        const int q = blockIdx.x % kTreeQueues;
        *sU = kTreeQueues * __hip_atomic_fetch_add(ctr + nctr - 2 - kTreeQueues + q, 1, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT) + q;
    }
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(*sU);
}

// after the waits of lane-held counters: the block's device-coherent loads
// of the data they guard come after them
__device__ __forceinline__ void treeAcquire() {
    __syncthreads();
    waitVmem();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// the block's counter updates have completed before it counts its exit; the
// last block out sees every other block's completed
__device__ __forceinline__ void treeExit(int32_t* ctr, int nctr, int* sU) {
    if (threadIdx.x == 0) {
        waitVmem();
        *sU = __hip_atomic_fetch_add(ctr + nctr - 2, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
              (int)gridDim.x - 1;
    }
    __syncthreads();
    if (__builtin_amdgcn_readfirstlane(*sU))
        for (int i = threadIdx.x; i < nctr - 1; i += blockDim.x) ctr[i] = 0;
}

// The upward pass of one right-hand side: S2M units (s2mPer consecutive
// pieces of the S2M order, a piece being a leaf item's direction chunk of
// BS; leaves largest first), then the fused M2M units (parent stem, run of
// parent theta rows) level by level from the finest parent level.  An M2M
// unit waits for every S2M or M2M unit that writes one of its children.
template <typename R, int BS>
__global__ __launch_bounds__(BS) void kTreeUp(TreeUp<R> tu) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int sU;
    int32_t* stalls = tu.ctr + tu.nctr - 1;
    // the previous unit: S2M pieces [sig0, sig0 + nsig) or, when sig0 < 0,
    // the counter -sig0 - 1 (an M2M stem)
    int sig0 = 0, nsig = 0;
    for (;;) {
        const int u = treeStep(tu.ctr, tu.nctr, nsig,
                               [&](int i) { return sig0 >= 0 ? tu.order[(sig0 + i) / tu.s2mChunks] : -sig0 - 1; },
                               &sU);
        if (u >= tu.total) break;
        if (u < tu.nS2M) {
            const int f0 = u * tu.s2mPer, f1 = min(f0 + tu.s2mPer, tu.nS2MPieces);
            for (int f = f0; f < f1; ++f) {
                const LeafItem it = tu.items[tu.order[f / tu.s2mChunks]];
                const int d0 = (f % tu.s2mChunks) * BS;
                if (d0 < it.ndir) s2mOne<R, true>(it, d0, tu.rp, tu.x, tu.perm, tu.M);
            }
            sig0 = f0;
            nsig = f1 - f0;
        } else {
            int l = 0;
            while (l + 1 < tu.nlev && u >= tu.lv[l + 1].unit0) ++l;
            const TreeUpLevel<R>& L = tu.lv[l];
            const int v = u - L.unit0;
            const int stem = v / L.nchunk, chunk = v - stem * L.nchunk;
            const int c0 = L.ptr[stem], nc = L.ptr[stem + 1] - c0;
            if ((int)threadIdx.x < nc) {
                const int c = L.childCtr[c0 + threadIdx.x];
                if (c >= 0) treeWait(tu.ctr + c, c < tu.nLeaf ? tu.s2mChunks : tu.lv[l > 0 ? l - 1 : 0].nchunk, stalls);
            }
            treeAcquire();
            m2mFused<R, BS, true>(stem, chunk * L.rows, L.rows, L.stemSlot, L.ptr, L.child, L.oct, L.nthP, L.nphP,
                                  L.mthC, L.mphC, tu.p, L.thNear, L.thW, L.phNear, L.phW, L.phaseUp, tu.M + L.MoffC,
                                  tu.M + L.MoffP, smem);
            sig0 = -(L.ctr0 + stem) - 1;
            nsig = 1;
        }
    }
    treeExit(tu.ctr, tu.nctr, &sU);
}

template <typename R>
void launchTreeUp(const TreeUp<R>& tu, int bs, size_t lds, int grid, hipStream_t st) {
    if (tu.total <= 0) return;
    if (bs == 192)
        hipLaunchKernelGGL((kTreeUp<R, 192>), dim3(grid), dim3(192), lds, st, tu);
    else if (bs == 256)
        hipLaunchKernelGGL((kTreeUp<R, 256>), dim3(grid), dim3(256), lds, st, tu);
    else
        throw Error(MLFMA_ERR_ARG, "tree walk block size must be 192 or 256");
}

template <typename R>
int treeUpOccupancy(int bs, size_t lds) {
    int n = 0;
    hipError_t e = bs == 192 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kTreeUp<R, 192>, 192, lds)
                             : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kTreeUp<R, 256>, 256, lds);
    if (e != hipSuccess) throw Error(MLFMA_ERR_HIP, "occupancy query of the tree walk");
    return n;
}

// ------------------------------------------------------------------ M2L
// L_t[d] = sum over the interaction list of T_off[d] M_s[d]  (src/fmm/node.cpp:105-134),
// quadrature weights folded into T.  A block is (level, sibling group, chunk
// of 256 directions): the up to eight children of one parent share most of
// their sources, so each distinct source field is read once and applied to
// every sibling whose list holds it, with that pair's operator (scalar
// branches on the uniform offset table).  One launch covers every level;
// grid z is the right-hand side (fields at stride lv.fs).  CONJ: conjugated
// operators (the adjoint pass, whose offsets name the reversed pairs).
template <typename R, bool CONJ>
__global__ __launch_bounds__(TB) void kM2LGroup(M2LGroups<R> lv) {
    int b = blockIdx.x, l = 0;
    while (l + 1 < lv.nlev && b >= lv.blk[l + 1]) ++l;
    b -= lv.blk[l];
    const int nch = lv.nch[l];
    // chunks of one item on different XCDs (measured faster for these two
    // kernels, whose batched fields do not fit one XCD's L2 anyway)
    const int item = xcdItem(b / nch, lv.per[l]);
    const int chunk = b - (b / nch) * nch;
    if (item >= lv.ngrp[l]) return;
    const int ndir = lv.ndir[l];
    const int d = chunk * TB + threadIdx.x;
    if (d >= ndir) return;
    const int64_t zo = (int64_t)blockIdx.z * lv.fs;
    const int32_t* src = lv.src[l];
    const int32_t* off = lv.off[l];
    const cplx<R>* T = lv.T[l] + d;
    const cplx<R>* M = lv.M[l] + zo + 2 * d;
    const int q0 = lv.ptr[l][item], q1 = lv.ptr[l][item + 1];
    cplx<R> acc[8][2];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k][0] = acc[k][1] = mk<R>(0, 0);
    for (int q = q0; q < q1; ++q) {
        const cplx<R>* m = M + (int64_t)src[q] * ndir * 2;
        const cplx<R> m0 = m[0], m1 = m[1];
        const int32_t* o8 = off + 8 * q;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int o = o8[k];
            if (o >= 0) {
                cplx<R> t = T[(int64_t)o * ndir];
                if (CONJ) t.im = -t.im;
                acc[k][0] = acc[k][0] + t * m0;
                acc[k][1] = acc[k][1] + t * m1;
            }
        }
    }
    const int32_t* tg = lv.tgt[l] + 8 * item;
    cplx<R>* L = lv.L[l] + zo + 2 * d;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int slot = tg[k];
        if (slot >= 0) {
            cplx<R>* out = L + (int64_t)slot * ndir * 2;
            out[0] = acc[k][0];
            out[1] = acc[k][1];
        }
    }
}

template <typename R>
void launchM2LGroup(M2LGroups<R> lv, bool conjT, hipStream_t st, int nb) {
    int total = 0;
    for (int l = 0; l < lv.nlev; ++l) {
        lv.nch[l] = (lv.ndir[l] + TB - 1) / TB;
        lv.per[l] = (lv.ngrp[l] + 7) / 8;
        lv.blk[l] = total;
        total += 8 * lv.per[l] * lv.nch[l];
    }
    if (total <= 0) return;
    const dim3 g(total, 1, nb);
    if (conjT)
        hipLaunchKernelGGL((kM2LGroup<R, true>), g, dim3(TB), 0, st, lv);
    else
        hipLaunchKernelGGL((kM2LGroup<R, false>), g, dim3(TB), 0, st, lv);
}

// ------------------------------------------------------------------ M2L per target
// L_t[d] = sum over the interaction list of T_off[d] M_s[d]  (src/fmm/node.cpp:105-134),
// quadrature weights folded into T.  One launch covers every level: a block
// is (level, target box, chunk of 256 directions); the (source, offset) list
// is read with uniform loads; within a level the targets of one XCD are
// spatially contiguous so their shared sources stay in that XCD's L2.  NB
// right-hand sides (fields at stride lv.fs) share every operator read; one
// right-hand side keeps two pairs in flight instead.
template <typename R, bool CONJ, int NB>
__global__ __launch_bounds__(TB) void kM2Lx(M2LLevels<R> lv) {
    int b = blockIdx.x, l = 0;
    while (l + 1 < lv.nlev && b >= lv.blk[l + 1]) ++l;
    b -= lv.blk[l];
    const int nch = lv.nch[l];
    // chunks of one item on different XCDs (measured faster for these two
    // kernels, whose batched fields do not fit one XCD's L2 anyway)
    const int item = xcdItem(b / nch, lv.per[l]);
    const int chunk = b - (b / nch) * nch;
    if (item >= lv.ntgt[l]) return;
    const int ndir = lv.ndir[l];
    const int d = chunk * TB + threadIdx.x;
    if (d >= ndir) return;
    const int32_t* src = lv.src[l];
    const int32_t* off = lv.off[l];
    const cplx<R>* T = lv.T[l];
    const cplx<R>* M = lv.M[l];
    const int q0 = lv.ptr[l][item], q1 = lv.ptr[l][item + 1];
    cplx<R>* out = lv.L[l] + ((int64_t)lv.tgt[l][item] * ndir + d) * 2;
    if constexpr (NB == 1) {
        cplx<R> a0 = mk<R>(0, 0), a1 = mk<R>(0, 0), b0 = mk<R>(0, 0), b1 = mk<R>(0, 0);
        int q = q0;
        for (; q + 1 < q1; q += 2) {
            cplx<R> t0 = T[(int64_t)off[q] * ndir + d];
            cplx<R> t1 = T[(int64_t)off[q + 1] * ndir + d];
            if (CONJ) {
                t0.im = -t0.im;
                t1.im = -t1.im;
            }
            const cplx<R>* m0 = M + ((int64_t)src[q] * ndir + d) * 2;
            const cplx<R>* m1 = M + ((int64_t)src[q + 1] * ndir + d) * 2;
            a0 = a0 + t0 * m0[0];
            a1 = a1 + t0 * m0[1];
            b0 = b0 + t1 * m1[0];
            b1 = b1 + t1 * m1[1];
        }
        if (q < q1) {
            cplx<R> t0 = T[(int64_t)off[q] * ndir + d];
            if (CONJ) t0.im = -t0.im;
            const cplx<R>* m0 = M + ((int64_t)src[q] * ndir + d) * 2;
            a0 = a0 + t0 * m0[0];
            a1 = a1 + t0 * m0[1];
        }
        out[0] = a0 + b0;
        out[1] = a1 + b1;
    } else {
        cplx<R> a[NB][2];
#pragma unroll
        for (int r = 0; r < NB; ++r) a[r][0] = a[r][1] = mk<R>(0, 0);
        for (int q = q0; q < q1; ++q) {
            cplx<R> t = T[(int64_t)off[q] * ndir + d];
            if (CONJ) t.im = -t.im;
            const cplx<R>* m = M + ((int64_t)src[q] * ndir + d) * 2;
#pragma unroll
            for (int r = 0; r < NB; ++r) {
                a[r][0] = a[r][0] + t * m[r * lv.fs];
                a[r][1] = a[r][1] + t * m[r * lv.fs + 1];
            }
        }
#pragma unroll
        for (int r = 0; r < NB; ++r) {
            out[r * lv.fs] = a[r][0];
            out[r * lv.fs + 1] = a[r][1];
        }
    }
}

// Single right-hand side: a block is (level, target box, chunk of 64
// directions).  Block b runs on XCD b mod 8; each XCD takes a contiguous
// eighth of the level's (chunk, target) units in chunk-major order, so it
// reads the operators and source fields of about one direction chunk (a
// level's operators are read by one or two XCDs instead of all eight), and
// consecutive blocks walk neighbouring targets, whose sources overlap in
// its L2.  The block's four waves split the target's pair list (wave w takes
// pairs w, w + 4, ...; MLFMA_M2L_PAIRS pairs in flight per wave), so that
// many independent pairs are in flight per direction instead of a long
// dependent walk of index and field loads; the
// wave partials are added in LDS in a fixed order.  Lanes past the level's
// direction count read a clamped direction and store nothing.  bid: the
// block's index in the M2L launch; red: 3 x 64 x 2 values of LDS.
template <typename R, bool CONJ>
__device__ __forceinline__ void m2lwBody(int bid, const M2LLevels<R>& lv, cplx<R> (*red)[64][2]) {
    int b = bid, l = 0;
    while (l + 1 < lv.nlev && b >= lv.blk[l + 1]) ++l;
    b -= lv.blk[l];
    const int nch = lv.nch[l];
    const int u = (b & 7) * lv.per[l] + (b >> 3);  // per[l] = ceil(units / 8)
    if (u >= lv.ntgt[l] * nch) return;            // uniform per block
    const int chunk = u / lv.ntgt[l];
    const int item = u - chunk * lv.ntgt[l];
    const int ndir = lv.ndir[l];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int d = chunk * 64 + lane;
    const int dc = d < ndir ? d : ndir - 1;
    const int32_t* src = lv.src[l];
    const int32_t* off = lv.off[l];
    const cplx<R>* T = lv.T[l] + dc;
    const cplx<R>* M = lv.M[l] + 2 * dc;
    const int q0 = lv.ptr[l][item], n = lv.ptr[l][item + 1] - q0;
    cplx<R> a0 = mk<R>(0, 0), a1 = mk<R>(0, 0), b0 = mk<R>(0, 0), b1 = mk<R>(0, 0);
    // The wave's pairs q0 + w + 4 i are taken 64 at a time: lane i loads the
    // (source, offset) of pair i once, and every pair's indices are then read
    // from that lane (v_readlane), so that a step of P pairs waits for one
    // round trip (the field loads), not two.  The even pairs of a step add
    // into a, the odd ones into b.
    constexpr int P = MLFMA_M2L_PAIRS;
    for (int base = 0; base < n; base += 256) {
        const int qi = base + w + 4 * lane;
        const int mySrc = qi < n ? src[q0 + qi] : 0;
        const int myOff = qi < n ? off[q0 + qi] : 0;
        const int cnt = min(64, max(0, (n - base - w + 3) / 4));
        int i = 0;
        for (; i + P <= cnt; i += P) {
            cplx<R> t[P], m[P][2];
#pragma unroll
            for (int j = 0; j < P; ++j) {
                const int sj = __builtin_amdgcn_readlane(mySrc, i + j);
                const int oj = __builtin_amdgcn_readlane(myOff, i + j);
                t[j] = T[(int64_t)oj * ndir];
                const cplx<R>* mp = M + (int64_t)sj * ndir * 2;
                m[j][0] = mp[0];
                m[j][1] = mp[1];
            }
#pragma unroll
            for (int j = 0; j < P; ++j) {
                if (CONJ) t[j].im = -t[j].im;
                if (j % 2 == 0) {
                    a0 = a0 + t[j] * m[j][0];
                    a1 = a1 + t[j] * m[j][1];
                } else {
                    b0 = b0 + t[j] * m[j][0];
                    b1 = b1 + t[j] * m[j][1];
                }
            }
        }
        for (; i < cnt; ++i) {
            const int sj = __builtin_amdgcn_readlane(mySrc, i);
            const int oj = __builtin_amdgcn_readlane(myOff, i);
            cplx<R> t0 = T[(int64_t)oj * ndir];
            if (CONJ) t0.im = -t0.im;
            const cplx<R>* m0 = M + (int64_t)sj * ndir * 2;
            a0 = a0 + t0 * m0[0];
            a1 = a1 + t0 * m0[1];
        }
    }
    if (w > 0) {
        red[w - 1][lane][0] = a0 + b0;
        red[w - 1][lane][1] = a1 + b1;
    }
    __syncthreads();
    if (w == 0 && d < ndir) {
        cplx<R>* out = lv.L[l] + ((int64_t)lv.tgt[l][item] * ndir + d) * 2;
        out[0] = ((a0 + b0) + red[0][lane][0]) + (red[1][lane][0] + red[2][lane][0]);
        out[1] = ((a1 + b1) + red[0][lane][1]) + (red[1][lane][1] + red[2][lane][1]);
    }
}

template <typename R, bool CONJ>
__global__ __launch_bounds__(TB) void kM2Lw(M2LLevels<R> lv) {
    __shared__ cplx<R> red[3][64][2];
    m2lwBody<R, CONJ>(blockIdx.x, lv, red);
}

template <typename R>
void launchM2Lx(M2LLevels<R> lv, bool conjT, hipStream_t st, int nb) {
    int total = 0;
    for (int l = 0; l < lv.nlev; ++l) {
        lv.nch[l] = (lv.ndir[l] + TB - 1) / TB;
        lv.per[l] = (lv.ntgt[l] + 7) / 8;
        lv.blk[l] = total;
        total += 8 * lv.per[l] * lv.nch[l];
    }
    if (total <= 0) return;
    if (nb == 1) {
        total = 0;
        for (int l = 0; l < lv.nlev; ++l) {
            lv.nch[l] = (lv.ndir[l] + 63) / 64;
            lv.per[l] = (lv.ntgt[l] * lv.nch[l] + 7) / 8;
            lv.blk[l] = total;
            total += 8 * lv.per[l];
        }
        if (conjT)
            hipLaunchKernelGGL((kM2Lw<R, true>), dim3(total), dim3(TB), 0, st, lv);
        else
            hipLaunchKernelGGL((kM2Lw<R, false>), dim3(total), dim3(TB), 0, st, lv);
        return;
    }
#define M2L_CASE(C, B) hipLaunchKernelGGL((kM2Lx<R, C, B>), dim3(total), dim3(TB), 0, st, lv)
    if (nb == 1) { if (conjT) M2L_CASE(true, 1); else M2L_CASE(false, 1); }
    else if (nb == 2) { if (conjT) M2L_CASE(true, 2); else M2L_CASE(false, 2); }
    else if (nb == 4) { if (conjT) M2L_CASE(true, 4); else M2L_CASE(false, 4); }
    else if (nb == 8) { if (conjT) M2L_CASE(true, 8); else M2L_CASE(false, 8); }
    else throw Error(MLFMA_ERR_ARG, "right-hand-side batch must be 1, 2, 4 or 8");
#undef M2L_CASE
}

// ------------------------------------------------------------------ M2L over target runs
// L_t[d] = sum over the interaction list of T_off[d] M_s[d]  (src/fmm/node.cpp:105-134).
// Block = (level, run of consecutive targets, chunk of W directions).  The
// run's distinct sources (at most 256 boxes for at most 128 targets; each is
// used by many of the run's targets) and the 316 operators
// are staged in LDS for the chunk, W directions x 2 components of a source
// being one contiguous 64-byte read; the targets' sums then read only LDS,
// so each source field and operator is read from the caches once per run
// and chunk instead of once per pair.  Lane (target, direction): each lane
// sums its target's pairs in list order.  Within a level, the units of one
// XCD are a contiguous range of chunks, so its L2 holds the lines (4
// directions of a source box in 128 bytes) that neighbouring chunks share.
// Grid z: right-hand side (fields at stride lv.fs).  CONJ: conjugated
// operators (the adjoint pass, packed with the reversed pairs' offsets).
template <typename R, bool CONJ>
__global__ __launch_bounds__(TB) void kM2LT(M2LTiles<R> lv) {
    constexpr int W = m2ltDirs<R>();
    extern __shared__ __align__(16) unsigned char smem[];
    int b = blockIdx.x, l = 0;
    while (l + 1 < lv.nlev && b >= lv.blk[l + 1]) ++l;
    b -= lv.blk[l];
    const int nrun = lv.nrun[l];
    const int u = xcdItem(b, lv.per[l]);
    if (u >= nrun * lv.nch[l]) return;  // uniform per block
    const int chunk = u / nrun, g = u - chunk * nrun;
    const int ndir = lv.ndir[l];
    const int d0 = chunk * W;
    const cplx<R>* M = lv.M[l] + blockIdx.z * lv.fs;
    cplx<R>* L = lv.L[l] + blockIdx.z * lv.fs;
    const int s0 = lv.rsrc[l][g], ns = lv.rsrc[l][g + 1] - s0;
    cplx<R>* sM = reinterpret_cast<cplx<R>*>(smem);  // ns x W x 2
    cplx<R>* sT = sM + (size_t)ns * W * 2;           // kM2LOffsets x W
    const int32_t* slots = lv.slot[l] + s0;
    for (int e = threadIdx.x; e < ns * W * 2; e += TB) {
        const int i = e / (2 * W), r = e - i * (2 * W);
        const int d = d0 + (r >> 1);
        sM[e] = d < ndir ? M[((int64_t)slots[i] * ndir + d) * 2 + (r & 1)] : mk<R>(0, 0);
    }
    const cplx<R>* T = lv.T[l];
    for (int e = threadIdx.x; e < kM2LOffsets * W; e += TB) {
        const int o = e / W, d = d0 + (e - o * W);
        cplx<R> t = d < ndir ? T[(int64_t)o * ndir + d] : mk<R>(0, 0);
        if (CONJ) t.im = -t.im;
        sT[e] = t;
    }
    __syncthreads();
    const int w = threadIdx.x % W;
    const int d = d0 + w;
    const int k1 = lv.run[l][g + 1];
    const int32_t* ptr = lv.ptr[l];
    const int32_t* pk = lv.pk[l];
    for (int k = lv.run[l][g] + threadIdx.x / W; k < k1; k += TB / W) {
        const int q0 = ptr[k], q1 = ptr[k + 1];
        cplx<R> a0 = mk<R>(0, 0), a1 = mk<R>(0, 0);
        int q = q0;
        for (; q + 3 < q1; q += 4) {  // four pairs' index and LDS loads in flight
            int v[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) v[j] = pk[q + j];
            cplx<R> tt[4], mm[4][2];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                tt[j] = sT[(v[j] & 511) * W + w];
                const cplx<R>* m = sM + ((v[j] >> 9) * W + w) * 2;
                mm[j][0] = m[0];
                mm[j][1] = m[1];
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                a0 = a0 + tt[j] * mm[j][0];
                a1 = a1 + tt[j] * mm[j][1];
            }
        }
        for (; q + 1 < q1; q += 2) {
            const int v0 = pk[q], v1 = pk[q + 1];
            const cplx<R> t0 = sT[(v0 & 511) * W + w], t1 = sT[(v1 & 511) * W + w];
            const cplx<R>* m0 = sM + ((v0 >> 9) * W + w) * 2;
            const cplx<R>* m1 = sM + ((v1 >> 9) * W + w) * 2;
            a0 = a0 + t0 * m0[0];
            a1 = a1 + t0 * m0[1];
            a0 = a0 + t1 * m1[0];
            a1 = a1 + t1 * m1[1];
        }
        if (q < q1) {
            const int v0 = pk[q];
            const cplx<R> t0 = sT[(v0 & 511) * W + w];
            const cplx<R>* m0 = sM + ((v0 >> 9) * W + w) * 2;
            a0 = a0 + t0 * m0[0];
            a1 = a1 + t0 * m0[1];
        }
        if (d < ndir) {
            cplx<R>* out = L + ((int64_t)lv.tgt[l][k] * ndir + d) * 2;
            out[0] = a0;
            out[1] = a1;
        }
    }
}

template <typename R>
void launchM2LT(M2LTiles<R> lv, bool conjT, hipStream_t st, int nb) {
    constexpr int W = m2ltDirs<R>();
    int total = 0;
    for (int l = 0; l < lv.nlev; ++l) {
        lv.nch[l] = (lv.ndir[l] + W - 1) / W;
        lv.per[l] = (lv.nrun[l] * lv.nch[l] + 7) / 8;
        lv.blk[l] = total;
        total += 8 * lv.per[l];
    }
    if (total <= 0) return;
    const dim3 g(total, 1, nb);
    const size_t lds = m2ltLds<R>(lv.maxSrc);
    if (conjT)
        hipLaunchKernelGGL((kM2LT<R, true>), g, dim3(TB), lds, st, lv);
    else
        hipLaunchKernelGGL((kM2LT<R, false>), g, dim3(TB), lds, st, lv);
}

// ------------------------------------------------------------------ fused L2L
// One block per child box whose parent has a non-empty interaction list
// (src/fmm/stem.cpp:109-131,186-244).  The gather-form anterpolation tables
// are staged in LDS; the parent local, phase-shifted to the child centre on
// the parent grid, is staged in LDS; the phi anterpolation goes to LDS; the
// theta anterpolation with the pole rule is added into the child local.
template <typename R, bool ADJ, int BS>
__global__ __launch_bounds__(BS) void kL2LFused(const int32_t* __restrict__ parent, const int32_t* __restrict__ child,
                                                const int32_t* __restrict__ oct, int nitems, int per, int mthP,
                                                int mphP, int nthC, int nphC, const int32_t* __restrict__ atPhPtr,
                                                const int32_t* __restrict__ atPhSrc, const double* __restrict__ atPhW,
                                                int nAtPh, const int32_t* __restrict__ atThPtr,
                                                const int32_t* __restrict__ atThSrc, const double* __restrict__ atThW,
                                                const int32_t* __restrict__ atThFlip, int nAtTh,
                                                const cplx<R>* __restrict__ phaseDn, const cplx<R>* __restrict__ Lp,
                                                cplx<R>* __restrict__ Lc, int64_t fs) {
    // ADJ: the adjoint of M2M for one child, conj(phase_up) (child grid) times
    // the anterpolated parent field; phaseDn then holds the child-grid phases
    extern __shared__ __align__(16) unsigned char smem[];
    const int item = xcdItem(blockIdx.x, per);
    if (item >= nitems) return;
    Lp += blockIdx.z * fs;  // right-hand side blockIdx.z
    Lc += blockIdx.z * fs;
    const int ndirP = mthP * mphP;
    const int nInner = mthP * nphC;
    cplx<R>* sP = reinterpret_cast<cplx<R>*>(smem);
    cplx<R>* sI = sP + 2 * ndirP;
    R* tPhW = reinterpret_cast<R*>(sI + 2 * nInner);
    R* tThW = tPhW + nAtPh;                       // pole flips folded into the sign
    int* tPhPtr = reinterpret_cast<int*>(tThW + nAtTh);
    int* tPhSrc = tPhPtr + nphC + 1;
    int* tThPtr = tPhSrc + nAtPh;
    int* tThSrc = tThPtr + nthC + 1;              // 2 * coarse theta + flip
    // Prologue: every global load of a lane (tables, the parent local and its
    // phases, the child local the result is added to) is issued before the
    // first LDS store; a loop with one store per iteration waits for each
    // load in turn.  Up to U elements per lane and array go this way, the
    // rest (large grids only) through the loops after it.
    constexpr int U = 3, UO = 2;
    const int t = threadIdx.x;
    const int nOut = nthC * nphC;
    const cplx<R>* Lb = Lp + (int64_t)parent[item] * ndirP * 2;
    const cplx<R>* ph = ADJ ? nullptr : phaseDn + (int64_t)oct[item] * ndirP;
    const cplx<R>* phC = ADJ ? phaseDn + (int64_t)oct[item] * nOut : nullptr;
    cplx<R>* out = Lc + (int64_t)child[item] * nOut * 2;
    double wPh[U] = {}, wTh[U] = {};
    int sPh[U] = {}, sTh[U] = {}, fTh[U] = {};
    cplx<R> f0[U] = {}, f1[U] = {}, fq[U] = {}, o0[UO] = {}, o1[UO] = {}, oq[UO] = {};
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int i = t + u * BS;
        if (i < nAtPh) {
            wPh[u] = atPhW[i];
            sPh[u] = atPhSrc[i];
        }
        if (i < nAtTh) {
            wTh[u] = atThW[i];
            sTh[u] = atThSrc[i];
            fTh[u] = atThFlip[i];
        }
        if (i < ndirP) {
            if (!ADJ) fq[u] = ph[i];
            f0[u] = Lb[2 * i];
            f1[u] = Lb[2 * i + 1];
        }
    }
    const int pPh = t <= nphC ? atPhPtr[t] : 0;
    const int pTh = t <= nthC ? atThPtr[t] : 0;
#pragma unroll
    for (int u = 0; u < UO; ++u) {
        const int e = t + u * BS;
        if (e < nOut) {
            o0[u] = out[2 * e];
            o1[u] = out[2 * e + 1];
            if (ADJ) oq[u] = phC[e];
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int i = t + u * BS;
        if (i < nAtPh) {
            tPhW[i] = (R)wPh[u];
            tPhSrc[i] = sPh[u];
        }
        if (i < nAtTh) {
            const bool flip = fTh[u] != 0;
            tThW[i] = (R)(flip ? -wTh[u] : wTh[u]);
            tThSrc[i] = 2 * sTh[u] + (flip ? 1 : 0);
        }
        if (i < ndirP) {
            sP[2 * i] = ADJ ? f0[u] : fq[u] * f0[u];
            sP[2 * i + 1] = ADJ ? f1[u] : fq[u] * f1[u];
        }
    }
    if (t <= nphC) tPhPtr[t] = pPh;
    if (t <= nthC) tThPtr[t] = pTh;
    for (int i = t + U * BS; i < nAtPh; i += BS) {
        tPhW[i] = (R)atPhW[i];
        tPhSrc[i] = atPhSrc[i];
    }
    for (int i = t + U * BS; i < nAtTh; i += BS) {
        const bool flip = atThFlip[i] != 0;
        tThW[i] = (R)(flip ? -atThW[i] : atThW[i]);
        tThSrc[i] = 2 * atThSrc[i] + (flip ? 1 : 0);
    }
    for (int i = t + BS; i <= nphC; i += BS) tPhPtr[i] = atPhPtr[i];
    for (int i = t + BS; i <= nthC; i += BS) tThPtr[i] = atThPtr[i];
    for (int d = t + U * BS; d < ndirP; d += BS) {
        const cplx<R> q = ADJ ? mk<R>(1, 0) : ph[d];
        sP[2 * d] = ADJ ? Lb[2 * d] : q * Lb[2 * d];
        sP[2 * d + 1] = ADJ ? Lb[2 * d + 1] : q * Lb[2 * d + 1];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nInner; e += BS) {
        const int ith = e / nphC, jph = e - ith * nphC;
        const cplx<R>* row = sP + 2 * ith * mphP;
        cplx<R> a0 = mk<R>(0, 0), a1 = mk<R>(0, 0);
#pragma unroll 4
        for (int q = tPhPtr[jph]; q < tPhPtr[jph + 1]; ++q) {
            const int d = 2 * tPhSrc[q];
            const R w = tPhW[q];
            a0 = a0 + w * row[d];
            a1 = a1 + w * row[d + 1];
        }
        sI[2 * e] = a0;
        sI[2 * e + 1] = a1;
    }
    __syncthreads();
    // theta anterpolation (pole rule) of output e, times the conjugated
    // child-grid phase for the adjoint
    auto anter = [&](int e, cplx<R>& a0, cplx<R>& a1, cplx<R> q) {
        const int jf = e / nphC, js = e - jf * nphC;
        const int jsf = js + (js < nphC / 2 ? nphC / 2 : -nphC / 2);
        a0 = mk<R>(0, 0);
        a1 = mk<R>(0, 0);
#pragma unroll 4
        for (int k = tThPtr[jf]; k < tThPtr[jf + 1]; ++k) {
            const int sf = tThSrc[k];
            const int d = 2 * ((sf >> 1) * nphC + ((sf & 1) ? jsf : js));
            const R w = tThW[k];
            a0 = a0 + w * sI[d];
            a1 = a1 + w * sI[d + 1];
        }
        if (ADJ) {
            q.im = -q.im;
            a0 = q * a0;
            a1 = q * a1;
        }
    };
#pragma unroll
    for (int u = 0; u < UO; ++u) {
        const int e = t + u * BS;
        if (e < nOut) {
            cplx<R> a0, a1;
            anter(e, a0, a1, oq[u]);
            out[2 * e] = o0[u] + a0;
            out[2 * e + 1] = o1[u] + a1;
        }
    }
    for (int e = t + UO * BS; e < nOut; e += BS) {
        cplx<R> a0, a1;
        anter(e, a0, a1, ADJ ? phC[e] : mk<R>(0, 0));
        out[2 * e] = out[2 * e] + a0;
        out[2 * e + 1] = out[2 * e + 1] + a1;
    }
}

template <typename R>
bool launchL2LFused(bool adjoint, int nitems, const int32_t* parent, const int32_t* child, const int32_t* oct, int mthP, int mphP,
                    int nthC, int nphC, const int32_t* atPhPtr, const int32_t* atPhSrc, const double* atPhW, int nAtPh,
                    const int32_t* atThPtr, const int32_t* atThSrc, const double* atThW, const int32_t* atThFlip,
                    int nAtTh, const cplx<R>* phaseDn, const cplx<R>* Lp, cplx<R>* Lc, size_t ldsMax,
                    hipStream_t st, int nb, int64_t fs) {
    if (nitems <= 0) return true;
    const size_t lds = sizeof(cplx<R>) * 2 * ((size_t)mthP * mphP + (size_t)mthP * nphC) +
                       sizeof(R) * ((size_t)nAtPh + nAtTh) +
                       sizeof(int) * ((size_t)nphC + 1 + nAtPh + nthC + 1 + nAtTh);
    if (lds > ldsMax) return false;
    const int per = (nitems + 7) / 8;
    const dim3 g(8 * per, 1, nb);
    // blocks whose LDS leaves room for one or two per CU, or levels with
    // few blocks each with a large grid, get more waves, so that each CU
    // still has enough loads in flight
#define L2L_LAUNCH(BSZ)                                                                                              \
    {                                                                                                               \
        if (adjoint)                                                                                                \
            hipLaunchKernelGGL((kL2LFused<R, true, BSZ>), g, dim3(BSZ), lds, st, parent, child, oct, nitems, per,     \
                               mthP, mphP, nthC, nphC, atPhPtr, atPhSrc, atPhW, nAtPh, atThPtr, atThSrc, atThW,      \
                               atThFlip, nAtTh, phaseDn, Lp, Lc, fs);                                               \
        else                                                                                                        \
            hipLaunchKernelGGL((kL2LFused<R, false, BSZ>), g, dim3(BSZ), lds, st, parent, child, oct, nitems, per,    \
                               mthP, mphP, nthC, nphC, atPhPtr, atPhSrc, atPhW, nAtPh, atThPtr, atThSrc, atThW,      \
                               atThFlip, nAtTh, phaseDn, Lp, Lc, fs);                                               \
    }
    // one block per CU by LDS, or few blocks each with work for 1024 threads
    if (2 * lds > ldsMax || (nitems < 1024 && mthP * nphC >= 1024))
        L2L_LAUNCH(1024)
    else if (3 * lds > ldsMax)
        L2L_LAUNCH(512)
    else
        L2L_LAUNCH(256)
#undef L2L_LAUNCH
    return true;
}


// ------------------------------------------------------------------ L2L adjoint
// The adjoint of L2L (src/fmm/stem.cpp:109-131,186-244): the parent gains
// conj(phase_down) (parent grid) times the interpolated child field, summed
// over the children that L2L feeds.  One block per (parent, run of parent
// theta rows); per child the theta interpolation (pole rule) goes to LDS and
// the phi interpolation, phase and sum stay in registers.
template <typename R, int EPT>
__global__ __launch_bounds__(TB) void kL2LAdj(const int32_t* __restrict__ parSlot, const int32_t* __restrict__ ptr,
                                              const int32_t* __restrict__ child, const int32_t* __restrict__ oct,
                                              int npar, int per, int rows, int nthP, int nphP, int mthC, int mphC,
                                              int p, const int32_t* __restrict__ thNear, const double* __restrict__ thW,
                                              const int32_t* __restrict__ phNear, const double* __restrict__ phW,
                                              const cplx<R>* __restrict__ phaseDn, const cplx<R>* __restrict__ Lc,
                                              cplx<R>* __restrict__ Lp) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int item = xcdItem(blockIdx.x, per);
    if (item >= npar) return;
    const int j0 = blockIdx.y * rows;
    const int nr = min(rows, nthP - j0);
    const int ndirC = mthC * mphC;
    const int ndirP = nthP * nphP;
    const int nInner = nr * mphC;
    const int nOut = nr * nphP;
    const int tp = 2 * p;
    cplx<R>* sC = reinterpret_cast<cplx<R>*>(smem);  // ndirC x 2
    cplx<R>* sI = sC + 2 * ndirC;                     // rows x mphC x 2
    R* tThW = reinterpret_cast<R*>(sI + (size_t)2 * rows * mphC);
    R* tPhW = tThW + rows * tp;
    int* tThNear = reinterpret_cast<int*>(tPhW + nphP * tp);
    int* tPhNear = tThNear + rows;
    for (int i = threadIdx.x; i < nr * tp; i += TB) tThW[i] = (R)thW[j0 * tp + i];
    for (int i = threadIdx.x; i < nphP * tp; i += TB) tPhW[i] = (R)phW[i];
    for (int i = threadIdx.x; i < nr; i += TB) tThNear[i] = thNear[j0 + i];
    for (int i = threadIdx.x; i < nphP; i += TB) tPhNear[i] = phNear[i];
    cplx<R> acc[EPT][2];
#pragma unroll
    for (int k = 0; k < EPT; ++k) acc[k][0] = acc[k][1] = mk<R>(0, 0);
    for (int c = ptr[item]; c < ptr[item + 1]; ++c) {
        __syncthreads();
        const cplx<R>* Cb = Lc + (int64_t)child[c] * ndirC * 2;
        for (int d = threadIdx.x; d < 2 * ndirC; d += TB) sC[d] = Cb[d];
        __syncthreads();
        for (int e = threadIdx.x; e < nInner; e += TB) {
            const int jr = e / mphC, iph = e - jr * mphC;
            const int near = tThNear[jr];
            const int iphf = iph + (iph < mphC / 2 ? mphC / 2 : -mphC / 2);
            cplx<R> a0 = mk<R>(0, 0), a1 = mk<R>(0, 0);
            for (int q = 0; q < tp; ++q) {
                const int i = near + 1 - p + q;
                const int f = i < 0 ? -i - 1 : (i >= mthC ? 2 * mthC - i - 1 : i);
                const bool flip = (i != f);
                const int dd = 2 * (f * mphC + (flip ? iphf : iph));
                const R w = flip ? -tThW[jr * tp + q] : tThW[jr * tp + q];
                a0 = a0 + w * sC[dd];
                a1 = a1 + w * sC[dd + 1];
            }
            sI[2 * e] = a0;
            sI[2 * e + 1] = a1;
        }
        __syncthreads();
        const cplx<R>* ph = phaseDn + (int64_t)oct[c] * ndirP + (int64_t)j0 * nphP;
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int e = threadIdx.x + k * TB;
            if (e < nOut) {
                const int jr = e / nphP, jph = e - jr * nphP;
                const int near = tPhNear[jph];
                const cplx<R>* in = sI + 2 * jr * mphC;
                cplx<R> a0 = mk<R>(0, 0), a1 = mk<R>(0, 0);
                for (int q = 0; q < tp; ++q) {
                    int i = near + 1 - p + q;
                    i = i < 0 ? i + mphC : (i >= mphC ? i - mphC : i);
                    const R w = tPhW[jph * tp + q];
                    a0 = a0 + w * in[2 * i];
                    a1 = a1 + w * in[2 * i + 1];
                }
                cplx<R> s = ph[e];
                s.im = -s.im;
                acc[k][0] = acc[k][0] + s * a0;
                acc[k][1] = acc[k][1] + s * a1;
            }
        }
    }
    cplx<R>* out = Lp + ((int64_t)parSlot[item] * nthP + j0) * nphP * 2;
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
        const int e = threadIdx.x + k * TB;
        if (e < nOut) {
            out[2 * e] = out[2 * e] + acc[k][0];
            out[2 * e + 1] = out[2 * e + 1] + acc[k][1];
        }
    }
}

template <typename R>
bool launchL2LAdj(int npar, const int32_t* parSlot, const int32_t* ptr, const int32_t* child, const int32_t* oct,
                  int nthP, int nphP, int mthC, int mphC, int p, const int32_t* thNear, const double* thW,
                  const int32_t* phNear, const double* phW, const cplx<R>* phaseDn, const cplx<R>* Lc, cplx<R>* Lp,
                  size_t ldsMax, hipStream_t st, int minBlocks) {
    if (npar <= 0) return true;
    if (minBlocks <= 0) minBlocks = MLFMA_M2M_MIN_BLOCKS;
    int rows = nthP;
    while (rows > 2 && (int64_t)npar * ((nthP + rows - 1) / rows) < minBlocks) rows = (rows + 1) / 2;
    while (rows > 1 && rows * nphP > 16 * TB) rows = (rows + 1) / 2;  // at most 16 outputs per thread
    const int tp = 2 * p;
    const size_t lds = sizeof(cplx<R>) * 2 * ((size_t)mthC * mphC + (size_t)rows * mphC) +
                       sizeof(R) * ((size_t)rows + nphP) * tp + sizeof(int) * ((size_t)rows + nphP);
    const int ept = (rows * nphP + TB - 1) / TB;
    if (lds > ldsMax || ept > 16) return false;
    const int per = (npar + 7) / 8;
    dim3 g(8 * per, (nthP + rows - 1) / rows);
#define ADJ_CASE(E)                                                                                                \
    hipLaunchKernelGGL((kL2LAdj<R, E>), g, dim3(TB), lds, st, parSlot, ptr, child, oct, npar, per, rows, nthP, nphP, \
                       mthC, mphC, p, thNear, thW, phNear, phW, phaseDn, Lc, Lp)
    if (ept <= 1) ADJ_CASE(1);
    else if (ept <= 2) ADJ_CASE(2);
    else if (ept <= 4) ADJ_CASE(4);
    else if (ept <= 8) ADJ_CASE(8);
    else ADJ_CASE(16);
#undef ADJ_CASE
    return true;
}

template <typename R>
__global__ void kZeroField(cplx<R>* __restrict__ f, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i < n) f[i] = mk<R>(0, 0);
}

template <typename R>
void launchZeroField(cplx<R>* f, int64_t n, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(kZeroField<R>, dim3((unsigned)((n + TB - 1) / TB)), dim3(TB), 0, st, f, n);
}

// ------------------------------------------------------------------ near field
// One block per work item (rows of one leaf x a run of its columns, values
// column-major with C k folded in).  Each value is read from HBM once and
// used twice, as the reference uses each stored pair entry
// (src/fmm/leaf.cpp:329-355): row sums Z x_cols for the item's rows and, for
// the pair columns, column sums Z^T x_rows.  Tiles of NT values are staged in
// LDS with a leading dimension made odd (conflict-free column reads); the
// partial sums go to a buffer that kNearSum folds in a fixed order.
static_assert(MLFMA_NEAR_LOADS == 5 || MLFMA_NEAR_LOADS == 6 || MLFMA_NEAR_LOADS == 8,
              "near-field loads per lane: 5, 6 or 8");
constexpr int NT = MLFMA_NEAR_LOADS * TB;  // values per tile

// V: the stored value type (double; float for the FP32 inner solves of the
// mixed-precision path, products and sums still in FP64).  bid: the item;
// sT, sxr, sxc: LDS of NT x (16 / sizeof(cplx<V>)) values, 128 and TB values.
template <bool CONJ, typename V, int NL = MLFMA_NEAR_LOADS>
__device__ __forceinline__ void nearSymBody(int bid, const NearItem* __restrict__ items,
                                            const int32_t* __restrict__ colsIn, const int64_t* __restrict__ perm,
                                            const cplx<V>* __restrict__ val, const cplx<double>* __restrict__ x,
                                            cplx<double>* __restrict__ part, cplx<V>* sT, cplx<double>* sxr,
                                            cplx<double>* sxc) {
    // Every load of a lane is 16 bytes: one FP64 value or two FP32 values,
    // so an FP32 tile holds twice the entries in the same 24 KiB of LDS (the
    // tile is kept in V).  30 KiB of LDS per block (five blocks per CU): the
    // row-sum reduction at the end reuses the tile's storage, the column sums
    // reduce by shuffles.
    constexpr int EPL = sizeof(V) == 4 ? 2 : 1;  // values per lane and load
    constexpr int NTV = NL * TB * EPL;
    const NearItem I = items[bid];
    const int nr = I.nr;        // 1..128
    const int ld = nr | 1;      // odd leading dimension in LDS
    const int cwMax = min(TB, NTV / ld);
    if (threadIdx.x < nr) sxr[threadIdx.x] = x[perm[I.rbeg + threadIdx.x]];
    const int G = TB / nr;
    const int r = threadIdx.x % nr, g = threadIdx.x / nr;
    const int t0 = threadIdx.x;
    const int ntile = (I.nc + cwMax - 1) / cwMax;
    auto tileCw = [&](int j) { return min(cwMax, I.nc - j * cwMax); };
    // Software pipeline: the loads of tile j+1 are issued before the LDS
    // work of tile j.  Indices are clamped into the tile so every load is
    // valid; the column indices of the x gathers run one tile ahead of the
    // gathers that use them, so that a tile's loads wait for one round trip.
    int ci = t0 < tileCw(0) ? colsIn[I.cptr + t0] : 0;
    V vr[NL][EPL], vi[NL][EPL];
    cplx<double> xc = mk(0.0, 0.0);
    auto issue = [&](int j) {
        const int c0 = j * cwMax, cw = tileCw(j), ne = cw * nr;
        const V* sd = reinterpret_cast<const V*>(val + I.voff + (int64_t)c0 * nr);
        if constexpr (EPL == 1) {
            // FP64 values are read once per matvec: non-temporal loads, so
            // that they do not evict the patterns and fields the other stages
            // re-read from L2 and the MALL
            using d2 = double __attribute__((ext_vector_type(2)));
            const d2* sv = reinterpret_cast<const d2*>(sd);
#pragma unroll
            for (int k = 0; k < NL; ++k) {
                const d2 v = __builtin_nontemporal_load(sv + min(t0 + k * TB, ne - 1));
                vr[k][0] = v.x;
                vi[k][0] = CONJ ? -v.y : v.y;
            }
        } else {
#pragma unroll
            for (int k = 0; k < NL; ++k)
#pragma unroll
                for (int e = 0; e < EPL; ++e) {
                    const int q = 2 * min((t0 + k * TB) * EPL + e, ne - 1);
                    vr[k][e] = sd[q];
                    vi[k][e] = CONJ ? -sd[q + 1] : sd[q + 1];
                }
        }
        if (t0 < cw) xc = x[ci];
        const int cn = c0 + cwMax;
        ci = t0 < min(cwMax, I.nc - cn) ? colsIn[I.cptr + cn + t0] : 0;
    };
    auto stage = [&](int j) {
        const int cw = tileCw(j), ne = cw * nr;
#pragma unroll
        for (int k = 0; k < NL; ++k)
#pragma unroll
            for (int e = 0; e < EPL; ++e) {
                const int q = (t0 + k * TB) * EPL + e;
                if (q < ne) {
                    const int c = q / nr;
                    sT[c * ld + (q - c * nr)] = cplx<V>{vr[k][e], vi[k][e]};
                }
            }
        if (t0 < cw) sxc[t0] = xc;
    };
    double re = 0, im = 0;
    auto compute = [&](int j) {
        const int c0 = j * cwMax, cw = tileCw(j);
        if (g < G)
            for (int c = g; c < cw; c += G) {
                const cplx<V> z = sT[c * ld + r];
                const cplx<double> xx = sxc[c];
                re += (double)z.re * xx.re - (double)z.im * xx.im;
                im += (double)z.re * xx.im + (double)z.im * xx.re;
            }
        // column sums of the pair columns: S lanes per column (a power of two
        // <= 64, so a column's lanes sit in one wave), each over a contiguous
        // run of rows, added by a butterfly in a fixed order
        if (c0 + cw > I.nself) {  // uniform over the block
            int S = 1;
            while (S < 64 && 2 * S * cw <= TB) S *= 2;
            const int cc = t0 / S, sg = t0 - cc * S;
            double cr = 0, cm = 0;
            const bool mine = cc < cw && c0 + cc >= I.nself;
            if (mine) {
                const int L = (nr + S - 1) / S;
                const int qa = min(nr, sg * L), qb = min(nr, qa + L);
                const cplx<V>* col = sT + cc * ld;
                for (int q = qa; q < qb; ++q) {
                    const cplx<V> z0 = col[q];
                    const cplx<double> x0 = sxr[q];
                    cr += (double)z0.re * x0.re - (double)z0.im * x0.im;
                    cm += (double)z0.re * x0.im + (double)z0.im * x0.re;
                }
            }
            for (int o = S / 2; o > 0; o >>= 1) {  // every lane of the wave takes part
                cr += __shfl_xor(cr, o, 64);
                cm += __shfl_xor(cm, o, 64);
            }
            if (mine && sg == 0) part[I.cpo + c0 + cc - I.nself] = mk(cr, cm);
        }
    };
    issue(0);
    for (int j = 0; j < ntile; ++j) {
        __syncthreads();  // the previous tile is consumed
        stage(j);
        __syncthreads();
        if (j + 1 < ntile) issue(j + 1);
        compute(j);
    }
    double* sre = reinterpret_cast<double*>(sT);  // the tile's storage, free now
    double* sim = sre + TB;
    __syncthreads();  // every lane is done with the last tile
    sre[threadIdx.x] = (g < G) ? re : 0.0;
    sim[threadIdx.x] = (g < G) ? im : 0.0;
    __syncthreads();
    if (threadIdx.x < nr) {
        double a = 0, b = 0;
        for (int q = 0; q < G; ++q) {
            a += sre[q * nr + threadIdx.x];
            b += sim[q * nr + threadIdx.x];
        }
        part[I.rpo + threadIdx.x] = mk(a, b);
    }
}

template <bool CONJ, typename V>
__global__ __launch_bounds__(TB) void kNearSym(const NearItem* __restrict__ items, const int32_t* __restrict__ colsIn,
                                               const int64_t* __restrict__ perm, const cplx<V>* __restrict__ val,
                                               const cplx<double>* __restrict__ x, cplx<double>* __restrict__ part) {
    __shared__ cplx<V> sT[NT * (sizeof(V) == 4 ? 2 : 1)];
    __shared__ cplx<double> sxr[128], sxc[TB];
    nearSymBody<CONJ, V>(blockIdx.x, items, colsIn, perm, val, x, part, sT, sxr, sxc);
}

// The same with a fixed number of blocks, each taking items blockIdx.x,
// blockIdx.x + gridDim.x, ...: launched beside the far-field chain, the near
// field then holds only that many of a CU's LDS and wave slots, so that the
// chain's blocks are not kept waiting for the whole near field to drain.
// With few waves per CU, each lane keeps more values in flight.
constexpr int NEAR_LOADS_P = 8;
constexpr int NTP = NEAR_LOADS_P * TB;
template <bool CONJ, typename V>
__global__ __launch_bounds__(TB) void kNearSymP(const NearItem* __restrict__ items, int nitems,
                                                const int32_t* __restrict__ colsIn, const int64_t* __restrict__ perm,
                                                const cplx<V>* __restrict__ val, const cplx<double>* __restrict__ x,
                                                cplx<double>* __restrict__ part) {
    __shared__ cplx<V> sT[NTP * (sizeof(V) == 4 ? 2 : 1)];
    __shared__ cplx<double> sxr[128], sxc[TB];
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        nearSymBody<CONJ, V, NEAR_LOADS_P>(it, items, colsIn, perm, val, x, part, sT, sxr, sxc);
        __syncthreads();  // the item's LDS reads are done before the next item's stores
    }
}

// The same for NB >= 2 right-hand sides (x column stride xs, partial sums of
// right-hand side r at part + r ps): each value is read once and used 2 NB
// times.  Items have at most 128 rows; tiles of 4 values per lane.
constexpr int NTB = 4 * TB;
constexpr int CWB = 64;

template <bool CONJ, int NB, typename V>
__global__ __launch_bounds__(TB) void kNearSymB(const NearItem* __restrict__ items, const int32_t* __restrict__ colsIn,
                                                const int64_t* __restrict__ perm, const cplx<V>* __restrict__ val,
                                                const cplx<double>* __restrict__ x, int64_t xs,
                                                cplx<double>* __restrict__ part, int64_t ps) {
    __shared__ cplx<double> sT[NTB];
    __shared__ cplx<double> sxr[NB][128];
    __shared__ cplx<double> sxc[NB][CWB];
    __shared__ double sre[TB], sim[TB];
    const NearItem I = items[blockIdx.x];
    const int nr = I.nr;  // 1..128
    const int ld = nr | 1;
    const int cwMax = min(CWB, NTB / ld);
    const int t0 = threadIdx.x;
    for (int e = t0; e < NB * nr; e += TB) {
        const int rh = e / nr, q = e - rh * nr;
        sxr[rh][q] = x[rh * xs + perm[I.rbeg + q]];
    }
    const int G = TB / nr;
    const int r = t0 % nr, g = t0 / nr;
    double re[NB], im[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) re[k] = im[k] = 0.0;
    double* sTd = reinterpret_cast<double*>(sT);
    // Software pipeline: the values of tile k+1 and its x gathers are issued
    // before the LDS work of tile k, and the column indices of the gathers
    // run one tile further ahead, so that every wait is for loads issued one
    // tile earlier.  Lane t0 gathers entries t0 and t0 + TB of the tile's
    // NB x cw column values (NB cw <= 2 TB).
    static_assert(NB * CWB <= 2 * TB, "two column gathers per lane");
    int c0 = 0, cw = min(cwMax, I.nc), ne = cw * nr;
    double v0r, v0i, v1r, v1i, v2r, v2i, v3r, v3i;
    cplx<double> xg0 = mk(0.0, 0.0), xg1 = mk(0.0, 0.0);
    int ia0 = 0, ia1 = 0;  // column indices (input order) of the next gathers
    auto colIdx = [&](int cc0, int ccw, int e) {
        return e < NB * ccw ? colsIn[I.cptr + cc0 + e % ccw] : 0;
    };
    auto loadVals = [&](int cc0, int nne) {
        const V* sd = reinterpret_cast<const V*>(val + I.voff + (int64_t)cc0 * nr);
        const int q0 = 2 * min(t0, nne - 1), q1 = 2 * min(t0 + TB, nne - 1);
        const int q2 = 2 * min(t0 + 2 * TB, nne - 1), q3 = 2 * min(t0 + 3 * TB, nne - 1);
        v0r = sd[q0];
        v0i = sd[q0 + 1];
        v1r = sd[q1];
        v1i = sd[q1 + 1];
        v2r = sd[q2];
        v2i = sd[q2 + 1];
        v3r = sd[q3];
        v3i = sd[q3 + 1];
    };
    auto gather = [&](int ccw) {
        if (t0 < NB * ccw) xg0 = x[(t0 / ccw) * xs + ia0];
        if (t0 + TB < NB * ccw) xg1 = x[((t0 + TB) / ccw) * xs + ia1];
    };
    ia0 = colIdx(0, cw, t0);
    ia1 = colIdx(0, cw, t0 + TB);
    loadVals(0, ne);
    gather(cw);
    {
        const int cn = cwMax, cwn = min(cwMax, I.nc - cn);
        ia0 = cwn > 0 ? colIdx(cn, cwn, t0) : 0;
        ia1 = cwn > 0 ? colIdx(cn, cwn, t0 + TB) : 0;
    }
    while (true) {
        __syncthreads();  // the previous tile is consumed
#define NEARB_ST(K, VR, VI)                                   \
    if (t0 + K * TB < ne) {                                   \
        const int e = t0 + K * TB;                            \
        const int c = e / nr;                                 \
        const int a = 2 * (c * ld + (e - c * nr));            \
        sTd[a] = VR;                                          \
        sTd[a + 1] = CONJ ? -(VI) : (VI);                     \
    }
        NEARB_ST(0, v0r, v0i) NEARB_ST(1, v1r, v1i) NEARB_ST(2, v2r, v2i) NEARB_ST(3, v3r, v3i)
#undef NEARB_ST
        if (t0 < NB * cw) sxc[t0 / cw][t0 % cw] = xg0;
        if (t0 + TB < NB * cw) sxc[(t0 + TB) / cw][(t0 + TB) % cw] = xg1;
        __syncthreads();
        const int c0n = c0 + cwMax;
        const bool more = c0n < I.nc;  // uniform over the block
        int cwn = 0, nen = 0;
        if (more) {
            cwn = min(cwMax, I.nc - c0n);
            nen = cwn * nr;
            loadVals(c0n, nen);
            gather(cwn);
            const int cn = c0n + cwMax, cwnn = min(cwMax, I.nc - cn);
            ia0 = cwnn > 0 ? colIdx(cn, cwnn, t0) : 0;
            ia1 = cwnn > 0 ? colIdx(cn, cwnn, t0 + TB) : 0;
        }
        if (g < G)
            for (int c = g; c < cw; c += G) {
                const cplx<double> z = sT[c * ld + r];
#pragma unroll
                for (int k = 0; k < NB; ++k) {
                    const cplx<double> xx = sxc[k][c];
                    re[k] += z.re * xx.re - z.im * xx.im;
                    im[k] += z.re * xx.im + z.im * xx.re;
                }
            }
        for (int e = t0; e < NB * cw; e += TB) {
            const int rh = e / cw, c = e - rh * cw;
            if (c0 + c < I.nself) continue;
            const cplx<double>* col = sT + c * ld;
            double cr = 0, ci = 0;
            for (int q = 0; q < nr; ++q) {
                const cplx<double> z = col[q], xx = sxr[rh][q];
                cr += z.re * xx.re - z.im * xx.im;
                ci += z.re * xx.im + z.im * xx.re;
            }
            part[rh * ps + I.cpo + c0 + c - I.nself] = mk(cr, ci);
        }
        if (!more) break;
        c0 = c0n;
        cw = cwn;
        ne = nen;
    }
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        __syncthreads();
        sre[t0] = (g < G) ? re[k] : 0.0;
        sim[t0] = (g < G) ? im[k] : 0.0;
        __syncthreads();
        if (t0 < nr) {
            double a = 0, b = 0;
            for (int q = 0; q < G; ++q) {
                a += sre[q * nr + t0];
                b += sim[q * nr + t0];
            }
            part[k * ps + I.rpo + t0] = mk(a, b);
        }
    }
}

// Sum of the near-field partials of tree position t, split over the 64 lanes
// of a wave (same lane assignment and butterfly every call: deterministic);
// partial j at part[j * stride] (stride 8: the interleaved partials of eight
// right-hand sides, part offset to the right-hand side).
__device__ __forceinline__ cplx<double> nearFold(const int64_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                                                 const cplx<double>* __restrict__ part, int64_t t, int lane,
                                                 int stride = 1) {
    double a = 0, b = 0;
    for (int64_t q = ptr[t] + lane; q < ptr[t + 1]; q += 64) {
        const cplx<double> v = part[(int64_t)idx[q] * stride];
        a += v.re;
        b += v.im;
    }
    return mk(wsum(a), wsum(b));
}

// y[k xs + perm[t]] (=|+=) near-field sum of right-hand side k from the
// interleaved partials (part[j 8 + k]: eight lanes per 128-byte line), eight
// right-hand sides, 16 lanes per listed row (two partials of the eight per
// round, four rows per wave; a wave per row, eight partials per round, took
// 0.105 ms for the 201,720 rows at C3, this 0.070 ms)
__global__ __launch_bounds__(TB) void kNearSum8(const int64_t* __restrict__ rows, int64_t nrows,
                                                const int64_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                                                const cplx<double>* __restrict__ part, const int64_t* __restrict__ perm,
                                                cplx<double>* __restrict__ y, int64_t xs, int accumulate) {
    const int lane = threadIdx.x & 63, sl = lane & 15;
    const int64_t w = 4 * ((int64_t)blockIdx.x * (TB / 64) + (threadIdx.x >> 6)) + (lane >> 4);
    if (w - (lane >> 4) >= nrows) return;  // uniform per wave
    const int64_t t = rows[w < nrows ? w : nrows - 1];
    double a = 0, b = 0;
    const int k = sl & 7;
    const int64_t q1 = ptr[t + 1];
    for (int64_t q = ptr[t] + (sl >> 3); q < q1; q += 2) {
        const cplx<double> v = part[(int64_t)idx[q] * 8 + k];
        a += v.re;
        b += v.im;
    }
    a += __shfl_xor(a, 8, 64);
    b += __shfl_xor(b, 8, 64);
    if (sl < 8 && w < nrows) {
        cplx<double>* o = y + k * xs + perm[t];
        const cplx<double> v = mk(a, b);
        *o = accumulate ? *o + v : v;
    }
}

// y[perm[t]] (=|+=) near-field sum, one wave per listed row (rows that have no
// far-field leaf; the far-field rows are finished by kL2Tx); NB right-hand
// sides (y column stride xs, partials at stride ps).
template <int NB>
__global__ __launch_bounds__(TB) void kNearSum(const int64_t* __restrict__ rows, int64_t nrows,
                                               const int64_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                                               const cplx<double>* __restrict__ part, int64_t ps,
                                               const int64_t* __restrict__ perm, cplx<double>* __restrict__ y,
                                               int64_t xs, int accumulate) {
    const int64_t w = (int64_t)blockIdx.x * (TB / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (w >= nrows) return;  // uniform per wave
    const int64_t t = rows[w];
    const int64_t o = perm[t];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        // eight right-hand sides: interleaved partials (kNearSymW)
        const cplx<double> v =
            NB == 8 ? nearFold(ptr, idx, part + k, t, lane, 8) : nearFold(ptr, idx, part + k * ps, t, lane);
        if (lane == 0) y[k * xs + o] = accumulate ? y[k * xs + o] + v : v;
    }
}

template <typename V>
void launchNearSym(bool conjVal, const NearItem* items, int nitems, const int32_t* colsIn, const int64_t* perm,
                   const cplx<V>* val, const cplx<double>* x, cplx<double>* part, hipStream_t st, int nb,
                   int64_t xs, int64_t ps, int grid) {
    if (nitems <= 0) return;
    if (nb == 1 && grid > 0) {
        // resident blocks whenever asked for (the kernel, not the item count,
        // fixes how each row's sum is grouped)
        const int g = std::min(grid, nitems);
        if (conjVal)
            hipLaunchKernelGGL((kNearSymP<true, V>), dim3(g), dim3(TB), 0, st, items, nitems, colsIn, perm, val, x,
                               part);
        else
            hipLaunchKernelGGL((kNearSymP<false, V>), dim3(g), dim3(TB), 0, st, items, nitems, colsIn, perm, val, x,
                               part);
        return;
    }
#define NEARB_CASE(C, B)                                                                                      \
    hipLaunchKernelGGL((kNearSymB<C, B, V>), dim3(nitems), dim3(TB), 0, st, items, colsIn, perm, val, x, xs, part, \
                       ps)
    if (nb == 1) {
        if (conjVal)
            hipLaunchKernelGGL((kNearSym<true, V>), dim3(nitems), dim3(TB), 0, st, items, colsIn, perm, val, x, part);
        else
            hipLaunchKernelGGL((kNearSym<false, V>), dim3(nitems), dim3(TB), 0, st, items, colsIn, perm, val, x, part);
    } else if (nb == 2) {
        if (conjVal) NEARB_CASE(true, 2); else NEARB_CASE(false, 2);
    } else if (nb == 4) {
        if (conjVal) NEARB_CASE(true, 4); else NEARB_CASE(false, 4);
    } else
        throw Error(MLFMA_ERR_ARG, "near field: batch of 1, 2 or 4 (eight: launchNearSymM8)");
#undef NEARB_CASE
}
// Eight right-hand sides on the matrix cores (v_mfma_f64_16x16x4f64: A
// 16 x 4, B 4 x 16, D 16 x 16; lane l holds A[l%16][l/16], B[l/16][l%16] and
// D[4i + l/16][l%16], i < 4).  The 16 columns of B are the real parts of the
// eight right-hand sides followed by their imaginary parts, so with
// B' = [-Im | Re] one accumulator takes Re Z . B + Im Z . B' = [Re Y | Im Y].
// x comes interleaved (xt[i 8 + k]) and the partial sums go out interleaved
// (part[j 8 + k]), so each gather and each row's or column's eight results
// are one 128-byte line.  Values are read from memory once.
//
// Wave-tiled: every wave of the block takes its own
// 16-column tiles of the item (tiles w, w + 4, ...) and, per tile, every
// 16-row tile of the pass (rows 0-63, then 64-127 for taller items): it
// loads the 16 x 16 block of values straight from memory in the MFMA A
// layout of the row sums (lane l: row l % 16, column 4 s + l / 16 of the
// block; four runs of 256 contiguous bytes per step), adds Z X_cols to the
// pass's row sums it keeps in registers, and for pair columns passes the
// block through its own LDS square to the transposed layout and adds
// Z^T X_rows to the tile's column sums (the second pass adds to the first's,
// which the same lane wrote).  No block barrier in the loop: the waves
// stream independently, the next block's values in flight during this
// block's products.  The row sums of the four waves are added in LDS in a
// fixed order at the end of each pass.  At most 168 registers per lane
// (three waves per SIMD: the compiler keeps two values in scratch).
template <bool CONJ, typename V>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(3))) void kNearSymW(const NearItem* __restrict__ items, const int32_t* __restrict__ colsIn,
                                                const int64_t* __restrict__ perm, const cplx<V>* __restrict__ val,
                                                const cplx<double>* __restrict__ xt,
                                                cplx<double>* __restrict__ part) {
    using d4 = double __attribute__((ext_vector_type(4)));
    constexpr int NB = 8, RT = 4;                  // right-hand sides, row tiles per pass
    __shared__ cplx<double> sxr[128 * NB];         // x of the item's rows, interleaved
    __shared__ double sZ[TB / 64][2][16][17];      // per wave: one 16 x 16 block (re, im), for the transpose
    __shared__ double sRed[RT * 16 * 16];          // row sums of a pass, [row][re 0-7 | im 0-7]
    const NearItem I = items[blockIdx.x];
    const int nr = I.nr, nc = I.nc;
    const int t0 = threadIdx.x, lane = t0 & 63;
    const int w = __builtin_amdgcn_readfirstlane(t0 >> 6);
    const int lm = lane & 15, lk = lane >> 4;
    const int rh = lm & 7;
    const bool reCol = lm < 8;
    for (int e = t0; e < NB * nr; e += TB) sxr[e] = xt[perm[I.rbeg + (e >> 3)] * NB + (e & 7)];
    __syncthreads();
    const int nrt = (nr + 15) >> 4, nct = (nc + 15) >> 4;
    const cplx<V>* Zb = val + I.voff;
    // lane's values of block (rt, ct), step s: row rt 16 + lm, column ct 16 + 4 s + lk (zero outside)
    auto loadBlock = [&](int rt, int ct, cplx<double>* z) {
        const int r = rt * 16 + lm;
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
            const int c = ct * 16 + 4 * s4 + lk;
            const bool ok = r < nr && c < nc;
            const cplx<V> v = Zb[(int64_t)min(c, nc - 1) * nr + min(r, nr - 1)];
            z[s4] = ok ? mk((double)v.re, CONJ ? -(double)v.im : (double)v.im) : mk(0.0, 0.0);
        }
    };
    for (int pass = 0; pass * RT < nrt; ++pass) {  // uniform
        const int rt0 = pass * RT, rt1 = min(nrt, rt0 + RT);
        d4 acc[RT];  // both products of a row tile into one accumulator
#pragma unroll
        for (int j = 0; j < RT; ++j) acc[j] = d4{0, 0, 0, 0};
        for (int ct = w; ct < nct; ct += TB / 64) {
            // B of the row sums: x of the tile's columns
            double bR[4], bRp[4];
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                const int c = ct * 16 + 4 * s4 + lk;
                const cplx<double> xv = c < nc ? xt[(int64_t)colsIn[I.cptr + c] * NB + rh] : mk(0.0, 0.0);
                bR[s4] = reCol ? xv.re : xv.im;
                bRp[s4] = reCol ? -xv.im : xv.re;
            }
            const bool pairs = ct * 16 + 15 >= I.nself;  // uniform per wave
            d4 cA = d4{0, 0, 0, 0};
            cplx<double> z[4], zn[4];
            loadBlock(rt0, ct, z);
#pragma unroll
            for (int j = 0; j < RT; ++j) {
                const int rt = rt0 + j;
                if (rt < rt1) {  // uniform
                    if (rt + 1 < rt1) loadBlock(rt + 1, ct, zn);
#pragma unroll
                    for (int s4 = 0; s4 < 4; ++s4) {
                        acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(z[s4].re, bR[s4], acc[j], 0, 0, 0);
                        acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(z[s4].im, bRp[s4], acc[j], 0, 0, 0);
                    }
                    if (pairs) {
                        // transpose through the wave's LDS square: element (row lm, column 4 s + lk)
#pragma unroll
                        for (int s4 = 0; s4 < 4; ++s4) {
                            sZ[w][0][lm][4 * s4 + lk] = z[s4].re;
                            sZ[w][1][lm][4 * s4 + lk] = z[s4].im;
                        }
                        __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): the wave's LDS writes are done
                        __builtin_amdgcn_wave_barrier();
#pragma unroll
                        for (int s4 = 0; s4 < 4; ++s4) {
                            // A: Z^T, element (column lm, row 4 s + lk) of the block; B: x of that row
                            const double ar = sZ[w][0][4 * s4 + lk][lm], ai = sZ[w][1][4 * s4 + lk][lm];
                            const int r = rt * 16 + 4 * s4 + lk;
                            const cplx<double> xv = r < nr ? sxr[r * NB + rh] : mk(0.0, 0.0);
                            const double b = reCol ? xv.re : xv.im, bp = reCol ? -xv.im : xv.re;
                            cA = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, b, cA, 0, 0, 0);
                            cA = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bp, cA, 0, 0, 0);
                        }
                        __builtin_amdgcn_s_waitcnt(0xc07f);  // the reads are done before the next block's writes
                        __builtin_amdgcn_wave_barrier();
                    }
#pragma unroll
                    for (int s4 = 0; s4 < 4; ++s4) z[s4] = zn[s4];
                }
            }
            if (pairs)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int c = ct * 16 + 4 * i + lk;
                    if (c < nc && c >= I.nself) {
                        double* o = reinterpret_cast<double*>(part + (I.cpo + c - I.nself) * NB + rh) + (reCol ? 0 : 1);
                        *o = pass ? *o + cA[i] : cA[i];
                    }
                }
        }
        // the pass's row sums of the four waves, added in wave order
        for (int wv = 0; wv < TB / 64; ++wv) {
            if (w == wv)
#pragma unroll
                for (int j = 0; j < RT; ++j)
                    if (rt0 + j < rt1)
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            const int row = 16 * j + 4 * i + lk;  // within the pass
                            const double v = acc[j][i];
                            sRed[row * 16 + lm] = wv ? sRed[row * 16 + lm] + v : v;
                        }
            __syncthreads();
        }
        const int prow = min(RT * 16, nr - RT * 16 * pass);
        for (int e = t0; e < prow * 16; e += TB) {
            const int row = RT * 16 * pass + (e >> 4), n = e & 15;
            double* o = reinterpret_cast<double*>(part + (I.rpo + row) * NB + (n & 7));
            o[n < 8 ? 0 : 1] = sRed[e];
        }
        __syncthreads();  // sRed is reused by the next pass
    }
}

template <typename V>
void launchNearSymM8(bool conjVal, const NearItem* items, int nitems, const int32_t* colsIn, const int64_t* perm,
                     const cplx<V>* val, const cplx<double>* xt, cplx<double>* part, hipStream_t st) {
    if (nitems <= 0) return;
    if (conjVal)
        hipLaunchKernelGGL((kNearSymW<true, V>), dim3(nitems), dim3(TB), 0, st, items, colsIn, perm, val, xt, part);
    else
        hipLaunchKernelGGL((kNearSymW<false, V>), dim3(nitems), dim3(TB), 0, st, items, colsIn, perm, val, xt, part);
}
template void launchNearSymM8<double>(bool, const NearItem*, int, const int32_t*, const int64_t*,
                                      const cplx<double>*, const cplx<double>*, cplx<double>*, hipStream_t);

// xt[i 8 + k] = x[k xs + i]: eight right-hand sides, each unknown's eight
// values in one 128-byte line
__global__ void kInterleave8(const cplx<double>* __restrict__ x, int64_t xs, int64_t n, cplx<double>* __restrict__ xt) {
    const int64_t e = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (e >= 8 * n) return;
    xt[e] = x[(e & 7) * xs + (e >> 3)];
}
void launchInterleave8(const cplx<double>* x, int64_t xs, int64_t n, cplx<double>* xt, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(kInterleave8, dim3((unsigned)((8 * n + TB - 1) / TB)), dim3(TB), 0, st, x, xs, n, xt);
}

template void launchNearSym<double>(bool, const NearItem*, int, const int32_t*, const int64_t*, const cplx<double>*,
                                    const cplx<double>*, cplx<double>*, hipStream_t, int, int64_t, int64_t, int);
template void launchNearSym<float>(bool, const NearItem*, int, const int32_t*, const int64_t*, const cplx<float>*,
                                   const cplx<double>*, cplx<double>*, hipStream_t, int, int64_t, int64_t, int);

// The per-target M2L of one right-hand side and the near-field pairs in one
// launch (stages in order on one stream; CONJ: the adjoint's conjugated
// operators and values): both read only inputs of the matvec
// (multipoles; x) and write disjoint outputs (locals; partial sums), and the
// cache-bound M2L leaves HBM bandwidth that the pair stream uses.  Runs of 8
// blocks (one per XCD) alternate between the two in proportion to their
// counts, so both run from the first wave on (the pair runs spread over the
// first gW of the gT runs: the pair items are longer than the M2L blocks, so
// ending them earlier shortens the launch's tail); an M2L block keeps its
// index modulo 8, hence its XCD, and the LDS is one buffer sized for the
// larger.
template <typename R, typename V, bool CONJ>
__global__ __launch_bounds__(TB) void kM2LNear(M2LLevels<R> lv, int gN, int gW, const NearItem* __restrict__ items,
                                               int nitems, const int32_t* __restrict__ colsIn,
                                               const int64_t* __restrict__ perm, const cplx<V>* __restrict__ val,
                                               const cplx<double>* __restrict__ x, cplx<double>* __restrict__ part) {
    constexpr int EPL = sizeof(V) == 4 ? 2 : 1;
    constexpr int SZN = NT * EPL * (int)sizeof(cplx<V>) + (128 + TB) * (int)sizeof(cplx<double>);
    constexpr int SZM = 3 * 64 * 2 * (int)sizeof(cplx<R>);
    __shared__ __align__(16) unsigned char sm[SZN > SZM ? SZN : SZM];
    const int g = blockIdx.x >> 3, xx = blockIdx.x & 7;
    const int f0 = (int)min((int64_t)gN, (int64_t)g * gN / gW);
    const int f1 = (int)min((int64_t)gN, (int64_t)(g + 1) * gN / gW);
    if (f1 > f0) {
        const int item = 8 * f0 + xx;
        if (item >= nitems) return;
        cplx<V>* sT = reinterpret_cast<cplx<V>*>(sm);
        cplx<double>* sxr = reinterpret_cast<cplx<double>*>(sm + NT * EPL * sizeof(cplx<V>));
        nearSymBody<CONJ, V>(item, items, colsIn, perm, val, x, part, sT, sxr, sxr + 128);
        return;
    }
    m2lwBody<R, CONJ>(8 * (g - f0) + xx, lv, reinterpret_cast<cplx<R>(*)[64][2]>(sm));
}

template <typename R, typename V>
void launchM2LNear(M2LLevels<R> lv, const NearItem* items, int nitems, const int32_t* colsIn, const int64_t* perm,
                   const cplx<V>* val, const cplx<double>* x, cplx<double>* part, hipStream_t st, bool conj,
                   int windowPct) {
    int total = 0;
    for (int l = 0; l < lv.nlev; ++l) {
        lv.nch[l] = (lv.ndir[l] + 63) / 64;
        lv.per[l] = (lv.ntgt[l] * lv.nch[l] + 7) / 8;
        lv.blk[l] = total;
        total += 8 * lv.per[l];
    }
    const int gM = total / 8, gN = (nitems + 7) / 8, gT = gM + gN;
    if (gT == 0) return;
    // gW >= max(gN, 1): at most one pair run per run, and no division by zero
    const int gW = std::max({gN, 1, (int)((int64_t)gT * std::min(100, std::max(1, windowPct)) / 100)});
    if (conj)
        hipLaunchKernelGGL((kM2LNear<R, V, true>), dim3(8 * gT), dim3(TB), 0, st, lv, gN, gW, items, nitems, colsIn,
                           perm, val, x, part);
    else
        hipLaunchKernelGGL((kM2LNear<R, V, false>), dim3(8 * gT), dim3(TB), 0, st, lv, gN, gW, items, nitems, colsIn,
                           perm, val, x, part);
}
template void launchM2LNear<double, double>(M2LLevels<double>, const NearItem*, int, const int32_t*, const int64_t*,
                                            const cplx<double>*, const cplx<double>*, cplx<double>*, hipStream_t, bool, int);
template void launchM2LNear<float, float>(M2LLevels<float>, const NearItem*, int, const int32_t*, const int64_t*,
                                          const cplx<float>*, const cplx<double>*, cplx<double>*, hipStream_t, bool, int);
template void launchM2LNear<float, double>(M2LLevels<float>, const NearItem*, int, const int32_t*, const int64_t*,
                                           const cplx<double>*, const cplx<double>*, cplx<double>*, hipStream_t, bool, int);

// ------------------------------------------------------------------ preconditioner
// y = P x with P block diagonal over the leaves (the inverses of their self
// blocks): one block per (leaf, right-hand side); the leaf's x values are
// gathered into LDS, each lane forms rows (column-major P: lanes read
// consecutive rows of a column).  The adjoint reads P^H.
template <bool ADJ>
__global__ __launch_bounds__(TB) void kBlockPrec(const PrecBlock* __restrict__ blocks, const cplx<double>* __restrict__ P,
                                                 const int64_t* __restrict__ perm, const cplx<double>* __restrict__ x,
                                                 cplx<double>* __restrict__ y, int64_t xs) {
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<double>* sx = reinterpret_cast<cplx<double>*>(smem);
    const PrecBlock B = blocks[blockIdx.x];
    x += blockIdx.y * xs;
    y += blockIdx.y * xs;
    const int s = B.cnt;
    for (int i = threadIdx.x; i < s; i += TB) sx[i] = x[perm[B.sbeg + i]];
    __syncthreads();
    const cplx<double>* Pb = P + B.off;
    for (int r = threadIdx.x; r < s; r += TB) {
        cplx<double> a = mk(0.0, 0.0);
        if (!ADJ) {
            for (int c = 0; c < s; ++c) a = a + Pb[(int64_t)c * s + r] * sx[c];
        } else {
            for (int c = 0; c < s; ++c) {
                const cplx<double> v = Pb[(int64_t)r * s + c];
                a = a + mk(v.re, -v.im) * sx[c];
            }
        }
        y[perm[B.sbeg + r]] = a;
    }
}

void launchBlockPrec(const PrecBlock* blocks, int nblocks, int maxcnt, const cplx<double>* P, const int64_t* perm,
                     const cplx<double>* x, cplx<double>* y, int nb, int64_t xs, bool adjoint, hipStream_t st) {
    if (nblocks <= 0) return;
    const dim3 g((unsigned)nblocks, nb);
    const size_t lds = (size_t)std::max(1, maxcnt) * sizeof(cplx<double>);
    if (adjoint)
        hipLaunchKernelGGL(kBlockPrec<true>, g, dim3(TB), lds, st, blocks, P, perm, x, y, xs);
    else
        hipLaunchKernelGGL(kBlockPrec<false>, g, dim3(TB), lds, st, blocks, P, perm, x, y, xs);
}

__global__ void kGatherValues(const cplx<double>* __restrict__ val, const int64_t* __restrict__ idx, int64_t n,
                              cplx<double>* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i < n) out[i] = val[idx[i]];
}
void launchGatherValues(const cplx<double>* val, const int64_t* idx, int64_t n, cplx<double>* out, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(kGatherValues, dim3((unsigned)((n + TB - 1) / TB)), dim3(TB), 0, st, val, idx, n, out);
}

// FP32 copy of the near-field values (the mixed-precision inner solves)
__global__ void kToFloat(const cplx<double>* __restrict__ in, cplx<float>* __restrict__ out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i < n) out[i] = cplx<float>{(float)in[i].re, (float)in[i].im};
}
void launchToFloat(const cplx<double>* in, cplx<float>* out, int64_t n, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(kToFloat, dim3((unsigned)((n + TB - 1) / TB)), dim3(TB), 0, st, in, out, n);
}

void launchNearSum(const int64_t* rows, int64_t nrows, const int64_t* sumPtr, const int32_t* sumIdx,
                   const cplx<double>* part, const int64_t* perm, cplx<double>* y, int accumulate, hipStream_t st,
                   int nb, int64_t xs, int64_t ps) {
    if (nrows <= 0) return;
    const int64_t nbk = (nrows + TB / 64 - 1) / (TB / 64);
#define NSUM_CASE(B)                                                                                          \
    hipLaunchKernelGGL(kNearSum<B>, dim3((unsigned)nbk), dim3(TB), 0, st, rows, nrows, sumPtr, sumIdx, part, ps, \
                       perm, y, xs, accumulate)
    if (nb == 1) NSUM_CASE(1);
    else if (nb == 2) NSUM_CASE(2);
    else if (nb == 4) NSUM_CASE(4);
    else if (nb == 8)  // interleaved partials of the eight-RHS near field
        hipLaunchKernelGGL(kNearSum8, dim3((unsigned)((nrows + 4 * (TB / 64) - 1) / (4 * (TB / 64)))), dim3(TB), 0, st,
                           rows, nrows, sumPtr, sumIdx, part, perm, y, xs, accumulate);
    else throw Error(MLFMA_ERR_ARG, "right-hand-side batch must be 1, 2, 4 or 8");
#undef NSUM_CASE
}

// ------------------------------------------------------------------ L2T
// y_s (=|+=) near_s + C k sum_d conj(rp_s[d]) . L[d]  (src/fmm/leaf.cpp:257-280,
// with the near-field partial sums of the row folded in).  One wave per source
// row of a far-field leaf; the row's leaf gives its local field.  NB
// right-hand sides share every pattern read (fields at stride fs, y at column
// stride xs, partials at stride ps).
template <typename R, int NB>
__global__ __launch_bounds__(TB) void kL2Tx(const int64_t* __restrict__ rowPos, const int32_t* __restrict__ rowItem,
                                            int64_t nrows, const LeafItem* __restrict__ items,
                                            const cplx<R>* __restrict__ rp, const cplx<R>* __restrict__ Lc, int64_t fs,
                                            cplx<double> Ck, const int64_t* __restrict__ sumPtr,
                                            const int32_t* __restrict__ sumIdx, const cplx<double>* __restrict__ part,
                                            int64_t ps, const int64_t* __restrict__ perm, cplx<double>* __restrict__ y,
                                            int64_t xs, int accumulate, int per) {
    // rows of one XCD are contiguous, so a leaf's local field is read by one L2
    const int64_t w = (int64_t)xcdItem(blockIdx.x, per) * (TB / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (w >= nrows) return;  // uniform per wave
    const int64_t t = rowPos[w];
    const LeafItem it = items[rowItem[w]];
    const cplx<R>* r = rp + it.rpOff + (t - it.sbeg) * it.ndir * 2;
    const cplx<R>* Lb = Lc + it.fOff;
    R re[NB], im[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) re[k] = im[k] = 0;
    for (int d = lane; d < it.ndir; d += 64) {
        const cplx<R> r0 = r[2 * d], r1 = r[2 * d + 1];
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const cplx<R> v = conjmul(r0, Lb[k * fs + 2 * d]) + conjmul(r1, Lb[k * fs + 2 * d + 1]);
            re[k] += v.re;
            im[k] += v.im;
        }
    }
    const int64_t o = perm[t];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const R a = wsum(re[k]), b = wsum(im[k]);
        const cplx<double> nearv = NB == 8 ? nearFold(sumPtr, sumIdx, part + k, t, lane, 8)
                                           : nearFold(sumPtr, sumIdx, part + k * ps, t, lane);
        if (lane == 0) {
            const cplx<double> v = nearv + Ck * mk((double)a, (double)b);
            y[k * xs + o] = accumulate ? y[k * xs + o] + v : v;
        }
    }
}

// One right-hand side, 16 lanes per row: the wave's four rows stream their
// patterns (two complex per lane and load: 256 contiguous bytes per row and
// instruction in FP64, 128 in FP32, four loads in flight) against the leaf's
// local; each row's sum and its near-field partials are reduced over its 16
// lanes.  Two rows per wave with lanes over the directions (three rounds of
// 64 over the leaf level's 162 directions, the last one part-idle, and a
// 64-lane reduction per row) took 217 / 142 us (FP64 / FP32) at 201,720
// RWGs against 204 / 111 us; eight lanes per row 209 / 110 us, eight loads
// in flight 214 / 117 us.
template <typename R>
__global__ __launch_bounds__(TB) void kL2TS(const int64_t* __restrict__ rowPos, const int32_t* __restrict__ rowItem,
                                            int64_t nrows, const LeafItem* __restrict__ items,
                                            const cplx<R>* __restrict__ rp, const cplx<R>* __restrict__ Lc,
                                            cplx<double> Ck, const int64_t* __restrict__ sumPtr,
                                            const int32_t* __restrict__ sumIdx, const cplx<double>* __restrict__ part,
                                            const int64_t* __restrict__ perm, cplx<double>* __restrict__ y,
                                            int accumulate, int per) {
    const int lane = threadIdx.x & 63, g = lane >> 4, s = lane & 15;
    const int64_t w4 = 4 * ((int64_t)xcdItem(blockIdx.x, per) * (TB / 64) + (threadIdx.x >> 6));
    if (w4 >= nrows) return;  // uniform per wave
    const int64_t row = w4 + g;
    const bool ok = row < nrows;
    const int64_t rr = ok ? row : nrows - 1;
    const int64_t t = rowPos[rr];
    const LeafItem I = items[rowItem[rr]];
    const int ne = 2 * I.ndir;  // complex entries of the row (even)
    const cplx<R>* r = rp + I.rpOff + (t - I.sbeg) * ne;
    const cplx<R>* L = Lc + I.fOff;
    const int64_t p0 = sumPtr[t], p1 = sumPtr[t + 1];
    double nre = 0, nim = 0;
    for (int64_t q = p0 + s; q < p1; q += 16) {
        const cplx<double> v = part[sumIdx[q]];
        nre += v.re;
        nim += v.im;
    }
    R re = 0, im = 0;
    constexpr int U = 4;
    int e = 2 * s;
    for (; e + 32 * (U - 1) < ne; e += 32 * U) {
        cplx<R> a[U][2], m[U][2];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            a[u][0] = r[e + 32 * u];
            a[u][1] = r[e + 32 * u + 1];
            m[u][0] = L[e + 32 * u];
            m[u][1] = L[e + 32 * u + 1];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const cplx<R> v = conjmul(a[u][0], m[u][0]) + conjmul(a[u][1], m[u][1]);
            re += v.re;
            im += v.im;
        }
    }
    for (; e < ne; e += 32) {
        const cplx<R> v = conjmul(r[e], L[e]) + conjmul(r[e + 1], L[e + 1]);
        re += v.re;
        im += v.im;
    }
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) {
        re += __shfl_xor(re, o, 64);
        im += __shfl_xor(im, o, 64);
        nre += __shfl_xor(nre, o, 64);
        nim += __shfl_xor(nim, o, 64);
    }
    if (s == 0 && ok) {
        const int64_t o = perm[t];
        const cplx<double> v = mk(nre, nim) + Ck * mk((double)re, (double)im);
        y[o] = accumulate ? y[o] + v : v;
    }
}

// Eight right-hand sides on the matrix cores: per far-field leaf
// Y = conj(P) L, P the leaf's patterns (rows x 2 ndir complex, streamed from
// HBM once), L its locals of the eight right-hand sides (2 ndir x 8, staged
// in LDS as FP64).  With B = [Re L | Im L] and B' = [Im L | -Re L] (16 real
// columns), Re P . B + Im P . B' = [Re Y | Im Y] (v_mfma_f64_16x16x4f64, the
// two products in separate accumulators); row tiles of 16 go to the waves.
// y already holds the near-field sums (kNearSum8 over the same rows);
// y += C k Y.  Block = one leaf; LDS 2 ndir x 8 x 16 bytes.
template <typename R>
__global__ __launch_bounds__(TB) void kL2TM(const LeafItem* __restrict__ items, int nitems, int per,
                                            const cplx<R>* __restrict__ rp, const cplx<R>* __restrict__ Lc, int64_t fs,
                                            cplx<double> Ck, const int64_t* __restrict__ perm,
                                            cplx<double>* __restrict__ y, int64_t xs) {
    using d4 = double __attribute__((ext_vector_type(4)));
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<double>* sL = reinterpret_cast<cplx<double>*>(smem);  // [K][8]
    const int item = xcdItem(blockIdx.x, per);
    if (item >= nitems) return;
    const LeafItem it = items[item];
    const int K = 2 * it.ndir;
    for (int e = threadIdx.x; e < 8 * K; e += TB) {
        const int r = e / K, kk = e - r * K;
        const cplx<R> v = Lc[r * fs + it.fOff + kk];
        sL[kk * 8 + r] = mk((double)v.re, (double)v.im);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lm = lane & 15, lk = lane >> 4;
    const int rh = lm & 7;
    const bool reCol = lm < 8;
    const int nrt = (it.cnt + 15) >> 4;
    for (int rt = w; rt < nrt; rt += 4) {  // uniform per wave
        const int row = 16 * rt + lm;
        const bool rowOk = row < it.cnt;
        const cplx<R>* pr = rp + it.rpOff + (int64_t)(rowOk ? row : 0) * K;
        d4 P = d4{0, 0, 0, 0}, Q = d4{0, 0, 0, 0};
        // software pipeline: the pattern loads of the next U steps are issued
        // before the products of these U (indices clamped, so every load is
        // valid; entries past K or the leaf's rows count as zero)
        constexpr int U = 8;
        cplx<R> abuf[U];
#pragma unroll
        for (int u = 0; u < U; ++u) abuf[u] = pr[min(4 * u + lk, K - 1)];
        for (int k0 = 0; k0 < K; k0 += 4 * U) {
            cplx<R> nbuf[U];
#pragma unroll
            for (int u = 0; u < U; ++u) nbuf[u] = pr[min(k0 + 4 * U + 4 * u + lk, K - 1)];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int kk = k0 + 4 * u + lk;
                const bool ok = kk < K;
                const cplx<double> l = sL[min(kk, K - 1) * 8 + rh];
                const double ar = (ok && rowOk) ? (double)abuf[u].re : 0.0;
                const double ai = (ok && rowOk) ? (double)abuf[u].im : 0.0;
                const double b = reCol ? l.re : l.im;
                const double bp = reCol ? l.im : -l.re;
                P = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, b, P, 0, 0, 0);
                Q = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bp, Q, 0, 0, 0);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) abuf[u] = nbuf[u];
        }
        // element i: row 16 rt + 4 i + lk, column lm (Re of RHS lm for
        // lm < 8, Im of RHS lm - 8 otherwise); lane l ^ 8 holds the other part
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const double v = P[i] + Q[i];
            const double u = __shfl_xor(v, 8, 64);
            const int r = 16 * rt + 4 * i + lk;
            if (reCol && r < it.cnt) {
                cplx<double>* o = y + rh * xs + perm[it.sbeg + r];
                *o = *o + Ck * mk(v, u);
            }
        }
    }
}

template <typename R>
bool launchL2TM8(const LeafItem* items, int nitems, int maxndir, const cplx<R>* rp, const cplx<R>* Lc, int64_t fs,
                 cplx<double> Ck, const int64_t* rowPos, int64_t nrows, const int64_t* sumPtr, const int32_t* sumIdx,
                 const cplx<double>* part, const int64_t* perm, cplx<double>* y, int64_t xs, int accumulate,
                 size_t ldsMax, hipStream_t st) {
    const size_t lds = (size_t)2 * maxndir * 8 * sizeof(cplx<double>);
    if (lds > ldsMax) return false;
    launchNearSum(rowPos, nrows, sumPtr, sumIdx, part, perm, y, accumulate, st, 8, xs, 0);
    if (nitems <= 0) return true;
    const int per = (nitems + 7) / 8;
    hipLaunchKernelGGL((kL2TM<R>), dim3(8 * per), dim3(TB), lds, st, items, nitems, per, rp, Lc, fs, Ck, perm, y, xs);
    return true;
}
template bool launchL2TM8<double>(const LeafItem*, int, int, const cplx<double>*, const cplx<double>*, int64_t,
                                  cplx<double>, const int64_t*, int64_t, const int64_t*, const int32_t*,
                                  const cplx<double>*, const int64_t*, cplx<double>*, int64_t, int, size_t,
                                  hipStream_t);
template bool launchL2TM8<float>(const LeafItem*, int, int, const cplx<float>*, const cplx<float>*, int64_t,
                                 cplx<double>, const int64_t*, int64_t, const int64_t*, const int32_t*,
                                 const cplx<double>*, const int64_t*, cplx<double>*, int64_t, int, size_t,
                                 hipStream_t);

template <typename R>
void launchL2Tx(const int64_t* rowPos, const int32_t* rowItem, int64_t nrows, const LeafItem* items, const cplx<R>* rp,
                const cplx<R>* Lc, cplx<double> Ck, const int64_t* sumPtr, const int32_t* sumIdx,
                const cplx<double>* part, const int64_t* perm, cplx<double>* y, int accumulate, hipStream_t st, int nb,
                int64_t fs, int64_t xs, int64_t ps) {
    if (nrows <= 0) return;
    if (nb == 1) {
        const int64_t nbk = (nrows + 4 * (TB / 64) - 1) / (4 * (TB / 64));
        if (nbk > INT32_MAX / 2) throw Error(MLFMA_ERR_ARG, "too many far-field rows");
        const int per = (int)((nbk + 7) / 8);
        hipLaunchKernelGGL((kL2TS<R>), dim3((unsigned)(8 * per)), dim3(TB), 0, st, rowPos, rowItem, nrows, items,
                           rp, Lc, Ck, sumPtr, sumIdx, part, perm, y, accumulate, per);
        return;
    }
    const int64_t nbk = (nrows + TB / 64 - 1) / (TB / 64);
    if (nbk > INT32_MAX / 2) throw Error(MLFMA_ERR_ARG, "too many far-field rows");
    const int per = (int)((nbk + 7) / 8);
#define L2T_CASE(B)                                                                                                   \
    hipLaunchKernelGGL((kL2Tx<R, B>), dim3((unsigned)(8 * per)), dim3(TB), 0, st, rowPos, rowItem, nrows, items, rp, \
                       Lc, fs, Ck, sumPtr, sumIdx, part, ps, perm, y, xs, accumulate, per)
    if (nb == 1) L2T_CASE(1);
    else if (nb == 2) L2T_CASE(2);
    else if (nb == 4) L2T_CASE(4);
    else if (nb == 8) L2T_CASE(8);
    else throw Error(MLFMA_ERR_ARG, "right-hand-side batch must be 1, 2, 4 or 8");
#undef L2T_CASE
}

// Direction-weighted disaggregation for the position gradient:
// out[i][a] = sum_d khat_d[a] conj(rp_s[d]) . F[d] for every source s of a
// far-field leaf (i = perm[s], input order), F the leaf's field; one wave per
// source row.  With F the forward locals L(I) this is the row derivative of
// L2T, with F the adjoint multipoles M'(lambda) the column derivative of S2M
// (a dipole pattern's position derivative is -ik khat times the pattern).
__global__ __launch_bounds__(TB) void kL2TGrad(const int64_t* __restrict__ rowPos, const int32_t* __restrict__ rowItem,
                                               int64_t nrows, const LeafItem* __restrict__ items,
                                               const LevelDev* __restrict__ lv, const cplx<double>* __restrict__ rp,
                                               const cplx<double>* __restrict__ F, const int64_t* __restrict__ perm,
                                               cplx<double>* __restrict__ out) {
    const int64_t w = (int64_t)blockIdx.x * (TB / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (w >= nrows) return;
    const int64_t t = rowPos[w];
    const LeafItem it = items[rowItem[w]];
    const double* kh = lv[it.level].khat;
    const cplx<double>* r = rp + it.rpOff + (t - it.sbeg) * it.ndir * 2;
    const cplx<double>* Fb = F + it.fOff;
    double re[3] = {0, 0, 0}, im[3] = {0, 0, 0};
    for (int d = lane; d < it.ndir; d += 64) {
        const cplx<double> v = conjmul(r[2 * d], Fb[2 * d]) + conjmul(r[2 * d + 1], Fb[2 * d + 1]);
        for (int a = 0; a < 3; ++a) {
            re[a] += kh[3 * d + a] * v.re;
            im[a] += kh[3 * d + a] * v.im;
        }
    }
    for (int a = 0; a < 3; ++a) {
        const double x = wsum(re[a]), y = wsum(im[a]);
        if (lane == 0) out[3 * perm[t] + a] = mk(x, y);
    }
}

void launchL2TGrad(const int64_t* rowPos, const int32_t* rowItem, int64_t nrows, const LeafItem* items,
                   const LevelDev* lv, const cplx<double>* rp, const cplx<double>* F, const int64_t* perm,
                   cplx<double>* out, hipStream_t st) {
    if (nrows <= 0) return;
    const int64_t nbk = (nrows + TB / 64 - 1) / (TB / 64);
    hipLaunchKernelGGL(kL2TGrad, dim3((unsigned)nbk), dim3(TB), 0, st, rowPos, rowItem, nrows, items, lv, rp, F, perm,
                       out);
}

#define INST(R)                                                                                                     \
    template void launchS2Mx<R>(const LeafItem*, const int32_t*, int, int, const cplx<R>*, const cplx<double>*,    \
                                const int64_t*, cplx<R>*, hipStream_t, int, int64_t, int64_t);                     \
    template bool launchM2MFused<R>(int, const int32_t*, const int32_t*, const int32_t*, const int32_t*, int, int, int, \
                                    int, int, const int32_t*, const double*, const int32_t*, const double*,         \
                                    const cplx<R>*, const cplx<R>*, cplx<R>*, size_t, hipStream_t, int, int64_t,    \
                                    int);                                                                      \
    template void launchM2LGroup<R>(M2LGroups<R>, bool, hipStream_t, int);                                          \
    template void launchM2Lx<R>(M2LLevels<R>, bool, hipStream_t, int);                                              \
    template void launchM2LT<R>(M2LTiles<R>, bool, hipStream_t, int);                                               \
    template bool launchL2LFused<R>(bool, int, const int32_t*, const int32_t*, const int32_t*, int, int, int, int,   \
                                    const int32_t*, const int32_t*, const double*, int, const int32_t*,             \
                                    const int32_t*, const double*, const int32_t*, int, const cplx<R>*,             \
                                    const cplx<R>*, cplx<R>*, size_t, hipStream_t, int, int64_t);                   \
    template void launchL2Tx<R>(const int64_t*, const int32_t*, int64_t, const LeafItem*, const cplx<R>*,           \
                                const cplx<R>*, cplx<double>, const int64_t*, const int32_t*, const cplx<double>*,  \
                                const int64_t*, cplx<double>*, int, hipStream_t, int, int64_t, int64_t, int64_t);   \
    template bool launchL2LAdj<R>(int, const int32_t*, const int32_t*, const int32_t*, const int32_t*, int, int, int, \
                                  int, int, const int32_t*, const double*, const int32_t*, const double*,           \
                                  const cplx<R>*, const cplx<R>*, cplx<R>*, size_t, hipStream_t, int);              \
    template void launchZeroField<R>(cplx<R>*, int64_t, hipStream_t);                                             \
    template int m2mRows<R>(int, int, int, int, int, int, size_t, int, size_t*);                                   \
    template void launchTreeUp<R>(const TreeUp<R>&, int, size_t, int, hipStream_t);                                \
    template int treeUpOccupancy<R>(int, size_t);
INST(double)
INST(float)
#undef INST

}  // namespace mlfma
