// solver.hip — GMRES on the device (Arnoldi with wave-reduction dot products),
// far field and the dense direct matvec.  Reference: src/solver.cpp:5-146.
#include <hip/hip_runtime.h>

#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "dense.hpp"
#include "kernels_mv.hpp"
#include "plan.hpp"

namespace mlfma {

constexpr int KB = 256;

__device__ __forceinline__ double waveSum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block-wide sum of a complex pair; result valid in thread 0.
__device__ __forceinline__ void blockSum2(double& a, double& b) {
    __shared__ double sa[KB / 64], sb[KB / 64];
    a = waveSum(a);
    b = waveSum(b);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
        sa[w] = a;
        sb[w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = 0;
        b = 0;
        for (int q = 0; q < KB / 64; ++q) {
            a += sa[q];
            b += sb[q];
        }
    }
}

// partial[i * nb + block] = sum over the block's rows of conj(Q_i) w; grid
// (row blocks, columns): one column per block, so the ncol reductions run in
// parallel instead of one after another
__global__ void kDotPartial(const cplx<double>* __restrict__ Q, int64_t ld, int ncol, const cplx<double>* __restrict__ w,
                            int64_t n, cplx<double>* __restrict__ partial) {
    const int64_t r = (int64_t)blockIdx.x * KB + threadIdx.x;
    const int nb = gridDim.x;
    const int i = blockIdx.y;
    double re = 0, im = 0;
    if (r < n) {
        const cplx<double> wv = w[r];
        const cplx<double> q = Q[(int64_t)i * ld + r];
        re = q.re * wv.re + q.im * wv.im;
        im = q.re * wv.im - q.im * wv.re;
    }
    blockSum2(re, im);
    if (threadIdx.x == 0) partial[(int64_t)i * nb + blockIdx.x] = cplx<double>{re, im};
}

// h[i] (+)= sum_b partial[i * nb + b]; one block per column
__global__ void kReduce(const cplx<double>* __restrict__ partial, int nb, cplx<double>* __restrict__ h, int add) {
    const int i = blockIdx.x;
    double re = 0, im = 0;
    for (int b = threadIdx.x; b < nb; b += KB) {
        re += partial[(int64_t)i * nb + b].re;
        im += partial[(int64_t)i * nb + b].im;
    }
    blockSum2(re, im);
    if (threadIdx.x == 0) {
        if (add) {
            h[i].re += re;
            h[i].im += im;
        } else
            h[i] = cplx<double>{re, im};
    }
}

// w -= Q h (ncol columns); partial[block] = sum |w|^2 of the block when norm != 0
__global__ void kUpdate(const cplx<double>* __restrict__ Q, int64_t ld, int ncol, const cplx<double>* __restrict__ h,
                        cplx<double>* __restrict__ w, int64_t n, cplx<double>* __restrict__ partial, int norm) {
    const int64_t r = (int64_t)blockIdx.x * KB + threadIdx.x;
    double s2 = 0;
    if (r < n) {
        cplx<double> a = w[r];
        for (int i = 0; i < ncol; ++i) {
            const cplx<double> q = Q[(int64_t)i * ld + r];
            const cplx<double> c = h[i];
            a.re -= q.re * c.re - q.im * c.im;
            a.im -= q.re * c.im + q.im * c.re;
        }
        w[r] = a;
        s2 = a.re * a.re + a.im * a.im;
    }
    if (norm) {
        double z = 0;
        blockSum2(s2, z);
        if (threadIdx.x == 0) partial[blockIdx.x] = cplx<double>{s2, 0.0};
    }
}

// The same three kernels for several lanes at once (grid z or y = lane): a
// lane is a vector segment of laneLen[k] values at laneBase[k] in every
// column (a right-hand side, or one scatterer's unknowns of a right-hand side
// in a forest plan); basis column i at Q + i ld, w at W, partials at
// partial + k ps, coefficients at h + k hs.  Blocks past a lane's length
// contribute zero.
// Dot products of the basis columns with w for several lanes at once: grid
// (row blocks, lanes, column groups of cg); a block reads its RB x KB rows
// of w once (RB rows per thread, KB apart) and then those rows of each of
// its columns, each column's sum reduced over the wave and then over the
// block's waves in LDS, in a fixed order: partial[k ps + i nbd + block]
// (nbd = gridDim.x).  The basis may be FP32 (the Krylov vectors of the
// solves with the FP32 far field); products and sums are FP64.  Dynamic
// LDS: cg (KB / 64) complex.
//   wnorm: the blocks of column group 0 also write |w|^2 of their rows as the
//          partial of column ncol (the norm before the pass);
//   reorth: per lane, 0 when the second pass is skipped: the block writes
//          zero partials without reading the basis.
constexpr int RB = 4;
constexpr int kDotCols = 8;  // columns per block of a one-vector solve
template <typename TQ>
__global__ __launch_bounds__(KB) void kDotManyB(const cplx<TQ>* __restrict__ Q, int64_t ld, int ncol,
                                                const cplx<double>* __restrict__ W,
                                                const int64_t* __restrict__ laneBase,
                                                const int64_t* __restrict__ laneLen, cplx<double>* __restrict__ partial,
                                                int64_t ps, int cg, int wnorm, const int* __restrict__ reorth) {
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<double>* sh = reinterpret_cast<cplx<double>*>(smem);  // [cg][KB / 64]
    constexpr int NW = KB / 64;
    const int nbd = gridDim.x, k = blockIdx.y;
    const int c0 = blockIdx.z * cg, c1 = min(ncol, c0 + cg);
    if (reorth && !reorth[k]) {  // uniform over the block
        for (int i = c0 + (int)threadIdx.x; i < c1; i += KB)
            partial[(int64_t)k * ps + (int64_t)i * nbd + blockIdx.x] = cplx<double>{0.0, 0.0};
        return;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t len = laneLen[k], base = laneBase[k];
    cplx<double> w[RB];
    int64_t off[RB];
    bool ok[RB];
#pragma unroll
    for (int t = 0; t < RB; ++t) {
        const int64_t r = ((int64_t)blockIdx.x * RB + t) * KB + threadIdx.x;
        ok[t] = r < len;
        off[t] = base + (ok[t] ? r : 0);
        w[t] = ok[t] ? W[off[t]] : cplx<double>{0.0, 0.0};
    }
#pragma unroll 2
    for (int i = c0; i < c1; ++i) {
        const cplx<TQ>* q = Q + (int64_t)i * ld;
        double re = 0, im = 0;
#pragma unroll
        for (int t = 0; t < RB; ++t) {
            const cplx<TQ> v = ok[t] ? q[off[t]] : cplx<TQ>{0, 0};
            const double vr = (double)v.re, vi = (double)v.im;
            re += vr * w[t].re + vi * w[t].im;
            im += vr * w[t].im - vi * w[t].re;
        }
        re = waveSum(re);
        im = waveSum(im);
        if (lane == 0) sh[(i - c0) * NW + wv] = cplx<double>{re, im};
    }
    __shared__ double shn[NW];
    const bool norm = wnorm && blockIdx.z == 0;  // uniform over the block
    if (norm) {
        double s2 = 0;
#pragma unroll
        for (int t = 0; t < RB; ++t) s2 += w[t].re * w[t].re + w[t].im * w[t].im;
        s2 = waveSum(s2);
        if (lane == 0) shn[wv] = s2;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < c1 - c0; i += KB) {
        double re = 0, im = 0;
        for (int v = 0; v < NW; ++v) {
            re += sh[i * NW + v].re;
            im += sh[i * NW + v].im;
        }
        partial[(int64_t)k * ps + (int64_t)(c0 + i) * nbd + blockIdx.x] = cplx<double>{re, im};
    }
    if (norm && threadIdx.x == 0) {
        double s2 = 0;
        for (int v = 0; v < NW; ++v) s2 += shn[v];
        partial[(int64_t)k * ps + (int64_t)ncol * nbd + blockIdx.x] = cplx<double>{s2, 0.0};
    }
}

// Reorthogonalization test of classical Gram-Schmidt (Daniel, Gragg,
// Kaufman and Stewart 1976): the second pass runs only when the first one
// cancelled more than half of |w|^2, |w_after|^2 < |w_before|^2 / 2.  One
// block per lane: |w_before|^2 from nb0 partials at before + k ps0 (stride
// sb between partials), |w_after|^2 from nb1 doubles or complex real parts
// at after + k ps1 (stride sa); reorth[k] = test, norm1[k] = |w_after|^2.
__global__ void kReorthTest(const double* __restrict__ before, int64_t ps0, int nb0, int sb,
                            const double* __restrict__ after, int64_t ps1, int nb1, int sa, int* __restrict__ reorth,
                            double* __restrict__ norm1) {
    const int k = blockIdx.x;
    double t0 = 0, t1 = 0;
    for (int b = threadIdx.x; b < nb0; b += KB) t0 += before[(int64_t)k * ps0 + (int64_t)b * sb];
    for (int b = threadIdx.x; b < nb1; b += KB) t1 += after[(int64_t)k * ps1 + (int64_t)b * sa];
    blockSum2(t0, t1);
    if (threadIdx.x == 0) {
        reorth[k] = t1 < 0.5 * t0 ? 1 : 0;
        norm1[k] = t1;
    }
}

__global__ void kReduceB(const cplx<double>* __restrict__ partial, int64_t ps, int nb, cplx<double>* __restrict__ h,
                         int64_t hs) {
    const int i = blockIdx.x, k = blockIdx.y;
    double re = 0, im = 0;
    for (int b = threadIdx.x; b < nb; b += KB) {
        re += partial[(int64_t)k * ps + (int64_t)i * nb + b].re;
        im += partial[(int64_t)k * ps + (int64_t)i * nb + b].im;
    }
    blockSum2(re, im);
    if (threadIdx.x == 0) h[(int64_t)k * hs + i] = cplx<double>{re, im};
}

// reorth: per lane, 0 when the pass is skipped (w stays; its norm partials
// are still written)
template <typename TQ>
__global__ void kUpdateB(const cplx<TQ>* __restrict__ Q, int64_t ld, int ncol, const cplx<double>* __restrict__ h,
                         int64_t hs, cplx<double>* __restrict__ W, const int64_t* __restrict__ laneBase,
                         const int64_t* __restrict__ laneLen, cplx<double>* __restrict__ partial, int64_t ps,
                         int norm, const int* __restrict__ reorth) {
    const int64_t r = (int64_t)blockIdx.x * KB + threadIdx.x;
    const int k = blockIdx.y;
    const bool skip = reorth && !reorth[k];
    double s2 = 0;
    if (r < laneLen[k] && skip) {
        const cplx<double> a = W[laneBase[k] + r];
        s2 = a.re * a.re + a.im * a.im;
    } else if (r < laneLen[k]) {
        const int64_t o = laneBase[k] + r;
        cplx<double> a = W[o];
        for (int i = 0; i < ncol; ++i) {
            const cplx<TQ> q = Q[(int64_t)i * ld + o];
            const double qr = (double)q.re, qi = (double)q.im;
            const cplx<double> c = h[(int64_t)k * hs + i];
            a.re -= qr * c.re - qi * c.im;
            a.im -= qr * c.im + qi * c.re;
        }
        W[o] = a;
        s2 = a.re * a.re + a.im * a.im;
    }
    if (norm) {
        double z = 0;
        blockSum2(s2, z);
        if (threadIdx.x == 0) partial[(int64_t)k * ps + blockIdx.x] = cplx<double>{s2, 0.0};
    }
}

// out = in / s for FP32 or FP64 out; widen: out = in (FP32 to FP64)
template <typename TQ>
__global__ void kScaleTo(const cplx<double>* __restrict__ in, double s, cplx<TQ>* __restrict__ out, int64_t n) {
    const int64_t r = (int64_t)blockIdx.x * KB + threadIdx.x;
    if (r < n) out[r] = cplx<TQ>{(TQ)(in[r].re / s), (TQ)(in[r].im / s)};
}
__global__ void kWiden(const cplx<float>* __restrict__ in, cplx<double>* __restrict__ out, int64_t n) {
    const int64_t r = (int64_t)blockIdx.x * KB + threadIdx.x;
    if (r < n) out[r] = cplx<double>{(double)in[r].re, (double)in[r].im};
}

// out = in / s (s real, on the host)
__global__ void kScale(const cplx<double>* __restrict__ in, double s, cplx<double>* __restrict__ out, int64_t n) {
    const int64_t r = (int64_t)blockIdx.x * KB + threadIdx.x;
    if (r < n) out[r] = cplx<double>{in[r].re / s, in[r].im / s};
}

// x = (accumulate ? x : 0) + Q y
template <typename TQ>
__global__ void kCombine(const cplx<TQ>* __restrict__ Q, int64_t ld, int ncol, const cplx<double>* __restrict__ y,
                         cplx<double>* __restrict__ x, int64_t n, int accumulate) {
    const int64_t r = (int64_t)blockIdx.x * KB + threadIdx.x;
    if (r >= n) return;
    cplx<double> a = accumulate ? x[r] : cplx<double>{0, 0};
    for (int i = 0; i < ncol; ++i) {
        const cplx<TQ> q = Q[(int64_t)i * ld + r];
        const double qr = (double)q.re, qi = (double)q.im;
        const cplx<double> c = y[i];
        a.re += qr * c.re - qi * c.im;
        a.im += qr * c.im + qi * c.re;
    }
    x[r] = a;
}

// r = b - a
__global__ void kSub(const cplx<double>* __restrict__ b, const cplx<double>* __restrict__ a, cplx<double>* __restrict__ r,
                     int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * KB + threadIdx.x;
    if (i < n) r[i] = cplx<double>{b[i].re - a[i].re, b[i].im - a[i].im};
}

// ---- CGS for the pipelined GMRES(m): the dot products of every column in
// one launch (kDotManyB over a single lane) and an update kernel that
// reduces the dot partials itself (one wave per coefficient, fixed order),
// so a CGS pass is two launches
// h_i = sum_b partial[i nbd + b] for i < ncol (block 0 also stores them to
// hout), then w -= Q h; normPart[block] = sum |w|^2 of the block's rows when
// norm != 0.  Dynamic LDS: ncol complex.
template <typename TQ>
__global__ void kUpdateR(const cplx<TQ>* __restrict__ Q, int64_t ld, int ncol,
                         const cplx<double>* __restrict__ partial, int nbd, cplx<double>* __restrict__ hout,
                         cplx<double>* __restrict__ w, int64_t n, double* __restrict__ normPart, int norm) {
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<double>* sh = reinterpret_cast<cplx<double>*>(smem);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int i = wv; i < ncol; i += KB / 64) {
        double re = 0, im = 0;
        for (int b = lane; b < nbd; b += 64) {
            re += partial[(int64_t)i * nbd + b].re;
            im += partial[(int64_t)i * nbd + b].im;
        }
        re = waveSum(re);
        im = waveSum(im);
        if (lane == 0) {
            sh[i] = cplx<double>{re, im};
            if (blockIdx.x == 0) hout[i] = cplx<double>{re, im};
        }
    }
    __syncthreads();
    const int64_t r = (int64_t)blockIdx.x * KB + threadIdx.x;
    double s2 = 0;
    if (r < n) {
        cplx<double> a = w[r];
        for (int i = 0; i < ncol; ++i) {
            const cplx<TQ> q = Q[(int64_t)i * ld + r];
            const double qr = (double)q.re, qi = (double)q.im;
            const cplx<double> c = sh[i];
            a.re -= qr * c.re - qi * c.im;
            a.im -= qr * c.im + qi * c.re;
        }
        w[r] = a;
        s2 = a.re * a.re + a.im * a.im;
    }
    if (norm) {
        double z = 0;
        blockSum2(s2, z);
        if (threadIdx.x == 0) normPart[blockIdx.x] = s2;
    }
}

// ---- GMRES(m) step on the device (no host round trip per iteration)
__device__ __forceinline__ cplx<double> cmul(cplx<double> a, cplx<double> b) {
    return cplx<double>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cplx<double> cconj(cplx<double> a) { return cplx<double>{a.re, -a.im}; }

// Column j of the Hessenberg matrix from the two CGS passes (hbuf: pass 1 at
// [0, ncol), pass 2 at [ncol + 1, 2 ncol + 1); |w|^2 summed from the nbn
// per-block partials of the second pass), rotated by the
// previous Givens rotations, the new rotation, and the rotated right-hand
// side g; stat[3 j .. 3 j + 2] = (|g_{j+1}| / bn, |w|, rho), also mirrored to
// the host-mapped statHost.  One block; the rotations run on thread 0.
// The Givens step of GMRES(m), called by every thread of one block with s2 = |w|^2
// (valid in thread 0): the block stages the new Hessenberg column (the two
// CGS passes added) and the previous rotations in LDS, then thread 0 applies
// the rotations from LDS (a chain of dependent global loads otherwise) and
// writes the column, the new rotation and the rotated right-hand side back.
// pass2: the second pass's coefficients when the caller holds them in LDS
// (at sh, so each is replaced in place by the column entry), else null (at
// hbuf + ncol + 1).
__device__ void givensStep(const cplx<double>* __restrict__ hbuf, double s2, int j, double bn,
                           cplx<double>* __restrict__ cs, cplx<double>* __restrict__ sn, cplx<double>* __restrict__ g,
                           cplx<double>* __restrict__ H, int ldH, double* __restrict__ stat,
                           double* __restrict__ statHost, cplx<double>* sh, const cplx<double>* pass2 = nullptr) {
    // sh: 3 (j + 2) complex of LDS: column, cs, sn
    const int ncol = j + 1;
    cplx<double>* hc = sh;
    cplx<double>* scs = sh + ncol + 1;
    cplx<double>* ssn = scs + ncol;
    for (int i = threadIdx.x; i < ncol; i += blockDim.x) {
        const cplx<double> b = pass2 ? pass2[i] : hbuf[ncol + 1 + i];
        hc[i] = cplx<double>{hbuf[i].re + b.re, hbuf[i].im + b.im};
        if (i < j) {
            scs[i] = cs[i];
            ssn[i] = sn[i];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double hn = sqrt(s2);
        hc[ncol] = cplx<double>{hn, 0.0};
        cplx<double> hi = hc[0];
        for (int i = 0; i < j; ++i) {
            const cplx<double> hi1 = hc[i + 1], c = scs[i], sv = ssn[i];
            const cplx<double> a = cmul(cconj(c), hi), b = cmul(cconj(sv), hi1);
            const cplx<double> cc = cmul(sv, hi), e = cmul(c, hi1);
            hc[i] = cplx<double>{a.re + b.re, a.im + b.im};
            hi = cplx<double>{-cc.re + e.re, -cc.im + e.im};
        }
        hc[j] = hi;
        const double rho = sqrt(hc[j].re * hc[j].re + hc[j].im * hc[j].im + hc[ncol].re * hc[ncol].re +
                                hc[ncol].im * hc[ncol].im);
        cs[j] = cplx<double>{hc[j].re / rho, hc[j].im / rho};
        sn[j] = cplx<double>{hc[ncol].re / rho, hc[ncol].im / rho};
        hc[j] = cplx<double>{rho, 0.0};
        hc[ncol] = cplx<double>{0.0, 0.0};
        const cplx<double> gj = g[j];
        const cplx<double> t = cmul(sn[j], gj);
        g[j + 1] = cplx<double>{-t.re, -t.im};
        g[j] = cmul(cconj(cs[j]), gj);
        const double est = hypot(g[j + 1].re, g[j + 1].im) / bn;
        stat[3 * j] = est;
        stat[3 * j + 1] = hn;
        stat[3 * j + 2] = rho;
        statHost[3 * j] = est;
        statHost[3 * j + 1] = hn;
        statHost[3 * j + 2] = rho;
    }
    __syncthreads();
    cplx<double>* h = H + (int64_t)j * ldH;
    for (int i = threadIdx.x; i <= ncol; i += blockDim.x) h[i] = hc[i];
}

// The second CGS pass (kUpdateR with the norm) and the Givens step in one
// launch: every block publishes its |w|^2 partial and counts itself in; the
// block that counts last sums the partials in the same order every time and
// runs the step with the pass-2 coefficients it reduced into its own LDS
// (every block forms the same ones).  No block waits for another.  The
// counter is back at zero when the kernel ends.  The partials are the only
// data handed between blocks: each goes to memory as an agent-scope store
// that completes (s_waitcnt vmcnt(0)) before its block's agent-scope count,
// and the last block reads them with agent-scope loads, so no block needs a
// device fence (which writes back or invalidates its XCD's whole L2).
struct GivensArgs {
    const cplx<double>* hbuf;
    int j;
    double bn;
    cplx<double>*cs, *sn, *g, *H;
    int ldH;
    double *stat, *statHost;
    unsigned int* counter;
};

template <typename TQ>
__global__ void kUpdateRGivens(const cplx<TQ>* __restrict__ Q, int64_t ld, int ncol,
                               const cplx<double>* __restrict__ partial, int nbd, cplx<double>* __restrict__ hout,
                               cplx<double>* __restrict__ w, int64_t n, double* __restrict__ normPart, GivensArgs ga,
                               const int* __restrict__ reorth, const double* __restrict__ norm1) {
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<double>* sh = reinterpret_cast<cplx<double>*>(smem);
    if (!*reorth) {  // uniform: the first pass was enough; w and its norm stand
        if (blockIdx.x != 0) return;
        for (int i = threadIdx.x; i < ncol; i += KB) {
            sh[i] = cplx<double>{0.0, 0.0};
            hout[i] = sh[i];
        }
        __syncthreads();
        givensStep(ga.hbuf, *norm1, ga.j, ga.bn, ga.cs, ga.sn, ga.g, ga.H, ga.ldH, ga.stat, ga.statHost, sh, sh);
        return;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int i = wv; i < ncol; i += KB / 64) {
        double re = 0, im = 0;
        for (int b = lane; b < nbd; b += 64) {
            re += partial[(int64_t)i * nbd + b].re;
            im += partial[(int64_t)i * nbd + b].im;
        }
        re = waveSum(re);
        im = waveSum(im);
        if (lane == 0) sh[i] = cplx<double>{re, im};
    }
    __syncthreads();
    const int64_t r = (int64_t)blockIdx.x * KB + threadIdx.x;
    double s2 = 0;
    if (r < n) {
        cplx<double> a = w[r];
        for (int i = 0; i < ncol; ++i) {
            const cplx<TQ> q = Q[(int64_t)i * ld + r];
            const double qr = (double)q.re, qi = (double)q.im;
            const cplx<double> c = sh[i];
            a.re -= qr * c.re - qi * c.im;
            a.im -= qr * c.im + qi * c.re;
        }
        w[r] = a;
        s2 = a.re * a.re + a.im * a.im;
    }
    double z = 0;
    blockSum2(s2, z);
    // thread 0 alone writes what other blocks read (the partial) and counts
    // the block in once the store has completed
    __shared__ unsigned int rank;
    if (threadIdx.x == 0) {
        if (blockIdx.x == 0)
            for (int i = 0; i < ncol; ++i) hout[i] = sh[i];
        __hip_atomic_store(normPart + blockIdx.x, s2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_s_waitcnt(0 | (0x7 << 4) | (0xf << 8));  // vmcnt(0)
        rank = __hip_atomic_fetch_add(ga.counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (rank != gridDim.x - 1) return;  // uniform over the block
    double t2 = 0, t0 = 0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += KB)
        t2 += __hip_atomic_load(normPart + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    blockSum2(t2, t0);
    if (threadIdx.x == 0) *ga.counter = 0;
    __syncthreads();
    givensStep(ga.hbuf, t2, ga.j, ga.bn, ga.cs, ga.sn, ga.g, ga.H, ga.ldH, ga.stat, ga.statHost, sh, sh);
}

// The Givens step for nrhs right-hand sides, one block each: lane k's CGS output at
// hbuf + k hs (pass 1 at [0, ncol), |w|^2 at [ncol], pass 2 at [ncol + 1,
// 2 ncol + 1)), its state at cs/sn + k (m + 1), g + k (m + 2), H + k (m + 1)^2,
// stat/statHost + k 3 (m + 1).  Lanes with on[k] == 0 are skipped.
__global__ void kGivensB(const cplx<double>* __restrict__ hbuf, int64_t hs, int j, int m,
                         const double* __restrict__ bn, const int* __restrict__ on, cplx<double>* __restrict__ cs,
                         cplx<double>* __restrict__ sn, cplx<double>* __restrict__ g, cplx<double>* __restrict__ H,
                         double* __restrict__ stat, double* __restrict__ statHost) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int k = blockIdx.x;
    if (!on[k]) return;  // uniform over the block
    const cplx<double>* hb = hbuf + (int64_t)k * hs;
    givensStep(hb, hb[j + 1].re, j, bn[k], cs + (int64_t)k * (m + 1), sn + (int64_t)k * (m + 1),
               g + (int64_t)k * (m + 2), H + (int64_t)k * (m + 1) * (m + 1), m + 1, stat + (int64_t)k * 3 * (m + 1),
               statHost + (int64_t)k * 3 * (m + 1), reinterpret_cast<cplx<double>*>(smem));
}

// Q_{j+1} lane k = W lane k / |w| (that lane's statistic), zero for lanes off
template <typename TQ>
__global__ void kScaleB(const cplx<double>* __restrict__ W, const double* __restrict__ stat, int j, int m,
                        const int* __restrict__ on, cplx<TQ>* __restrict__ Qn, const int64_t* __restrict__ laneBase,
                        const int64_t* __restrict__ laneLen, cplx<double>* __restrict__ wide) {
    const int64_t r = (int64_t)blockIdx.x * KB + threadIdx.x;
    const int k = blockIdx.y;
    if (r >= laneLen[k]) return;
    const int64_t o = laneBase[k] + r;
    cplx<TQ> q{0, 0};
    if (on[k]) {
        const double v = stat[(int64_t)k * 3 * (m + 1) + 3 * j + 1];
        q = cplx<TQ>{(TQ)(W[o].re / v), (TQ)(W[o].im / v)};
    }
    Qn[o] = q;
    if (wide) wide[o] = cplx<double>{(double)q.re, (double)q.im};
}
__global__ void kInitGB(cplx<double>* __restrict__ g, const double* __restrict__ beta, int m2, int nrhs) {
    const int i = blockIdx.x * KB + threadIdx.x;
    if (i >= m2 * nrhs) return;
    const int k = i / m2, q = i - k * m2;
    g[i] = cplx<double>{q == 0 ? beta[k] : 0.0, 0.0};
}

// out column c = sum_j Q column j * P(j, c) for c < nout (P column-major,
// leading dimension ldp, FP64): the basis recombination of a deflated
// restart.  One thread per row, eight output columns at a time; sums in FP64.
template <typename TQ>
__global__ void kRecombine(const cplx<TQ>* __restrict__ Q, int64_t ld, int nin, const cplx<double>* __restrict__ P,
                           int ldp, int nout, cplx<TQ>* __restrict__ out, int64_t ldo, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int c0 = 0; c0 < nout; c0 += 8) {
        const int cn = min(8, nout - c0);
        double ar[8] = {}, ai[8] = {};
        for (int j = 0; j < nin; ++j) {
            const cplx<TQ> q = Q[(int64_t)j * ld + i];
            const double qr = (double)q.re, qi = (double)q.im;
#pragma unroll
            for (int c = 0; c < 8; ++c)
                if (c < cn) {
                    const cplx<double> p = P[(int64_t)(c0 + c) * ldp + j];
                    ar[c] += qr * p.re - qi * p.im;
                    ai[c] += qr * p.im + qi * p.re;
                }
        }
        for (int c = 0; c < cn; ++c) {
            cplx<TQ> v;
            v.re = (TQ)ar[c];
            v.im = (TQ)ai[c];
            out[(int64_t)(c0 + c) * ldo + i] = v;
        }
    }
}

// out = in / *s (s on the device)
// wide (FP32 basis): also the stored vector widened to FP64, the input of
// the next step's matvec
template <typename TQ>
__global__ void kScaleDev(const cplx<double>* __restrict__ in, const double* __restrict__ s,
                         cplx<TQ>* __restrict__ out, int64_t n, cplx<double>* __restrict__ wide) {
    const int64_t r = (int64_t)blockIdx.x * KB + threadIdx.x;
    const double v = *s;
    if (r < n) {
        const cplx<TQ> q{(TQ)(in[r].re / v), (TQ)(in[r].im / v)};
        out[r] = q;
        if (wide) wide[r] = cplx<double>{(double)q.re, (double)q.im};
    }
}

// g = (beta, 0, ..., 0), m + 1 entries
__global__ void kInitG(cplx<double>* __restrict__ g, double beta, int m1) {
    const int i = blockIdx.x * KB + threadIdx.x;
    if (i < m1) g[i] = cplx<double>{i == 0 ? beta : 0.0, 0.0};
}

// Krylov workspace and helpers bound to one plan
struct Krylov {
    DevicePlan& d;
    hipStream_t st;
    int64_t n;
    int nb;
    cplx<double>* partial = nullptr;
    cplx<double>* h = nullptr;  // device column
    // len: the vector length it works on (a lane of a batched solve); the
    // plan's n by default
    Krylov(DevicePlan& dp, hipStream_t s, int maxcol, int64_t len = -1)
        : d(dp), st(s), n(len < 0 ? dp.hp.src.n : len) {
        nb = (int)((n + KB - 1) / KB);
        partial = (cplx<double>*)d.alloc((size_t)nb * (maxcol + 2) * sizeof(cplx<double>));
        h = (cplx<double>*)d.alloc((size_t)2 * (maxcol + 2) * sizeof(cplx<double>));
    }
    ~Krylov() {
        d.release(partial);
        d.release(h);
    }
    double norm(const cplx<double>* v) {
        hipLaunchKernelGGL(kUpdate, dim3(nb), dim3(KB), 0, st, (const cplx<double>*)nullptr, (int64_t)0, 0,
                           (const cplx<double>*)nullptr, const_cast<cplx<double>*>(v), n, partial, 1);
        hipLaunchKernelGGL(kReduce, dim3(1), dim3(KB), 0, st, partial, nb, h, 0);
        cplx<double> r;
        hipCheck(hipMemcpyAsync(&r, h, sizeof(r), hipMemcpyDeviceToHost, st), "norm copy");
        hipCheck(hipStreamSynchronize(st), "norm");
        return std::sqrt(r.re);
    }
    // classical Gram-Schmidt applied twice against Q[:, 0..ncol); returns h and |w|
    void orthogonalize(const cplx<double>* Q, int ncol, cplx<double>* w, std::vector<std::complex<double>>& hout,
                       double& wnorm) {
        for (int pass = 0; pass < 2; ++pass) {
            hipLaunchKernelGGL(kDotPartial, dim3(nb, ncol), dim3(KB), 0, st, Q, n, ncol, (const cplx<double>*)w, n, partial);
            hipLaunchKernelGGL(kReduce, dim3(ncol), dim3(KB), 0, st, partial, nb, h + (pass ? ncol + 1 : 0), 0);
            hipLaunchKernelGGL(kUpdate, dim3(nb), dim3(KB), 0, st, Q, n, ncol, (const cplx<double>*)(h + (pass ? ncol + 1 : 0)),
                               w, n, partial + (size_t)nb * (ncol + 1), pass);
        }
        // |w|^2 of the final vector is in partial[nb*(ncol+1) ...]
        hipLaunchKernelGGL(kReduce, dim3(1), dim3(KB), 0, st, partial + (size_t)nb * (ncol + 1), nb, h + ncol, 0);
        std::vector<cplx<double>> hh(2 * ncol + 1);
        hipCheck(hipMemcpyAsync(hh.data(), h, hh.size() * sizeof(cplx<double>), hipMemcpyDeviceToHost, st), "h copy");
        hipCheck(hipStreamSynchronize(st), "arnoldi");
        hout.resize(ncol);
        for (int i = 0; i < ncol; ++i) hout[i] = std::complex<double>(hh[i].re + hh[ncol + 1 + i].re, hh[i].im + hh[ncol + 1 + i].im);
        wnorm = std::sqrt(hh[ncol].re);
    }
    // the two CGS passes of orthogonalize() queued without a host sync, for
    // columns Q + i ld; readOrtho() after a stream sync returns h and |w|
    void queueOrtho(const cplx<double>* Q, int64_t ld, int ncol, cplx<double>* w, cplx<double>* hhost) {
        for (int pass = 0; pass < 2; ++pass) {
            hipLaunchKernelGGL(kDotPartial, dim3(nb, ncol), dim3(KB), 0, st, Q, ld, ncol, (const cplx<double>*)w, n, partial);
            hipLaunchKernelGGL(kReduce, dim3(ncol), dim3(KB), 0, st, partial, nb, h + (pass ? ncol + 1 : 0), 0);
            hipLaunchKernelGGL(kUpdate, dim3(nb), dim3(KB), 0, st, Q, ld, ncol,
                               (const cplx<double>*)(h + (pass ? ncol + 1 : 0)), w, n,
                               partial + (size_t)nb * (ncol + 1), pass);
        }
        hipLaunchKernelGGL(kReduce, dim3(1), dim3(KB), 0, st, partial + (size_t)nb * (ncol + 1), nb, h + ncol, 0);
        hipCheck(hipMemcpyAsync(hhost, h, (2 * ncol + 1) * sizeof(cplx<double>), hipMemcpyDeviceToHost, st),
                 "h copy");
    }
    static void readOrtho(const cplx<double>* hh, int ncol, std::vector<std::complex<double>>& hout, double& wnorm) {
        hout.resize(ncol);
        for (int i = 0; i < ncol; ++i)
            hout[i] = std::complex<double>(hh[i].re + hh[ncol + 1 + i].re, hh[i].im + hh[ncol + 1 + i].im);
        wnorm = std::sqrt(hh[ncol].re);
    }
    void scale(const cplx<double>* in, double s, cplx<double>* out) {
        hipLaunchKernelGGL(kScale, dim3(nb), dim3(KB), 0, st, in, s, out, n);
    }
    template <typename TQ>
    void combine(const cplx<TQ>* Q, int ncol, const std::vector<std::complex<double>>& y, cplx<double>* x, int acc,
                 int64_t ld = 0) {
        std::vector<cplx<double>> yy(ncol);
        for (int i = 0; i < ncol; ++i) yy[i] = cplx<double>{y[i].real(), y[i].imag()};
        hipCheck(hipMemcpyAsync(h, yy.data(), ncol * sizeof(cplx<double>), hipMemcpyHostToDevice, st), "y upload");
        hipLaunchKernelGGL(kCombine<TQ>, dim3(nb), dim3(KB), 0, st, Q, ld ? ld : n, ncol, (const cplx<double>*)h, x, n,
                           acc);
        hipCheck(hipStreamSynchronize(st), "combine");
    }
    template <typename TQ>
    void scaleTo(const cplx<double>* in, double s, cplx<TQ>* out) {
        hipLaunchKernelGGL(kScaleTo<TQ>, dim3(nb), dim3(KB), 0, st, in, s, out, n);
    }
    void sub(const cplx<double>* b, const cplx<double>* a, cplx<double>* r) {
        hipLaunchKernelGGL(kSub, dim3(nb), dim3(KB), 0, st, b, a, r, n);
    }
};

using zc = std::complex<double>;

// Back substitution with the upper-triangular R stored by columns (H[j][i]).
static std::vector<zc> backSolve(const std::vector<std::vector<zc>>& H, const std::vector<zc>& g, int m) {
    std::vector<zc> y(m);
    for (int i = m - 1; i >= 0; --i) {
        zc s = g[i];
        for (int j = i + 1; j < m; ++j) s -= H[j][i] * y[j];
        y[i] = s / H[i][i];
    }
    return y;
}

struct SolveOut {
    int iters = 0;
    int status = MLFMA_OK;
    double est = 0;
};

// Reference-compatible GMRES (src/solver.cpp:116-146): unrestarted, complex
// Givens rotation t = sqrt(z^2 + w^2) without conjugates (src/math.h:168-172),
// stops when |g_{k+1}|/g0 <= tol or after maxit iterations.
static SolveOut gmresCompat(DevicePlan& d, Krylov& K, const cplx<double>* b, cplx<double>* x, int maxit, double tol,
                            double* hist) {
    SolveOut out;
    const int64_t n = d.hp.src.n;
    cplx<double>* Q = (cplx<double>*)d.alloc((size_t)n * (maxit + 1) * sizeof(cplx<double>));
    cplx<double>* w = (cplx<double>*)d.alloc((size_t)n * sizeof(cplx<double>));
    const double g0 = K.norm(b);
    if (!(g0 > 0) || !std::isfinite(g0)) {
        hipCheck(hipMemsetAsync(x, 0, n * sizeof(cplx<double>), K.st), "zero x");
        hipCheck(hipStreamSynchronize(K.st), "sync");
        d.release(Q);
        d.release(w);
        out.status = MLFMA_ERR_BREAKDOWN;
        return out;
    }
    K.scale(b, g0, Q);
    std::vector<std::vector<zc>> H;
    std::vector<zc> g(maxit + 1, 0.0), cs(maxit), sn(maxit);
    g[0] = g0;
    int it = 0;
    for (;;) {
        const int j = it;
        d.applySolveOp(Q + (size_t)j * n, w, 0, K.st, false, false);
        std::vector<zc> h;
        double hn;
        K.orthogonalize(Q, j + 1, w, h, hn);
        K.scale(w, hn, Q + (size_t)(j + 1) * n);
        h.push_back(hn);
        for (int i = 0; i < j; ++i) {
            const zc a = cs[i] * h[i] + sn[i] * h[i + 1];
            h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1];
            h[i] = a;
        }
        const zc t = std::sqrt(h[j] * h[j] + h[j + 1] * h[j + 1]);
        cs[j] = h[j] / t;
        sn[j] = h[j + 1] / t;
        h[j] = cs[j] * h[j] + sn[j] * h[j + 1];
        h[j + 1] = 0.0;
        H.push_back(h);
        g[j + 1] = -sn[j] * g[j];
        g[j] = cs[j] * g[j];
        ++it;
        const double res = std::abs(g[it]) / g0;
        if (hist) hist[it - 1] = res;
        out.est = res;
        if (!std::isfinite(res)) {
            out.status = MLFMA_ERR_BREAKDOWN;
            break;
        }
        if (!(res > tol && it < maxit)) break;
    }
    out.iters = it;
    if (out.status == MLFMA_OK && out.est > tol) out.status = MLFMA_ERR_NOT_CONVERGED;
    const std::vector<zc> y = backSolve(H, g, it);
    K.combine(Q, it, y, x, 0);
    d.release(Q);
    d.release(w);
    return out;
}

// Restarted GMRES(m) with unitary complex Givens rotations; convergence is
// checked on the true residual at every restart.  With fp32 the far field of
// every matvec runs in FP32 (used as the inner solver of the refinement loop).
// Device state of one pipelined GMRES(m) cycle: rotations, rotated
// right-hand side, rotated Hessenberg columns, per-step statistics (device
// and host-mapped) and one event per step in flight.
constexpr int kWindow = 4;
struct GivensWork {
    DevicePlan& d;
    cplx<double>*cs = nullptr, *sn = nullptr, *g = nullptr, *H = nullptr;
    double* stat = nullptr;      // device
    double* statHost = nullptr;  // host-mapped (pinned)
    double* statDev = nullptr;   // device address of statHost
    std::array<hipEvent_t, kWindow> ev{};
    cplx<double>* dotPart = nullptr;  // wide dot partials, (m + 1) x nbd
    double* normPart = nullptr;       // |w|^2 per update block
    unsigned int* counter = nullptr;  // blocks done in kUpdateRGivens (zero between launches)
    int nbd = 0, nbu = 0;
    int64_t *lane0 = nullptr, *laneN = nullptr;  // the vector as the one lane of kDotManyB
    int* reorth = nullptr;                       // the second CGS pass runs (kReorthTest)
    double* norm1 = nullptr;                     // |w|^2 after the first pass
    GivensWork(DevicePlan& dp, int m) : d(dp) {
        const int64_t n = d.hp.src.n;
        nbu = (int)((n + KB - 1) / KB);
        nbd = (int)((n + (int64_t)KB * RB - 1) / ((int64_t)KB * RB));
        lane0 = d.upload(std::vector<int64_t>{0});
        laneN = d.upload(std::vector<int64_t>{n});
        reorth = (int*)d.alloc(sizeof(int));
        norm1 = (double*)d.alloc(sizeof(double));
        dotPart = (cplx<double>*)d.alloc((size_t)nbd * (m + 1) * sizeof(cplx<double>));
        normPart = (double*)d.alloc((size_t)nbu * sizeof(double));
        counter = (unsigned int*)d.alloc(sizeof(unsigned int));
        hipCheck(hipMemset(counter, 0, sizeof(unsigned int)), "hipMemset");
        cs = (cplx<double>*)d.alloc((size_t)(m + 1) * sizeof(cplx<double>));
        sn = (cplx<double>*)d.alloc((size_t)(m + 1) * sizeof(cplx<double>));
        g = (cplx<double>*)d.alloc((size_t)(m + 2) * sizeof(cplx<double>));
        H = (cplx<double>*)d.alloc((size_t)(m + 1) * (m + 1) * sizeof(cplx<double>));
        stat = (double*)d.alloc((size_t)3 * (m + 1) * sizeof(double));
        hipCheck(hipHostMalloc((void**)&statHost, (size_t)3 * (m + 1) * sizeof(double), hipHostMallocMapped),
                 "hipHostMalloc");
        hipCheck(hipHostGetDevicePointer((void**)&statDev, statHost, 0), "hipHostGetDevicePointer");
        for (auto& e : ev) hipCheck(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    }
    ~GivensWork() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (statHost) (void)hipHostFree(statHost);
        for (void* p : {(void*)cs, (void*)sn, (void*)g, (void*)H, (void*)stat, (void*)dotPart, (void*)normPart,
                        (void*)counter, (void*)lane0, (void*)laneN, (void*)reorth, (void*)norm1})
            d.release(p);
    }
};

// lastResidual = false (the inner solves of FGMRES): when maxit is reached the
// solve ends without the residual of the final iterate (one matvec fewer)
// With the FP32 far field (fp32) the basis is stored in FP32, as in the
// batched solver (TQ = float; each basis vector widened to FP64 for its
// matvec, products and sums in FP64).
// deflate = k > 0: deflated restarts (GMRES-DR, dense.hpp): a cycle that runs
// all m steps without meeting the tolerance hands its k harmonic Ritz vectors
// of smallest modulus and its residual direction to the next cycle, which
// starts at step k with its first k Hessenberg columns already triangular
// (identity rotations) and its first k inputs Zk kept on the device.  The
// stopping test still uses the true residual at every restart; a restart
// whose small dense problem fails (or a cycle that ended early) restarts
// plainly from the residual.
template <typename TQ>
static SolveOut gmresRestartT(DevicePlan& d, Krylov& K, const cplx<double>* b, cplx<double>* x, int m, int maxit,
                              double tol, bool fp32, bool x0zero, double* hist, int histOff, bool lastResidual,
                              int deflate) {
    constexpr bool narrow = std::is_same<TQ, float>::value;
    SolveOut out;
    const int64_t n = d.hp.src.n;
    m = std::max(1, std::min(m, maxit));
    const int kd = deflate > 0 && deflate + 1 < m ? deflate : 0;
    cplx<TQ>* Q = (cplx<TQ>*)d.alloc((size_t)n * (m + 1) * sizeof(cplx<TQ>));
    cplx<double>* w = (cplx<double>*)d.alloc((size_t)n * sizeof(cplx<double>));
    cplx<double>* x64 = narrow ? (cplx<double>*)d.alloc((size_t)n * sizeof(cplx<double>)) : nullptr;
    // right preconditioning (d.solvePrec): steps apply Z P, and x += P Q y
    cplx<double>* pt = d.solvePrec ? (cplx<double>*)d.alloc((size_t)n * sizeof(cplx<double>)) : nullptr;
    // deflation: the kept inputs, the next cycle's first basis vectors
    // (written beside Q, then copied in) and the two small matrices
    cplx<TQ>* Zk = kd ? (cplx<TQ>*)d.alloc((size_t)n * kd * sizeof(cplx<TQ>)) : nullptr;
    cplx<TQ>* Vn = kd ? (cplx<TQ>*)d.alloc((size_t)n * (kd + 1) * sizeof(cplx<TQ>)) : nullptr;
    cplx<double>* dP = kd ? (cplx<double>*)d.alloc((size_t)(2 * m + 1) * (kd + 1) * sizeof(cplx<double>)) : nullptr;
    auto releaseAll = [&]() {
        for (void* p : {(void*)Q, (void*)w, (void*)x64, (void*)pt, (void*)Zk, (void*)Vn, (void*)dP})
            if (p) d.release(p);
    };
    GivensWork G(d, m);
    const double bn = K.norm(b);
    if (!(bn > 0) || !std::isfinite(bn)) {
        hipCheck(hipMemsetAsync(x, 0, n * sizeof(cplx<double>), K.st), "zero x");
        hipCheck(hipStreamSynchronize(K.st), "sync");
        releaseAll();
        out.status = bn == 0 ? MLFMA_OK : MLFMA_ERR_NONFINITE;
        return out;
    }
    if (x0zero) hipCheck(hipMemsetAsync(x, 0, n * sizeof(cplx<double>), K.st), "zero x");
    int total = 0;
    double res = 1.0;
    bool cycleMet = false;  // the last cycle stopped on its residual estimate
    dense::DrState dr;      // the deflated start of the next cycle (dr.k > 0)
    for (;;) {
        if (!lastResidual && (total >= maxit || cycleMet)) break;
        // r = b - A x
        double beta;
        if (x0zero && total == 0) {
            hipCheck(hipMemcpyAsync(w, b, n * sizeof(cplx<double>), hipMemcpyDeviceToDevice, K.st), "copy");
            beta = bn;
        } else {
            d.applySolveOp(x, w, 0, K.st, fp32, false);
            K.sub(b, w, w);
            beta = K.norm(w);
        }
        res = beta / bn;
        out.est = res;
        if (res <= tol || total >= maxit) break;
        // k0: the cycle's first step (the deflated columns before it)
        const int k0 = dr.k;
        if (k0 > 0) {
            // Q[0..k0] and Zk hold the deflated block (written at the
            // restart); its triangular columns, identity rotations and the
            // rotated right-hand side go to the Givens state
            std::vector<cplx<double>> Hh((size_t)k0 * (m + 1), cplx<double>{0, 0}), gh(m + 2, cplx<double>{0, 0});
            std::vector<cplx<double>> ch(k0, cplx<double>{1, 0}), sh(k0, cplx<double>{0, 0});
            for (int c = 0; c < k0; ++c)
                for (int i = 0; i <= c; ++i)
                    Hh[(size_t)c * (m + 1) + i] = cplx<double>{dr.Rk(i, c).real(), dr.Rk(i, c).imag()};
            for (int i = 0; i <= k0; ++i) gh[i] = cplx<double>{dr.g[i].real(), dr.g[i].imag()};
            hipCheck(hipMemcpyAsync(G.H, Hh.data(), Hh.size() * sizeof(cplx<double>), hipMemcpyHostToDevice, K.st),
                     "H upload");
            hipCheck(hipMemcpyAsync(G.g, gh.data(), gh.size() * sizeof(cplx<double>), hipMemcpyHostToDevice, K.st),
                     "g upload");
            hipCheck(hipMemcpyAsync(G.cs, ch.data(), k0 * sizeof(cplx<double>), hipMemcpyHostToDevice, K.st),
                     "cs upload");
            hipCheck(hipMemcpyAsync(G.sn, sh.data(), k0 * sizeof(cplx<double>), hipMemcpyHostToDevice, K.st),
                     "sn upload");
            hipCheck(hipStreamSynchronize(K.st), "deflated start");
        } else
            K.scaleTo<TQ>(w, beta, Q);
        // Pipelined cycle: each step (matvec, two CGS passes, the Givens update
        // and the scaling of the next basis vector) is queued without a host
        // round trip; up to kWindow steps are in flight while the host reads
        // the residual estimates of finished steps (host-mapped) and stops
        // queueing at the first one that meets the stopping test.  Steps queued
        // beyond it only append columns, so the solution uses exactly the
        // columns the synchronous algorithm would.
        if (k0 == 0)
            hipLaunchKernelGGL(kInitG, dim3((m + 1 + KB - 1) / KB), dim3(KB), 0, K.st, G.g, beta, m + 1);
        int jq = k0, jc = k0, ncolUsed = -1;
        // each event record costs the stream a few microseconds, so one event
        // covers two steps (and the last step queued before the host waits);
        // cover[s % kWindow]: the event slot that covers step s
        std::array<int, kWindow> cover{};
        int uncovered = k0;
        while (true) {
            while (ncolUsed < 0 && jq < m && total + (jq - jc) < maxit && jq - jc < kWindow) {
                // FP32 basis: x64 holds column jq widened (the first column's
                // here, the others' from the scaling of the step before)
                const cplx<double>* qj;
                if constexpr (narrow) {
                    if (jq == k0)
                        hipLaunchKernelGGL(kWiden, dim3(K.nb), dim3(KB), 0, K.st,
                                           (const cplx<float>*)(Q + (size_t)k0 * n), x64, n);
                    qj = x64;
                } else
                    qj = Q + (size_t)jq * n;
                if (pt) {
                    d.applyPrecond(qj, pt, 1, n, d.solveAdjoint, K.st);
                    d.applySolveOp(pt, w, 0, K.st, fp32, false);
                } else
                    d.applySolveOp(qj, w, 0, K.st, fp32, false);
                const int ncol = jq + 1;
                // classical Gram-Schmidt; the second pass (which ends with
                // the Givens step) reorthogonalizes only when the first one
                // cancelled more than half of |w|^2 (kReorthTest)
                for (int pass = 0; pass < 2; ++pass) {
                    // one vector: column groups of kDotCols fill the device
                    hipLaunchKernelGGL(kDotManyB<TQ>, dim3(G.nbd, 1, (ncol + kDotCols - 1) / kDotCols), dim3(KB),
                                       (size_t)kDotCols * (KB / 64) * sizeof(cplx<double>), K.st, (const cplx<TQ>*)Q,
                                       n, ncol, (const cplx<double>*)w, (const int64_t*)G.lane0,
                                       (const int64_t*)G.laneN, G.dotPart, (int64_t)0, kDotCols, pass == 0 ? 1 : 0,
                                       pass ? (const int*)G.reorth : nullptr);
                    if (pass == 0) {
                        hipLaunchKernelGGL(kUpdateR<TQ>, dim3(G.nbu), dim3(KB), ncol * sizeof(cplx<double>), K.st,
                                           (const cplx<TQ>*)Q, n, ncol, (const cplx<double>*)G.dotPart, G.nbd,
                                           K.h, w, n, G.normPart, 1);
                        hipLaunchKernelGGL(kReorthTest, dim3(1), dim3(KB), 0, K.st,
                                           (const double*)(G.dotPart + (size_t)ncol * G.nbd), (int64_t)0, G.nbd, 2,
                                           (const double*)G.normPart, (int64_t)0, G.nbu, 1, G.reorth, G.norm1);
                    } else
                        hipLaunchKernelGGL(kUpdateRGivens<TQ>, dim3(G.nbu), dim3(KB),
                                           (3 * ncol + 3) * sizeof(cplx<double>), K.st, (const cplx<TQ>*)Q, n, ncol,
                                           (const cplx<double>*)G.dotPart, G.nbd, K.h + ncol + 1, w, n, G.normPart,
                                           GivensArgs{K.h, jq, bn, G.cs, G.sn, G.g, G.H, m + 1, G.stat, G.statDev,
                                                      G.counter},
                                           (const int*)G.reorth, (const double*)G.norm1);
                }
                hipLaunchKernelGGL(kScaleDev<TQ>, dim3(K.nb), dim3(KB), 0, K.st, (const cplx<double>*)w,
                                   (const double*)(G.stat + 3 * jq + 1), Q + (size_t)(jq + 1) * n, n, x64);
                const bool more = jq + 1 < m && total + (jq + 1 - jc) < maxit && jq + 1 - jc < kWindow;
                if ((jq & 1) || !more) {
                    hipCheck(hipEventRecord(G.ev[jq % kWindow], K.st), "event");
                    for (int q = uncovered; q <= jq; ++q) cover[q % kWindow] = jq % kWindow;
                    uncovered = jq + 1;
                }
                ++jq;
            }
            if (jc == jq) break;
            hipCheck(hipEventSynchronize(G.ev[cover[jc % kWindow]]), "gmres step");
            const double est = G.statHost[3 * jc], hn = G.statHost[3 * jc + 1], rho = G.statHost[3 * jc + 2];
            if (rho == 0) {
                out.status = MLFMA_ERR_BREAKDOWN;
                ncolUsed = jc;
                break;
            }
            ++total;
            if (hist) hist[histOff + total - 1] = est;
            ++jc;
            if (est <= tol || hn == 0) {
                ncolUsed = jc;
                cycleMet = true;
                break;
            }
            if (total >= maxit) {
                ncolUsed = jc;
                break;
            }
        }
        if (ncolUsed < 0) ncolUsed = jc;
        hipCheck(hipStreamSynchronize(K.st), "gmres cycle");  // steps queued past the stop
        const int j = ncolUsed;
        std::vector<cplx<double>> Hh((size_t)j * (m + 1)), gh(j + 1);
        if (j > 0) {
            hipCheck(hipMemcpy(Hh.data(), G.H, Hh.size() * sizeof(cplx<double>), hipMemcpyDeviceToHost), "H copy");
            hipCheck(hipMemcpy(gh.data(), G.g, gh.size() * sizeof(cplx<double>), hipMemcpyDeviceToHost), "g copy");
            std::vector<std::vector<zc>> H(j);
            for (int c = 0; c < j; ++c) {
                H[c].resize(c + 1);
                for (int i = 0; i <= c; ++i) H[c][i] = zc(Hh[(size_t)c * (m + 1) + i].re, Hh[(size_t)c * (m + 1) + i].im);
            }
            std::vector<zc> g(j);
            for (int i = 0; i < j; ++i) g[i] = zc(gh[i].re, gh[i].im);
            const std::vector<zc> y = backSolve(H, g, j);
            // dst (+)= Z y: the kept inputs Zk for the first k0 columns, the
            // basis vectors after them
            auto addTo = [&](cplx<double>* dst, int acc) {
                if (k0 > 0) {
                    K.combine<TQ>(Zk, k0, std::vector<zc>(y.begin(), y.begin() + k0), dst, acc);
                    if (j > k0)
                        K.combine<TQ>(Q + (size_t)k0 * n, j - k0, std::vector<zc>(y.begin() + k0, y.end()), dst, 1);
                } else
                    K.combine<TQ>(Q, j, y, dst, acc);
            };
            if (pt) {  // x += P (Z y), through w
                addTo(w, 0);
                d.applyPrecond(w, pt, 1, n, d.solveAdjoint, K.st);
                K.combine(pt, 1, std::vector<zc>{1.0}, x, 1);
            } else
                addTo(x, 1);
        }
        if (out.status != MLFMA_OK) break;
        // deflated restart after a full cycle that did not meet the tolerance
        dense::DrState nd;
        bool kept = false;
        if (kd && j == m && !cycleMet && total < maxit) {
            std::vector<cplx<double>> ch(m), sh(m);
            hipCheck(hipMemcpy(ch.data(), G.cs, m * sizeof(cplx<double>), hipMemcpyDeviceToHost), "cs copy");
            hipCheck(hipMemcpy(sh.data(), G.sn, m * sizeof(cplx<double>), hipMemcpyDeviceToHost), "sn copy");
            dense::Mat R(m, m);
            std::vector<zc> cs(m), sn(m);
            for (int c = 0; c < m; ++c) {
                for (int i = 0; i <= c; ++i) R(i, c) = zc(Hh[(size_t)c * (m + 1) + i].re, Hh[(size_t)c * (m + 1) + i].im);
                cs[c] = zc(ch[c].re, ch[c].im);
                sn[c] = zc(sh[c].re, sh[c].im);
            }
            kept = dense::drRestart(R, cs, sn, zc(gh[m].re, gh[m].im), k0, dr.Qhat, kd, nd) && nd.defect < 1e-8;
        }
        if (kept) {
            // Zk = Q[0..m) Pk, then Q[0..kd] = Q[0..m] U (through Vn)
            std::vector<cplx<double>> ph((size_t)m * kd + (size_t)(m + 1) * (kd + 1));
            for (int c = 0; c < kd; ++c)
                for (int i = 0; i < m; ++i) ph[(size_t)c * m + i] = cplx<double>{nd.Pk(i, c).real(), nd.Pk(i, c).imag()};
            cplx<double>* pu = ph.data() + (size_t)m * kd;
            for (int c = 0; c <= kd; ++c)
                for (int i = 0; i <= m; ++i)
                    pu[(size_t)c * (m + 1) + i] = cplx<double>{nd.U(i, c).real(), nd.U(i, c).imag()};
            hipCheck(hipMemcpyAsync(dP, ph.data(), ph.size() * sizeof(cplx<double>), hipMemcpyHostToDevice, K.st),
                     "restart upload");
            hipLaunchKernelGGL(kRecombine<TQ>, dim3(K.nb), dim3(KB), 0, K.st, (const cplx<TQ>*)Q, n, m,
                               (const cplx<double>*)dP, m, kd, Zk, n, n);
            hipLaunchKernelGGL(kRecombine<TQ>, dim3(K.nb), dim3(KB), 0, K.st, (const cplx<TQ>*)Q, n, m + 1,
                               (const cplx<double>*)(dP + (size_t)m * kd), m + 1, kd + 1, Vn, n, n);
            hipCheck(hipMemcpyAsync(Q, Vn, (size_t)n * (kd + 1) * sizeof(cplx<TQ>), hipMemcpyDeviceToDevice, K.st),
                     "restart basis");
            hipCheck(hipStreamSynchronize(K.st), "deflated restart");
            dr = std::move(nd);
        } else
            dr = dense::DrState{};
    }
    out.iters = total;
    if (out.status == MLFMA_OK && out.est > tol) out.status = MLFMA_ERR_NOT_CONVERGED;
    releaseAll();
    return out;
}

static SolveOut gmresRestart(DevicePlan& d, Krylov& K, const cplx<double>* b, cplx<double>* x, int m, int maxit,
                             double tol, bool fp32, bool x0zero, double* hist, int histOff, bool lastResidual = true,
                             int deflate = 0) {
    if (fp32)
        return gmresRestartT<float>(d, K, b, x, m, maxit, tol, fp32, x0zero, hist, histOff, lastResidual, deflate);
    return gmresRestartT<double>(d, K, b, x, m, maxit, tol, fp32, x0zero, hist, histOff, lastResidual, deflate);
}

// y = Z x for nrhs column vectors (stride n) through the batched kernels
// (fp32: the FP32 far field).
static void applyMany(DevicePlan& d, const cplx<double>* x, cplx<double>* y, int nrhs, hipStream_t st,
                      bool fp32 = false) {
    const int64_t n = d.hp.src.n;
    int r = 0;
    while (r < nrhs) {
        const int left = nrhs - r;
        const int nb = left >= 8 ? 8 : (left >= 4 ? 4 : (left >= 2 ? 2 : 1));
        d.applyOp(x + (size_t)r * n, y + (size_t)r * n, 0, st, fp32, false, nb, n);
        r += nb;
    }
    if (d.shifted()) launchAxpy(x, d.solveShift, y, (int64_t)nrhs * n, st);
}

// Lanes of a batched solve: nvec vectors of the plan (column stride n) cut
// into the plan's segments (the scatterers of a forest plan, else one
// segment): lane k is segment k % nseg of vector k / nseg.
struct Lanes {
    DevicePlan& d;
    int nvec = 0, nseg = 0, count = 0;
    int64_t n = 0, maxLen = 0;
    std::vector<int64_t> base, len;
    int64_t *dBase = nullptr, *dLen = nullptr;
    Lanes(DevicePlan& dp, int nv) : d(dp), nvec(nv), n(dp.hp.src.n) {
        const std::vector<int64_t> seg = dp.hp.forest ? dp.hp.segOff : std::vector<int64_t>{0, n};
        nseg = (int)seg.size() - 1;
        count = nvec * nseg;
        for (int v = 0; v < nvec; ++v)
            for (int sg = 0; sg < nseg; ++sg) {
                base.push_back((int64_t)v * n + seg[sg]);
                len.push_back(seg[sg + 1] - seg[sg]);
                maxLen = std::max(maxLen, seg[sg + 1] - seg[sg]);
            }
        dBase = d.upload(base);
        dLen = d.upload(len);
    }
    ~Lanes() {
        d.release(dBase);
        d.release(dLen);
    }
    int blocks() const { return (int)std::max<int64_t>(1, (maxLen + KB - 1) / KB); }
};

// CGS2 of every lane at once: one launch per kernel for all of them; hhost
// receives, per lane at stride hs, the layout of Krylov::queueOrtho (pass 1,
// |w|^2, pass 2).
struct BatchOrtho {
    DevicePlan& d;
    const Lanes& LN;
    int nb, nbd;  // row blocks of the update (KB rows) and of the dot products (RB KB rows)
    int64_t ps, hs;
    cplx<double>*partial = nullptr, *h = nullptr;
    int* reorth = nullptr;     // per lane: the second CGS pass runs
    double* norm1 = nullptr;   // per lane: |w|^2 after the first pass
    BatchOrtho(DevicePlan& dp, const Lanes& lanes, int m) : d(dp), LN(lanes) {
        nb = LN.blocks();
        nbd = (nb + RB - 1) / RB;
        ps = (int64_t)nb * (m + 3);
        hs = 2 * (int64_t)m + 3;
        partial = (cplx<double>*)d.alloc((size_t)LN.count * ps * sizeof(cplx<double>));
        h = (cplx<double>*)d.alloc((size_t)LN.count * hs * sizeof(cplx<double>));
        reorth = (int*)d.alloc((size_t)LN.count * sizeof(int));
        norm1 = (double*)d.alloc((size_t)LN.count * sizeof(double));
    }
    ~BatchOrtho() {
        d.release(partial);
        d.release(h);
        d.release(reorth);
        d.release(norm1);
    }
    // CGS with the second pass per lane only when the first one cancelled
    // more than half of |w|^2 (kReorthTest); a skipped pass leaves zero
    // coefficients, so the layout of hhost is the same
    template <typename TQ>
    void queue(const cplx<TQ>* Q, int64_t ld, int ncol, cplx<double>* W, cplx<double>* hhost, hipStream_t st) {
        const int nl = LN.count;
        cplx<double>* normPart = partial + (size_t)nb * (ncol + 1);
        for (int pass = 0; pass < 2; ++pass) {
            hipLaunchKernelGGL(kDotManyB<TQ>, dim3(nbd, nl), dim3(KB), (size_t)ncol * (KB / 64) * sizeof(cplx<double>),
                               st, Q, ld, ncol, (const cplx<double>*)W, (const int64_t*)LN.dBase,
                               (const int64_t*)LN.dLen, partial, ps, ncol, pass == 0 ? 1 : 0,
                               pass ? (const int*)reorth : nullptr);
            // pass 1 also sums |w_before|^2 (column ncol) into h[ncol]
            hipLaunchKernelGGL(kReduceB, dim3(pass ? ncol : ncol + 1, nl), dim3(KB), 0, st,
                               (const cplx<double>*)partial, ps, nbd, h + (pass ? ncol + 1 : 0), hs);
            hipLaunchKernelGGL(kUpdateB<TQ>, dim3(nb, nl), dim3(KB), 0, st, Q, ld, ncol,
                               (const cplx<double>*)(h + (pass ? ncol + 1 : 0)), hs, W, (const int64_t*)LN.dBase,
                               (const int64_t*)LN.dLen, normPart, ps, 1, pass ? (const int*)reorth : nullptr);
            if (pass == 0)
                hipLaunchKernelGGL(kReorthTest, dim3(nl), dim3(KB), 0, st, (const double*)(h + ncol), 2 * hs, 1, 2,
                                   (const double*)normPart, 2 * ps, nb, 2, reorth, norm1);
        }
        hipLaunchKernelGGL(kReduceB, dim3(1, nl), dim3(KB), 0, st, (const cplx<double>*)(partial + (size_t)nb * (ncol + 1)),
                           ps, nb, h + ncol, hs);
        if (hhost)
            hipCheck(hipMemcpyAsync(hhost, h, (size_t)nl * hs * sizeof(cplx<double>), hipMemcpyDeviceToHost, st),
                     "h copy");
    }
};

// Device state of one pipelined batched GMRES(m) cycle (per right-hand side:
// rotations, rotated right-hand side, rotated columns, step statistics).
struct GivensBatch {
    DevicePlan& d;
    int nrhs, m;
    cplx<double>*cs = nullptr, *sn = nullptr, *g = nullptr, *H = nullptr;
    double *stat = nullptr, *statHost = nullptr, *statDev = nullptr, *bn = nullptr, *beta = nullptr;
    int* on = nullptr;
    std::array<hipEvent_t, kWindow> ev{};
    GivensBatch(DevicePlan& dp, int nr, int m_) : d(dp), nrhs(nr), m(m_) {
        cs = (cplx<double>*)d.alloc((size_t)nrhs * (m + 1) * sizeof(cplx<double>));
        sn = (cplx<double>*)d.alloc((size_t)nrhs * (m + 1) * sizeof(cplx<double>));
        g = (cplx<double>*)d.alloc((size_t)nrhs * (m + 2) * sizeof(cplx<double>));
        H = (cplx<double>*)d.alloc((size_t)nrhs * (m + 1) * (m + 1) * sizeof(cplx<double>));
        stat = (double*)d.alloc((size_t)nrhs * 3 * (m + 1) * sizeof(double));
        bn = (double*)d.alloc((size_t)nrhs * sizeof(double));
        beta = (double*)d.alloc((size_t)nrhs * sizeof(double));
        on = (int*)d.alloc((size_t)nrhs * sizeof(int));
        hipCheck(hipHostMalloc((void**)&statHost, (size_t)nrhs * 3 * (m + 1) * sizeof(double), hipHostMallocMapped),
                 "hipHostMalloc");
        hipCheck(hipHostGetDevicePointer((void**)&statDev, statHost, 0), "hipHostGetDevicePointer");
        for (auto& e : ev) hipCheck(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    }
    ~GivensBatch() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (statHost) (void)hipHostFree(statHost);
        for (void* p : {(void*)cs, (void*)sn, (void*)g, (void*)H, (void*)stat, (void*)bn, (void*)beta, (void*)on})
            d.release(p);
    }
};

// f(i) for i in [0, count) on up to 16 host threads
template <typename F>
static void parallelFor(int count, F f) {
    const int nt = std::max(1, std::min({count, 16, (int)std::thread::hardware_concurrency()}));
    if (nt <= 1) {
        for (int i = 0; i < count; ++i) f(i);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&]() {
            for (int i = next++; i < count; i = next++) f(i);
        });
    for (auto& t : th) t.join();
}

// GMRES(m) on every lane in lockstep: one batched matvec per step over the
// block of the vectors' current Krylov columns (basis stored [step][vector][n]),
// per lane classical Gram-Schmidt twice and unitary Givens rotations as in
// gmresRestart.  A lane that has converged stops updating (its part of the
// block is zero) until the others finish the cycle.
// With the FP32 far field (fp32) the Krylov basis is stored in FP32 (TQ =
// float): these solves are preconditioners or refinement steps whose
// operator carries FP32 rounding anyway, and the basis sweeps of the
// orthogonalization move half the bytes; each basis vector is widened to
// FP64 for its matvec, and the products and sums stay FP64.
// deflate = k > 0: deflated restarts per lane as in gmresRestartT (the host
// problems of the lanes on parallel threads, the recombination launched per
// lane).  The lanes stay in lockstep: once any unfinished lane holds a
// deflated start, the cycle starts at step k for all, and a lane without one
// (it stopped early on its estimate, or its restart problem failed) restarts
// plainly inside that cycle: zero basis vectors and kept inputs in columns
// 0..k-1 with identity Hessenberg columns and a zero right-hand side there
// (their coefficients come out exactly zero), its residual direction in
// column k; its next deflated restart uses columns k..m only.
template <typename TQ>
static std::vector<SolveOut> gmresBatchT(DevicePlan& d, std::vector<std::unique_ptr<Krylov>>& K, hipStream_t st,
                                         const cplx<double>* B, cplx<double>* X, const Lanes& LN, int m, int maxit,
                                         const std::vector<double>& tolv, double* hist, bool fp32, int histOff,
                                         int histLd, bool lastResidual, int deflate) {
    constexpr bool narrow = std::is_same<TQ, float>::value;
    if (!histLd) histLd = maxit;
    const int64_t n = d.hp.src.n;
    const int nrhs = LN.count, nvec = LN.nvec;
    m = std::max(1, std::min(m, maxit));
    const int64_t ld = (int64_t)nvec * n;
    cplx<TQ>* Q = (cplx<TQ>*)d.alloc((size_t)ld * (m + 1) * sizeof(cplx<TQ>));
    cplx<double>* W = (cplx<double>*)d.alloc((size_t)ld * sizeof(cplx<double>));
    // the FP64 copy of the basis vector a step multiplies (FP32 basis)
    cplx<double>* X64 = narrow ? (cplx<double>*)d.alloc((size_t)ld * sizeof(cplx<double>)) : nullptr;
    // right preconditioning (d.solvePrec): steps apply Z P, and X += P Q y
    cplx<double>* PT = d.solvePrec ? (cplx<double>*)d.alloc((size_t)ld * sizeof(cplx<double>)) : nullptr;
    BatchOrtho BO(d, LN, m);
    GivensBatch GB(d, nrhs, m);
    const int kd = deflate > 0 && deflate + 1 < m ? deflate : 0;
    const size_t pSz = (size_t)m * kd + (size_t)(m + 1) * (kd + 1);  // per lane: Pk, then U
    cplx<TQ>* Zk = kd ? (cplx<TQ>*)d.alloc((size_t)ld * kd * sizeof(cplx<TQ>)) : nullptr;
    cplx<TQ>* Vn = kd ? (cplx<TQ>*)d.alloc((size_t)ld * (kd + 1) * sizeof(cplx<TQ>)) : nullptr;
    cplx<double>* dP = kd ? (cplx<double>*)d.alloc(pSz * nrhs * sizeof(cplx<double>)) : nullptr;
    std::vector<dense::DrState> dr(nrhs);
    std::vector<SolveOut> out(nrhs);
    std::vector<double> bn(nrhs);
    std::vector<char> done(nrhs, 0);
    for (int r = 0; r < nrhs; ++r) {
        bn[r] = K[r]->norm(B + LN.base[r]);
        if (!std::isfinite(bn[r])) throw Error(MLFMA_ERR_NONFINITE, "non-finite b");
        if (bn[r] == 0) done[r] = 1;
    }
    hipCheck(hipMemsetAsync(X, 0, (size_t)ld * sizeof(cplx<double>), st), "zero x");
    int total = 0;
    bool cycleMet = false;  // every lane of the last cycle stopped on its estimate
    for (;;) {
        if (!lastResidual && (total >= maxit || cycleMet)) break;
        // residuals r = b - A x
        if (total == 0)
            hipCheck(hipMemcpyAsync(W, B, (size_t)ld * sizeof(cplx<double>), hipMemcpyDeviceToDevice, st), "copy");
        else {
            applyMany(d, X, W, nvec, st, fp32);
            for (int r = 0; r < nrhs; ++r) K[r]->sub(B + LN.base[r], W + LN.base[r], W + LN.base[r]);
        }
        std::vector<double> beta(nrhs, 0.0);
        bool all = true;
        for (int r = 0; r < nrhs; ++r) {
            if (bn[r] == 0) continue;
            beta[r] = total == 0 ? bn[r] : K[r]->norm(W + LN.base[r]);
            out[r].est = beta[r] / bn[r];
            done[r] = out[r].est <= tolv[r];
            all = all && done[r];
        }
        if (all || total >= maxit) break;
        // the cycle's first step: k when an unfinished lane has a deflated
        // start; the other unfinished lanes start plainly at column k
        int k0 = 0;
        for (int r = 0; r < nrhs; ++r)
            if (!done[r] && dr[r].k == kd && kd) k0 = kd;
        std::vector<char> embed(nrhs, 0);
        for (int r = 0; r < nrhs; ++r) {
            embed[r] = k0 > 0 && !done[r] && dr[r].k != kd;
            if (k0 == 0 || embed[r]) dr[r] = dense::DrState{};
        }
        // Pipelined cycle, as in gmresRestart: every step (batched matvec,
        // CGS2 of all lanes, per-lane Givens update, scaling) is queued without
        // a host round trip, up to kWindow steps ahead of the host, which reads
        // each lane's residual estimate from host-mapped memory and stops the
        // lane at its first step that meets the test.  A lane's later steps
        // only append columns; the solution uses exactly its columns up to the
        // stop.
        std::vector<int> steps(nrhs, 0);
        std::vector<char> active(nrhs, 0);
        std::vector<int> onh(nrhs, 0);
        for (int r = 0; r < nrhs; ++r) {
            active[r] = !done[r];
            onh[r] = active[r];
            if (k0 == 0) {
                if (active[r]) K[r]->scaleTo<TQ>(W + LN.base[r], beta[r], Q + LN.base[r]);
                else hipCheck(hipMemsetAsync(Q + LN.base[r], 0, LN.len[r] * sizeof(cplx<TQ>), st), "zero lane");
            } else if (!active[r] || embed[r]) {
                for (int c = 0; c <= k0; ++c)
                    hipCheck(hipMemsetAsync(Q + (size_t)c * ld + LN.base[r], 0, LN.len[r] * sizeof(cplx<TQ>), st),
                             "zero lane");
                if (embed[r]) {
                    for (int c = 0; c < k0; ++c)
                        hipCheck(hipMemsetAsync(Zk + (size_t)c * ld + LN.base[r], 0, LN.len[r] * sizeof(cplx<TQ>),
                                                st), "zero lane");
                    K[r]->scaleTo<TQ>(W + LN.base[r], beta[r], Q + (size_t)k0 * ld + LN.base[r]);
                }
            }
        }
        hipCheck(hipMemcpyAsync(GB.on, onh.data(), nrhs * sizeof(int), hipMemcpyHostToDevice, st), "lanes");
        hipCheck(hipMemcpyAsync(GB.bn, bn.data(), nrhs * sizeof(double), hipMemcpyHostToDevice, st), "norms");
        hipCheck(hipMemcpyAsync(GB.beta, beta.data(), nrhs * sizeof(double), hipMemcpyHostToDevice, st), "betas");
        hipLaunchKernelGGL(kInitGB, dim3(((m + 2) * nrhs + KB - 1) / KB), dim3(KB), 0, st, GB.g, (const double*)GB.beta,
                           m + 2, nrhs);
        std::vector<cplx<double>> hH, hg, hc, hs;  // deflated starts (alive until the sync below)
        if (k0 > 0) {
            // per active lane: its triangular columns, identity rotations and
            // rotated right-hand side (Q[0..k0] and Zk hold its block); only
            // each lane's first k0 columns, k0 rotations and k0 + 1 entries
            // of g travel (the steps write the rest; kInitGB zeroed g)
            const size_t hc0 = (size_t)k0 * (m + 1);
            hH.assign((size_t)nrhs * hc0, cplx<double>{0, 0});
            hg.assign((size_t)nrhs * (k0 + 1), cplx<double>{0, 0});
            hc.assign((size_t)nrhs * k0, cplx<double>{1, 0});
            hs.assign((size_t)nrhs * k0, cplx<double>{0, 0});
            for (int r = 0; r < nrhs; ++r) {
                if (!active[r]) continue;
                cplx<double>* Hr = hH.data() + (size_t)r * hc0;
                if (embed[r]) {  // plain start at column k0
                    for (int c = 0; c < k0; ++c) Hr[(size_t)c * (m + 1) + c] = cplx<double>{1, 0};
                    hg[(size_t)r * (k0 + 1) + k0] = cplx<double>{beta[r], 0};
                    continue;
                }
                for (int c = 0; c < k0; ++c)
                    for (int i = 0; i <= c; ++i)
                        Hr[(size_t)c * (m + 1) + i] = cplx<double>{dr[r].Rk(i, c).real(), dr[r].Rk(i, c).imag()};
                for (int i = 0; i <= k0; ++i)
                    hg[(size_t)r * (k0 + 1) + i] = cplx<double>{dr[r].g[i].real(), dr[r].g[i].imag()};
            }
            const size_t z = sizeof(cplx<double>);
            hipCheck(hipMemcpy2DAsync(GB.H, (size_t)(m + 1) * (m + 1) * z, hH.data(), hc0 * z, hc0 * z, nrhs,
                                      hipMemcpyHostToDevice, st), "H upload");
            hipCheck(hipMemcpy2DAsync(GB.g, (size_t)(m + 2) * z, hg.data(), (k0 + 1) * z, (k0 + 1) * z, nrhs,
                                      hipMemcpyHostToDevice, st), "g upload");
            hipCheck(hipMemcpy2DAsync(GB.cs, (size_t)(m + 1) * z, hc.data(), k0 * z, k0 * z, nrhs,
                                      hipMemcpyHostToDevice, st), "cs upload");
            hipCheck(hipMemcpy2DAsync(GB.sn, (size_t)(m + 1) * z, hs.data(), k0 * z, k0 * z, nrhs,
                                      hipMemcpyHostToDevice, st), "sn upload");
            hipCheck(hipStreamSynchronize(st), "deflated start");
        }
        bool breakdown = false;
        int jq = k0, jc = k0;
        bool stop = false;
        while (true) {
            while (!stop && jq < m && total + (jq - jc) < maxit && jq - jc < kWindow) {
                // FP32 basis: X64 holds column jq widened (the first column's
                // here, the others' from the scaling of the step before)
                const cplx<double>* qj;
                if constexpr (narrow) {
                    if (jq == k0)
                        hipLaunchKernelGGL(kWiden, dim3((unsigned)((ld + KB - 1) / KB)), dim3(KB), 0, st,
                                           (const cplx<float>*)(Q + (size_t)k0 * ld), X64, ld);
                    qj = X64;
                } else
                    qj = Q + (size_t)jq * ld;
                if (PT) {
                    d.applyPrecond(qj, PT, nvec, n, false, st);
                    applyMany(d, PT, W, nvec, st, fp32);
                } else
                    applyMany(d, qj, W, nvec, st, fp32);
                BO.queue<TQ>(Q, ld, jq + 1, W, nullptr, st);
                hipLaunchKernelGGL(kGivensB, dim3(nrhs), dim3(64), (3 * jq + 6) * sizeof(cplx<double>), st,
                                   (const cplx<double>*)BO.h, BO.hs, jq, m,
                                   (const double*)GB.bn, (const int*)GB.on, GB.cs, GB.sn, GB.g, GB.H, GB.stat,
                                   GB.statDev);
                hipLaunchKernelGGL(kScaleB<TQ>, dim3(LN.blocks(), nrhs), dim3(KB), 0, st, (const cplx<double>*)W,
                                   (const double*)GB.stat, jq, m, (const int*)GB.on, Q + (size_t)(jq + 1) * ld,
                                   (const int64_t*)LN.dBase, (const int64_t*)LN.dLen, X64);
                hipCheck(hipEventRecord(GB.ev[jq % kWindow], st), "event");
                ++jq;
            }
            if (jc == jq) break;
            hipCheck(hipEventSynchronize(GB.ev[jc % kWindow]), "gmres step");
            bool any = false;
            for (int r = 0; r < nrhs; ++r) {
                if (!active[r]) continue;
                const double* sr = GB.statHost + (size_t)r * 3 * (m + 1) + 3 * jc;
                const double est = sr[0], hn = sr[1], rho = sr[2];
                if (rho == 0) {
                    out[r].status = MLFMA_ERR_BREAKDOWN;
                    active[r] = 0;
                    breakdown = true;
                    continue;
                }
                steps[r] = jc + 1;
                if (hist) hist[(size_t)r * histLd + histOff + total] = est;
                ++out[r].iters;
                if (est <= tolv[r] || hn == 0)
                    active[r] = 0;
                else
                    any = true;
            }
            ++total;
            ++jc;
            if (!any) cycleMet = true;
            if (!any || total >= maxit) stop = true;
            if (stop) break;  // steps queued past the stop are not read
        }
        hipCheck(hipStreamSynchronize(st), "gmres cycle");
        // lanes that ran the whole cycle without meeting their tolerance
        // restart deflated: their Givens state, read with the update below
        std::vector<dense::Mat> Rl(nrhs);
        std::vector<std::vector<zc>> csl(nrhs), snl(nrhs);
        std::vector<zc> gml(nrhs);
        std::vector<char> want(nrhs, 0);
        for (int r = 0; r < nrhs; ++r)
            want[r] = kd && steps[r] == m && active[r] && !breakdown && total < maxit &&
                      (!embed[r] || kd + 1 < m - k0);
        for (int r = 0; r < nrhs; ++r)
            if (steps[r] > 0) {
                const int j = steps[r];
                std::vector<cplx<double>> Hh((size_t)j * (m + 1)), gh(j + 1);
                hipCheck(hipMemcpy(Hh.data(), GB.H + (size_t)r * (m + 1) * (m + 1), Hh.size() * sizeof(cplx<double>),
                                   hipMemcpyDeviceToHost),
                         "H copy");
                hipCheck(hipMemcpy(gh.data(), GB.g + (size_t)r * (m + 2), gh.size() * sizeof(cplx<double>),
                                   hipMemcpyDeviceToHost),
                         "g copy");
                std::vector<std::vector<zc>> H(j);
                for (int c = 0; c < j; ++c) {
                    H[c].resize(c + 1);
                    for (int i = 0; i <= c; ++i)
                        H[c][i] = zc(Hh[(size_t)c * (m + 1) + i].re, Hh[(size_t)c * (m + 1) + i].im);
                }
                std::vector<zc> g(j);
                for (int i = 0; i < j; ++i) g[i] = zc(gh[i].re, gh[i].im);
                const std::vector<zc> y = backSolve(H, g, j);
                cplx<double>* dst = PT ? W + LN.base[r] : X + LN.base[r];
                if (k0 > 0) {  // the first k0 inputs are the kept vectors Zk
                    K[r]->combine<TQ>(Zk + LN.base[r], k0, std::vector<zc>(y.begin(), y.begin() + k0), dst,
                                      PT ? 0 : 1, ld);
                    if (j > k0)
                        K[r]->combine<TQ>(Q + (size_t)k0 * ld + LN.base[r], j - k0,
                                          std::vector<zc>(y.begin() + k0, y.end()), dst, 1, ld);
                } else
                    K[r]->combine<TQ>(Q + LN.base[r], j, y, dst, PT ? 0 : 1, ld);
                if (want[r]) {
                    // an embedded plain start: its cycle is columns e..m
                    const int e = embed[r] ? k0 : 0, mm = m - e;
                    std::vector<cplx<double>> ch(m), sh(m);
                    hipCheck(hipMemcpy(ch.data(), GB.cs + (size_t)r * (m + 1), m * sizeof(cplx<double>),
                                       hipMemcpyDeviceToHost), "cs copy");
                    hipCheck(hipMemcpy(sh.data(), GB.sn + (size_t)r * (m + 1), m * sizeof(cplx<double>),
                                       hipMemcpyDeviceToHost), "sn copy");
                    Rl[r] = dense::Mat(mm, mm);
                    csl[r].resize(mm);
                    snl[r].resize(mm);
                    for (int c = 0; c < mm; ++c) {
                        for (int i = 0; i <= c; ++i) Rl[r](i, c) = H[c + e][i + e];
                        csl[r][c] = zc(ch[c + e].re, ch[c + e].im);
                        snl[r][c] = zc(sh[c + e].re, sh[c + e].im);
                    }
                    gml[r] = zc(gh[m].re, gh[m].im);
                }
            } else if (PT)
                hipCheck(hipMemsetAsync(W + LN.base[r], 0, LN.len[r] * sizeof(cplx<double>), st), "zero lane");
        if (PT) {  // X += P (Q y) of every lane, through W
            d.applyPrecond(W, PT, nvec, n, false, st);
            for (int r = 0; r < nrhs; ++r)
                if (steps[r] > 0) K[r]->combine(PT + LN.base[r], 1, std::vector<zc>{1.0}, X + LN.base[r], 1);
        }
        if (breakdown) break;
        if (kd) {
            // the lanes' restart problems on host threads, then per kept lane
            // Zk = Q[0..m) Pk and Q[0..kd] = Q[0..m] U (through Vn)
            std::vector<dense::DrState> nd(nrhs);
            std::vector<char> kept(nrhs, 0);
            parallelFor(nrhs, [&](int r) {
                if (want[r])
                    kept[r] = dense::drRestart(Rl[r], csl[r], snl[r], gml[r], embed[r] ? 0 : k0, dr[r].Qhat, kd,
                                               nd[r]) &&
                              nd[r].defect < 1e-8;
            });
            std::vector<cplx<double>> ph(pSz * nrhs);
            bool any = false;
            for (int r = 0; r < nrhs; ++r) {
                if (!kept[r]) {
                    dr[r] = dense::DrState{};
                    continue;
                }
                any = true;
                const int mm = m - (embed[r] ? k0 : 0);  // the lane's cycle columns
                cplx<double>* pk = ph.data() + pSz * r;
                cplx<double>* pu = pk + (size_t)mm * kd;
                for (int c = 0; c < kd; ++c)
                    for (int i = 0; i < mm; ++i)
                        pk[(size_t)c * mm + i] = cplx<double>{nd[r].Pk(i, c).real(), nd[r].Pk(i, c).imag()};
                for (int c = 0; c <= kd; ++c)
                    for (int i = 0; i <= mm; ++i)
                        pu[(size_t)c * (mm + 1) + i] = cplx<double>{nd[r].U(i, c).real(), nd[r].U(i, c).imag()};
            }
            if (any) {
                hipCheck(hipMemcpyAsync(dP, ph.data(), ph.size() * sizeof(cplx<double>), hipMemcpyHostToDevice, st),
                         "restart upload");
                for (int r = 0; r < nrhs; ++r) {
                    if (!kept[r]) continue;
                    const int64_t len = LN.len[r];
                    const dim3 g((unsigned)((len + KB - 1) / KB));
                    const int e = embed[r] ? k0 : 0, mm = m - e;
                    const cplx<TQ>* q0 = Q + (size_t)e * ld + LN.base[r];
                    const cplx<double>* pk = dP + pSz * r;
                    hipLaunchKernelGGL(kRecombine<TQ>, g, dim3(KB), 0, st, q0, ld, mm, pk, mm, kd, Zk + LN.base[r],
                                       ld, len);
                    hipLaunchKernelGGL(kRecombine<TQ>, g, dim3(KB), 0, st, q0, ld, mm + 1, pk + (size_t)mm * kd,
                                       mm + 1, kd + 1, Vn + LN.base[r], ld, len);
                }
                for (int r = 0; r < nrhs; ++r) {
                    if (!kept[r]) continue;
                    for (int c = 0; c <= kd; ++c)
                        hipCheck(hipMemcpyAsync(Q + (size_t)c * ld + LN.base[r], Vn + (size_t)c * ld + LN.base[r],
                                                LN.len[r] * sizeof(cplx<TQ>), hipMemcpyDeviceToDevice, st),
                                 "restart basis");
                    dr[r] = std::move(nd[r]);
                }
                hipCheck(hipStreamSynchronize(st), "deflated restart");
            }
        }
    }
    for (int r = 0; r < nrhs; ++r)
        if (out[r].status == MLFMA_OK && out[r].est > tolv[r]) out[r].status = MLFMA_ERR_NOT_CONVERGED;
    for (void* p : {(void*)Q, (void*)W, (void*)X64, (void*)PT, (void*)Zk, (void*)Vn, (void*)dP})
        if (p) d.release(p);
    return out;
}

static std::vector<SolveOut> gmresBatch(DevicePlan& d, std::vector<std::unique_ptr<Krylov>>& K, hipStream_t st,
                                        const cplx<double>* B, cplx<double>* X, const Lanes& LN, int m, int maxit,
                                        const std::vector<double>& tolv, double* hist, bool fp32 = false,
                                        int histOff = 0, int histLd = 0, bool lastResidual = true,
                                        int deflate = 0) {
    if (fp32)
        return gmresBatchT<float>(d, K, st, B, X, LN, m, maxit, tolv, hist, fp32, histOff, histLd, lastResidual,
                                  deflate);
    return gmresBatchT<double>(d, K, st, B, X, LN, m, maxit, tolv, hist, fp32, histOff, histLd, lastResidual,
                               deflate);
}

// Batched iterative refinement (MLFMA_GMRES_IR): FP64 residuals of all
// right-hand sides (batched FP64 matvec) around batched inner GMRES(m) solves
// with the FP32 far field (at most 30 m steps each), at most 30 corrections;
// each right-hand side's inner tolerance is max(1e-4, tol / (2 residual)).
static std::vector<SolveOut> gmresBatchIR(DevicePlan& d, std::vector<std::unique_ptr<Krylov>>& K, hipStream_t st,
                                          const cplx<double>* B, cplx<double>* X, const Lanes& LN, int m,
                                          int maxit, double tol, double* hist, int deflate) {
    const int64_t n = d.hp.src.n;
    const int nrhs = LN.count, nvec = LN.nvec;
    const int64_t ld = (int64_t)nvec * n;
    cplx<double>* R = (cplx<double>*)d.alloc((size_t)ld * sizeof(cplx<double>));
    cplx<double>* D = (cplx<double>*)d.alloc((size_t)ld * sizeof(cplx<double>));
    std::vector<SolveOut> out(nrhs);
    std::vector<double> bn(nrhs);
    for (int r = 0; r < nrhs; ++r) bn[r] = K[r]->norm(B + LN.base[r]);
    hipCheck(hipMemsetAsync(X, 0, (size_t)ld * sizeof(cplx<double>), st), "zero x");
    int total = 0;
    for (int outer = 0;; ++outer) {
        // the residuals of the current iterates, also after the last correction
        if (outer == 0)
            hipCheck(hipMemcpyAsync(R, B, (size_t)ld * sizeof(cplx<double>), hipMemcpyDeviceToDevice, st), "copy");
        else {
            applyMany(d, X, R, nvec, st, false);
            for (int r = 0; r < nrhs; ++r) K[r]->sub(B + LN.base[r], R + LN.base[r], R + LN.base[r]);
        }
        bool all = true;
        std::vector<double> itol(nrhs, 1.0);
        for (int r = 0; r < nrhs; ++r) {
            const double res = bn[r] > 0 ? K[r]->norm(R + LN.base[r]) / bn[r] : 0.0;
            out[r].est = res;
            if (res <= tol) {  // converged: its inner right-hand side is zero
                hipCheck(hipMemsetAsync(R + LN.base[r], 0, LN.len[r] * sizeof(cplx<double>), st), "zero lane");
                continue;
            }
            all = false;
            itol[r] = std::max(1e-4, 0.5 * tol / res);
        }
        if (all || total >= maxit || outer == 30) break;
        // each correction's inner solve at most 30 m steps: a lane that
        // stalls then cannot spend the whole budget of the others' corrections
        std::vector<SolveOut> in = gmresBatch(d, K, st, R, D, LN, m, std::min(maxit - total, 30 * m), itol, hist,
                                              true, total, maxit, true, deflate);
        int steps = 0;
        for (int r = 0; r < nrhs; ++r) {
            out[r].iters += in[r].iters;
            steps = std::max(steps, in[r].iters);
            if (in[r].status == MLFMA_ERR_BREAKDOWN) out[r].status = MLFMA_ERR_BREAKDOWN;
            // X += D
            std::vector<zc> one{1.0};
            K[r]->combine(D + LN.base[r], 1, one, X + LN.base[r], 1);
        }
        total += std::max(steps, 1);
    }
    for (int r = 0; r < nrhs; ++r)
        if (out[r].status == MLFMA_OK && out[r].est > tol) out[r].status = MLFMA_ERR_NOT_CONVERGED;
    d.release(R);
    d.release(D);
    return out;
}

// Flexible GMRES (Saad 1993) in FP64 whose preconditioner is an inner GMRES
// solve with the FP32 far field: for each lane (right-hand side) the outer
// step j applies z_j = M_j^-1 v_j (inner GMRES(mIn) from zero to a relative
// innerTol, at most mIn iterations; the map changes from step to step, which
// is why the outer method stores the z_j), w = Z z_j with the FP64 matvec,
// classical Gram-Schmidt twice against v_0..v_j, and a unitary Givens
// rotation; x = x_0 + sum_j y_j z_j at the end of each cycle of mOut steps.
// Every cycle starts from the true FP64 residual, and the solve ends on it.
// One right-hand side runs the pipelined single-vector GMRES as the inner
// solver; several run the batched one (one batched matvec per inner step)
// with one batched FP64 matvec and one batched CGS2 per outer step.
// iterations counts matvecs: the inner FP32 ones and the outer FP64 ones.
static std::vector<SolveOut> fgmres(DevicePlan& d, std::vector<std::unique_ptr<Krylov>>& K, hipStream_t st,
                                    const cplx<double>* B, cplx<double>* X, const Lanes& LN, int mOut, int mIn,
                                    int maxit, double tol, double innerTol, double* hist, int histLd) {
    const int64_t n = d.hp.src.n;
    const int nrhs = LN.count, nvec = LN.nvec;
    const int64_t ld = (int64_t)nvec * n;
    mOut = std::max(1, mOut);
    mIn = std::max(1, mIn);
    cplx<double>* V = (cplx<double>*)d.alloc((size_t)ld * (mOut + 1) * sizeof(cplx<double>));
    cplx<double>* Zb = (cplx<double>*)d.alloc((size_t)ld * mOut * sizeof(cplx<double>));
    cplx<double>* W = (cplx<double>*)d.alloc((size_t)ld * sizeof(cplx<double>));
    std::unique_ptr<BatchOrtho> BO;
    if (nrhs > 1) BO.reset(new BatchOrtho(d, LN, mOut));
    std::vector<cplx<double>> hh(nrhs > 1 ? (size_t)nrhs * BO->hs : 0);
    std::vector<SolveOut> out(nrhs);
    std::vector<double> bn(nrhs);
    std::vector<int> used(nrhs, 0);  // iterations (matvecs) per lane
    for (int r = 0; r < nrhs; ++r) {
        bn[r] = K[r]->norm(B + LN.base[r]);
        if (!std::isfinite(bn[r])) throw Error(MLFMA_ERR_NONFINITE, "non-finite b");
    }
    hipCheck(hipMemsetAsync(X, 0, (size_t)ld * sizeof(cplx<double>), st), "zero x");
    // the FP64 operator (one right-hand side: Z or, for an adjoint solve, Z^H)
    auto op64 = [&](const cplx<double>* x, cplx<double>* y) {
        if (nrhs == 1)
            d.applySolveOp(x, y, 0, st, false, false);
        else
            applyMany(d, x, y, nvec, st, false);
    };
    int total = 0;  // matvecs of the lockstep (the most any lane has used)
    for (int cycle = 0;; ++cycle) {
        // true FP64 residuals R = B - Z X (into W)
        if (cycle == 0)
            hipCheck(hipMemcpyAsync(W, B, (size_t)ld * sizeof(cplx<double>), hipMemcpyDeviceToDevice, st), "copy");
        else {
            op64(X, W);
            for (int r = 0; r < nrhs; ++r) K[r]->sub(B + LN.base[r], W + LN.base[r], W + LN.base[r]);
        }
        std::vector<double> beta(nrhs, 0.0);
        std::vector<char> active(nrhs, 0);
        bool anyActive = false;
        for (int r = 0; r < nrhs; ++r) {
            if (bn[r] == 0 || out[r].status == MLFMA_ERR_BREAKDOWN) continue;
            beta[r] = cycle == 0 ? bn[r] : K[r]->norm(W + LN.base[r]);
            out[r].est = beta[r] / bn[r];
            active[r] = out[r].est > tol && beta[r] > 0;
            anyActive = anyActive || active[r];
        }
        if (!anyActive || total >= maxit) break;
        for (int r = 0; r < nrhs; ++r) {
            if (active[r])
                K[r]->scale(W + LN.base[r], beta[r], V + LN.base[r]);
            else
                hipCheck(hipMemsetAsync(V + LN.base[r], 0, LN.len[r] * sizeof(cplx<double>), st), "zero lane");
        }
        // per lane: Hessenberg columns (rotated), rotations, rotated rhs g
        std::vector<std::vector<std::vector<zc>>> H(nrhs);
        std::vector<std::vector<zc>> cs(nrhs), sn(nrhs), g(nrhs);
        std::vector<int> steps(nrhs, 0);
        for (int r = 0; r < nrhs; ++r) g[r] = {zc(beta[r], 0.0)};
        for (int j = 0; j < mOut && total < maxit; ++j) {
            // z_j = M^-1 v_j: inner FP32-far-field GMRES on every lane at once
            // (lanes that are done carry zero right-hand sides)
            cplx<double>* Vj = V + (size_t)j * ld;
            cplx<double>* Zj = Zb + (size_t)j * ld;
            const int budget = std::max(1, std::min(mIn, maxit - total - 1));
            int innerIts = 0;
            if (nrhs == 1) {
                SolveOut in = gmresRestart(d, *K[0], Vj, Zj, budget, budget, innerTol, true, true, hist, total, false);
                innerIts = in.iters;
                used[0] += in.iters;
            } else {
                std::vector<double> itol(nrhs, innerTol);
                std::vector<SolveOut> in =
                    gmresBatch(d, K, st, Vj, Zj, LN, budget, budget, itol, hist, true, total, histLd, false);
                for (int r = 0; r < nrhs; ++r) {
                    innerIts = std::max(innerIts, in[r].iters);
                    if (active[r]) used[r] += in[r].iters;
                }
            }
            total += innerIts;
            // w = Z z_j (FP64) and CGS2 against v_0..v_j
            op64(Zj, W);
            ++total;
            std::vector<std::vector<zc>> hcol(nrhs);
            std::vector<double> wn(nrhs, 0.0);
            if (nrhs == 1) {
                K[0]->orthogonalize(V, j + 1, W, hcol[0], wn[0]);
            } else {
                BO->queue(V, ld, j + 1, W, hh.data(), st);
                hipCheck(hipStreamSynchronize(st), "fgmres ortho");
                for (int r = 0; r < nrhs; ++r) Krylov::readOrtho(hh.data() + (size_t)r * BO->hs, j + 1, hcol[r], wn[r]);
            }
            bool any = false;
            for (int r = 0; r < nrhs; ++r) {
                cplx<double>* Vn = V + (size_t)(j + 1) * ld + LN.base[r];
                if (!active[r]) {
                    hipCheck(hipMemsetAsync(Vn, 0, LN.len[r] * sizeof(cplx<double>), st), "zero lane");
                    continue;
                }
                ++used[r];
                std::vector<zc>& h = hcol[r];
                h.push_back(zc(wn[r], 0.0));
                for (int i = 0; i < j; ++i) {  // previous rotations
                    const zc a = cs[r][i] * h[i] + sn[r][i] * h[i + 1];
                    h[i + 1] = -std::conj(sn[r][i]) * h[i] + cs[r][i] * h[i + 1];
                    h[i] = a;
                }
                const double na = std::abs(h[j]), t = std::hypot(na, wn[r]);
                zc c, s;
                if (t == 0) {
                    out[r].status = MLFMA_ERR_BREAKDOWN;
                    active[r] = 0;
                    hipCheck(hipMemsetAsync(Vn, 0, LN.len[r] * sizeof(cplx<double>), st), "zero lane");
                    continue;
                }
                if (na == 0) {
                    c = 0.0;
                    s = 1.0;
                } else {
                    c = na / t;
                    s = (h[j] / na) * wn[r] / t;
                }
                cs[r].push_back(c);
                sn[r].push_back(s);
                h[j] = c * h[j] + s * h[j + 1];
                h.pop_back();
                H[r].push_back(h);
                const zc gj = g[r][j];
                g[r][j] = c * gj;
                g[r].push_back(-std::conj(s) * gj);
                steps[r] = j + 1;
                const double est = std::abs(g[r][j + 1]) / bn[r];
                out[r].est = est;
                if (hist) hist[(size_t)r * histLd + std::min(total, std::max(histLd, 1)) - 1] = est;
                if (est <= tol || wn[r] == 0) {
                    active[r] = 0;
                    hipCheck(hipMemsetAsync(Vn, 0, LN.len[r] * sizeof(cplx<double>), st), "zero lane");
                } else {
                    K[r]->scale(W + LN.base[r], wn[r], Vn);
                    any = true;
                }
            }
            if (!any) break;
        }
        // x += Z y per lane
        for (int r = 0; r < nrhs; ++r)
            if (steps[r] > 0) {
                const std::vector<zc> y = backSolve(H[r], g[r], steps[r]);
                K[r]->combine(Zb + LN.base[r], steps[r], y, X + LN.base[r], 1, ld);
            }
    }
    for (int r = 0; r < nrhs; ++r) {
        out[r].iters = used[r];
        if (out[r].status == MLFMA_OK && out[r].est > tol) out[r].status = MLFMA_ERR_NOT_CONVERGED;
    }
    d.release(V);
    d.release(Zb);
    d.release(W);
    return out;
}

// Mixed precision by iterative refinement (MLFMA_GMRES_IR): FP64 residuals
// (FP64 matvec) around restarted inner GMRES(m) solves whose far field is
// FP32, at most 20 corrections; no outer Krylov space is kept.
static SolveOut gmresIR(DevicePlan& d, Krylov& K, const cplx<double>* b, cplx<double>* x, int m, int maxit, double tol,
                        double* hist, int deflate) {
    SolveOut out;
    const int64_t n = d.hp.src.n;
    cplx<double>* r = (cplx<double>*)d.alloc((size_t)n * sizeof(cplx<double>));
    cplx<double>* dx = (cplx<double>*)d.alloc((size_t)n * sizeof(cplx<double>));
    cplx<double>* ax = (cplx<double>*)d.alloc((size_t)n * sizeof(cplx<double>));
    const double bn = K.norm(b);
    hipCheck(hipMemsetAsync(x, 0, n * sizeof(cplx<double>), K.st), "zero x");
    int total = 0;
    double res = 1.0;
    const double inner = 1e-4;
    for (int outer = 0;; ++outer) {
        // the residual of the current iterate, also after the last correction
        if (outer == 0) {
            hipCheck(hipMemcpyAsync(r, b, n * sizeof(cplx<double>), hipMemcpyDeviceToDevice, K.st), "copy");
            res = 1.0;
        } else {
            d.applySolveOp(x, ax, 0, K.st, false, false);
            K.sub(b, ax, r);
            res = K.norm(r) / bn;
        }
        out.est = res;
        if (res <= tol || total >= maxit || outer == 20) break;
        const double itol = std::max(inner, 0.5 * tol / res);
        SolveOut s = gmresRestart(d, K, r, dx, m, maxit - total, itol, true, true, hist, total, true, deflate);
        total += s.iters;
        // x += dx
        std::vector<zc> one{1.0};
        K.combine(dx, 1, one, x, 1);
        if (s.status == MLFMA_ERR_BREAKDOWN) {
            out.status = s.status;
            break;
        }
    }
    out.iters = total;
    if (out.status == MLFMA_OK && res > tol) out.status = MLFMA_ERR_NOT_CONVERGED;
    d.release(r);
    d.release(dx);
    d.release(ax);
    return out;
}

// Seconds of `iters` GMRES(iters) iterations from x = 0 on the plan's
// excitation: what mlfma_autotune times to choose the stage order of
// one-right-hand-side solves.
double timeSolveSteps(DevicePlan& d, int iters) {
    const int64_t n = d.hp.src.n;
    hipStream_t st = d.stream;
    d.ensureRhs(1);
    std::vector<cplx<double>> bh(n);
    for (int64_t i = 0; i < n; ++i) bh[i] = cplx<double>{-d.Vin[i].real(), -d.Vin[i].imag()};
    cplx<double>* bt = (cplx<double>*)d.alloc(n * sizeof(cplx<double>));
    cplx<double>* xt = (cplx<double>*)d.alloc(n * sizeof(cplx<double>));
    hipCheck(hipMemcpyAsync(bt, bh.data(), n * sizeof(cplx<double>), hipMemcpyHostToDevice, st), "b upload");
    double secs = 0;
    {
        Krylov K(d, st, iters + 2);
        hipCheck(hipStreamSynchronize(st), "autotune");
        const auto t0 = std::chrono::steady_clock::now();
        // the steps the plan's one-vector solves run: with the FP32 far
        // field on an FP32 plan (the inner solves of the mixed method)
        gmresRestart(d, K, bt, xt, iters, iters, 1e-300, d.useFp32, true, nullptr, 0);
        hipCheck(hipStreamSynchronize(st), "autotune");
        secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    d.release(bt);
    d.release(xt);
    return secs;
}

}  // namespace mlfma

using namespace mlfma;

#define API_BEGIN try {
#define API_END                                  \
    }                                            \
    catch (const Error& e) {                     \
        g_lastError = e.what();                  \
        return e.code;                           \
    }                                            \
    catch (const std::bad_alloc&) {              \
        g_lastError = "host allocation failed";  \
        return MLFMA_ERR_OOM;                    \
    }                                            \
    catch (const std::exception& e) {            \
        g_lastError = e.what();                  \
        return MLFMA_ERR_ARG;                    \
    }

static DevicePlan& Pl(mlfma_plan* p) {
    if (!p || !p->d) throw Error(MLFMA_ERR_ARG, "null plan");
    if (p->d->hostOnly) throw Error(MLFMA_ERR_STATE, "host-only plan (device = -1) has no device data");
    return *p->d;
}

extern "C" {

int mlfma_solve(mlfma_plan* plan, const double* b, double* x, const mlfma_solve_opts* opts, mlfma_solve_report* rep,
                double* hist) {
    API_BEGIN
    DevicePlan& d = Pl(plan);
    if (d.hp.forest)
        throw Error(MLFMA_ERR_STATE, "a forest plan solves every scatterer separately: use mlfma_solve_batch");
    if (!x || !opts) throw Error(MLFMA_ERR_ARG, "null argument");
    if (!opts->use_voltage && !b) throw Error(MLFMA_ERR_ARG, "b is NULL and use_voltage = 0");
    if (opts->maxit < 1 || !(opts->tol > 0)) throw Error(MLFMA_ERR_ARG, "maxit must be >= 1 and tol > 0");
    if (opts->restart > 2048) throw Error(MLFMA_ERR_ARG, "restart must be <= 2048");
    const int method = opts->method & 0xff;
    if (opts->deflate < 0 || opts->deflate > 64) throw Error(MLFMA_ERR_ARG, "deflate must be in [0, 64]");
    if (opts->deflate && method != MLFMA_GMRES_RESTART && method != MLFMA_GMRES_IR)
        throw Error(MLFMA_ERR_ARG, "deflated restarts are for MLFMA_GMRES_RESTART and MLFMA_GMRES_IR");
    if (opts->deflate && opts->deflate + 1 >= (opts->restart > 0 ? opts->restart : 50))
        throw Error(MLFMA_ERR_ARG, "deflate must be smaller than restart - 1");
    struct AdjointMode {  // cleared however the solve ends
        DevicePlan& d;
        ~AdjointMode() {
            d.solveAdjoint = false;
            d.solveInOrder = false;
            d.solvePrec = false;
            d.solveShift = cplx<double>{0, 0};
        }
    } adjointMode{d};
    if (!std::isfinite(opts->shift_re) || !std::isfinite(opts->shift_im))
        throw Error(MLFMA_ERR_NONFINITE, "non-finite shift");
    d.solveShift = cplx<double>{opts->shift_re, opts->shift_im};
    if (opts->precond < 0 || opts->precond > 1) throw Error(MLFMA_ERR_ARG, "unknown preconditioner");
    if (opts->precond && method == MLFMA_GMRES_COMPAT)
        throw Error(MLFMA_ERR_ARG, "the reference-compatible GMRES takes no preconditioner");
    if (opts->precond) {
        hipCheck(hipSetDevice(d.device), "hipSetDevice");
        d.buildPrecond();
        d.solvePrec = true;
    }
    d.solveAdjoint = (opts->method & MLFMA_SOLVE_ADJOINT) != 0;
    // one right-hand side: the matvecs run their stages in order when the
    // autotuner found that overlap does not pay (2.7% per iteration at 30k)
    d.solveInOrder = d.solveSerial;
    if (opts->method & ~(0xff | MLFMA_SOLVE_ADJOINT | MLFMA_SOLVE_DEVICE))
        throw Error(MLFMA_ERR_ARG, "unknown solve method bits");
    const bool devPtrs = (opts->method & MLFMA_SOLVE_DEVICE) != 0;
    if ((method == MLFMA_FGMRES_IR || method == MLFMA_GMRES_IR) && !d.useFp32)
        throw Error(MLFMA_ERR_STATE, "mixed-precision solves need an FP32 plan");
    hipCheck(hipSetDevice(d.device), "hipSetDevice");
    const auto t0 = std::chrono::steady_clock::now();
    const int64_t n = d.hp.src.n;
    hipStream_t st = d.stream;
    d.ensureRhs(1);
    // Krylov vectors stay in input order: the matvec kernels apply the tree
    // permutation themselves
    cplx<double>* bt = (cplx<double>*)d.alloc(n * sizeof(cplx<double>));
    cplx<double>* xtree = (cplx<double>*)d.alloc(n * sizeof(cplx<double>));
    if (devPtrs && !opts->use_voltage) {
        // b on the device: ready when the call is made (the solve runs on the
        // plan's stream); a non-finite b shows as a non-finite norm
        hipCheck(hipMemcpyAsync(bt, b, n * sizeof(cplx<double>), hipMemcpyDeviceToDevice, st), "b copy");
    } else {
        std::vector<cplx<double>> bh(n);
        for (int64_t i = 0; i < n; ++i) {
            if (opts->use_voltage) {
                bh[i].re = -d.Vin[i].real();
                bh[i].im = -d.Vin[i].imag();
            } else {
                bh[i].re = b[2 * i];
                bh[i].im = b[2 * i + 1];
            }
            if (!std::isfinite(bh[i].re) || !std::isfinite(bh[i].im)) {
                d.release(bt);
                d.release(xtree);
                throw Error(MLFMA_ERR_NONFINITE, "non-finite b");
            }
        }
        hipCheck(hipMemcpyAsync(bt, bh.data(), n * sizeof(cplx<double>), hipMemcpyHostToDevice, st), "b upload");
    }
    const int mOut = opts->outer_restart > 0 ? opts->outer_restart : 30;
    Krylov K(d, st, std::max(std::max(opts->maxit, opts->restart), mOut) + 2);
    SolveOut o;
    if (method == MLFMA_GMRES_COMPAT)
        o = gmresCompat(d, K, bt, xtree, opts->maxit, opts->tol, hist);
    else if (method == MLFMA_GMRES_RESTART)
        o = gmresRestart(d, K, bt, xtree, opts->restart > 0 ? opts->restart : 50, opts->maxit, opts->tol, false, true,
                         hist, 0, true, opts->deflate);
    else if (method == MLFMA_FGMRES_IR) {
        std::vector<std::unique_ptr<Krylov>> Kv;
        Kv.emplace_back(new Krylov(d, st, std::max(opts->restart > 0 ? opts->restart : 50, mOut) + 2));
        const Lanes one(d, 1);
        o = fgmres(d, Kv, st, bt, xtree, one, mOut, opts->restart > 0 ? opts->restart : 50, opts->maxit, opts->tol,
                   opts->inner_tol > 0 ? opts->inner_tol : 1e-1, hist, opts->maxit)[0];
    } else if (method == MLFMA_GMRES_IR)
        o = gmresIR(d, K, bt, xtree, opts->restart > 0 ? opts->restart : 50, opts->maxit, opts->tol, hist,
                    opts->deflate);
    else
        throw Error(MLFMA_ERR_ARG, "unknown solve method");
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    // true residual with the FP64 matvec
    d.applySolveOp(xtree, d.yt, 0, st, false, false);
    K.sub(bt, d.yt, d.yt);
    const double rn = K.norm(d.yt), bn = K.norm(bt);
    hipCheck(hipMemcpyAsync(x, xtree, n * sizeof(cplx<double>), devPtrs ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                            st),
             "x download");
    hipCheck(hipStreamSynchronize(st), "solve");
    d.release(bt);
    d.release(xtree);
    if (rep) {
        rep->iterations = o.iters;
        rep->status = o.status;
        rep->resid_est = o.est;
        rep->resid_true = bn > 0 ? rn / bn : 0.0;
        rep->seconds = secs;
    }
    if (o.status == MLFMA_ERR_BREAKDOWN) {
        g_lastError = "GMRES breakdown: a zero rotation (the Krylov space stopped growing); x and the report hold "
                      "the last iterate";
        return MLFMA_ERR_BREAKDOWN;
    }
    return MLFMA_OK;
    API_END
}

int mlfma_farfield(mlfma_plan* plan, const double* I, const double* dirs, int64_t ndir, double* E, int32_t ptr_mode,
                   void* stream) {
    API_BEGIN
    DevicePlan& d = Pl(plan);
    if (!I || !dirs || !E || ndir < 0) throw Error(MLFMA_ERR_ARG, "bad farfield arguments");
    if (ptr_mode != MLFMA_PTR_HOST && ptr_mode != MLFMA_PTR_DEVICE) throw Error(MLFMA_ERR_ARG, "bad ptr_mode");
    if (ndir > INT32_MAX) throw Error(MLFMA_ERR_ARG, "too many directions");
    if (ndir == 0) return MLFMA_OK;
    hipCheck(hipSetDevice(d.device), "hipSetDevice");
    const int64_t n = d.hp.src.n;
    hipStream_t st = stream ? (hipStream_t)stream : d.stream;
    const size_t need = (size_t)ndir * farfieldChunks(n, ndir) * 3;
    if (need > d.farPartLen) {
        if (d.farPart) {
            hipCheck(hipStreamSynchronize(st), "farfield");
            d.release(d.farPart);
        }
        d.farPart = (cplx<double>*)d.alloc(need * sizeof(cplx<double>));
        d.farPartLen = need;
    }
    if (ptr_mode == MLFMA_PTR_DEVICE) {
        launchFarfield(d.src, n, (const cplx<double>*)I, d.perm, dirs, ndir, d.hp.k, d.Ck, d.farPart,
                       (cplx<double>*)E, st);
        hipCheck(hipGetLastError(), "farfield launch");
        if (!(d.flags & MLFMA_FLAG_ASYNC)) hipCheck(hipStreamSynchronize(st), "farfield");
        return MLFMA_OK;
    }
    d.ensureRhs(1);
    hipCheck(hipMemcpyAsync(d.xs, I, n * sizeof(cplx<double>), hipMemcpyHostToDevice, st), "I upload");
    double* dd = (double*)d.alloc((size_t)ndir * 3 * sizeof(double));
    cplx<double>* dE = (cplx<double>*)d.alloc((size_t)ndir * 3 * sizeof(cplx<double>));
    hipCheck(hipMemcpyAsync(dd, dirs, (size_t)ndir * 3 * sizeof(double), hipMemcpyHostToDevice, st), "dirs upload");
    launchFarfield(d.src, n, d.xs, d.perm, dd, ndir, d.hp.k, d.Ck, d.farPart, dE, st);
    hipCheck(hipGetLastError(), "farfield launch");
    hipCheck(hipMemcpyAsync(E, dE, (size_t)ndir * 3 * sizeof(cplx<double>), hipMemcpyDeviceToHost, st), "E download");
    hipCheck(hipStreamSynchronize(st), "farfield");
    d.release(dd);
    d.release(dE);
    return MLFMA_OK;
    API_END
}

int mlfma_solve_batch(mlfma_plan* plan, const double* b, double* x, int32_t nrhs, const mlfma_solve_opts* opts,
                      mlfma_solve_report* reps, double* hist) {
    API_BEGIN
    DevicePlan& d = Pl(plan);
    if (!b || !x || !opts || nrhs < 1) throw Error(MLFMA_ERR_ARG, "bad batch solve arguments");
    if (opts->deflate < 0 || opts->deflate > 64) throw Error(MLFMA_ERR_ARG, "deflate must be in [0, 64]");
    if (opts->deflate && opts->method == MLFMA_FGMRES_IR)
        throw Error(MLFMA_ERR_ARG, "deflated restarts are for MLFMA_GMRES_RESTART and MLFMA_GMRES_IR");
    if (opts->deflate && opts->deflate + 1 >= (opts->restart > 0 ? opts->restart : 50))
        throw Error(MLFMA_ERR_ARG, "deflate must be smaller than restart - 1");
    if (opts->method != MLFMA_GMRES_RESTART && opts->method != MLFMA_FGMRES_IR && opts->method != MLFMA_GMRES_IR)
        throw Error(MLFMA_ERR_ARG, "batched solves use MLFMA_GMRES_RESTART, MLFMA_FGMRES_IR or MLFMA_GMRES_IR");
    if (opts->method != MLFMA_GMRES_RESTART && !d.useFp32)
        throw Error(MLFMA_ERR_STATE, "mixed-precision solves need an FP32 plan");
    if (opts->maxit < 1 || !(opts->tol > 0)) throw Error(MLFMA_ERR_ARG, "maxit must be >= 1 and tol > 0");
    if (opts->restart > 2048) throw Error(MLFMA_ERR_ARG, "restart must be <= 2048");
    if (d.hp.staleBoxes > 0 && !(d.flags & MLFMA_FLAG_FRESH_LOCALS))
        throw Error(MLFMA_ERR_STATE, "batched solves need a linear operator (MLFMA_FLAG_FRESH_LOCALS)");
    if (!std::isfinite(opts->shift_re) || !std::isfinite(opts->shift_im))
        throw Error(MLFMA_ERR_NONFINITE, "non-finite shift");
    if (opts->precond < 0 || opts->precond > 1) throw Error(MLFMA_ERR_ARG, "unknown preconditioner");
    const int64_t n = d.hp.src.n;
    for (size_t i = 0; i < (size_t)n * nrhs * 2; ++i)
        if (!std::isfinite(b[i])) throw Error(MLFMA_ERR_NONFINITE, "non-finite b");
    hipCheck(hipSetDevice(d.device), "hipSetDevice");
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = d.stream;
    struct SolveMode {  // cleared however the solve ends
        DevicePlan& d;
        ~SolveMode() {
            d.solvePrec = false;
            d.solveShift = cplx<double>{0, 0};
        }
    } solveMode{d};
    d.solveShift = cplx<double>{opts->shift_re, opts->shift_im};
    if (opts->precond) {
        d.buildPrecond();
        d.solvePrec = true;
    }
    d.ensureBatch(nrhs >= 8 ? 8 : (nrhs >= 4 ? 4 : (nrhs >= 2 ? 2 : 1)));
    // one lane per right-hand side and scatterer (segment) of the plan
    const Lanes LN(d, nrhs);
    const size_t bytes = (size_t)n * nrhs * sizeof(cplx<double>);
    cplx<double>* B = (cplx<double>*)d.alloc(bytes);
    cplx<double>* X = (cplx<double>*)d.alloc(bytes);
    cplx<double>* R = (cplx<double>*)d.alloc(bytes);
    hipCheck(hipMemcpyAsync(B, b, bytes, hipMemcpyHostToDevice, st), "b upload");
    const int m = opts->restart > 0 ? opts->restart : 50;
    const int mOut = opts->outer_restart > 0 ? opts->outer_restart : 30;
    std::vector<std::unique_ptr<Krylov>> K;
    for (int k = 0; k < LN.count; ++k)
        K.emplace_back(new Krylov(d, st, std::max(std::min(m, opts->maxit), mOut) + 2, LN.len[k]));
    std::vector<SolveOut> o =
        opts->method == MLFMA_FGMRES_IR
            ? fgmres(d, K, st, B, X, LN, mOut, m, opts->maxit, opts->tol,
                     opts->inner_tol > 0 ? opts->inner_tol : 1e-1, hist, opts->maxit)
        : opts->method == MLFMA_GMRES_IR
            ? gmresBatchIR(d, K, st, B, X, LN, m, opts->maxit, opts->tol, hist, opts->deflate)
            : gmresBatch(d, K, st, B, X, LN, m, opts->maxit, std::vector<double>(LN.count, opts->tol), hist, false,
                         0, 0, true, opts->deflate);
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    // true residuals with the FP64 matvec, per lane
    applyMany(d, X, R, nrhs, st);
    for (int k = 0; k < LN.count; ++k) {
        K[k]->sub(B + LN.base[k], R + LN.base[k], R + LN.base[k]);
        const double rn = K[k]->norm(R + LN.base[k]), bn = K[k]->norm(B + LN.base[k]);
        if (reps) {
            reps[k].iterations = o[k].iters;
            reps[k].status = o[k].status;
            reps[k].resid_est = o[k].est;
            reps[k].resid_true = bn > 0 ? rn / bn : 0.0;
            reps[k].seconds = secs;
        }
    }
    hipCheck(hipMemcpyAsync(x, X, bytes, hipMemcpyDeviceToHost, st), "x download");
    hipCheck(hipStreamSynchronize(st), "batch solve");
    K.clear();
    d.release(B);
    d.release(X);
    d.release(R);
    int worst = MLFMA_OK;
    for (auto& e : o)
        if (e.status != MLFMA_OK) worst = e.status;
    return worst == MLFMA_ERR_NOT_CONVERGED ? MLFMA_OK : worst;
    API_END
}

// The far field of mlfma_farfield on device pointers without allocating:
// the source-range partial sums go to this device's fixed array (kFarPartCap
// triples), the same kernels and chunking as mlfma_farfield.  Nothing is
// allocated, freed or synchronised: the work is queued on the stream.
int mlfma_farfield_noalloc(mlfma_plan* plan, const double* I, const double* dirs, int64_t ndir, double* E,
                           void* stream) {
    API_BEGIN
    DevicePlan& d = Pl(plan);
    if (!I || !dirs || !E || ndir < 0) throw Error(MLFMA_ERR_ARG, "bad farfield arguments");
    if (ndir > kFarPartCap)
        throw Error(MLFMA_ERR_ARG, "mlfma_farfield_noalloc takes at most " + std::to_string(kFarPartCap) +
                                       " directions (its partial-sum array is fixed); use mlfma_farfield");
    if (ndir == 0) return MLFMA_OK;
    hipCheck(hipSetDevice(d.device), "hipSetDevice");
    cplx<double>* part = farfieldScratch(d.device);
    if (!part) throw Error(MLFMA_ERR_STATE, "the far-field partial-sum array of this device is not set up");
    const int64_t n = d.hp.src.n;
    if ((int64_t)farfieldChunks(n, ndir) * ndir > kFarPartCap)
        throw Error(MLFMA_ERR_STATE, "far-field chunking exceeds the partial-sum array");
    hipStream_t st = stream ? (hipStream_t)stream : d.stream;
    launchFarfield(d.src, n, (const cplx<double>*)I, d.perm, dirs, ndir, d.hp.k, d.Ck, part, (cplx<double>*)E, st);
    hipCheck(hipGetLastError(), "farfield launch");
    return MLFMA_OK;
    API_END
}

int mlfma_farfield_adjoint(mlfma_plan* plan, const double* E, const double* dirs, int64_t ndir, double* I,
                           int32_t ptr_mode, void* stream) {
    API_BEGIN
    DevicePlan& d = Pl(plan);
    if (!I || !dirs || !E || ndir < 0) throw Error(MLFMA_ERR_ARG, "bad farfield arguments");
    if (ptr_mode != MLFMA_PTR_HOST && ptr_mode != MLFMA_PTR_DEVICE) throw Error(MLFMA_ERR_ARG, "bad ptr_mode");
    if (ndir > INT32_MAX) throw Error(MLFMA_ERR_ARG, "too many directions");
    hipCheck(hipSetDevice(d.device), "hipSetDevice");
    const int64_t n = d.hp.src.n;
    hipStream_t st = stream ? (hipStream_t)stream : d.stream;
    const cplx<double> Ckc{d.Ck.re, -d.Ck.im};
    if (ptr_mode == MLFMA_PTR_DEVICE) {
        launchFarfieldAdj(d.src, n, (const cplx<double>*)E, dirs, (int)ndir, d.hp.k, Ckc, d.perm, (cplx<double>*)I,
                          st);
        hipCheck(hipGetLastError(), "farfield adjoint launch");
        if (!(d.flags & MLFMA_FLAG_ASYNC)) hipCheck(hipStreamSynchronize(st), "farfield adjoint");
        return MLFMA_OK;
    }
    d.ensureRhs(1);
    const int64_t nd = std::max<int64_t>(ndir, 1);
    double* dd = (double*)d.alloc((size_t)nd * 3 * sizeof(double));
    cplx<double>* dE = (cplx<double>*)d.alloc((size_t)nd * 3 * sizeof(cplx<double>));
    if (ndir > 0) {
        hipCheck(hipMemcpyAsync(dd, dirs, (size_t)ndir * 3 * sizeof(double), hipMemcpyHostToDevice, st), "dirs upload");
        hipCheck(hipMemcpyAsync(dE, E, (size_t)ndir * 3 * sizeof(cplx<double>), hipMemcpyHostToDevice, st), "E upload");
    }
    launchFarfieldAdj(d.src, n, dE, dd, (int)ndir, d.hp.k, Ckc, d.perm, d.xs, st);
    hipCheck(hipGetLastError(), "farfield adjoint launch");
    hipCheck(hipMemcpyAsync(I, d.xs, n * sizeof(cplx<double>), hipMemcpyDeviceToHost, st), "I download");
    hipCheck(hipStreamSynchronize(st), "farfield adjoint");
    d.release(dd);
    d.release(dE);
    return MLFMA_OK;
    API_END
}

// grad[i][a] = -2 Re(lambda^H (dZ/dr_i^a) I) for dipole sources.  Z = near +
// C k R^H A S depends on the positions through the near entries and the
// patterns rp_s[d] = P_d phat_s e^{i k khat_d.(c - r_s)} of R and S (the
// translations depend on box centres only), so with L = A S I the forward
// locals and M' = A^H R lambda the adjoint multipoles,
//   row part    conj(lambda_i) C k (ik)  sum_d khat_d conj(rp_i[d]) . L[d]
//   column part I_i C k (-ik) conj(sum_d khat_d conj(rp_i[d]) . M'[d])
// and the near part from the Green's-function gradient of each stored entry.
int mlfma_position_gradient(mlfma_plan* plan, const double* I, const double* lambda, double* grad) {
    API_BEGIN
    DevicePlan& d = Pl(plan);
    if (!I || !lambda || !grad) throw Error(MLFMA_ERR_ARG, "null argument");
    if (d.hp.src.kind != MLFMA_SRC_DIPOLE || d.src.physical)
        throw Error(MLFMA_ERR_ARG, "position gradients are provided for dipole sources only");
    if (d.hp.staleBoxes > 0 && !(d.flags & MLFMA_FLAG_FRESH_LOCALS))
        throw Error(MLFMA_ERR_STATE,
                    "the operator keeps stale locals (boxes with an empty interaction list) and is not linear; "
                    "set MLFMA_FLAG_FRESH_LOCALS for a position gradient");
    hipCheck(hipSetDevice(d.device), "hipSetDevice");
    const int64_t n = d.hp.src.n;
    hipStream_t st = d.stream;
    d.ensureRhs(1);
    cplx<double>* dI = (cplx<double>*)d.alloc((size_t)n * sizeof(cplx<double>));
    cplx<double>* dLam = (cplx<double>*)d.alloc((size_t)n * sizeof(cplx<double>));
    cplx<double>* dY = (cplx<double>*)d.alloc((size_t)n * sizeof(cplx<double>));
    cplx<double>* gRow = (cplx<double>*)d.alloc((size_t)n * 3 * sizeof(cplx<double>));
    cplx<double>* gCol = (cplx<double>*)d.alloc((size_t)n * 3 * sizeof(cplx<double>));
    double* gNear = (double*)d.alloc((size_t)n * 3 * sizeof(double));
    std::vector<int64_t> ip(n);
    for (int64_t t = 0; t < n; ++t) ip[d.hp.tree.perm[t]] = t;
    int64_t* iperm = (int64_t*)d.alloc((size_t)n * sizeof(int64_t));
    hipCheck(hipMemcpyAsync(iperm, ip.data(), n * sizeof(int64_t), hipMemcpyHostToDevice, st), "iperm upload");
    hipCheck(hipMemcpyAsync(dI, I, n * sizeof(cplx<double>), hipMemcpyHostToDevice, st), "I upload");
    hipCheck(hipMemcpyAsync(dLam, lambda, n * sizeof(cplx<double>), hipMemcpyHostToDevice, st), "lambda upload");
    hipCheck(hipMemsetAsync(gRow, 0, (size_t)n * 3 * sizeof(cplx<double>), st), "clear");
    hipCheck(hipMemsetAsync(gCol, 0, (size_t)n * 3 * sizeof(cplx<double>), st), "clear");
    hipCheck(hipMemsetAsync(gNear, 0, (size_t)n * 3 * sizeof(double), st), "clear");
    d.applyOp(dI, dY, 0, st, false, false);  // forward locals L(I) stay in the local field
    launchL2TGrad(d.farRowPos, d.farRowItem, d.nFarRows, d.leafItems, d.levelDev, d.rp64, d.L64, d.perm, gRow, st);
    d.applyOpAdjoint(dLam, dY, 0, st, false);  // adjoint multipoles M'(lambda) stay in the multipole field
    launchL2TGrad(d.farRowPos, d.farRowItem, d.nFarRows, d.leafItems, d.levelDev, d.rp64, d.M64, d.perm, gCol, st);
    launchNearGradDip(d.src, d.nearItems, d.nNearItems, d.nearColsIn, d.perm, iperm, dI, dLam, d.hp.k, d.Ck, gNear, st);
    hipCheck(hipGetLastError(), "position gradient launch");
    std::vector<std::complex<double>> row(3 * n), col(3 * n);
    std::vector<double> nearg(3 * n);
    hipCheck(hipMemcpyAsync(row.data(), gRow, 3 * n * sizeof(cplx<double>), hipMemcpyDeviceToHost, st), "download");
    hipCheck(hipMemcpyAsync(col.data(), gCol, 3 * n * sizeof(cplx<double>), hipMemcpyDeviceToHost, st), "download");
    hipCheck(hipMemcpyAsync(nearg.data(), gNear, 3 * n * sizeof(double), hipMemcpyDeviceToHost, st), "download");
    hipCheck(hipStreamSynchronize(st), "position gradient");
    for (void* p : {(void*)dI, (void*)dLam, (void*)dY, (void*)gRow, (void*)gCol, (void*)gNear, (void*)iperm})
        d.release(p);
    const std::complex<double> Ck(d.Ck.re, d.Ck.im), ik(0.0, d.hp.k);
    const std::complex<double>* Ih = (const std::complex<double>*)I;
    const std::complex<double>* Lh = (const std::complex<double>*)lambda;
    for (int64_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            const std::complex<double> v =
                std::conj(Lh[i]) * Ck * ik * row[3 * i + a] - Ih[i] * Ck * ik * std::conj(col[3 * i + a]);
            grad[3 * i + a] = nearg[3 * i + a] - 2.0 * v.real();
        }
    return MLFMA_OK;
    API_END
}

int mlfma_matvec_direct(mlfma_plan* plan, const double* x, double* y, int32_t ptr_mode, void* stream) {
    API_BEGIN
    DevicePlan& d = Pl(plan);
    if (!x || !y) throw Error(MLFMA_ERR_ARG, "null argument");
    hipCheck(hipSetDevice(d.device), "hipSetDevice");
    const int64_t n = d.hp.src.n;
    hipStream_t st = stream ? (hipStream_t)stream : d.stream;
    d.ensureRhs(1);
    const cplx<double>* xin = (const cplx<double>*)x;
    cplx<double>* yout = (cplx<double>*)y;
    if (ptr_mode == MLFMA_PTR_HOST) {
        hipCheck(hipMemcpyAsync(d.xs, x, n * sizeof(cplx<double>), hipMemcpyHostToDevice, st), "x upload");
        xin = d.xs;
        yout = d.ys;
    } else if (ptr_mode != MLFMA_PTR_DEVICE)
        throw Error(MLFMA_ERR_ARG, "bad ptr_mode");
    launchGather(xin, d.perm, d.xt, n, 1, st);
    launchDirect(d.src, d.xt, n, d.hp.k, d.Ck, d.yt, st);
    hipCheck(hipGetLastError(), "direct launch");
    launchScatter(d.yt, d.perm, yout, n, 1, 0, st);
    if (ptr_mode == MLFMA_PTR_HOST)
        hipCheck(hipMemcpyAsync(y, d.ys, n * sizeof(cplx<double>), hipMemcpyDeviceToHost, st), "y download");
    hipCheck(hipStreamSynchronize(st), "direct");
    return MLFMA_OK;
    API_END
}

}  // extern "C"
