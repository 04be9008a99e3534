// cplx_ops.hpp — device complex arithmetic on interleaved (re, im) pairs.
#pragma once

#include "kernels.hpp"

namespace mlfma {

template <typename R>
__device__ __forceinline__ cplx<R> mk(R a, R b) {
    cplx<R> c;
    c.re = a;
    c.im = b;
    return c;
}
template <typename R>
__device__ __forceinline__ cplx<R> operator+(cplx<R> a, cplx<R> b) { return mk(a.re + b.re, a.im + b.im); }
template <typename R>
__device__ __forceinline__ cplx<R> operator-(cplx<R> a, cplx<R> b) { return mk(a.re - b.re, a.im - b.im); }
template <typename R>
__device__ __forceinline__ cplx<R> operator*(cplx<R> a, cplx<R> b) {
    return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename R>
__device__ __forceinline__ cplx<R> operator*(R s, cplx<R> a) { return mk(s * a.re, s * a.im); }
template <typename R>
__device__ __forceinline__ cplx<R> operator*(cplx<R> a, R s) { return mk(s * a.re, s * a.im); }
template <typename R>
__device__ __forceinline__ cplx<R> conjmul(cplx<R> a, cplx<R> b) {  // conj(a) * b
    return mk(a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re);
}
template <typename R>
__device__ __forceinline__ cplx<R> cdiv(cplx<R> a, R s) { return mk(a.re / s, a.im / s); }
__device__ __forceinline__ cplx<double> cexpi(double a) {
    double s, c;
    sincos(a, &s, &c);
    return mk(c, s);
}
template <typename R, typename S>
__device__ __forceinline__ cplx<R> cast(cplx<S> a) { return mk((R)a.re, (R)a.im); }


__device__ __forceinline__ double wsum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wsum(float v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block index -> work item such that the items of each XCD are contiguous:
// workgroups are dispatched round-robin over the 8 XCDs, so block b runs on
// XCD b % 8; with a grid of 8 * per blocks XCD x gets items [x*per, (x+1)*per).
__device__ __forceinline__ int xcdItem(int b, int per) { return (b & 7) * per + (b >> 3); }

// Device-coherent loads and stores (relaxed atomics at agent scope: on
// gfx950 they carry sc1 and go past the XCD's own L2, which is not coherent
// with the other XCDs' L2), for values handed from one block to another
// within a launch without agent-scope fences, which write back and
// invalidate the whole L2.
template <typename R, bool COH>
__device__ __forceinline__ cplx<R> ldc(const cplx<R>* p) {
    if constexpr (COH)
        return mk<R>(__hip_atomic_load(&p->re, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                     __hip_atomic_load(&p->im, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    else
        return *p;
}
template <typename R, bool COH>
__device__ __forceinline__ void stc(cplx<R>* p, cplx<R> v) {
    if constexpr (COH) {
        __hip_atomic_store(&p->re, v.re, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&p->im, v.im, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        *p = v;
    }
}

// The value lane l (wave-uniform) holds, read by every lane (two v_readlane).
__device__ __forceinline__ double readlaneD(double v, int l) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffull), l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

}  // namespace mlfma
