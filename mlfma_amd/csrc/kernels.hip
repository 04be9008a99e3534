// kernels.hip — HIP kernels for gfx950 (MI355X) of the MLFMA matvec and its
// setup generators.  Wave64 throughout; every kernel is a streaming pass whose
// roofline is HBM (or L2 for the M2L source reads), see the launchers for the
// per-item work each block does.
#include <algorithm>
#include <mutex>
#include <vector>

#include "kernels.hpp"
#include "cplx_ops.hpp"

namespace mlfma {

constexpr int BS = 256;
constexpr double kPiD = 3.14159265358979323846;

// ------------------------------------------------------------------ source kernels

// Closed-form integral of e^{i k.r} times the RWG basis over its two triangles
// (the analytic branch of src/sources/rwg.cpp:46-86, including its small-angle
// branches at |alpha|, |beta|, |gamma| < 1e-6).
__device__ void rwgPlaneWave(const RwgDev& r, const TriDev* tris, const double kv[3], cplx<double> out[3]) {
    const cplx<double> I = mk(0.0, 1.0);
    out[0] = out[1] = out[2] = mk(0.0, 0.0);
    for (int t = 0; t < 2; ++t) {
        const TriDev& T = tris[r.tri[t]];
        const double alpha = kv[0] * T.D0[0] + kv[1] * T.D0[1] + kv[2] * T.D0[2];
        const double beta = -(kv[0] * T.D2[0] + kv[1] * T.D2[1] + kv[2] * T.D2[2]);
        const double gamma = alpha - beta;
        const double asq = alpha * alpha, bsq = beta * beta;
        const cplx<double> eA = cexpi(alpha), eB = cexpi(beta);
        const bool za = fabs(alpha) < 1e-6, zb = fabs(beta) < 1e-6, zg = fabs(gamma) < 1e-6;
        const cplx<double> one = mk(1.0, 0.0);
        const cplx<double> f0a = za ? mk(0.0, -1.0) : cdiv(one - eA, alpha);
        const cplx<double> f0b = zb ? mk(0.0, -1.0) : cdiv(one - eB, beta);
        const cplx<double> f1a = za ? mk(-0.5, 0.0) : cdiv(one - mk(1.0, -alpha) * eA, asq);
        const cplx<double> f1b = zb ? mk(-0.5, 0.0) : cdiv(one - mk(1.0, -beta) * eB, bsq);
        double A[3];
        for (int i = 0; i < 3; ++i) A[i] = T.X0[i] - r.xpm[t][i];
        cplx<double> rv[3];
        if (zg) {
            const cplx<double> f2 = cdiv(eA * mk(asq - 2.0, 2.0 * alpha) + mk(2.0, 0.0), 2.0 * alpha * asq);
            const cplx<double> if2 = I * f2;
            for (int i = 0; i < 3; ++i) rv[i] = mk(-f1a.re * A[i], -f1a.im * A[i]) - (T.D0[i] - T.D2[i]) * if2;
        } else {
            const cplx<double> I0 = cdiv(f0a - f0b, gamma);
            const cplx<double> I1 = I * (I0 + f1a);
            const cplx<double> I2 = mk(0.0, -1.0) * (I0 + f1b);
            for (int i = 0; i < 3; ++i)
                rv[i] = A[i] * I0 + cdiv(T.D0[i] * I1 - T.D2[i] * I2, gamma);
        }
        const cplx<double> ph = cexpi(kv[0] * T.X0[0] + kv[1] * T.X0[1] + kv[2] * T.X0[2]);
        const double s = t ? -1.0 : 1.0;
        for (int i = 0; i < 3; ++i) out[i] = out[i] + s * (ph * rv[i]);
    }
    for (int i = 0; i < 3; ++i) out[i] = r.leng * out[i];
}

// Radiation pattern of one source about X along k khat, as (theta, phi)
// components (src/fmm/leaf.cpp:185-207; dipole.h:47-51; rwg.h:23-25).
__device__ void radiation(const SrcDev& s, int64_t t, const double X[3], const double kv[3], const double* M,
                          cplx<double>& o0, cplx<double>& o1) {
    cplx<double> r[3];
    if (s.kind == 0) {
        const DipoleDev& d = s.dip[t];
        const cplx<double> e =
            cexpi(kv[0] * (X[0] - d.pos[0]) + kv[1] * (X[1] - d.pos[1]) + kv[2] * (X[2] - d.pos[2]));
        for (int i = 0; i < 3; ++i) r[i] = d.phat[i] * e;
    } else {
        cplx<double> f[3];
        rwgPlaneWave(s.rwg[t], s.tri, kv, f);
        const cplx<double> e = cexpi(kv[0] * X[0] + kv[1] * X[1] + kv[2] * X[2]);
        for (int i = 0; i < 3; ++i) r[i] = e * mk(f[i].re, -f[i].im);
    }
    o0 = M[0] * r[0] + M[1] * r[1] + M[2] * r[2];
    o1 = M[3] * r[0] + M[4] * r[1] + M[5] * r[2];
}

__global__ void kRadpat(SrcDev s, const LeafItem* __restrict__ items, const LevelDev* __restrict__ lv, double k,
                        cplx<double>* __restrict__ rp) {
    const LeafItem it = items[blockIdx.y];
    const int d = blockIdx.x * BS + threadIdx.x;
    if (d >= it.ndir) return;
    const LevelDev L = lv[it.level];
    const double kv[3] = {k * L.khat[3 * d], k * L.khat[3 * d + 1], k * L.khat[3 * d + 2]};
    const double X[3] = {it.cx, it.cy, it.cz};
    const double* M = L.thph + 6 * d;
    for (int q = 0; q < it.cnt; ++q) {
        cplx<double> a, b;
        radiation(s, it.sbeg + q, X, kv, M, a, b);
        const int64_t o = it.rpOff + ((int64_t)q * it.ndir + d) * 2;
        rp[o] = a;
        rp[o + 1] = b;
    }
}

void launchRadpat(const SrcDev& s, const LeafItem* items, int nitems, const LevelDev* lv, double k,
                  cplx<double>* rp, int maxndir, hipStream_t st) {
    if (nitems <= 0) return;
    dim3 g((maxndir + BS - 1) / BS, nitems);
    hipLaunchKernelGGL(kRadpat, g, dim3(BS), 0, st, s, items, lv, k, rp);
}

// Galerkin entry <f_obs, G f_src> without the C k factor
// (src/sources/dipole.h:62-70, src/sources/rwg.cpp:92-139): the dyadic Green's
// function e^{ikr}/r [(1 + i/kr - 1/kr^2) I - (1 + 3i/kr - 3/kr^2) rr]
// (src/math.h:132-142) contracted with the two test/basis vectors.
__device__ __forceinline__ cplx<double> greenContract(const double ro[3], const double rs[3], const double vo[3],
                                                      const double vs[3], double k) {
    const double X[3] = {ro[0] - rs[0], ro[1] - rs[1], ro[2] - rs[2]};
    const double r = sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
    const double kr = k * r, inv = 1.0 / kr, inv2 = inv * inv;
    const cplx<double> e = cexpi(kr);
    const double ir = 1.0 / r;
    const double vovs = vo[0] * vs[0] + vo[1] * vs[1] + vo[2] * vs[2];
    const double vor = (vo[0] * X[0] + vo[1] * X[1] + vo[2] * X[2]) * ir;
    const double vsr = (vs[0] * X[0] + vs[1] * X[1] + vs[2] * X[2]) * ir;
    // a (vo.vs) - b (vo.r)(vs.r) with a = 1 - inv2 + i inv, b = 1 - 3 inv2 + 3 i inv
    const double pr = vor * vsr;
    const cplx<double> c = mk((1.0 - inv2) * vovs - (1.0 - 3.0 * inv2) * pr, inv * vovs - 3.0 * inv * pr);
    return (ir)*(e * c);
}

// Gradient of greenContract with respect to ro (minus that for rs):
// E = A (vo.vs) - B (vo.R^)(vs.R^), A = e^{ikr}(1/r - 1/(k^2 r^3) + i/(k r^2)),
// B = e^{ikr}(1/r - 3/(k^2 r^3) + 3i/(k r^2)), R = ro - rs, so
// grad E = A' R^ (vo.vs) - B' R^ (vo.R^)(vs.R^)
//          - B [(vo - (vo.R^)R^)(vs.R^) + (vo.R^)(vs - (vs.R^)R^)] / r.
__device__ void greenGrad(const double ro[3], const double rs[3], const double vo[3], const double vs[3], double k,
                          cplx<double> g[3]) {
    const double X[3] = {ro[0] - rs[0], ro[1] - rs[1], ro[2] - rs[2]};
    const double r = sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
    const double ir = 1.0 / r, ir2 = ir * ir, ir3 = ir2 * ir, ir4 = ir2 * ir2;
    const double ik2 = 1.0 / (k * k), ikk = 1.0 / k;
    const cplx<double> e = cexpi(k * r);
    const double h[3] = {X[0] * ir, X[1] * ir, X[2] * ir};
    const cplx<double> A = e * mk(ir - ik2 * ir3, ikk * ir2);
    const cplx<double> B = e * mk(ir - 3.0 * ik2 * ir3, 3.0 * ikk * ir2);
    const cplx<double> Ad = mk(0.0, k) * A + e * mk(-ir2 + 3.0 * ik2 * ir4, -2.0 * ikk * ir3);
    const cplx<double> Bd = mk(0.0, k) * B + e * mk(-ir2 + 9.0 * ik2 * ir4, -6.0 * ikk * ir3);
    const double vovs = vo[0] * vs[0] + vo[1] * vs[1] + vo[2] * vs[2];
    const double voh = vo[0] * h[0] + vo[1] * h[1] + vo[2] * h[2];
    const double vsh = vs[0] * h[0] + vs[1] * h[1] + vs[2] * h[2];
    for (int a = 0; a < 3; ++a) {
        const double t = ((vo[a] - voh * h[a]) * vsh + voh * (vs[a] - vsh * h[a])) * ir;
        g[a] = (h[a] * vovs) * Ad - (h[a] * voh * vsh) * Bd - t * B;
    }
}

// Near-field part of the dipole position gradient, -2 Re(lambda^H dZ I), one
// block per near item; for an entry (a row, b column) with w = conj(l_a) I_b
// (+ conj(l_b) I_a for pair columns, applied both ways) the row source gains
// -2 Re(C k grad E . w) and the column source the opposite (E depends on
// r_a - r_b).  Accumulated with FP64 atomics into grad (input order, n x 3);
// colsIn are the items' columns in input order, iperm maps input order to tree
// positions.
__global__ void kNearGradDip(SrcDev s, const NearItem* __restrict__ items, const int32_t* __restrict__ colsIn,
                             const int64_t* __restrict__ perm, const int64_t* __restrict__ iperm,
                             const cplx<double>* __restrict__ I,
                             const cplx<double>* __restrict__ lam, double k, cplx<double> Ck, double* __restrict__ grad) {
    const NearItem it = items[blockIdx.x];
    const int64_t ne = (int64_t)it.nr * it.nc;
    for (int64_t e = threadIdx.x; e < ne; e += BS) {
        const int64_t c = e / it.nr, r = e - c * it.nr;
        const int64_t ib = colsIn[it.cptr + c];
        const int64_t a = it.rbeg + r, b = iperm[ib];
        if (a == b) continue;
        const DipoleDev& da = s.dip[a];
        const DipoleDev& db = s.dip[b];
        if (da.pos[0] == db.pos[0] && da.pos[1] == db.pos[1] && da.pos[2] == db.pos[2]) continue;
        const int64_t ia = perm[a];
        cplx<double> w = mk(lam[ia].re, -lam[ia].im) * I[ib];
        if (c >= it.nself) w = w + mk(lam[ib].re, -lam[ib].im) * I[ia];
        cplx<double> g[3];
        greenGrad(da.pos, db.pos, da.phat, db.phat, k, g);
        for (int q = 0; q < 3; ++q) {
            const double v = -2.0 * ((Ck * g[q]) * w).re;
            atomicAdd(&grad[3 * ia + q], v);
            atomicAdd(&grad[3 * ib + q], -v);
        }
    }
}

void launchNearGradDip(const SrcDev& s, const NearItem* items, int nitems, const int32_t* colsIn, const int64_t* perm,
                       const int64_t* iperm, const cplx<double>* I, const cplx<double>* lam, double k, cplx<double> Ck,
                       double* grad, hipStream_t st) {
    if (nitems <= 0) return;
    hipLaunchKernelGGL(kNearGradDip, dim3(nitems), dim3(BS), 0, st, s, items, colsIn, perm, iperm, I, lam, k, Ck,
                       grad);
}

// ---- physical near field (no counterpart in the reference, which drops
// coincident triangles and keeps the hypersingular dyadic kernel).
//
// Mixed-potential Galerkin entry with the reference's normalisation
// (f = +-(l / 2A) rho, quadrature weights summing to 1/2, so the area factors
// cancel):  Z_mn = l_m l_n sum_{T,T'} s s' sum_q sum_q' w w' (rho.rho' - 4/k^2)
// e^{ikR}/R, equal to the reference's dyadic form for separated triangles
// (integration by parts over both triangles of each RWG).  For triangles
// closer than twice the longer edge the 1/R part of the inner integral is
// done in closed form (Wilton, Rao, Glisson, IEEE TAP 32 (1984) 276-281;
// Graglia, IEEE TAP 41 (1993) 1448-1455) and only (e^{ikR} - 1)/R is
// sampled.

// I = int_T dS'/|r - r'|,  Iv = int_T (r' - r)/|r - r'| dS'
__device__ void triPotentials(const double r[3], const TriDev& T, double& I, double Iv[3]) {
    double v[3][3];
    for (int a = 0; a < 3; ++a) {
        v[0][a] = T.X0[a];
        v[1][a] = T.X0[a] + T.D0[a];
        v[2][a] = T.X0[a] - T.D2[a];
    }
    const double e1[3] = {v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2]};
    const double e2[3] = {v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2]};
    double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int a = 0; a < 3; ++a) n[a] /= nn;
    const double d = n[0] * (r[0] - v[0][0]) + n[1] * (r[1] - v[0][1]) + n[2] * (r[2] - v[0][2]);
    const double ad = fabs(d);
    const double rho[3] = {r[0] - d * n[0], r[1] - d * n[1], r[2] - d * n[2]};
    double scale = 0;
    I = 0;
    double Ip[3] = {0, 0, 0};
    for (int i = 0; i < 3; ++i) {
        const double* a = v[i];
        const double* b = v[(i + 1) % 3];
        double l[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const double len = sqrt(l[0] * l[0] + l[1] * l[1] + l[2] * l[2]);
        scale = fmax(scale, len);
        for (int c = 0; c < 3; ++c) l[c] /= len;
        const double u[3] = {l[1] * n[2] - l[2] * n[1], l[2] * n[0] - l[0] * n[2], l[0] * n[1] - l[1] * n[0]};
        const double lp = (b[0] - rho[0]) * l[0] + (b[1] - rho[1]) * l[1] + (b[2] - rho[2]) * l[2];
        const double lm = (a[0] - rho[0]) * l[0] + (a[1] - rho[1]) * l[1] + (a[2] - rho[2]) * l[2];
        const double P0 = (a[0] - rho[0]) * u[0] + (a[1] - rho[1]) * u[1] + (a[2] - rho[2]) * u[2];
        const double R02 = P0 * P0 + d * d;
        const double Rp = sqrt(lp * lp + R02), Rm = sqrt(lm * lm + R02);
        // log((R+ + l+)/(R- + l-)); R + l = R0^2 / (R - l) for l < 0 avoids
        // the cancellation near the edge's line
        double lg = 0;
        if (R02 > 1e-28 * scale * scale) {
            const double ap = lp >= 0 ? Rp + lp : R02 / (Rp - lp);
            const double am = lm >= 0 ? Rm + lm : R02 / (Rm - lm);
            lg = log(ap / am);
        } else if (lm > 0 || lp < 0)  // on the edge's line, outside the segment
            lg = lm > 0 ? log(lp / lm) : log(lm / lp);
        // (on the segment itself R0 = 0 multiplies the logarithm)
        double at = 0;
        if (fabs(P0) > 1e-14 * scale)
            at = atan(P0 * lp / (R02 + ad * Rp)) - atan(P0 * lm / (R02 + ad * Rm));
        I += P0 * lg - ad * at;
        const double w = 0.5 * (R02 * lg + lp * Rp - lm * Rm);
        for (int c = 0; c < 3; ++c) Ip[c] += w * u[c];
    }
    for (int c = 0; c < 3; ++c) Iv[c] = Ip[c] - d * n[c] * I;
}

__device__ cplx<double> entryPhys(const SrcDev& s, int64_t o, int64_t q, double k) {
    const RwgDev& a = s.rwg[o];
    const RwgDev& b = s.rwg[q];
    const double k2 = 4.0 / (k * k);
    cplx<double> acc = mk(0.0, 0.0);
    for (int ot = 0; ot < 2; ++ot) {
        const TriDev& TO = s.tri[a.tri[ot]];
        for (int st = 0; st < 2; ++st) {
            const TriDev& TS = s.tri[b.tri[st]];
            const double sg = (ot ^ st) ? -1.0 : 1.0;
            // closeness: centroid distance against the longer of the two triangles' edges
            double co[3], cs[3], eo = 0, es = 0;
            for (int c = 0; c < 3; ++c) {
                co[c] = TO.X0[c] + (TO.D0[c] - TO.D2[c]) / 3.0;
                cs[c] = TS.X0[c] + (TS.D0[c] - TS.D2[c]) / 3.0;
            }
            auto edge = [](const TriDev& T) {
                const double e0 = T.D0[0] * T.D0[0] + T.D0[1] * T.D0[1] + T.D0[2] * T.D0[2];
                const double e2 = T.D2[0] * T.D2[0] + T.D2[1] * T.D2[1] + T.D2[2] * T.D2[2];
                const double e1x = T.D0[0] + T.D2[0], e1y = T.D0[1] + T.D2[1], e1z = T.D0[2] + T.D2[2];
                return sqrt(fmax(fmax(e0, e2), e1x * e1x + e1y * e1y + e1z * e1z));
            };
            eo = edge(TO);
            es = edge(TS);
            const double dc = sqrt((co[0] - cs[0]) * (co[0] - cs[0]) + (co[1] - cs[1]) * (co[1] - cs[1]) +
                                   (co[2] - cs[2]) * (co[2] - cs[2]));
            const bool close = dc < 2.0 * fmax(eo, es);
            double areaS = 0;
            if (close) {
                const double cx = TS.D0[1] * TS.D2[2] - TS.D0[2] * TS.D2[1];
                const double cy = TS.D0[2] * TS.D2[0] - TS.D0[0] * TS.D2[2];
                const double cz = TS.D0[0] * TS.D2[1] - TS.D0[1] * TS.D2[0];
                areaS = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
            }
            cplx<double> part = mk(0.0, 0.0);
            for (int i = 0; i < TO.nq; ++i) {
                const double ro[3] = {TO.q[i][0], TO.q[i][1], TO.q[i][2]};
                const double vo[3] = {ro[0] - a.xpm[ot][0], ro[1] - a.xpm[ot][1], ro[2] - a.xpm[ot][2]};
                cplx<double> inner = mk(0.0, 0.0);
                for (int j = 0; j < TS.nq; ++j) {
                    const double rs[3] = {TS.q[j][0], TS.q[j][1], TS.q[j][2]};
                    const double vs[3] = {rs[0] - b.xpm[st][0], rs[1] - b.xpm[st][1], rs[2] - b.xpm[st][2]};
                    const double dx = ro[0] - rs[0], dy = ro[1] - rs[1], dz = ro[2] - rs[2];
                    const double R = sqrt(dx * dx + dy * dy + dz * dz);
                    const double h = vo[0] * vs[0] + vo[1] * vs[1] + vo[2] * vs[2] - k2;
                    cplx<double> g;
                    if (close) {  // (e^{ikR} - 1)/R, its R -> 0 limit ik
                        if (R > 0) {
                            const double sh = sin(0.5 * k * R);
                            g = mk(-2.0 * sh * sh / R, sin(k * R) / R);
                        } else
                            g = mk(0.0, k);
                    } else {
                        const cplx<double> e = cexpi(k * R);
                        g = mk(e.re / R, e.im / R);
                    }
                    inner = inner + (TS.q[j][3] * h) * g;
                }
                if (close) {  // int (rho_o . rho_s - 4/k^2)/R dS' / (2 A')
                    double I, Iv[3];
                    triPotentials(ro, TS, I, Iv);
                    const double rx[3] = {ro[0] - b.xpm[st][0], ro[1] - b.xpm[st][1], ro[2] - b.xpm[st][2]};
                    double sing = -k2 * I;
                    for (int c = 0; c < 3; ++c) sing += vo[c] * (Iv[c] + rx[c] * I);
                    inner = inner + mk(sing / (2.0 * areaS), 0.0);
                }
                part = part + TO.q[i][3] * inner;
            }
            acc = acc + sg * part;
        }
    }
    return (a.leng * b.leng) * acc;
}

__device__ cplx<double> entry(const SrcDev& s, int64_t o, int64_t q, double k) {
    if (s.physical) return entryPhys(s, o, q, k);
    if (s.kind == 0) {
        const DipoleDev& a = s.dip[o];
        const DipoleDev& b = s.dip[q];
        if (a.pos[0] == b.pos[0] && a.pos[1] == b.pos[1] && a.pos[2] == b.pos[2]) return mk(0.0, 0.0);
        return greenContract(a.pos, b.pos, a.phat, b.phat, k);
    }
    const RwgDev& a = s.rwg[o];
    const RwgDev& b = s.rwg[q];
    if (a.ctr[0] == b.ctr[0] && a.ctr[1] == b.ctr[1] && a.ctr[2] == b.ctr[2]) return mk(0.0, 0.0);
    cplx<double> acc = mk(0.0, 0.0);
    for (int ot = 0; ot < 2; ++ot) {
        const TriDev& TO = s.tri[a.tri[ot]];
        for (int st = 0; st < 2; ++st) {
            if (a.tri[ot] == b.tri[st]) continue;  // coincident triangles are skipped
            const TriDev& TS = s.tri[b.tri[st]];
            const double sg = (ot ^ st) ? -1.0 : 1.0;
            cplx<double> part = mk(0.0, 0.0);
            for (int i = 0; i < TO.nq; ++i) {
                const double ro[3] = {TO.q[i][0], TO.q[i][1], TO.q[i][2]};
                const double vo[3] = {ro[0] - a.xpm[ot][0], ro[1] - a.xpm[ot][1], ro[2] - a.xpm[ot][2]};
                cplx<double> inner = mk(0.0, 0.0);
                for (int j = 0; j < TS.nq; ++j) {
                    const double rs[3] = {TS.q[j][0], TS.q[j][1], TS.q[j][2]};
                    const double vs[3] = {rs[0] - b.xpm[st][0], rs[1] - b.xpm[st][1], rs[2] - b.xpm[st][2]};
                    inner = inner + TS.q[j][3] * greenContract(ro, rs, vo, vs, k);
                }
                part = part + TO.q[i][3] * inner;
            }
            acc = acc + sg * part;
        }
    }
    return (a.leng * b.leng) * acc;
}

// One thread per stored entry.  Pair columns hold entry(row, column) with the
// head leaf's source as the observer (src/fmm/leaf.cpp:115-128); self columns
// hold the reference's strict lower triangle mirrored, entry(later, earlier)
// in leaf order (src/fmm/leaf.cpp:361-383), and zero on the diagonal.
__global__ void kNearBuild(SrcDev s, const NearItem* __restrict__ items, int nitems,
                           const int64_t* __restrict__ cols, int64_t nnz, double k, cplx<double> Ck,
                           cplx<double>* __restrict__ val) {
    const int64_t e = (int64_t)blockIdx.x * BS + threadIdx.x;
    if (e >= nnz) return;
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {  // last item with voff <= e (items are in value order)
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].voff <= e)
            lo = mid;
        else
            hi = mid - 1;
    }
    const NearItem I = items[lo];
    const int64_t loc = e - I.voff;
    const int64_t c = loc / I.nr, r = loc - c * I.nr;
    const int64_t row = I.rbeg + r, col = cols[I.cptr + c];
    cplx<double> v;
    if (c < I.nself)
        v = (row == col && !s.physical) ? mk(0.0, 0.0) : entry(s, row > col ? row : col, row > col ? col : row, k);
    else
        v = entry(s, row, col, k);
    val[e] = Ck * v;
}

void launchNearBuild(const SrcDev& s, const NearItem* items, int nitems, const int64_t* cols, int64_t nnz, double k,
                     cplx<double> Ck, cplx<double>* val, hipStream_t st) {
    if (nnz <= 0 || nitems <= 0) return;
    const int64_t nb = (nnz + BS - 1) / BS;
    hipLaunchKernelGGL(kNearBuild, dim3((unsigned)nb), dim3(BS), 0, st, s, items, nitems, cols, nnz, k, Ck, val);
}

// ------------------------------------------------------------------ translation operators

// alpha_d(psi_i) = (ik/4pi) sum_{l<=L} i^l (2l+1) h_l(k d a) P_l(cos psi_i),
// psi_i = pi i/(nps-1) (src/fmm/tables.cpp:90-127); one thread per (distance, psi).
__global__ void kAlpha(int nps, int L, double k, double a, const double* __restrict__ dists,
                       cplx<double>* __restrict__ alpha) {
    const int i = blockIdx.x * BS + threadIdx.x;
    const int di = blockIdx.y;
    if (i >= nps) return;
    const double x = cos(kPiD * i / (double)(nps - 1));
    const double kr = k * dists[di] * a;
    // spherical Hankel h_l^(1)(kr) by upward recurrence (src/math.cpp:68-84)
    const cplx<double> e = cexpi(kr);
    cplx<double> hm = mk(e.im / kr, -e.re / kr);  // -i e^{ix}/x
    cplx<double> h = mk(-(e.re * (1.0 / kr) - e.im * (1.0 / (kr * kr))), -(e.im * (1.0 / kr) + e.re * (1.0 / (kr * kr))));
    double pm = 1.0, p = x;  // P_0, P_1
    // l = 0 term: i^0 * 1 * h_0 * P_0
    cplx<double> sum = hm;
    if (L >= 1) sum = sum + 3.0 * (mk(0.0, 1.0) * h) * p;
    for (int l = 1; l < L; ++l) {
        const cplx<double> hn = ((2.0 * l + 1.0) / kr) * h - hm;
        hm = h;
        h = hn;
        const double pn = ((2.0 * l + 1) * x * p - l * pm) / (double)(l + 1);
        pm = p;
        p = pn;
        const int m = (l + 1) & 3;
        const cplx<double> il = m == 0 ? mk(1.0, 0.0) : (m == 1 ? mk(0.0, 1.0) : (m == 2 ? mk(-1.0, 0.0) : mk(0.0, -1.0)));
        sum = sum + (2.0 * (l + 1) + 1.0) * (il * h) * p;
    }
    alpha[(int64_t)di * nps + i] = mk(0.0, k / (4.0 * kPiD)) * sum;
}

void launchAlpha(int nps, int L, double k, double nodeLeng, const double* dists, cplx<double>* alpha, hipStream_t st) {
    dim3 g((nps + BS - 1) / BS, 15);
    hipLaunchKernelGGL(kAlpha, g, dim3(BS), 0, st, nps, L, k, nodeLeng, dists, alpha);
}

// T(dir) = sum_q alpha[reflect(near+1-p+q)] l_q(psi) on the uniform psi grid with
// the half-sample index reflection of src/fmm/tables.cpp:200-213, then times the
// M2L quadrature weight of its direction (src/fmm/node.cpp:122-132).
__global__ void kTransl(int ndir, int nps, int p, const double* __restrict__ psi, const int32_t* __restrict__ nearIdx,
                        const int32_t* __restrict__ distIdx, const cplx<double>* __restrict__ alpha,
                        const double* __restrict__ dirWeight, cplx<double>* __restrict__ T) {
    const int d = blockIdx.x * BS + threadIdx.x;
    const int o = blockIdx.y;
    if (d >= ndir) return;
    const int64_t e = (int64_t)o * ndir + d;
    const double ps = psi[e];
    const int near = nearIdx[e];
    const cplx<double>* al = alpha + (int64_t)distIdx[o] * nps;
    const double h = kPiD / (double)(nps - 1);
    cplx<double> acc = mk(0.0, 0.0);
    for (int q = 0; q < 2 * p; ++q) {
        const int iq = near + 1 - p + q;
        const double xq = kPiD * iq / (double)(nps - 1);
        double w = 1.0;
        for (int j = 0; j < 2 * p; ++j) {
            if (j == q) continue;
            const double xj = kPiD * (near + 1 - p + j) / (double)(nps - 1);
            w *= (ps - xj) / (xq - xj);
        }
        const int f = iq < 0 ? -iq - 1 : (iq >= nps ? 2 * nps - iq - 1 : iq);
        acc = acc + w * al[f];
    }
    (void)h;
    T[e] = dirWeight[d] * acc;
}

void launchTransl(int ndir, int nps, int p, const double* psi, const int32_t* nearIdx, const int32_t* distIdx,
                  const cplx<double>* alpha, const double* dirWeight, cplx<double>* T, hipStream_t st) {
    dim3 g((ndir + BS - 1) / BS, 316);
    hipLaunchKernelGGL(kTransl, g, dim3(BS), 0, st, ndir, nps, p, psi, nearIdx, distIdx, alpha, dirWeight, T);
}

// ------------------------------------------------------------------ excitation and far field

// V = -E0 e^{ik.r} (p0 x . pol) for dipoles (the reference evaluates the voltage
// with the default moment p0 x, src/sources/dipole.h:11-16,29-32) and
// V = -E0 sum_i F_i pol_i for RWG (src/sources/rwg.h:18-21).
__global__ void kVoltage(SrcDev s, int64_t n, double w0, double w1, double w2, double p0, double p1, double p2,
                         double amp, cplx<double>* __restrict__ V) {
    const int64_t t = (int64_t)blockIdx.x * BS + threadIdx.x;
    if (t >= n) return;
    cplx<double> v;
    if (s.kind == 0) {
        const DipoleDev& d = s.dip[t];
        const cplx<double> e = cexpi(w0 * d.pos[0] + w1 * d.pos[1] + w2 * d.pos[2]);
        v = (-amp * 1.0e-8 * p0) * e;
    } else {
        const double kv[3] = {w0, w1, w2};
        cplx<double> f[3];
        rwgPlaneWave(s.rwg[t], s.tri, kv, f);
        v = (-amp) * (p0 * f[0] + p1 * f[1] + p2 * f[2]);
    }
    V[t] = v;
}

void launchVoltage(const SrcDev& s, int64_t n, const double* wv, const double* pol, double amp, cplx<double>* V,
                   hipStream_t st) {
    const int64_t nb = (n + BS - 1) / BS;
    hipLaunchKernelGGL(kVoltage, dim3((unsigned)nb), dim3(BS), 0, st, s, n, wv[0], wv[1], wv[2], pol[0], pol[1],
                       pol[2], amp, V);
}

// E(rhat) = C k (I - rr) sum_n I_n f_n(k rhat) with f_n = conj(F_n) (RWG) or
// e^{-ik rhat.r} p (dipole) (src/fmm/node.cpp:147-178).  The source sum of
// each direction is split into nch ranges of the tree order (grid (ndir,
// nch)); block (d, c) writes its three partial sums to part[(d nch + c) 3 + i]
// and kFarfieldReduce adds them in range order, so the result does not depend
// on scheduling.  I is read in input order through perm.
__global__ void kFarfieldPart(SrcDev s, int64_t n, int64_t chunk, const cplx<double>* __restrict__ I,
                              const int64_t* __restrict__ perm, const double* __restrict__ dirs, double k,
                              cplx<double>* __restrict__ part) {
    const int64_t d = blockIdx.x;
    const int c = blockIdx.y, nch = gridDim.y;
    const double rh[3] = {dirs[3 * d], dirs[3 * d + 1], dirs[3 * d + 2]};
    const double kv[3] = {k * rh[0], k * rh[1], k * rh[2]};
    const int64_t t0 = (int64_t)c * chunk;
    const int64_t t1 = t0 + chunk < n ? t0 + chunk : n;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t t = t0 + threadIdx.x; t < t1; t += BS) {
        cplx<double> f[3];
        if (s.kind == 0) {
            const DipoleDev& dp = s.dip[t];
            const cplx<double> e = cexpi(-(dp.pos[0] * kv[0] + dp.pos[1] * kv[1] + dp.pos[2] * kv[2]));
            for (int i = 0; i < 3; ++i) f[i] = dp.phat[i] * e;
        } else {
            rwgPlaneWave(s.rwg[t], s.tri, kv, f);
            for (int i = 0; i < 3; ++i) f[i] = mk(f[i].re, -f[i].im);
        }
        const cplx<double> It = I[perm[t]];
        for (int i = 0; i < 3; ++i) {
            const cplx<double> v = It * f[i];
            acc[2 * i] += v.re;
            acc[2 * i + 1] += v.im;
        }
    }
    __shared__ double red[BS / 64][6];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int j = 0; j < 6; ++j) {
        const double v = wsum(acc[j]);
        if (lane == 0) red[w][j] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int i = threadIdx.x;
        double re = 0, im = 0;
        for (int q = 0; q < BS / 64; ++q) {
            re += red[q][2 * i];
            im += red[q][2 * i + 1];
        }
        part[(d * nch + c) * 3 + i] = mk(re, im);
    }
}

__global__ void kFarfieldReduce(int64_t ndir, int nch, const cplx<double>* __restrict__ part,
                                const double* __restrict__ dirs, cplx<double> Ck, cplx<double>* __restrict__ E) {
    const int64_t d = (int64_t)blockIdx.x * BS + threadIdx.x;
    if (d >= ndir) return;
    cplx<double> a[3] = {mk(0.0, 0.0), mk(0.0, 0.0), mk(0.0, 0.0)};
    for (int c = 0; c < nch; ++c)
        for (int i = 0; i < 3; ++i) a[i] = a[i] + part[(d * nch + c) * 3 + i];
    const double rh[3] = {dirs[3 * d], dirs[3 * d + 1], dirs[3 * d + 2]};
    for (int i = 0; i < 3; ++i) {
        cplx<double> v = mk(0.0, 0.0);
        for (int j = 0; j < 3; ++j) v = v + ((i == j ? 1.0 : 0.0) - rh[i] * rh[j]) * a[j];
        E[3 * d + i] = Ck * v;
    }
}

// Source ranges per direction: enough blocks to fill the device (about 4096
// in all), at least 1024 sources per range, at most kFarPartCap partial
// triples in all (the size of the partial-sum buffer).
int farfieldChunks(int64_t n, int64_t ndir) {
    if (n <= 0 || ndir <= 0) return 1;
    int64_t nch = (4096 + ndir - 1) / ndir;
    nch = std::min<int64_t>(nch, (n + 1023) / 1024);
    nch = std::min<int64_t>(nch, kFarPartCap / ndir);
    nch = std::max<int64_t>(nch, 1);
    const int64_t chunk = (n + nch - 1) / nch;
    return (int)((n + chunk - 1) / chunk);
}

// Far-field partial sums of mlfma_farfield_noalloc (callers that cannot
// allocate device memory): one array per device, sized for the most partial
// triples farfieldChunks uses (kFarPartCap x 3 complex = 3 MiB), held until the
// process ends.  Host code never names the array in a launch or copy: it uses
// the address hipGetSymbolAddress gives on each device.
static __device__ __align__(256) cplx<double> scratch[kFarPartCap * 3];

namespace {
std::mutex g_scratchMutex;
std::vector<cplx<double>*> g_scratchPtr;  // by device ordinal
}  // namespace

hipError_t initFarfieldScratch(int device) {
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (device < 0) return hipErrorInvalidDevice;
    if ((size_t)device >= g_scratchPtr.size()) g_scratchPtr.resize((size_t)device + 1, nullptr);
    if (g_scratchPtr[device]) return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;  // the caller makes the device current
    cplx<double>* ptr = nullptr;
    e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratchPtr[device] = ptr;
    return hipSuccess;
}

cplx<double>* farfieldScratch(int device) {
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    return device >= 0 && (size_t)device < g_scratchPtr.size() ? g_scratchPtr[device] : nullptr;
}

void launchFarfield(const SrcDev& s, int64_t n, const cplx<double>* I, const int64_t* perm, const double* dirs,
                    int64_t ndir, double k, cplx<double> Ck, cplx<double>* part, cplx<double>* E, hipStream_t st) {
    if (ndir <= 0) return;
    const int nch = farfieldChunks(n, ndir);
    const int64_t chunk = n > 0 ? (n + nch - 1) / nch : 1;
    hipLaunchKernelGGL(kFarfieldPart, dim3((unsigned)ndir, nch), dim3(BS), 0, st, s, n, chunk, I, perm, dirs, k, part);
    hipLaunchKernelGGL(kFarfieldReduce, dim3((unsigned)((ndir + BS - 1) / BS)), dim3(BS), 0, st, ndir, nch,
                       (const cplx<double>*)part, dirs, Ck, E);
}

// Adjoint of kFarfield: I_n = conj(C k) sum_d f_n(k rhat_d)^H (I - rr) E_d,
// one thread per source (tree order); directions staged in LDS a block at a
// time with their projected, conjugate-scaled fields.
__global__ void kFarfieldAdj(SrcDev s, int64_t n, const cplx<double>* __restrict__ E, const double* __restrict__ dirs,
                             int ndir, double k, cplx<double> Ckc, const int64_t* __restrict__ perm,
                             cplx<double>* __restrict__ I) {
    __shared__ double sr[BS][3];
    __shared__ cplx<double> se[BS][3];
    const int64_t t = (int64_t)blockIdx.x * BS + threadIdx.x;
    cplx<double> acc = mk(0.0, 0.0);
    for (int d0 = 0; d0 < ndir; d0 += BS) {
        __syncthreads();
        const int d = d0 + threadIdx.x;
        if (d < ndir) {
            const double rh[3] = {dirs[3 * d], dirs[3 * d + 1], dirs[3 * d + 2]};
            for (int i = 0; i < 3; ++i) {
                cplx<double> v = mk(0.0, 0.0);
                for (int j = 0; j < 3; ++j) v = v + ((i == j ? 1.0 : 0.0) - rh[i] * rh[j]) * E[3 * d + j];
                se[threadIdx.x][i] = Ckc * v;
                sr[threadIdx.x][i] = rh[i];
            }
        }
        __syncthreads();
        if (t < n) {
            const int nd = min(BS, ndir - d0);
            for (int q = 0; q < nd; ++q) {
                const double kv[3] = {k * sr[q][0], k * sr[q][1], k * sr[q][2]};
                cplx<double> f[3];
                if (s.kind == 0) {
                    const DipoleDev& dp = s.dip[t];
                    const cplx<double> e = cexpi(-(dp.pos[0] * kv[0] + dp.pos[1] * kv[1] + dp.pos[2] * kv[2]));
                    for (int i = 0; i < 3; ++i) f[i] = dp.phat[i] * e;
                } else {
                    rwgPlaneWave(s.rwg[t], s.tri, kv, f);
                    for (int i = 0; i < 3; ++i) f[i] = mk(f[i].re, -f[i].im);
                }
                for (int i = 0; i < 3; ++i) acc = acc + mk(f[i].re, -f[i].im) * se[q][i];
            }
        }
    }
    if (t < n) I[perm[t]] = acc;
}

void launchFarfieldAdj(const SrcDev& s, int64_t n, const cplx<double>* E, const double* dirs, int ndir, double k,
                       cplx<double> Ckc, const int64_t* perm, cplx<double>* I, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(kFarfieldAdj, dim3((unsigned)((n + BS - 1) / BS)), dim3(BS), 0, st, s, n, E, dirs, ndir, k, Ckc,
                       perm, I);
}

// y += s x over n complex values (the shift of a shifted solve)
__global__ void kAxpy(const cplx<double>* __restrict__ x, cplx<double> s, cplx<double>* __restrict__ y, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
    if (i < n) y[i] = y[i] + s * x[i];
}
void launchAxpy(const cplx<double>* x, cplx<double> s, cplx<double>* y, int64_t n, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(kAxpy, dim3((unsigned)((n + BS - 1) / BS)), dim3(BS), 0, st, x, s, y, n);
}

// ------------------------------------------------------------------ matvec stages

__global__ void kGather(const cplx<double>* __restrict__ x, const int64_t* __restrict__ perm,
                        cplx<double>* __restrict__ xt, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * BS + threadIdx.x;
    const int64_t r = blockIdx.y;
    if (t >= n) return;
    xt[r * n + t] = x[r * n + perm[t]];
}

void launchGather(const cplx<double>* x, const int64_t* perm, cplx<double>* xt, int64_t n, int nrhs, hipStream_t st) {
    dim3 g((unsigned)((n + BS - 1) / BS), nrhs);
    hipLaunchKernelGGL(kGather, g, dim3(BS), 0, st, x, perm, xt, n);
}

__global__ void kScatter(const cplx<double>* __restrict__ yt, const int64_t* __restrict__ perm,
                         cplx<double>* __restrict__ y, int64_t n, int acc) {
    const int64_t t = (int64_t)blockIdx.x * BS + threadIdx.x;
    const int64_t r = blockIdx.y;
    if (t >= n) return;
    const int64_t i = r * n + perm[t];
    const cplx<double> v = yt[r * n + t];
    y[i] = acc ? y[i] + v : v;
}

void launchScatter(const cplx<double>* yt, const int64_t* perm, cplx<double>* y, int64_t n, int nrhs, int accumulate,
                   hipStream_t st) {
    dim3 g((unsigned)((n + BS - 1) / BS), nrhs);
    hipLaunchKernelGGL(kScatter, g, dim3(BS), 0, st, yt, perm, y, n, accumulate);
}

// M2M theta pass (src/fmm/stem.cpp:72-103,141-171): for one parent, the sum over
// its non-empty children of the phase-shifted child field interpolated in theta
// onto the parent's theta nodes, on the child's phi grid.  Across a pole the
// stencil reads the reflected theta node at phi + pi with both components negated.
template <typename R>
__global__ void kM2MTheta(const int32_t* __restrict__ ptr, const int32_t* __restrict__ child,
                          const int32_t* __restrict__ oct, int nthP, int mthC, int mphC, int p,
                          const int32_t* __restrict__ thNear, const double* __restrict__ thW,
                          const cplx<R>* __restrict__ phaseUp, const cplx<R>* __restrict__ Mc,
                          cplx<R>* __restrict__ inner) {
    const int item = blockIdx.y;
    const int e = blockIdx.x * BS + threadIdx.x;
    if (e >= nthP * mphC) return;
    const int jth = e / mphC, iph = e - jth * mphC;
    const int ndirC = mthC * mphC;
    const int near = thNear[jth];
    cplx<R> a0 = mk<R>(0, 0), a1 = mk<R>(0, 0);
    for (int c = ptr[item]; c < ptr[item + 1]; ++c) {
        const cplx<R>* Mb = Mc + (int64_t)child[c] * ndirC * 2;
        const cplx<R>* ph = phaseUp + (int64_t)oct[c] * ndirC;
        for (int q = 0; q < 2 * p; ++q) {
            const int i = near + 1 - p + q;
            const int f = i < 0 ? -i - 1 : (i >= mthC ? 2 * mthC - i - 1 : i);
            const bool flip = (i != f);
            const int ip = flip ? (iph + mphC / 2) % mphC : iph;
            const int d = f * mphC + ip;
            const R w = (R)(flip ? -thW[jth * 2 * p + q] : thW[jth * 2 * p + q]);
            const cplx<R> s = ph[d];
            a0 = a0 + w * (s * Mb[2 * d]);
            a1 = a1 + w * (s * Mb[2 * d + 1]);
        }
    }
    cplx<R>* o = inner + ((int64_t)item * nthP * mphC + e) * 2;
    o[0] = a0;
    o[1] = a1;
}

template <typename R>
void launchM2MTheta(int nitems, const int32_t* /*stemSlot*/, const int32_t* ptr, const int32_t* child,
                    const int32_t* oct, int nthP, int mthC, int mphC, int p, const int32_t* thNear, const double* thW,
                    const cplx<R>* phaseUp, const cplx<R>* Mc, cplx<R>* inner, hipStream_t st) {
    if (nitems <= 0) return;
    dim3 g((nthP * mphC + BS - 1) / BS, nitems);
    hipLaunchKernelGGL(kM2MTheta<R>, g, dim3(BS), 0, st, ptr, child, oct, nthP, mthC, mphC, p, thNear, thW, phaseUp,
                       Mc, inner);
}

// M2M phi pass (src/fmm/stem.cpp:173-183): periodic Lagrange interpolation in phi.
template <typename R>
__global__ void kM2MPhi(const int32_t* __restrict__ stemSlot, int nthP, int nphP, int mphC, int p,
                        const int32_t* __restrict__ phNear, const double* __restrict__ phW,
                        const cplx<R>* __restrict__ inner, cplx<R>* __restrict__ Mp) {
    const int item = blockIdx.y;
    const int e = blockIdx.x * BS + threadIdx.x;
    if (e >= nthP * nphP) return;
    const int jth = e / nphP, jph = e - jth * nphP;
    const cplx<R>* in = inner + (int64_t)item * nthP * mphC * 2 + (int64_t)jth * mphC * 2;
    const int near = phNear[jph];
    cplx<R> a0 = mk<R>(0, 0), a1 = mk<R>(0, 0);
    for (int q = 0; q < 2 * p; ++q) {
        int i = near + 1 - p + q;
        i = i < 0 ? i + mphC : (i >= mphC ? i - mphC : i);
        const R w = (R)phW[jph * 2 * p + q];
        a0 = a0 + w * in[2 * i];
        a1 = a1 + w * in[2 * i + 1];
    }
    cplx<R>* o = Mp + ((int64_t)stemSlot[item] * nthP * nphP + e) * 2;
    o[0] = a0;
    o[1] = a1;
}

template <typename R>
void launchM2MPhi(int nitems, const int32_t* stemSlot, int nthP, int nphP, int mphC, int p, const int32_t* phNear,
                  const double* phW, const cplx<R>* inner, cplx<R>* Mp, hipStream_t st) {
    if (nitems <= 0) return;
    dim3 g((nthP * nphP + BS - 1) / BS, nitems);
    hipLaunchKernelGGL(kM2MPhi<R>, g, dim3(BS), 0, st, stemSlot, nthP, nphP, mphC, p, phNear, phW, inner, Mp);
}

template <typename R>
__global__ void kZeroSlots(const int32_t* __restrict__ slots, int ndir, cplx<R>* __restrict__ Lc) {
    const int e = blockIdx.x * BS + threadIdx.x;
    if (e >= 2 * ndir) return;
    Lc[(int64_t)slots[blockIdx.y] * ndir * 2 + e] = mk<R>(0, 0);
}

template <typename R>
void launchZeroSlots(int n, const int32_t* slots, int ndir, cplx<R>* Lc, hipStream_t st) {
    if (n <= 0) return;
    dim3 g((2 * ndir + BS - 1) / BS, n);
    hipLaunchKernelGGL(kZeroSlots<R>, g, dim3(BS), 0, st, slots, ndir, Lc);
}

// L2L phi pass (transpose of the M2M phi pass, src/fmm/stem.cpp:109-131,202-216):
// phase shift of the parent local on the parent grid, then the gather form of
// the phi anterpolation onto the child's phi nodes.
template <typename R>
__global__ void kL2LPhi(const int32_t* __restrict__ parent, const int32_t* __restrict__ oct, int mthP, int mphP,
                        int nphC, const int32_t* __restrict__ atPtr, const int32_t* __restrict__ atSrc,
                        const double* __restrict__ atW, const cplx<R>* __restrict__ phaseDn,
                        const cplx<R>* __restrict__ Lp, cplx<R>* __restrict__ inner) {
    const int item = blockIdx.y;
    const int e = blockIdx.x * BS + threadIdx.x;
    if (e >= mthP * nphC) return;
    const int ith = e / nphC, jph = e - ith * nphC;
    const int ndirP = mthP * mphP;
    const cplx<R>* Lb = Lp + (int64_t)parent[item] * ndirP * 2;
    // phaseDn == nullptr: no phase (the first pass of the adjoint of M2M)
    const cplx<R>* ph = phaseDn ? phaseDn + (int64_t)oct[item] * ndirP : nullptr;
    cplx<R> a0 = mk<R>(0, 0), a1 = mk<R>(0, 0);
    for (int q = atPtr[jph]; q < atPtr[jph + 1]; ++q) {
        const int d = ith * mphP + atSrc[q];
        const R w = (R)atW[q];
        if (ph) {
            const cplx<R> s = ph[d];
            a0 = a0 + w * (s * Lb[2 * d]);
            a1 = a1 + w * (s * Lb[2 * d + 1]);
        } else {
            a0 = a0 + w * Lb[2 * d];
            a1 = a1 + w * Lb[2 * d + 1];
        }
    }
    cplx<R>* o = inner + ((int64_t)item * mthP * nphC + e) * 2;
    o[0] = a0;
    o[1] = a1;
}

template <typename R>
void launchL2LPhi(int nitems, const int32_t* parent, const int32_t* oct, int mthP, int mphP, int nphC,
                  const int32_t* atPtr, const int32_t* atSrc, const double* atW, const cplx<R>* phaseDn,
                  const cplx<R>* Lp, cplx<R>* inner, hipStream_t st) {
    if (nitems <= 0) return;
    dim3 g((mthP * nphC + BS - 1) / BS, nitems);
    hipLaunchKernelGGL(kL2LPhi<R>, g, dim3(BS), 0, st, parent, oct, mthP, mphP, nphC, atPtr, atSrc, atW, phaseDn, Lp,
                       inner);
}

// L2L theta pass (src/fmm/stem.cpp:218-243): gather form of the theta
// anterpolation with the pole reflection, added into the child's local.
template <typename R>
__global__ void kL2LTheta(const int32_t* __restrict__ child, int mthP, int nthC, int nphC,
                          const int32_t* __restrict__ atPtr, const int32_t* __restrict__ atSrc,
                          const double* __restrict__ atW, const int32_t* __restrict__ atFlip,
                          const cplx<R>* __restrict__ inner, cplx<R>* __restrict__ Lc,
                          const cplx<R>* __restrict__ postPhase, const int32_t* __restrict__ oct) {
    const int item = blockIdx.y;
    const int e = blockIdx.x * BS + threadIdx.x;
    if (e >= nthC * nphC) return;
    const int jf = e / nphC, js = e - jf * nphC;
    const cplx<R>* in = inner + (int64_t)item * mthP * nphC * 2;
    cplx<R> a0 = mk<R>(0, 0), a1 = mk<R>(0, 0);
    const int jsf = (js + nphC / 2) % nphC;
    for (int q = atPtr[jf]; q < atPtr[jf + 1]; ++q) {
        const bool flip = atFlip[q] != 0;
        const R w = (R)(flip ? -atW[q] : atW[q]);
        const int d = atSrc[q] * nphC + (flip ? jsf : js);
        a0 = a0 + w * in[2 * d];
        a1 = a1 + w * in[2 * d + 1];
    }
    if (postPhase) {  // the adjoint of M2M: conj of the child-grid up-phase
        cplx<R> s = postPhase[(int64_t)oct[item] * nthC * nphC + e];
        s.im = -s.im;
        a0 = s * a0;
        a1 = s * a1;
    }
    cplx<R>* o = Lc + ((int64_t)child[item] * nthC * nphC + e) * 2;
    o[0] = o[0] + a0;
    o[1] = o[1] + a1;
}

template <typename R>
void launchL2LTheta(int nitems, const int32_t* child, int mthP, int nthC, int nphC, const int32_t* atPtr,
                    const int32_t* atSrc, const double* atW, const int32_t* atFlip, const cplx<R>* inner, cplx<R>* Lc,
                    hipStream_t st, const cplx<R>* postPhase, const int32_t* oct) {
    if (nitems <= 0) return;
    dim3 g((nthC * nphC + BS - 1) / BS, nitems);
    hipLaunchKernelGGL(kL2LTheta<R>, g, dim3(BS), 0, st, child, mthP, nthC, nphC, atPtr, atSrc, atW, atFlip, inner,
                       Lc, postPhase, oct);
}

// Dense matvec over all pairs with the near-field kernel (the reference's
// evalDirect path, src/mlfma.cpp:84-107); one thread per row, for small n.
__global__ void kDirect(SrcDev s, const cplx<double>* __restrict__ x, int64_t n, double k, cplx<double> Ck,
                        cplx<double>* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
    if (i >= n) return;
    cplx<double> acc = mk(0.0, 0.0);
    for (int64_t j = 0; j < n; ++j) {
        if (j == i && !s.physical) continue;
        acc = acc + entry(s, i, j, k) * x[j];
    }
    y[i] = Ck * acc;
}

void launchDirect(const SrcDev& s, const cplx<double>* xt, int64_t n, double k, cplx<double> Ck, cplx<double>* yt,
                  hipStream_t st) {
    hipLaunchKernelGGL(kDirect, dim3((unsigned)((n + BS - 1) / BS)), dim3(BS), 0, st, s, xt, n, k, Ck, yt);
}

// explicit instantiations
#define INST(R)                                                                                                   \
    template void launchM2MTheta<R>(int, const int32_t*, const int32_t*, const int32_t*, const int32_t*, int, int,  \
                                    int, int, const int32_t*, const double*, const cplx<R>*, const cplx<R>*,       \
                                    cplx<R>*, hipStream_t);                                                       \
    template void launchM2MPhi<R>(int, const int32_t*, int, int, int, int, const int32_t*, const double*,          \
                                  const cplx<R>*, cplx<R>*, hipStream_t);                                         \
    template void launchZeroSlots<R>(int, const int32_t*, int, cplx<R>*, hipStream_t);                             \
    template void launchL2LPhi<R>(int, const int32_t*, const int32_t*, int, int, int, const int32_t*,              \
                                  const int32_t*, const double*, const cplx<R>*, const cplx<R>*, cplx<R>*,        \
                                  hipStream_t);                                                                   \
    template void launchL2LTheta<R>(int, const int32_t*, int, int, int, const int32_t*, const int32_t*,            \
                                    const double*, const int32_t*, const cplx<R>*, cplx<R>*, hipStream_t,         \
                                    const cplx<R>*, const int32_t*);
INST(double)
INST(float)
#undef INST

}  // namespace mlfma
