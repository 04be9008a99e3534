// plan.hip — device plan and the C-ABI of include/mlfma_c.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <array>
#include <map>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "host.hpp"
#include "kernels.hpp"
#include "kernels_mv.hpp"
#include "plan.hpp"

#include <type_traits>

namespace mlfma {

thread_local std::string g_lastError;

void hipCheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(MLFMA_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

static const double kPi = std::acos(-1.0);

void* DevicePlan::alloc(size_t bytes) {
    if (bytes == 0) bytes = 16;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) throw Error(MLFMA_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    allocs.push_back(p);
    bytesHeld += (int64_t)bytes;
    return p;
}

void DevicePlan::release(void* p) {
    auto it = std::find(allocs.begin(), allocs.end(), p);
    if (it != allocs.end()) {
        (void)hipFree(p);
        allocs.erase(it);
    }
}

void DevicePlan::clearGraphs() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
}

// Probe events live in a ring; a pair is read (its elapsed time added) when
// its slot is needed again or when the caller asks for the total.
void DevicePlan::probeBegin(int group, hipStream_t s) {
    if (probeGroup != group) return;
    if (probeRing.empty()) {
        probeRing.resize(256);
        for (auto& p : probeRing) {
            hipCheck(hipEventCreate(&p.a), "hipEventCreate");
            hipCheck(hipEventCreate(&p.b), "hipEventCreate");
        }
    }
    if (probePending == probeRing.size()) probeDrain(probeRing.size() - 1);
    hipCheck(hipEventRecord(probeRing[probeHead].a, s), "event");
}

void DevicePlan::probeEnd(int group, hipStream_t s) {
    if (probeGroup != group) return;
    hipCheck(hipEventRecord(probeRing[probeHead].b, s), "event");
    probeHead = (probeHead + 1) % probeRing.size();
    ++probePending;
}

// read the oldest pending pairs until at most `keep` remain
void DevicePlan::probeDrain(size_t keep) {
    while (probePending > keep) {
        const size_t i = (probeHead + probeRing.size() - probePending) % probeRing.size();
        hipCheck(hipEventSynchronize(probeRing[i].b), "probe");
        float t = 0;
        hipCheck(hipEventElapsedTime(&t, probeRing[i].a, probeRing[i].b), "hipEventElapsedTime");
        probeMs += t;
        ++probeLaunches;
        --probePending;
    }
}

DevicePlan::~DevicePlan() {
    if (hostOnly) return;
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    if (nearStream) (void)hipStreamSynchronize(nearStream);
    if (m2lStream) (void)hipStreamSynchronize(m2lStream);
    clearGraphs();
    for (void* p : allocs) (void)hipFree(p);
    for (auto& p : probeRing) {
        if (p.a) (void)hipEventDestroy(p.a);
        if (p.b) (void)hipEventDestroy(p.b);
    }
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {evFork, evJoin, evM2L, evNear, evNearB, evJoinB})
        if (e) (void)hipEventDestroy(e);
    if (nearStream) (void)hipStreamDestroy(nearStream);
    if (m2lStream) (void)hipStreamDestroy(m2lStream);
    if (stream) (void)hipStreamDestroy(stream);
}

template <typename T>
T* DevicePlan::upload(const std::vector<T>& v) {
    T* d = (T*)alloc(v.size() * sizeof(T));
    if (!v.empty()) hipCheck(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
    return d;
}
template int64_t* DevicePlan::upload<int64_t>(const std::vector<int64_t>&);  // used by the solver

static std::vector<double> inodeDistances() {  // 15 distinct |d| of the M2L offsets (src/math.cpp:86-107)
    std::vector<double> d;
    for (int z = 0; z <= 3; ++z)
        for (int y = 0; y <= 3; ++y)
            for (int x = 0; x <= 3; ++x)
                if (std::max(x, std::max(y, z)) > 1)
                    d.push_back(std::sqrt((double)x * x + (double)y * y + (double)z * z));
    std::sort(d.begin(), d.end());
    d.erase(std::unique(d.begin(), d.end()), d.end());
    return d;
}

template <typename R>
static std::vector<cplx<R>> toR(const std::vector<cplx<double>>& v) {
    std::vector<cplx<R>> o(v.size());
    for (size_t i = 0; i < v.size(); ++i) {
        o[i].re = (R)v[i].re;
        o[i].im = (R)v[i].im;
    }
    return o;
}

// flags of the events that fork and join the plan's side streams
#ifndef MLFMA_SYNC_EVENT_FLAGS
#define MLFMA_SYNC_EVENT_FLAGS hipEventDisableTiming
#endif

void DevicePlan::build() {
    HostPlan& h = hp;
    hipCheck(hipSetDevice(device), "hipSetDevice");
    // this device's address of the fixed far-field partial-sum array (once per device)
    hipCheck(initFarfieldScratch(device), "hipGetSymbolAddress (far-field partial sums)");
    hipCheck(hipStreamCreate(&stream), "hipStreamCreate");
    {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && v > 0)
            ldsMax = (size_t)v;
        hipCheck(hipDeviceGetAttribute(&numCU, hipDeviceAttributeMultiprocessorCount, device), "device attribute");
    }
    // stage events need no system-scope fence: they order work of one stream
    for (auto& e : ev) hipCheck(hipEventCreateWithFlags(&e, hipEventDisableSystemFence), "hipEventCreate");
    // the near field runs beside the far-field chain on a second stream
    // (default priority), forked and joined with events
    hipCheck(hipStreamCreateWithFlags(&nearStream, hipStreamNonBlocking), "hipStreamCreate");
    hipCheck(hipEventCreateWithFlags(&evFork, MLFMA_SYNC_EVENT_FLAGS), "hipEventCreate");
    hipCheck(hipEventCreateWithFlags(&evJoin, MLFMA_SYNC_EVENT_FLAGS), "hipEventCreate");
    hipCheck(hipStreamCreateWithFlags(&m2lStream, hipStreamNonBlocking), "hipStreamCreate");
    hipCheck(hipEventCreateWithFlags(&evM2L, MLFMA_SYNC_EVENT_FLAGS), "hipEventCreate");
    hipCheck(hipEventCreateWithFlags(&evNear, MLFMA_SYNC_EVENT_FLAGS), "hipEventCreate");
    hipCheck(hipEventCreateWithFlags(&evNearB, MLFMA_SYNC_EVENT_FLAGS), "hipEventCreate");
    hipCheck(hipEventCreateWithFlags(&evJoinB, MLFMA_SYNC_EVENT_FLAGS), "hipEventCreate");
    const int64_t n = h.src.n;
    const Tree& t = h.tree;
    const int ML = t.maxLevel;
    useFp32 = (h.cfg.precision == MLFMA_FP32);

    // sources in tree order
    src.kind = h.src.kind;
    src.physical = (flags & MLFMA_FLAG_PHYSICAL) ? 1 : 0;
    if (h.src.kind == MLFMA_SRC_DIPOLE) {
        std::vector<DipoleDev> d(n);
        for (int64_t q = 0; q < n; ++q) {
            const int64_t i = t.perm[q];
            d[q].pos[0] = h.src.pos[i].x; d[q].pos[1] = h.src.pos[i].y; d[q].pos[2] = h.src.pos[i].z;
            d[q].phat[0] = h.src.phat[i].x; d[q].phat[1] = h.src.phat[i].y; d[q].phat[2] = h.src.phat[i].z;
        }
        src.dip = upload(d);
    } else {
        std::vector<RwgDev> r(n);
        for (int64_t q = 0; q < n; ++q) {
            const int64_t i = t.perm[q];
            r[q].ctr[0] = h.src.ctr[i].x; r[q].ctr[1] = h.src.ctr[i].y; r[q].ctr[2] = h.src.ctr[i].z;
            r[q].leng = h.src.leng[i];
            for (int s = 0; s < 2; ++s) {
                r[q].tri[s] = h.src.tri[s][i];
                r[q].xpm[s][0] = h.src.xpm[s][i].x;
                r[q].xpm[s][1] = h.src.xpm[s][i].y;
                r[q].xpm[s][2] = h.src.xpm[s][i].z;
            }
        }
        std::vector<TriDev> tr(h.src.tris.size());
        for (size_t q = 0; q < tr.size(); ++q) {
            const TriQuad& T = h.src.tris[q];
            std::memset(&tr[q], 0, sizeof(TriDev));
            for (int a = 0; a < 3; ++a) {
                tr[q].X0[a] = T.X[0][a];
                tr[q].D0[a] = T.D0[a];
                tr[q].D2[a] = T.D2[a];
            }
            tr[q].nq = T.nq;
            for (int j = 0; j < T.nq; ++j)
                for (int a = 0; a < 4; ++a) tr[q].q[j][a] = T.q[j][a];
        }
        src.rwg = upload(r);
        src.tri = upload(tr);
    }
    perm = upload(t.perm);

    // angular data per level
    levels.assign(ML + 1, LevelBufs());
    std::vector<LevelDev> lvd(ML + 1);
    for (int l = 0; l <= ML; ++l) {
        const Grid& g = h.grids[l];
        LevelBufs& L = levels[l];
        L.nth = g.nth;
        L.nph = g.nph;
        L.ndir = g.ndir();
        L.khat.resize((size_t)3 * L.ndir);
        std::vector<double> thph((size_t)6 * L.ndir);
        for (int i = 0; i < g.nth; ++i)
            for (int j = 0; j < g.nph; ++j) {
                const int d = i * g.nph + j;
                const double th = g.theta[i], ph = g.phi[j];
                L.khat[3 * d] = std::sin(th) * std::cos(ph);
                L.khat[3 * d + 1] = std::sin(th) * std::sin(ph);
                L.khat[3 * d + 2] = std::cos(th);
                double* M = &thph[6 * d];
                M[0] = std::cos(th) * std::cos(ph);
                M[1] = std::cos(th) * std::sin(ph);
                M[2] = -std::sin(th);
                M[3] = -std::sin(ph);
                M[4] = std::cos(ph);
                M[5] = 0.0;
            }
        lvd[l].nth = g.nth;
        lvd[l].nph = g.nph;
        lvd[l].ndir = g.ndir();
        lvd[l].khat = upload(L.khat);
        lvd[l].thph = upload(thph);
    }
    levelDev = upload(lvd);

    // coefficient storage: levels >= 2, non-empty boxes only
    int64_t off = 0;
    for (int l = 2; l <= ML; ++l) {
        levels[l].Moff = off;
        off += (int64_t)h.boxes[l] * levels[l].ndir * 2;
    }
    fieldLen = off;
    M64 = (cplx<double>*)alloc(std::max<int64_t>(fieldLen, 1) * sizeof(cplx<double>));
    L64 = (cplx<double>*)alloc(std::max<int64_t>(fieldLen, 1) * sizeof(cplx<double>));
    hipCheck(hipMemset(M64, 0, std::max<int64_t>(fieldLen, 1) * sizeof(cplx<double>)), "memset");
    hipCheck(hipMemset(L64, 0, std::max<int64_t>(fieldLen, 1) * sizeof(cplx<double>)), "memset");
    if (useFp32) {
        M32 = (cplx<float>*)alloc(std::max<int64_t>(fieldLen, 1) * sizeof(cplx<float>));
        L32 = (cplx<float>*)alloc(std::max<int64_t>(fieldLen, 1) * sizeof(cplx<float>));
        hipCheck(hipMemset(M32, 0, std::max<int64_t>(fieldLen, 1) * sizeof(cplx<float>)), "memset");
        hipCheck(hipMemset(L32, 0, std::max<int64_t>(fieldLen, 1) * sizeof(cplx<float>)), "memset");
    }

    // far-field leaves and their radiation patterns
    std::vector<LeafItem> items;
    maxLeafNdir = 0;
    for (const auto& f : h.farLeaves) {
        const Node& nd = t.nodes[f.node];
        LeafItem it;
        it.sbeg = f.sbeg;
        it.cnt = (int32_t)f.cnt;
        it.ndir = levels[f.level].ndir;
        it.rpOff = 2 * f.rpOff;
        it.fOff = levels[f.level].Moff + (int64_t)f.slot * it.ndir * 2;
        it.level = f.level;
        it.pad = 0;
        it.cx = nd.center.x;
        it.cy = nd.center.y;
        it.cz = nd.center.z;
        items.push_back(it);
        maxLeafNdir = std::max(maxLeafNdir, it.ndir);
    }
    nLeafItems = (int)items.size();
    leafItems = upload(items);
    {
        std::vector<int32_t> order(nLeafItems);
        for (int i = 0; i < nLeafItems; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
            return (int64_t)items[a].cnt * items[a].ndir > (int64_t)items[b].cnt * items[b].ndir;
        });
        s2mOrder = upload(order);
    }
    {
        std::vector<int64_t> rows;
        std::vector<int32_t> ritem;
        for (int i = 0; i < nLeafItems; ++i)
            for (int q = 0; q < items[i].cnt; ++q) {
                rows.push_back(items[i].sbeg + q);
                ritem.push_back(i);
            }
        nFarRows = (int64_t)rows.size();
        farRowPos = upload(rows);
        farRowItem = upload(ritem);
        // rows that no far-field leaf covers get their near-field sum alone
        std::vector<char> far(h.src.n, 0);
        for (int64_t q : rows) far[q] = 1;
        std::vector<int64_t> only;
        for (int64_t q = 0; q < h.src.n; ++q)
            if (!far[q]) only.push_back(q);
        nNearOnlyRows = (int64_t)only.size();
        nearOnlyRows = upload(only);
    }
    rp64 = (cplx<double>*)alloc(std::max<int64_t>(2 * h.radpatEntries, 1) * sizeof(cplx<double>));
    launchRadpat(src, leafItems, nLeafItems, levelDev, h.k, rp64, maxLeafNdir, stream);
    hipCheck(hipGetLastError(), "radpat launch");

    // translation operators (levels >= 2), quadrature weights folded in
    const std::vector<double> dists = inodeDistances();
    double* dDists = upload(dists);
    for (int l = 2; l <= ML; ++l) {
        LevelBufs& L = levels[l];
        const Grid& g = h.grids[l];
        const int nps = (int)std::floor(h.cfg.over_interp * (g.nth - 1));
        const int p = h.cfg.interp_order;
        if (nps < 2 * p || nps < 2) throw Error(MLFMA_ERR_ARG, "over_interp too small for the psi stencil");
        L.nps = nps;
        std::vector<double> psi((size_t)316 * L.ndir);
        std::vector<int32_t> nearIdx((size_t)316 * L.ndir), distIdx(316);
        for (int o = 0; o < 316; ++o) {
            int dx, dy, dz;
            offsetVector(o, dx, dy, dz);
            const double r = std::sqrt((double)dx * dx + (double)dy * dy + (double)dz * dz);
            const double rh[3] = {dx / r, dy / r, dz / r};
            int di = 0;
            for (size_t q = 0; q < dists.size(); ++q)
                if (std::fabs(dists[q] - r) < 1e-9) di = (int)q;
            distIdx[o] = di;
            for (int d = 0; d < L.ndir; ++d) {
                const double c = L.khat[3 * d] * rh[0] + L.khat[3 * d + 1] * rh[1] + L.khat[3 * d + 2] * rh[2];
                const double ps = std::acos(c);
                psi[(size_t)o * L.ndir + d] = ps;
                nearIdx[(size_t)o * L.ndir + d] = (int32_t)std::floor((nps - 1) * ps / kPi);
            }
        }
        std::vector<double> dirW(L.ndir);
        const double phw = 2.0 * kPi / static_cast<double>(g.nph);
        for (int i = 0; i < g.nth; ++i)
            for (int j = 0; j < g.nph; ++j) dirW[i * g.nph + j] = g.weight[i] * phw;
        L.dirW = dirW;
        double* dPsi = upload(psi);
        int32_t* dNear = upload(nearIdx);
        int32_t* dDist = upload(distIdx);
        double* dW = upload(dirW);
        cplx<double>* dAlpha = (cplx<double>*)alloc((size_t)15 * nps * sizeof(cplx<double>));
        L.T64 = (cplx<double>*)alloc((size_t)316 * L.ndir * sizeof(cplx<double>));
        const double a = h.cfg.root_leng / std::pow(2.0, l);
        launchAlpha(nps, g.L, h.k, a, dDists, dAlpha, stream);
        launchTransl(L.ndir, nps, p, dPsi, dNear, dDist, dAlpha, dW, L.T64, stream);
        hipCheck(hipGetLastError(), "translation launch");
        hipCheck(hipStreamSynchronize(stream), "translation");
        release(dPsi);
        release(dNear);
        release(dDist);
        release(dW);
        release(dAlpha);
    }
    release(dDists);

    // phase tables: M2M on the child grid, L2L on the parent grid
    for (int l = 2; l <= ML; ++l) {
        LevelBufs& L = levels[l];
        const double hl = h.cfg.root_leng / std::pow(2.0, l) / 2.0;  // half side of a level-l box
        std::vector<cplx<double>> up((size_t)8 * L.ndir), dn;
        for (int o = 0; o < 8; ++o) {
            const double pm[3] = {(o & 1) ? 1.0 : -1.0, (o & 2) ? 1.0 : -1.0, (o & 4) ? 1.0 : -1.0};
            for (int d = 0; d < L.ndir; ++d) {
                const double kv[3] = {L.khat[3 * d] * h.k, L.khat[3 * d + 1] * h.k, L.khat[3 * d + 2] * h.k};
                // parent centre minus this (child) centre = -hl * pm
                const double ph = kv[0] * (-hl * pm[0]) + kv[1] * (-hl * pm[1]) + kv[2] * (-hl * pm[2]);
                up[(size_t)o * L.ndir + d] = cplx<double>{std::cos(ph), std::sin(ph)};
            }
        }
        if (l < ML) {
            const double hc = h.cfg.root_leng / std::pow(2.0, l + 1) / 2.0;  // child half side
            dn.resize((size_t)8 * L.ndir);
            for (int o = 0; o < 8; ++o) {
                const double pm[3] = {(o & 1) ? 1.0 : -1.0, (o & 2) ? 1.0 : -1.0, (o & 4) ? 1.0 : -1.0};
                for (int d = 0; d < L.ndir; ++d) {
                    const double kv[3] = {L.khat[3 * d] * h.k, L.khat[3 * d + 1] * h.k, L.khat[3 * d + 2] * h.k};
                    const double ph = kv[0] * (hc * pm[0]) + kv[1] * (hc * pm[1]) + kv[2] * (hc * pm[2]);
                    dn[(size_t)o * L.ndir + d] = cplx<double>{std::cos(ph), std::sin(ph)};
                }
            }
        }
        L.phUp64 = upload(up);
        if (!dn.empty()) L.phDn64 = upload(dn);
        if (useFp32) {
            L.phUp32 = upload(toR<float>(up));
            if (!dn.empty()) L.phDn32 = upload(toR<float>(dn));
            std::vector<cplx<double>> Th((size_t)316 * L.ndir);
            hipCheck(hipMemcpy(Th.data(), L.T64, Th.size() * sizeof(cplx<double>), hipMemcpyDeviceToHost), "T copy");
            L.T32 = upload(toR<float>(Th));
        }
    }

    // interpolation tables and worklists
    int64_t innerLen = 1;
    for (int l = 2; l <= ML; ++l) {
        LevelBufs& L = levels[l];
        if (l < ML) {
            const InterpTable& I = h.interp[l];
            L.thNear = upload(I.thNear);
            L.thW = upload(I.thW);
            L.phNear = upload(I.phNear);
            L.phW = upload(I.phW);
            L.atPhPtr = upload(I.atPhPtr);
            L.atPhSrc = upload(I.atPhSrc);
            L.atPhW = upload(I.atPhW);
            L.atThPtr = upload(I.atThPtr);
            L.atThSrc = upload(I.atThSrc);
            L.atThW = upload(I.atThW);
            L.atThFlip = upload(I.atThFlip);
            L.nAtPh = (int)I.atPhSrc.size();
            L.nAtTh = (int)I.atThSrc.size();
            // M2M into level l: stems with non-empty children
            std::vector<int32_t> stem, ptr{0}, ch, oc;
            const auto& P = h.m2mPtr[l];
            for (size_t s = 0; s + 1 < P.size(); ++s) {
                if (P[s + 1] == P[s]) continue;
                if (P[s + 1] - P[s] > 8) throw Error(MLFMA_ERR_STATE, "a box with more than eight children");
                stem.push_back((int32_t)s);
                for (int q = P[s]; q < P[s + 1]; ++q) {
                    ch.push_back(h.m2mChild[l][q]);
                    oc.push_back(h.m2mOct[l][q]);
                }
                ptr.push_back((int32_t)ch.size());
            }
            L.nstem = (int)stem.size();
            L.stemSlot = upload(stem);
            L.mPtr = upload(ptr);
            L.mChild = upload(ch);
            L.mOct = upload(oc);
            {
                // per child (parent slot, child slot, octant): the adjoint of M2M
                std::vector<int32_t> ap;
                for (size_t q = 0; q + 1 < ptr.size(); ++q)
                    for (int c = ptr[q]; c < ptr[q + 1]; ++c) ap.push_back(stem[q]);
                L.m2mAdjParent = upload(ap);
                L.nm2mChild = (int)ap.size();
            }
            innerLen = std::max<int64_t>(innerLen, (int64_t)L.nstem * L.nth * levels[l + 1].nph * 2);
            // the two-pass adjoint of M2M holds one anterpolated field per child
            innerLen = std::max<int64_t>(innerLen, (int64_t)L.nm2mChild * L.nth * levels[l + 1].nph * 2);
        }
        // M2L targets with a non-empty interaction list; stale boxes otherwise
        std::vector<int32_t> tgt, ptr{0}, sr, of, stale;
        const auto& P = h.m2lPtr[l];
        for (size_t s = 0; s + 1 < P.size(); ++s) {
            if (P[s + 1] == P[s]) {
                stale.push_back((int32_t)s);
                continue;
            }
            tgt.push_back((int32_t)s);
            for (int q = P[s]; q < P[s + 1]; ++q) {
                sr.push_back(h.m2lSrc[l][q]);
                of.push_back(h.m2lOff[l][q]);
            }
            ptr.push_back((int32_t)sr.size());
        }
        L.ntgt = (int)tgt.size();
        L.tgtSlot = upload(tgt);
        L.lPtr = upload(ptr);
        L.lSrc = upload(sr);
        L.lOff = upload(of);
        {
            // the adjoint pass reads each pair reversed: operator of the negated offset
            std::vector<int32_t> neg(of.size());
            for (size_t q = 0; q < of.size(); ++q) {
                int dx, dy, dz;
                offsetVector(of[q], dx, dy, dz);
                neg[q] = offsetIndex(-dx, -dy, -dz);
            }
            L.lOffNeg = upload(neg);
            // runs of consecutive targets for the run-tiled M2L (kM2LT): a
            // run grows while its distinct sources stay within
            // kM2LRunSources (their LDS) and it has at most kM2LRunTargets
            // targets; pairs packed as (source index in the run) << 9 | offset
            std::vector<int32_t> runT{0}, runS{0}, slots, pk(of.size()), pkn(of.size());
            std::unordered_map<int32_t, int32_t> loc;
            int maxSrc = 0;
            size_t r0 = 0;  // first target of the open run
            auto closeRun = [&](size_t k) {
                std::vector<int32_t> srcs;
                srcs.reserve(loc.size());
                for (auto& kv : loc) srcs.push_back(kv.first);
                std::sort(srcs.begin(), srcs.end());
                for (size_t i = 0; i < srcs.size(); ++i) loc[srcs[i]] = (int32_t)i;
                for (size_t t2 = r0; t2 < k; ++t2)
                    for (int q = ptr[t2]; q < ptr[t2 + 1]; ++q) {
                        pk[q] = loc[sr[q]] << 9 | of[q];
                        pkn[q] = loc[sr[q]] << 9 | neg[q];
                    }
                slots.insert(slots.end(), srcs.begin(), srcs.end());
                maxSrc = std::max(maxSrc, (int)srcs.size());
                runT.push_back((int32_t)k);
                runS.push_back((int32_t)slots.size());
                loc.clear();
                r0 = k;
            };
            for (size_t k = 0; k < tgt.size(); ++k) {
                int fresh = 0;
                for (int q = ptr[k]; q < ptr[k + 1]; ++q) fresh += loc.count(sr[q]) ? 0 : 1;
                if (k > r0 && ((int)(loc.size() + fresh) > kM2LRunSources || (int)(k - r0) >= kM2LRunTargets))
                    closeRun(k);
                for (int q = ptr[k]; q < ptr[k + 1]; ++q) loc.emplace(sr[q], 0);
            }
            if (tgt.size() > r0) closeRun(tgt.size());
            for (int32_t v : of)
                if (v < 0 || v >= kM2LOffsets) throw Error(MLFMA_ERR_STATE, "M2L offset out of range");
            if (maxSrc >= (1 << 22)) throw Error(MLFMA_ERR_STATE, "M2L run too large to pack");
            L.nrun = (int)runT.size() - 1;
            L.runMaxSrc = maxSrc;
            L.runTgt = upload(runT);
            L.runSrc = upload(runS);
            L.runSlot = upload(slots);
            L.lPk = upload(pk);
            L.lPkNeg = upload(pkn);
        }
        {
            // sibling groups for M2L: the targets of level l under one parent
            // share most of their sources; per group, every distinct source
            // with its offset to each of the (up to) eight sibling targets
            std::unordered_map<int, int> gid;  // parent node -> group
            std::vector<std::array<int32_t, 8>> gt;
            std::vector<std::map<int32_t, std::array<int32_t, 8>>> gs;
            for (size_t q = 0; q < tgt.size(); ++q) {
                const int slot = tgt[q];
                const Node& nd = t.nodes[h.slotNode[l][slot]];
                auto it = gid.find(nd.parent);
                int g;
                if (it == gid.end()) {
                    g = (int)gt.size();
                    gid.emplace(nd.parent, g);
                    gt.emplace_back();
                    gt.back().fill(-1);
                    gs.emplace_back();
                } else
                    g = it->second;
                gt[g][nd.branch] = slot;
                for (int e = ptr[q]; e < ptr[q + 1]; ++e) {
                    auto ins = gs[g].emplace(sr[e], std::array<int32_t, 8>{});
                    if (ins.second) ins.first->second.fill(-1);
                    ins.first->second[nd.branch] = of[e];
                }
            }
            std::vector<int32_t> gtf, gp{0}, gsrc, goff, goffn;
            for (size_t g = 0; g < gt.size(); ++g) {
                for (int k = 0; k < 8; ++k) gtf.push_back(gt[g][k]);
                for (auto& kv : gs[g]) {
                    gsrc.push_back(kv.first);
                    for (int k = 0; k < 8; ++k) {
                        const int o = kv.second[k];
                        goff.push_back(o);
                        int dx = 0, dy = 0, dz = 0;
                        if (o >= 0) offsetVector(o, dx, dy, dz);
                        goffn.push_back(o >= 0 ? offsetIndex(-dx, -dy, -dz) : -1);
                    }
                }
                gp.push_back((int32_t)gsrc.size());
            }
            L.ngrp = (int)gt.size();
            L.nGrpEntries = (int64_t)gsrc.size();
            L.nPairs = (int64_t)sr.size();
            L.gTgt = upload(gtf);
            L.gPtr = upload(gp);
            L.gSrc = upload(gsrc);
            L.gOff = upload(goff);
            L.gOffNeg = upload(goffn);
        }
        L.nstale = (int)stale.size();
        L.staleSlots = upload(stale);
        // L2L into level l (from l-1)
        if (l >= 3) {
            L.nl2l = (int)h.l2lChild[l].size();
            L.l2lChild = upload(h.l2lChild[l]);
            L.l2lParent = upload(h.l2lParent[l]);
            L.l2lOct = upload(h.l2lOct[l]);
            innerLen = std::max<int64_t>(innerLen, (int64_t)L.nl2l * levels[l - 1].nth * L.nph * 2);
            // the same items grouped by parent (the adjoint of L2L runs per parent)
            std::vector<int> order(L.nl2l);
            for (int q = 0; q < L.nl2l; ++q) order[q] = q;
            std::stable_sort(order.begin(), order.end(),
                             [&](int a, int b) { return h.l2lParent[l][a] < h.l2lParent[l][b]; });
            std::vector<int32_t> ps, pp{0}, pc, po;
            for (int q = 0; q < L.nl2l; ++q) {
                const int it = order[q];
                if (ps.empty() || ps.back() != h.l2lParent[l][it]) {
                    if (!ps.empty()) pp.push_back((int32_t)pc.size());
                    ps.push_back(h.l2lParent[l][it]);
                }
                pc.push_back(h.l2lChild[l][it]);
                po.push_back(h.l2lOct[l][it]);
            }
            if (!ps.empty()) pp.push_back((int32_t)pc.size());
            L.nl2lPar = (int)ps.size();
            L.l2lParSlot = upload(ps);
            L.l2lParPtr = upload(pp);
            L.l2lParChild = upload(pc);
            L.l2lParOct = upload(po);
        }
    }
    inner64 = (cplx<double>*)alloc(innerLen * sizeof(cplx<double>));
    if (useFp32) inner32 = (cplx<float>*)alloc(innerLen * sizeof(cplx<float>));

    // Completion counters of the upward tree walk (kTreeUp): one per far-field
    // leaf item (its S2M units), then one per stem of each M2M level, finest
    // parent level first; per M2M child entry the counter of the unit group
    // that writes that child's multipole (-1: nothing in the walk writes it)
    if (ML - 2 > kTreeMaxLevels) throw Error(MLFMA_ERR_STATE, "too many levels for the tree walk");
    {
        std::vector<std::vector<int32_t>> leafOf(ML + 1), stemOf(ML + 1);
        for (int l = 0; l <= ML; ++l) {
            leafOf[l].assign(h.slotNode[l].size(), -1);
            stemOf[l].assign(h.slotNode[l].size(), -1);
        }
        for (size_t i = 0; i < h.farLeaves.size(); ++i) leafOf[h.farLeaves[i].level][h.farLeaves[i].slot] = (int32_t)i;
        treeCtr0.assign(ML + 1, 0);
        int base = nLeafItems;
        for (int l = ML - 1; l >= 2; --l) {
            treeCtr0[l] = base;
            const auto& P = h.m2mPtr[l];
            int k = 0;
            for (size_t s = 0; s + 1 < P.size(); ++s)
                if (P[s + 1] != P[s]) stemOf[l][s] = k++;
            if (k != levels[l].nstem) throw Error(MLFMA_ERR_STATE, "tree walk stems out of step");
            base += k;
        }
        for (int l = ML - 1; l >= 2; --l) {
            std::vector<int32_t> cc;
            const auto& P = h.m2mPtr[l];
            for (size_t s = 0; s + 1 < P.size(); ++s)
                for (int q = P[s]; q < P[s + 1]; ++q) {
                    const int c = h.m2mChild[l][q];
                    const int li = leafOf[l + 1][c];
                    const int si = l + 1 <= ML - 1 ? stemOf[l + 1][c] : -1;
                    if (li >= 0 && si >= 0) throw Error(MLFMA_ERR_STATE, "a box both leaf and stem");
                    cc.push_back(li >= 0 ? li : (si >= 0 ? treeCtr0[l + 1] + si : -1));
                }
            levels[l].mChildCtr = upload(cc);
        }
        treeCtrLen = base + kTreeQueues + 2;
        treeCtr = (int32_t*)alloc((size_t)treeCtrLen * sizeof(int32_t));
        hipCheck(hipMemset(treeCtr, 0, (size_t)treeCtrLen * sizeof(int32_t)), "memset");
    }

    // near field
    std::vector<NearItem> nitemsH(h.nearItems.size());
    for (size_t q = 0; q < nitemsH.size(); ++q) {
        const auto& I = h.nearItems[q];
        if (I.nr < 1 || I.nr > 128 || I.nc < 1 || I.nself < 0 || I.nself > I.nc)
            throw Error(MLFMA_ERR_STATE, "near-field item out of range");
        nitemsH[q] = NearItem{I.rbeg, I.nr, I.nc, I.nself, 0, I.cptr, I.voff, I.rpo, I.cpo};
    }
    // the build kernel locates an entry's item by binary search over voff, so
    // its list stays in value order
    for (size_t q = 0; q < nitemsH.size(); ++q) {
        const int64_t end = nitemsH[q].voff + (int64_t)nitemsH[q].nr * nitemsH[q].nc;
        if ((q > 0 && nitemsH[q].voff < nitemsH[q - 1].voff) || end > h.nearNnz)
            throw Error(MLFMA_ERR_STATE, "near-field items out of value order");
    }
    nNearItems = (int)nitemsH.size();
    NearItem* itemsByValue = upload(nitemsH);
    int64_t* colsTree = upload(h.nearCols);
    {
        std::vector<int32_t> ci(h.nearCols.size());
        for (size_t q = 0; q < ci.size(); ++q) ci[q] = (int32_t)t.perm[h.nearCols[q]];
        nearColsIn = upload(ci);
    }
    nearSumPtr = upload(h.nearSumPtr);
    nearSumIdx = upload(h.nearSumIdx);
    nearPart = (cplx<double>*)alloc(std::max<int64_t>(h.nearParts, 1) * sizeof(cplx<double>));
    nearVal = (cplx<double>*)alloc(std::max<int64_t>(h.nearNnz, 1) * sizeof(cplx<double>));
    const std::complex<double> C = std::complex<double>(0.0, -1.0) * 299792458.0 * 1.256637E-6 / (4.0 * kPi);
    const std::complex<double> Ckc = C * h.k;
    Ck = cplx<double>{Ckc.real(), Ckc.imag()};
    launchNearBuild(src, itemsByValue, nNearItems, colsTree, h.nearNnz, h.k, Ck, nearVal, stream);
    hipCheck(hipGetLastError(), "near build launch");
    if (useFp32) {  // FP32 copy for the FP32 far-field path (entries computed in FP64)
        nearVal32 = (cplx<float>*)alloc(std::max<int64_t>(h.nearNnz, 1) * sizeof(cplx<float>));
        launchToFloat(nearVal, nearVal32, h.nearNnz, stream);
        hipCheck(hipGetLastError(), "near copy launch");
    }
    hipCheck(hipStreamSynchronize(stream), "near build");
    release(itemsByValue);
    release(colsTree);
    // the apply kernel takes one block per item: longest items first
    std::stable_sort(nitemsH.begin(), nitemsH.end(), [](const NearItem& a, const NearItem& b) {
        return (int64_t)a.nr * a.nc > (int64_t)b.nr * b.nc;
    });
    nearItems = upload(nitemsH);

    // radiation patterns in FP32 for the mixed-precision far field
    if (useFp32) {
        std::vector<cplx<double>> tmp(2 * h.radpatEntries);
        hipCheck(hipStreamSynchronize(stream), "radpat");
        if (!tmp.empty())
            hipCheck(hipMemcpy(tmp.data(), rp64, tmp.size() * sizeof(cplx<double>), hipMemcpyDeviceToHost), "rp copy");
        rp32 = upload(toR<float>(tmp));
    }

    // excitation
    cplx<double>* dV = (cplx<double>*)alloc(n * sizeof(cplx<double>));
    const double wv[3] = {h.pw.wavenum * h.pw.khat[0], h.pw.wavenum * h.pw.khat[1], h.pw.wavenum * h.pw.khat[2]};
    launchVoltage(src, n, wv, h.pw.pol, h.pw.amplitude, dV, stream);
    hipCheck(hipGetLastError(), "voltage launch");
    std::vector<cplx<double>> vt(n);
    hipCheck(hipStreamSynchronize(stream), "setup");
    hipCheck(hipMemcpy(vt.data(), dV, n * sizeof(cplx<double>), hipMemcpyDeviceToHost), "voltage copy");
    release(dV);
    Vin.resize(n);
    for (int64_t q = 0; q < n; ++q) Vin[t.perm[q]] = std::complex<double>(vt[q].re, vt[q].im);

    ensureRhs(1);
}

// Gauss-Jordan inversion with partial pivoting of a column-major s x s
// complex block in place; false when a pivot falls below tiny.
static bool invertBlock(std::complex<double>* A, int s, double tiny) {
    std::vector<int> piv(s);
    for (int k = 0; k < s; ++k) {
        int p = k;
        double best = std::abs(A[(size_t)k * s + k]);
        for (int i = k + 1; i < s; ++i) {
            const double v = std::abs(A[(size_t)k * s + i]);
            if (v > best) {
                best = v;
                p = i;
            }
        }
        if (!(best > tiny)) return false;
        piv[k] = p;
        if (p != k)
            for (int j = 0; j < s; ++j) std::swap(A[(size_t)j * s + k], A[(size_t)j * s + p]);
        const std::complex<double> d = 1.0 / A[(size_t)k * s + k];
        A[(size_t)k * s + k] = 1.0;
        for (int j = 0; j < s; ++j) A[(size_t)j * s + k] *= d;
        for (int i = 0; i < s; ++i) {
            if (i == k) continue;
            const std::complex<double> f = A[(size_t)k * s + i];
            if (f == 0.0) continue;
            A[(size_t)k * s + i] = 0.0;
            for (int j = 0; j < s; ++j) A[(size_t)j * s + i] -= f * A[(size_t)j * s + k];
        }
    }
    for (int k = s - 1; k >= 0; --k)  // undo the row swaps as column swaps of the inverse
        if (piv[k] != k)
            for (int i = 0; i < s; ++i) std::swap(A[(size_t)k * s + i], A[(size_t)piv[k] * s + i]);
    return true;
}

// The leaf block-Jacobi preconditioner: the self block of every non-empty
// leaf (its stored near-field entries, C k folded in) is gathered on the
// device, inverted on the host (leaves in parallel) and uploaded.  A block
// that is singular to working precision (e.g. a one-source leaf of the
// reference's kernel, whose diagonal is zero) is inverted with a small
// diagonal shift, and an all-zero block gets a scaled identity.
void DevicePlan::buildPrecond() {
    if (precVal) return;
    const HostPlan& h = hp;
    const Tree& t = h.tree;
    std::vector<PrecBlock> blocks;
    std::vector<int64_t> idx;
    size_t item = 0;
    int64_t off = 0;
    for (int id = 0; id < (int)t.nodes.size(); ++id) {
        const Node& nd = t.nodes[id];
        if (!nd.leaf() || nd.count() == 0) continue;
        const int64_t cnt = nd.count();
        if (cnt > 8192) throw Error(MLFMA_ERR_ARG, "the block preconditioner takes leaves of at most 8192 sources");
        blocks.push_back(PrecBlock{nd.sbeg, (int32_t)cnt, 0, off});
        const size_t base = idx.size();
        idx.resize(base + (size_t)cnt * cnt, -1);
        // this leaf's items come next in the item list; the first has c0 = 0
        if (item >= h.nearItems.size() || h.nearItems[item].rbeg != nd.sbeg)
            throw Error(MLFMA_ERR_STATE, "near-field items out of leaf order");
        const int64_t cbase = h.nearItems[item].cptr;
        while (item < h.nearItems.size() && h.nearItems[item].rbeg >= nd.sbeg && h.nearItems[item].rbeg < nd.send) {
            const auto& it = h.nearItems[item];
            const int64_t c0 = it.cptr - cbase, r0 = it.rbeg - nd.sbeg;
            for (int j = 0; j < it.nself; ++j)
                for (int i = 0; i < it.nr; ++i)
                    idx[base + (size_t)(c0 + j) * cnt + (r0 + i)] = it.voff + (int64_t)j * it.nr + i;
            ++item;
        }
        for (size_t q = base; q < idx.size(); ++q)
            if (idx[q] < 0) throw Error(MLFMA_ERR_STATE, "a self-block entry is missing from the near field");
        off += cnt * cnt;
        precMaxCnt = std::max(precMaxCnt, (int)cnt);
    }
    nPrecBlocks = (int)blocks.size();
    int64_t* dIdx = upload(idx);
    cplx<double>* dS = (cplx<double>*)alloc(std::max<int64_t>(off, 1) * sizeof(cplx<double>));
    launchGatherValues(nearVal, dIdx, off, dS, stream);
    hipCheck(hipGetLastError(), "gather launch");
    std::vector<std::complex<double>> S(off);
    hipCheck(hipStreamSynchronize(stream), "gather");
    if (off) hipCheck(hipMemcpy(S.data(), dS, off * sizeof(cplx<double>), hipMemcpyDeviceToHost), "block download");
    release(dIdx);
    // a scale for all-zero blocks: the median of the blocks' largest entries
    std::vector<double> mx(blocks.size(), 0.0);
    for (size_t b = 0; b < blocks.size(); ++b)
        for (int64_t q = 0; q < (int64_t)blocks[b].cnt * blocks[b].cnt; ++q)
            mx[b] = std::max(mx[b], std::abs(S[blocks[b].off + q]));
    std::vector<double> nz;
    for (double v : mx)
        if (v > 0) nz.push_back(v);
    std::sort(nz.begin(), nz.end());
    const double typical = nz.empty() ? 1.0 : nz[nz.size() / 2];
    std::atomic<size_t> next{0};
    std::atomic<int> regularised{0};
    auto work = [&]() {
        std::vector<std::complex<double>> A;
        for (size_t b; (b = next++) < blocks.size();) {
            const int c = blocks[b].cnt;
            std::complex<double>* B = S.data() + blocks[b].off;
            if (mx[b] == 0) {
                for (int64_t q = 0; q < (int64_t)c * c; ++q) B[q] = 0.0;
                for (int i = 0; i < c; ++i) B[(size_t)i * c + i] = 1.0 / typical;
                ++regularised;
                continue;
            }
            A.assign(B, B + (size_t)c * c);
            if (invertBlock(A.data(), c, 1e-13 * mx[b])) {
                std::copy(A.begin(), A.end(), B);
                continue;
            }
            ++regularised;
            for (double shift : {1e-8, 1e-5, 1e-2, 1.0}) {
                A.assign(B, B + (size_t)c * c);
                for (int i = 0; i < c; ++i) A[(size_t)i * c + i] += shift * mx[b];
                if (invertBlock(A.data(), c, 1e-13 * mx[b])) {
                    std::copy(A.begin(), A.end(), B);
                    break;
                }
            }
        }
    };
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned i = 1; i < nt; ++i) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
    precRegularised = regularised.load();
    if (off) hipCheck(hipMemcpy(dS, S.data(), off * sizeof(cplx<double>), hipMemcpyHostToDevice), "block upload");
    precVal = dS;
    precBlocks = upload(blocks);
}

void DevicePlan::applyPrecond(const cplx<double>* x, cplx<double>* y, int nb, int64_t xs, bool adjoint,
                              hipStream_t st) {
    launchBlockPrec(precBlocks, nPrecBlocks, precMaxCnt, precVal, perm, x, y, nb, xs, adjoint, st);
}

void DevicePlan::ensureRhs(int nrhs) {
    if (nrhs <= rhsCap) return;
    if (xt) release(xt);
    if (yt) release(yt);
    if (xs) release(xs);
    if (ys) release(ys);
    const int64_t n = hp.src.n;
    xt = (cplx<double>*)alloc((size_t)n * nrhs * sizeof(cplx<double>));
    yt = (cplx<double>*)alloc((size_t)n * nrhs * sizeof(cplx<double>));
    xs = (cplx<double>*)alloc((size_t)n * nrhs * sizeof(cplx<double>));
    ys = (cplx<double>*)alloc((size_t)n * nrhs * sizeof(cplx<double>));
    rhsCap = nrhs;
}

// One matvec y (=|+=) Z x on input-order device vectors (single right-hand side).
// M2L of every level: levels whose sibling targets share most of their
// sources (at least three list entries per distinct source) and that fill the
// device go through the sibling-group kernel, the others through the
// per-target kernel.  adjoint: conjugated operators of the reversed pairs.
// nearX (one right-hand side): the near-field pairs of x (conjugated values
// for the adjoint) go into the launch of the per-target levels.
template <typename R>
bool DevicePlan::launchM2L(const cplx<R>* in, cplx<R>* out, bool adjoint, hipStream_t st, int nb, int lmin, int lmax,
                           const cplx<double>* nearX) {
    const int ML = hp.tree.maxLevel;
    if (ML - 1 > kMaxLevels) throw Error(MLFMA_ERR_ARG, "too many tree levels");
    M2LGroups<R> g;
    M2LLevels<R> t;
    M2LTiles<R> r;
    g.fs = t.fs = r.fs = fieldLen;
    for (int l = std::max(2, lmin); l <= std::min(ML, lmax); ++l) {
        LevelBufs& L = levels[l];
        if (L.ntgt == 0) continue;
        // kernel per level: the target runs unless tuned or forced otherwise
        int mode = L.nrun > 0 ? 3 : 2;
        if (l < (int)tuneM2L.size() && tuneM2L[l]) mode = tuneM2L[l];  // autotuned choice
        if (mode == 3 && L.nrun == 0) mode = 2;
        if (flags & MLFMA_FLAG_M2L_GROUP) mode = 1;
        if (flags & MLFMA_FLAG_M2L_TARGET) mode = 2;
        const bool group = mode == 1;
        if (mode == 3) {
            const int k = r.nlev++;
            r.ndir[k] = L.ndir;
            r.nrun[k] = L.nrun;
            r.tgt[k] = L.tgtSlot;
            r.ptr[k] = L.lPtr;
            r.pk[k] = adjoint ? L.lPkNeg : L.lPk;
            r.run[k] = L.runTgt;
            r.rsrc[k] = L.runSrc;
            r.slot[k] = L.runSlot;
            r.T[k] = L.T<R>();
            r.M[k] = in + L.Moff;
            r.L[k] = out + L.Moff;
            r.maxSrc = std::max(r.maxSrc, L.runMaxSrc);
        } else if (group) {
            const int k = g.nlev++;
            g.ndir[k] = L.ndir;
            g.ngrp[k] = L.ngrp;
            g.tgt[k] = L.gTgt;
            g.ptr[k] = L.gPtr;
            g.src[k] = L.gSrc;
            g.off[k] = adjoint ? L.gOffNeg : L.gOff;
            g.T[k] = L.T<R>();
            g.M[k] = in + L.Moff;
            g.L[k] = out + L.Moff;
        } else {
            const int k = t.nlev++;
            t.ndir[k] = L.ndir;
            t.ntgt[k] = L.ntgt;
            t.tgt[k] = L.tgtSlot;
            t.ptr[k] = L.lPtr;
            t.src[k] = L.lSrc;
            t.off[k] = adjoint ? L.lOffNeg : L.lOff;
            t.T[k] = L.T<R>();
            t.M[k] = in + L.Moff;
            t.L[k] = out + L.Moff;
        }
    }
    if (g.nlev) {
        probeBegin(MLFMA_GROUP_M2L_SIBLING, st);
        launchM2LGroup<R>(g, adjoint, st, nb);
        probeEnd(MLFMA_GROUP_M2L_SIBLING, st);
    }
    if (r.nlev) {
        if (m2ltLds<R>(r.maxSrc) > ldsMax) throw Error(MLFMA_ERR_STATE, "M2L run sources exceed the LDS");
        probeBegin(MLFMA_GROUP_M2L_RUNS, st);
        launchM2LT<R>(r, adjoint, st, nb);
        probeEnd(MLFMA_GROUP_M2L_RUNS, st);
    }
    if (nearX && t.nlev && nb == 1 && nNearItems > 0) {
        // the FP32 far field reads the FP32 copy of the near-field values
        // (the adjoint the FP64 values)
        probeBegin(MLFMA_GROUP_M2L_NEAR, st);
        if constexpr (std::is_same<R, float>::value)
            if (nearVal32 && !adjoint) {
                launchM2LNear<R, float>(t, nearItems, nNearItems, nearColsIn, perm, nearVal32, nearX, nearPart, st,
                                        false, nearWindow);
                probeEnd(MLFMA_GROUP_M2L_NEAR, st);
                return true;
            }
        launchM2LNear<R, double>(t, nearItems, nNearItems, nearColsIn, perm, nearVal, nearX, nearPart, st, adjoint,
                                 nearWindow);
        probeEnd(MLFMA_GROUP_M2L_NEAR, st);
        return true;
    }
    if (t.nlev) {
        probeBegin(MLFMA_GROUP_M2L_TARGET, st);
        launchM2Lx<R>(t, adjoint, st, nb);
        probeEnd(MLFMA_GROUP_M2L_TARGET, st);
    }
    return false;
}

template <typename R>
void DevicePlan::apply(const cplx<double>* x, cplx<double>* y, int accumulate, hipStream_t st, bool timed, int nb,
                       int64_t xs) {
    // nb right-hand sides (1, 2, 4 or 8) with column stride xs; their fields
    // at stride fs in the field arrays, their near-field partials at stride ps
    if (nb > batchCap) throw Error(MLFMA_ERR_STATE, "right-hand-side batch larger than the plan's buffers");
    const int ML = hp.tree.maxLevel;
    const int p = hp.cfg.interp_order;
    const int64_t fs = fieldLen, ps = hp.nearParts;
    cplx<R>* M = fieldM<R>();
    cplx<R>* Lc = fieldL<R>();
    cplx<R>* inner = scratch<R>();
    const cplx<R>* rp = radpat<R>();
    // the near-field pairs depend on x only: fork them onto the near stream
    // so they overlap the far-field chain, before S2M (nearOverlap 1), after
    // it (2: the bandwidth-bound pair products then run beside the
    // latency-bound tree stages) or after the M2M chain (3: the M2M launches,
    // with few blocks at coarse levels, then run uncontended); in order (0)
    // they go into the per-target M2L launch of one right-hand side, which
    // leaves them HBM bandwidth, or else (stages timed, batches, no
    // per-target level) follow the L2L
    if (nb == 8) launchInterleave8(x, xs, hp.src.n, xInner, st);  // before any fork: every near launch reads it
    const bool serial = timed || (flags & MLFMA_FLAG_ONE_STREAM) || solveInOrder;
    const int overlap = serial ? 0 : nearOverlap;
    // M2L of the finest level needs only the leaf multipoles, and its locals
    // are first read by the L2L into that level: it runs on its own stream
    // beside the M2M chain and the coarser M2L
    const bool splitM2L = !serial && m2lSplit && ML >= 2 && levels[ML].ntgt > 0;
    // Each event record or wait on st costs the chain a few microseconds, so
    // both side streams fork from one event (evFork after S2M, or after the
    // M2M chain in mode 3; the near field forked before S2M has its own) and
    // join through one: the M2L stream waits for the near field before it
    // records evM2L.
    // beside the chain the work items [0, nearCut) fork at the overlap point
    // and the rest after the M2L (every item writes its own partial sums)
    const int nearCut = overlap && nearSplit < 100 ? (int)((int64_t)nNearItems * nearSplit / 100) : nNearItems;
    auto launchNear = [&](hipStream_t s, int i0, int cnt) {
        if (cnt <= 0) return;
        const NearItem* items = nearItems + i0;
        probeBegin(MLFMA_GROUP_NEAR, s);
        struct End {
            DevicePlan* d;
            hipStream_t s;
            ~End() { d->probeEnd(MLFMA_GROUP_NEAR, s); }
        } end{this, s};
        // eight right-hand sides: the matrix-core kernel on x interleaved,
        // with the FP64 values also under the FP32 far field (the FP32-value
        // instance needs more registers than the three-waves-per-SIMD budget
        // and runs slower: 0.62 against 0.44 ms per pass at 201,720 RWGs)
        if (nb == 8) {
            launchNearSymM8<double>(false, items, cnt, nearColsIn, perm, nearVal, xInner, nearPart, s);
            return;
        }
        // the FP32 far field (mixed-precision inner solves) reads the FP32
        // copy of the near-field values
        // beside the chain, a fixed number of blocks (nearBlocks per CU)
        const int grid = s == nearStream ? nearGrid() : 0;
        if (std::is_same<R, float>::value && nearVal32)
            launchNearSym<float>(false, items, cnt, nearColsIn, perm, nearVal32, x, nearPart, s, nb, xs,
                                 ps, grid);
        else
            launchNearSym<double>(false, items, cnt, nearColsIn, perm, nearVal, x, nearPart, s, nb, xs, ps,
                                  grid);
    };
    if (overlap == 1) {
        hipCheck(hipEventRecord(evNear, st), "event");
        hipCheck(hipStreamWaitEvent(nearStream, evNear, 0), "wait");
        launchNear(nearStream, 0, nearCut);
        hipCheck(hipEventRecord(evJoin, nearStream), "event");
    }
    if (timed) hipCheck(hipEventRecord(ev[0], st), "event");
    // one right-hand side, stages not timed: S2M and the M2M levels as one
    // persistent walk when chosen (false: a level's fields exceed the LDS)
    bool walkedUp = false;
    if (nb == 1 && !timed && (treeWalk & 1)) {
        probeBegin(MLFMA_GROUP_TREE_UP, st);
        walkedUp = launchTreeUpWalk<R>(x, st);
        probeEnd(MLFMA_GROUP_TREE_UP, st);
    }
    if (!walkedUp) {
        probeBegin(MLFMA_GROUP_S2M, st);
        launchS2Mx<R>(leafItems, s2mOrder, nLeafItems, maxLeafNdir, rp, x, perm, M, st, nb, xs, fs);
        probeEnd(MLFMA_GROUP_S2M, st);
    }
    const bool nearSide = overlap == 2 || overlap == 3;
    auto forkSides = [&]() {
        if (!nearSide && !splitM2L) return;
        hipCheck(hipEventRecord(evFork, st), "event");
        if (nearSide) {
            hipCheck(hipStreamWaitEvent(nearStream, evFork, 0), "wait");
            launchNear(nearStream, 0, nearCut);
            hipCheck(hipEventRecord(evJoin, nearStream), "event");
        }
        if (splitM2L) {
            hipCheck(hipStreamWaitEvent(m2lStream, evFork, 0), "wait");
            launchM2L<R>(M, Lc, false, m2lStream, nb, ML, ML);
            if (overlap) hipCheck(hipStreamWaitEvent(m2lStream, evJoin, 0), "wait");
            hipCheck(hipEventRecord(evM2L, m2lStream), "event");
        }
    };
    if (overlap != 3) forkSides();
    // st waits once for the side streams: before the L2L into the finest
    // level when M2L was split (the near field is joined through it), else
    // before the near-field sums
    bool sidePending = splitM2L || overlap;
    auto joinSide = [&]() {
        if (!sidePending) return;
        hipCheck(hipStreamWaitEvent(st, splitM2L ? evM2L : evJoin, 0), "wait");
        sidePending = false;
    };
    if (timed) hipCheck(hipEventRecord(ev[1], st), "event");
    for (int l = ML - 1; l >= 2 && !walkedUp; --l) {
        LevelBufs& P = levels[l];
        LevelBufs& C = levels[l + 1];
        probeBegin(MLFMA_GROUP_M2M, st);
        if (!launchM2MFused<R>(P.nstem, P.stemSlot, P.mPtr, P.mChild, P.mOct, P.nth, P.nph, C.nth, C.nph, p, P.thNear,
                               P.thW, P.phNear, P.phW, C.phUp<R>(), M + C.Moff, M + P.Moff, ldsMax, st, nb, fs,
                               minBlocksAt(l))) {
            for (int r = 0; r < nb; ++r) {
                launchM2MTheta<R>(P.nstem, P.stemSlot, P.mPtr, P.mChild, P.mOct, P.nth, C.nth, C.nph, p, P.thNear,
                                  P.thW, C.phUp<R>(), M + r * fs + C.Moff, inner, st);
                launchM2MPhi<R>(P.nstem, P.stemSlot, P.nth, P.nph, C.nph, p, P.phNear, P.phW, inner,
                                M + r * fs + P.Moff, st);
            }
        }
        probeEnd(MLFMA_GROUP_M2M, st);
    }
    if (overlap == 3) forkSides();
    if (timed) hipCheck(hipEventRecord(ev[2], st), "event");
    // stages in order: the near-field pairs share the per-target M2L launch
    // (not when stages are timed, so that each has its own events)
    const bool nearFused =
        launchM2L<R>(M, Lc, false, st, nb, 2, splitM2L ? ML - 1 : ML, !overlap && !timed ? x : nullptr);
    if (flags & MLFMA_FLAG_FRESH_LOCALS)
        for (int r = 0; r < nb; ++r)
            for (int l = 2; l <= ML; ++l) {
                LevelBufs& L = levels[l];
                launchZeroSlots<R>(L.nstale, L.staleSlots, L.ndir, Lc + r * fs + L.Moff, st);
            }
    if (timed) hipCheck(hipEventRecord(ev[3], st), "event");
    const bool nearLate = nearCut < nNearItems;
    if (nearLate) {
        hipCheck(hipEventRecord(evNearB, st), "event");
        hipCheck(hipStreamWaitEvent(nearStream, evNearB, 0), "wait");
        launchNear(nearStream, nearCut, nNearItems - nearCut);
        hipCheck(hipEventRecord(evJoinB, nearStream), "event");
    }
    for (int l = 3; l <= ML; ++l) {
        LevelBufs& P = levels[l - 1];
        LevelBufs& C = levels[l];
        if (l == ML) joinSide();
        probeBegin(MLFMA_GROUP_L2L, st);
        if (!launchL2LFused<R>(false, C.nl2l, C.l2lParent, C.l2lChild, C.l2lOct, P.nth, P.nph, C.nth, C.nph, P.atPhPtr,
                               P.atPhSrc, P.atPhW, P.nAtPh, P.atThPtr, P.atThSrc, P.atThW, P.atThFlip, P.nAtTh,
                               P.phDn<R>(), Lc + P.Moff, Lc + C.Moff, ldsMax, st, nb, fs)) {
            for (int r = 0; r < nb; ++r) {
                launchL2LPhi<R>(C.nl2l, C.l2lParent, C.l2lOct, P.nth, P.nph, C.nph, P.atPhPtr, P.atPhSrc, P.atPhW,
                                P.phDn<R>(), Lc + r * fs + P.Moff, inner, st);
                launchL2LTheta<R>(C.nl2l, C.l2lChild, P.nth, C.nth, C.nph, P.atThPtr, P.atThSrc, P.atThW, P.atThFlip,
                                  inner, Lc + r * fs + C.Moff, st);
            }
        }
        probeEnd(MLFMA_GROUP_L2L, st);
    }
    joinSide();
    if (nearLate) hipCheck(hipStreamWaitEvent(st, evJoinB, 0), "wait");
    if (timed) hipCheck(hipEventRecord(ev[4], st), "event");
    if (!overlap && !nearFused) launchNear(st, 0, nNearItems);
    if (timed) hipCheck(hipEventRecord(ev[5], st), "event");
    launchNearSum(nearOnlyRows, nNearOnlyRows, nearSumPtr, nearSumIdx, nearPart, perm, y, accumulate, st, nb, xs, ps);
    if (timed) hipCheck(hipEventRecord(ev[6], st), "event");
    probeBegin(MLFMA_GROUP_L2T, st);
    if (nb != 8 || !launchL2TM8<R>(leafItems, nLeafItems, maxLeafNdir, rp, Lc, fs, Ck, farRowPos, nFarRows,
                                   nearSumPtr, nearSumIdx, nearPart, perm, y, xs, accumulate, ldsMax, st))
        launchL2Tx<R>(farRowPos, farRowItem, nFarRows, leafItems, rp, Lc, Ck, nearSumPtr, nearSumIdx, nearPart, perm,
                      y, accumulate, st, nb, fs, xs, ps);
    probeEnd(MLFMA_GROUP_L2T, st);
    if (timed) hipCheck(hipEventRecord(ev[7], st), "event");
}

int DevicePlan::treeGrid(bool fp32, int bs, size_t lds) {
    for (const auto& o : treeOcc)
        if (o[0] == (int)fp32 && o[1] == bs && o[2] == (int)lds) return o[3] * numCU;
    const int n = fp32 ? treeUpOccupancy<float>(bs, lds) : treeUpOccupancy<double>(bs, lds);
    if (n <= 0) throw Error(MLFMA_ERR_STATE, "tree walk does not fit a compute unit");
    treeOcc.push_back({(int)fp32, bs, (int)lds, n});
    return n * numCU;
}

// The upward walk of one right-hand side (kTreeUp): the worklist is the S2M
// units (kTreeS2MPer leaf pieces each, in the S2M order) and then, per parent
// level from the finest, its M2M units with the row runs of the staged M2M
// launch.
template <typename R>
bool DevicePlan::launchTreeUpWalk(const cplx<double>* x, hipStream_t st) {
    const int ML = hp.tree.maxLevel;
    const int p = hp.cfg.interp_order;
    TreeUp<R> tu;
    const int bs = maxLeafNdir <= 192 ? 192 : 256;
    tu.nLeaf = nLeafItems;
    tu.s2mChunks = std::max(1, (maxLeafNdir + bs - 1) / bs);
    tu.nS2MPieces = nLeafItems * tu.s2mChunks;
    tu.s2mPer = kTreeS2MPer;
    tu.nS2M = (tu.nS2MPieces + tu.s2mPer - 1) / tu.s2mPer;
    tu.p = p;
    tu.ctr = treeCtr;
    tu.nctr = treeCtrLen;
    tu.items = leafItems;
    tu.order = s2mOrder;
    tu.rp = radpat<R>();
    tu.x = x;
    tu.perm = perm;
    tu.M = fieldM<R>();
    size_t lds = 0;
    int unit = tu.nS2M;
    for (int l = ML - 1; l >= 2; --l) {
        const LevelBufs& P = levels[l];
        const LevelBufs& C = levels[l + 1];
        TreeUpLevel<R>& L = tu.lv[tu.nlev++];
        size_t ll = 0;
        const int rows = P.nstem > 0 ? m2mRows<R>(P.nstem, P.nth, P.nph, C.nth, C.nph, p, ldsMax, minBlocksAt(l), &ll) : 1;
        if (rows == 0) return false;
        lds = std::max(lds, ll);
        L.nstem = P.nstem;
        L.rows = rows;
        L.nchunk = (P.nth + rows - 1) / rows;
        L.unit0 = unit;
        unit += P.nstem * L.nchunk;
        L.nthP = P.nth;
        L.nphP = P.nph;
        L.mthC = C.nth;
        L.mphC = C.nph;
        L.ctr0 = treeCtr0[l];
        L.stemSlot = P.stemSlot;
        L.ptr = P.mPtr;
        L.child = P.mChild;
        L.oct = P.mOct;
        L.childCtr = P.mChildCtr;
        L.thNear = P.thNear;
        L.phNear = P.phNear;
        L.thW = P.thW;
        L.phW = P.phW;
        L.phaseUp = C.phUp<R>();
        L.MoffC = C.Moff;
        L.MoffP = P.Moff;
    }
    tu.total = unit;
    // every ticket queue that has units has a block
    const int grid = std::max(std::min(treeGrid(std::is_same<R, float>::value, bs, lds), tu.total),
                              std::min(tu.total, kTreeQueues));
    launchTreeUp<R>(tu, bs, lds, grid, st);
    return true;
}

// y (=|+=) Z^H x.  Z = near + C k R^H A S with S (S2M) the radiation
// patterns, A = L2L . M2L . M2M and R^H (L2T) the conjugated patterns, so
// Z^H = conj(near) + conj(C k) S^H A^H R: the patterns aggregate x into the
// leaf locals (the S2M kernel), the adjoint of L2L carries them up, M2L with
// conjugated operators of the reversed pairs maps locals to multipoles, the
// adjoint of M2M carries those down, and the L2T kernel with conj(C k)
// disaggregates them.  This is the adjoint of the linear operator: boxes
// with an empty interaction list contribute nothing (fresh locals).
template <typename R>
void DevicePlan::applyAdjoint(const cplx<double>* x, cplx<double>* y, int accumulate, hipStream_t st) {
    const int ML = hp.tree.maxLevel;
    const int p = hp.cfg.interp_order;
    cplx<R>* M = fieldM<R>();
    cplx<R>* Lc = fieldL<R>();
    const cplx<R>* rp = radpat<R>();
    const bool overlap = nearOverlap != 0 && !(flags & MLFMA_FLAG_ONE_STREAM) && !solveInOrder;
    if (overlap) {
        hipCheck(hipEventRecord(evFork, st), "event");
        hipCheck(hipStreamWaitEvent(nearStream, evFork, 0), "wait");
        launchNearSym<double>(true, nearItems, nNearItems, nearColsIn, perm, nearVal, x, nearPart, nearStream, 1, 0, 0,
                              nearGrid());
        hipCheck(hipEventRecord(evJoin, nearStream), "event");
    }
    launchZeroField<R>(M, fieldLen, st);
    launchZeroField<R>(Lc, fieldLen, st);
    launchS2Mx<R>(leafItems, s2mOrder, nLeafItems, maxLeafNdir, rp, x, perm, Lc, st);
    for (int l = ML; l >= 3; --l) {
        LevelBufs& P = levels[l - 1];
        LevelBufs& C = levels[l];
        if (!launchL2LAdj<R>(C.nl2lPar, C.l2lParSlot, C.l2lParPtr, C.l2lParChild, C.l2lParOct, P.nth, P.nph, C.nth,
                             C.nph, p, P.thNear, P.thW, P.phNear, P.phW, P.phDn<R>(), Lc + C.Moff, Lc + P.Moff, ldsMax,
                             st, minBlocksAt(l - 1)))
            throw Error(MLFMA_ERR_ARG, "adjoint of the L2L into level " + std::to_string(l) +
                                           ": the level's fields exceed the LDS of one block (the adjoint has no "
                                           "two-pass path; use fewer digits or smaller leaf boxes)");
    }
    // adjoint locals in, adjoint multipoles out; in order, with the
    // near-field pairs in the same launch
    const bool nearFused = launchM2L<R>(Lc, M, true, st, 1, 2, ML, overlap ? nullptr : x);
    for (int l = 2; l < ML; ++l) {
        LevelBufs& P = levels[l];
        LevelBufs& C = levels[l + 1];
        if (!launchL2LFused<R>(true, P.nm2mChild, P.m2mAdjParent, P.mChild, P.mOct, P.nth, P.nph, C.nth, C.nph,
                               P.atPhPtr, P.atPhSrc, P.atPhW, P.nAtPh, P.atThPtr, P.atThSrc, P.atThW, P.atThFlip,
                               P.nAtTh, C.phUp<R>(), M + P.Moff, M + C.Moff, ldsMax, st)) {
            // two passes through global memory when the parent grid does not
            // fit one block's LDS: phi anterpolation of the parent per child,
            // then theta anterpolation, conj child-grid phase and the sum
            cplx<R>* inner = scratch<R>();
            launchL2LPhi<R>(P.nm2mChild, P.m2mAdjParent, P.mOct, P.nth, P.nph, C.nph, P.atPhPtr, P.atPhSrc, P.atPhW,
                            nullptr, M + P.Moff, inner, st);
            launchL2LTheta<R>(P.nm2mChild, P.mChild, P.nth, C.nth, C.nph, P.atThPtr, P.atThSrc, P.atThW, P.atThFlip,
                              inner, M + C.Moff, st, C.phUp<R>(), P.mOct);
        }
    }
    if (overlap)
        hipCheck(hipStreamWaitEvent(st, evJoin, 0), "wait");
    else if (!nearFused)
        launchNearSym<double>(true, nearItems, nNearItems, nearColsIn, perm, nearVal, x, nearPart, st);
    launchNearSum(nearOnlyRows, nNearOnlyRows, nearSumPtr, nearSumIdx, nearPart, perm, y, accumulate, st);
    const cplx<double> Ckc{Ck.re, -Ck.im};
    launchL2Tx<R>(farRowPos, farRowItem, nFarRows, leafItems, rp, M, Ckc, nearSumPtr, nearSumIdx, nearPart, perm, y,
                  accumulate, st);
}

void DevicePlan::applyOpAdjoint(const cplx<double>* x, cplx<double>* y, int accumulate, hipStream_t st, bool fp32) {
    if (hp.staleBoxes > 0 && !(flags & MLFMA_FLAG_FRESH_LOCALS))
        throw Error(MLFMA_ERR_STATE,
                    "the operator keeps stale locals (boxes with an empty interaction list) and is not linear; "
                    "set MLFMA_FLAG_FRESH_LOCALS for an adjoint");
    if (fp32) {
        if (!useFp32) throw Error(MLFMA_ERR_STATE, "plan was not built for FP32");
        applyAdjoint<float>(x, y, accumulate, st);
    } else
        applyAdjoint<double>(x, y, accumulate, st);
}

void DevicePlan::applyOp(const cplx<double>* x, cplx<double>* y, int accumulate, hipStream_t st, bool fp32, bool timed,
                         int nb, int64_t xs) {
    if (fp32) {
        if (!useFp32) throw Error(MLFMA_ERR_STATE, "plan was not built for FP32");
        apply<float>(x, y, accumulate, st, timed, nb, xs);
    } else
        apply<double>(x, y, accumulate, st, timed, nb, xs);
}

// Grow the field arrays and near-field partials to hold nb right-hand sides.
void DevicePlan::ensureBatch(int nb) {
    if (nb <= batchCap) return;
    hipCheck(hipDeviceSynchronize(), "batch resize");
    clearGraphs();  // captured graphs hold the old field addresses
    const int64_t len = std::max<int64_t>(fieldLen, 1) * nb;
    auto regrow = [&](auto*& ptr, size_t elem) {
        using T = std::remove_reference_t<decltype(*ptr)>;
        if (ptr) release(ptr);
        ptr = (T*)alloc(len * elem);
        hipCheck(hipMemset(ptr, 0, len * elem), "memset");
    };
    regrow(M64, sizeof(cplx<double>));
    regrow(L64, sizeof(cplx<double>));
    if (useFp32) {
        regrow(M32, sizeof(cplx<float>));
        regrow(L32, sizeof(cplx<float>));
    }
    release(nearPart);
    nearPart = (cplx<double>*)alloc(std::max<int64_t>(hp.nearParts, 1) * nb * sizeof(cplx<double>));
    if (nb >= 8 && !xInner) xInner = (cplx<double>*)alloc(std::max<int64_t>(hp.src.n, 1) * 8 * sizeof(cplx<double>));
    batchCap = nb;
}

// Right-hand sides go through in batches of 8, 4, 2, 1 when the operator is
// linear; a plan that keeps stale locals (the reference's state carried from
// one matvec to the next) applies them one after another.
void DevicePlan::matvec(const cplx<double>* x, cplx<double>* y, int nrhs, int accumulate, hipStream_t st, bool fp32) {
    const int64_t n = hp.src.n;
    const bool timed = (flags & MLFMA_FLAG_STAGE_TIMES) != 0;
    const bool linear = hp.staleBoxes == 0 || (flags & MLFMA_FLAG_FRESH_LOCALS);
    // MLFMA_FLAG_GRAPH: replay a captured launch sequence for these buffers.
    // The capture runs on the plan's own stream (torch's default stream cannot
    // be captured) and the graph is launched on st.  The graph holds kernel
    // launches only (no memset or host work), its buffers are sized first.
    if ((flags & MLFMA_FLAG_GRAPH) && !timed && linear && probeGroup == 0) {
        for (size_t i = 0; i < graphs.size(); ++i) {
            GraphEntry& g = graphs[i];
            if (g.x == x && g.y == y && g.nrhs == nrhs && g.accumulate == accumulate && g.fp32 == fp32) {
                hipCheck(hipGraphLaunch(g.exec, st), "hipGraphLaunch");
                std::rotate(graphs.begin() + i, graphs.begin() + i + 1, graphs.end());
                return;
            }
        }
        ensureBatch(nrhs >= 8 ? 8 : (nrhs >= 4 ? 4 : (nrhs >= 2 ? 2 : 1)));
        // the capture stream must not start before st's earlier work
        hipCheck(hipEventRecord(ev[8], st), "event");
        hipCheck(hipStreamWaitEvent(stream, ev[8], 0), "wait");
        hipCheck(hipStreamSynchronize(stream), "capture");
        hipGraph_t graph = nullptr;
        hipCheck(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
            int r = 0;
            while (r < nrhs) {
                const int left = nrhs - r;
                const int nb = left >= 8 ? 8 : (left >= 4 ? 4 : (left >= 2 ? 2 : 1));
                applyOp(x + (size_t)r * n, y + (size_t)r * n, accumulate, stream, fp32, false, nb, n);
                r += nb;
            }
        } catch (...) {
            (void)hipStreamEndCapture(stream, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        hipCheck(hipStreamEndCapture(stream, &graph), "hipStreamEndCapture");
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        hipCheck(e, "hipGraphInstantiate");
        if (graphs.size() >= 4) {
            (void)hipGraphExecDestroy(graphs.front().exec);
            graphs.erase(graphs.begin());
        }
        graphs.push_back(GraphEntry{x, y, nrhs, accumulate, fp32, exec});
        hipCheck(hipGraphLaunch(exec, st), "hipGraphLaunch");
        return;
    }
    int r = 0;
    while (r < nrhs) {
        const int left = nrhs - r;
        const int nb = !linear ? 1 : (left >= 8 ? 8 : (left >= 4 ? 4 : (left >= 2 ? 2 : 1)));
        if (nb > 1) ensureBatch(nb);
        applyOp(x + (size_t)r * n, y + (size_t)r * n, accumulate, st, fp32, timed && r == 0, nb, n);
        r += nb;
    }
    hipCheck(hipGetLastError(), "matvec launch");
    if (timed) timesValid = true;
}

}  // namespace mlfma

// ====================================================================== C-ABI
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

static DevicePlan& P(mlfma_plan* p) {
    if (!p || !p->d) throw Error(MLFMA_ERR_ARG, "null plan");
    return *p->d;
}
static const DevicePlan& P(const mlfma_plan* p) {
    if (!p || !p->d) throw Error(MLFMA_ERR_ARG, "null plan");
    return *p->d;
}
static void needDevice(const DevicePlan& d) {
    if (d.hostOnly) throw Error(MLFMA_ERR_STATE, "host-only plan (device = -1) has no device data");
}

extern "C" {

const char* mlfma_last_error(void) { return g_lastError.c_str(); }

int mlfma_plan_create(const mlfma_geometry* geom, const mlfma_planewave* pw, const mlfma_config* cfg,
                      mlfma_plan** out) {
    API_BEGIN
    if (!geom || !pw || !cfg || !out) throw Error(MLFMA_ERR_ARG, "null argument");
    *out = nullptr;
    auto plan = std::make_unique<mlfma_plan>();
    plan->d = std::make_unique<DevicePlan>();
    DevicePlan& d = *plan->d;
    if ((cfg->flags & MLFMA_FLAG_PHYSICAL) && geom->kind != MLFMA_SRC_RWG)
        throw Error(MLFMA_ERR_ARG, "MLFMA_FLAG_PHYSICAL needs RWG sources");
    buildHostPlan(d.hp, *geom, *pw, *cfg);
    d.flags = cfg->flags;
    d.hostOnly = cfg->device < 0;
    d.device = cfg->device;
    if (!d.hostOnly) d.build();
    *out = plan.release();
    return MLFMA_OK;
    API_END
}

int mlfma_plan_create_forest(const mlfma_geometry* geoms, const mlfma_config* cfgs, int32_t count,
                             const mlfma_planewave* pw, mlfma_plan** out) {
    API_BEGIN
    if (!geoms || !cfgs || !pw || !out || count < 1) throw Error(MLFMA_ERR_ARG, "bad forest arguments");
    *out = nullptr;
    auto plan = std::make_unique<mlfma_plan>();
    plan->d = std::make_unique<DevicePlan>();
    DevicePlan& d = *plan->d;
    std::vector<HostPlan> parts(count);
    for (int s = 0; s < count; ++s) {
        if (cfgs[s].device != cfgs[0].device) throw Error(MLFMA_ERR_ARG, "the scatterers of a forest share one device");
        if ((cfgs[s].flags & MLFMA_FLAG_PHYSICAL) && geoms[s].kind != MLFMA_SRC_RWG)
            throw Error(MLFMA_ERR_ARG, "MLFMA_FLAG_PHYSICAL needs RWG sources");
        buildHostPlan(parts[s], geoms[s], *pw, cfgs[s]);
    }
    mergeHostPlans(d.hp, parts);
    parts.clear();
    d.flags = cfgs[0].flags;
    d.hostOnly = cfgs[0].device < 0;
    d.device = cfgs[0].device;
    if (!d.hostOnly) d.build();
    *out = plan.release();
    return MLFMA_OK;
    API_END
}

int mlfma_plan_segments(const mlfma_plan* plan, int64_t* offsets, int32_t cap, int32_t* count) {
    API_BEGIN
    const DevicePlan& d = P(plan);
    if (!count) throw Error(MLFMA_ERR_ARG, "null argument");
    const HostPlan& h = d.hp;
    std::vector<int64_t> seg = h.forest ? h.segOff : std::vector<int64_t>{0, h.src.n};
    *count = (int32_t)seg.size() - 1;
    if (offsets)
        for (int32_t i = 0; i < std::min<int32_t>(cap, (int32_t)seg.size()); ++i) offsets[i] = seg[i];
    return MLFMA_OK;
    API_END
}

int mlfma_plan_destroy(mlfma_plan* plan) {
    API_BEGIN
    delete plan;
    return MLFMA_OK;
    API_END
}

int mlfma_plan_info(const mlfma_plan* plan, mlfma_info* o) {
    API_BEGIN
    const DevicePlan& d = P(plan);
    if (!o) throw Error(MLFMA_ERR_ARG, "null info");
    std::memset(o, 0, sizeof(*o));
    const HostPlan& h = d.hp;
    o->n = h.src.n;
    o->max_level = h.tree.maxLevel;
    o->num_nodes = (int64_t)h.tree.nodes.size();
    for (const auto& nd : h.tree.nodes)
        if (nd.leaf()) ++o->num_leaves;
    for (int l = 0; l <= h.tree.maxLevel && l < 16; ++l) {
        o->nth[l] = h.grids[l].nth;
        o->nph[l] = h.grids[l].nph;
        o->boxes[l] = h.boxes[l];
        o->m2l_pairs[l] = l < (int)h.m2lPairs.size() ? h.m2lPairs[l] : 0;
        if (o->m2l_pairs[l] > 0) ++o->num_levels_m2l;
    }
    o->near_nnz = h.nearNnz;
    o->radpat_entries = h.radpatEntries;
    o->stale_boxes = h.staleBoxes;
    o->truncated_l2l = h.truncatedL2L;
    o->psi_merged = 0;
    o->device_bytes = d.bytesHeld;
    return MLFMA_OK;
    API_END
}

int mlfma_voltage(const mlfma_plan* plan, double* v) {
    API_BEGIN
    const DevicePlan& d = P(plan);
    needDevice(d);
    if (!v) throw Error(MLFMA_ERR_ARG, "null output");
    for (size_t i = 0; i < d.Vin.size(); ++i) {
        v[2 * i] = d.Vin[i].real();
        v[2 * i + 1] = d.Vin[i].imag();
    }
    return MLFMA_OK;
    API_END
}

int mlfma_excitation(mlfma_plan* plan, const mlfma_planewave* pw, double* v) {
    API_BEGIN
    DevicePlan& d = P(plan);
    needDevice(d);
    if (!pw || !v) throw Error(MLFMA_ERR_ARG, "null argument");
    const double kn = std::sqrt(pw->khat[0] * pw->khat[0] + pw->khat[1] * pw->khat[1] + pw->khat[2] * pw->khat[2]);
    if (!(kn > 0) || std::abs(pw->wavenum - d.hp.pw.wavenum) > 1e-12 * d.hp.pw.wavenum)
        throw Error(MLFMA_ERR_ARG, "the plane wave needs a direction and the plan's wavenumber");
    hipCheck(hipSetDevice(d.device), "hipSetDevice");
    const int64_t n = d.hp.src.n;
    d.ensureRhs(1);
    const double wv[3] = {pw->wavenum * pw->khat[0] / kn, pw->wavenum * pw->khat[1] / kn,
                          pw->wavenum * pw->khat[2] / kn};
    launchVoltage(d.src, n, wv, pw->pol, pw->amplitude, d.yt, d.stream);
    hipCheck(hipGetLastError(), "voltage launch");
    std::vector<cplx<double>> vt(n);
    hipCheck(hipMemcpyAsync(vt.data(), d.yt, n * sizeof(cplx<double>), hipMemcpyDeviceToHost, d.stream), "V copy");
    hipCheck(hipStreamSynchronize(d.stream), "excitation");
    for (int64_t q = 0; q < n; ++q) {  // tree order -> input order
        const int64_t i = d.hp.tree.perm[q];
        v[2 * i] = vt[q].re;
        v[2 * i + 1] = vt[q].im;
    }
    return MLFMA_OK;
    API_END
}

int mlfma_matvec(mlfma_plan* plan, const double* x, double* y, int32_t nrhs, int32_t accumulate, int32_t ptr_mode,
                 void* stream) {
    API_BEGIN
    DevicePlan& d = P(plan);
    needDevice(d);
    if (!x || !y || nrhs < 1) throw Error(MLFMA_ERR_ARG, "bad matvec arguments");
    hipCheck(hipSetDevice(d.device), "hipSetDevice");
    hipStream_t st = stream ? (hipStream_t)stream : d.stream;
    const int64_t n = d.hp.src.n;
    const bool fp32 = d.useFp32;
    if (ptr_mode == MLFMA_PTR_DEVICE) {
        d.matvec((const cplx<double>*)x, (cplx<double>*)y, nrhs, accumulate, st, fp32);
        if (!(d.flags & MLFMA_FLAG_ASYNC)) hipCheck(hipStreamSynchronize(st), "matvec");
    } else if (ptr_mode == MLFMA_PTR_HOST) {
        d.ensureRhs(nrhs);
        const size_t bytes = (size_t)n * nrhs * sizeof(cplx<double>);
        hipCheck(hipMemcpyAsync(d.xs, x, bytes, hipMemcpyHostToDevice, st), "x upload");
        if (accumulate) hipCheck(hipMemcpyAsync(d.ys, y, bytes, hipMemcpyHostToDevice, st), "y upload");
        d.matvec(d.xs, d.ys, nrhs, accumulate, st, fp32);
        hipCheck(hipMemcpyAsync(y, d.ys, bytes, hipMemcpyDeviceToHost, st), "y download");
        hipCheck(hipStreamSynchronize(st), "matvec");
    } else
        throw Error(MLFMA_ERR_ARG, "bad ptr_mode");
    return MLFMA_OK;
    API_END
}

int mlfma_matvec_adjoint(mlfma_plan* plan, const double* x, double* y, int32_t nrhs, int32_t accumulate,
                         int32_t ptr_mode, void* stream) {
    API_BEGIN
    DevicePlan& d = P(plan);
    needDevice(d);
    if (!x || !y || nrhs < 1) throw Error(MLFMA_ERR_ARG, "bad matvec arguments");
    hipCheck(hipSetDevice(d.device), "hipSetDevice");
    hipStream_t st = stream ? (hipStream_t)stream : d.stream;
    const int64_t n = d.hp.src.n;
    const bool fp32 = d.useFp32;
    auto run = [&](const cplx<double>* xd, cplx<double>* yd) {
        for (int r = 0; r < nrhs; ++r) d.applyOpAdjoint(xd + (size_t)r * n, yd + (size_t)r * n, accumulate, st, fp32);
        hipCheck(hipGetLastError(), "adjoint matvec launch");
    };
    if (ptr_mode == MLFMA_PTR_DEVICE) {
        run((const cplx<double>*)x, (cplx<double>*)y);
        if (!(d.flags & MLFMA_FLAG_ASYNC)) hipCheck(hipStreamSynchronize(st), "adjoint matvec");
    } else if (ptr_mode == MLFMA_PTR_HOST) {
        d.ensureRhs(nrhs);
        const size_t bytes = (size_t)n * nrhs * sizeof(cplx<double>);
        hipCheck(hipMemcpyAsync(d.xs, x, bytes, hipMemcpyHostToDevice, st), "x upload");
        if (accumulate) hipCheck(hipMemcpyAsync(d.ys, y, bytes, hipMemcpyHostToDevice, st), "y upload");
        run(d.xs, d.ys);
        hipCheck(hipMemcpyAsync(y, d.ys, bytes, hipMemcpyDeviceToHost, st), "y download");
        hipCheck(hipStreamSynchronize(st), "adjoint matvec");
    } else
        throw Error(MLFMA_ERR_ARG, "bad ptr_mode");
    return MLFMA_OK;
    API_END
}

int mlfma_stage_times(const mlfma_plan* plan, double* ms) {
    API_BEGIN
    const DevicePlan& d = P(plan);
    needDevice(d);
    if (!d.timesValid) throw Error(MLFMA_ERR_STATE, "no matvec has run with MLFMA_FLAG_STAGE_TIMES set");
    auto el = [&](int a, int b) {
        float t = 0;
        hipCheck(hipEventElapsedTime(&t, d.ev[a], d.ev[b]), "hipEventElapsedTime");
        return (double)t;
    };
    ms[0] = el(0, 1);  // S2M
    ms[1] = el(1, 2);  // M2M
    ms[2] = el(2, 3);  // M2L
    ms[3] = el(3, 4);  // L2L
    ms[4] = el(6, 7);  // L2T
    ms[5] = el(4, 6);  // near field (pair products + fold)
    ms[6] = el(0, 7);  // total (first right-hand side)
    return MLFMA_OK;
    API_END
}

int mlfma_autotune(mlfma_plan* plan, int32_t nrhs, double* ms_before, double* ms_after) {
    API_BEGIN
    DevicePlan& d = P(plan);
    needDevice(d);
    if (nrhs != 1 && nrhs != 2 && nrhs != 4 && nrhs != 8) throw Error(MLFMA_ERR_ARG, "nrhs must be 1, 2, 4 or 8");
    const bool linear = d.hp.staleBoxes == 0 || (d.flags & MLFMA_FLAG_FRESH_LOCALS);
    if (nrhs > 1 && !linear)
        throw Error(MLFMA_ERR_STATE, "batched tuning needs a linear operator (MLFMA_FLAG_FRESH_LOCALS)");
    hipCheck(hipSetDevice(d.device), "hipSetDevice");
    const uint32_t saved = d.flags;
    hipCheck(hipStreamSynchronize(d.stream), "autotune");
    d.clearGraphs();
    d.flags = (saved & ~(MLFMA_FLAG_GRAPH | MLFMA_FLAG_ASYNC | MLFMA_FLAG_M2L_GROUP | MLFMA_FLAG_M2L_TARGET)) |
              MLFMA_FLAG_STAGE_TIMES;
    const int64_t n = d.hp.src.n;
    d.ensureRhs(nrhs);
    if (nrhs > 1) d.ensureBatch(nrhs);
    hipCheck(hipMemset(d.xt, 0, (size_t)n * nrhs * sizeof(cplx<double>)), "x");
    // mean device time of events a -> b over reps matvecs after one warmup
    // (the matvec the plan's callers run: the FP32 far field on an FP32 plan)
    auto timeStage = [&](int a, int b, int reps) {
        d.matvec(d.xt, d.yt, nrhs, 0, d.stream, d.useFp32);
        double acc = 0;
        for (int r = 0; r < reps; ++r) {
            d.matvec(d.xt, d.yt, nrhs, 0, d.stream, d.useFp32);
            hipCheck(hipStreamSynchronize(d.stream), "autotune");
            float t = 0;
            hipCheck(hipEventElapsedTime(&t, d.ev[a], d.ev[b]), "hipEventElapsedTime");
            acc += t;
        }
        return acc / reps;
    };
    // mean device time of whole matvecs (no stage events) over reps
    auto timeWhole = [&](int reps) {
        const uint32_t f = d.flags;
        d.flags &= ~MLFMA_FLAG_STAGE_TIMES;
        d.matvec(d.xt, d.yt, nrhs, 0, d.stream, d.useFp32);
        hipCheck(hipEventRecord(d.ev[0], d.stream), "event");
        for (int r = 0; r < reps; ++r) d.matvec(d.xt, d.yt, nrhs, 0, d.stream, d.useFp32);
        hipCheck(hipEventRecord(d.ev[7], d.stream), "event");
        hipCheck(hipStreamSynchronize(d.stream), "autotune");
        d.flags = f;
        float t = 0;
        hipCheck(hipEventElapsedTime(&t, d.ev[0], d.ev[7]), "hipEventElapsedTime");
        return (double)t / reps;
    };
    try {
        const double before = timeWhole(10);
        // M2M: block target for the parent row split, level by level (the
        // split re-sums the children in every block, so few large parents
        // and many small ones want different targets)
        d.tuneMinBlocks.assign(d.hp.tree.maxLevel + 1, 0);
        for (int l = d.hp.tree.maxLevel - 1; l >= 2; --l) {
            if (d.levels[l].nstem == 0) continue;
            double best = 1e300;
            int bm = 0;
            for (int mb : {128, 256, 512, 1024, 2048, 4096}) {
                d.tuneMinBlocks[l] = mb;
                const double t = timeStage(1, 2, 4);
                if (t < best) {
                    best = t;
                    bm = mb;
                }
            }
            d.tuneMinBlocks[l] = bm;
        }
        // M2L: sibling-group (1), per-target (2) or target-run (3) kernel
        // per level, every combination when there are at most five such
        // levels (a level at a time from the previous choice oscillated
        // between two settings)
        d.tuneM2L.assign(d.hp.tree.maxLevel + 1, 0);
        {
            std::vector<int> lv;
            for (int l = 2; l <= d.hp.tree.maxLevel; ++l)
                if (d.levels[l].ntgt > 0) lv.push_back(l);
            auto modes = [&](int l) { return d.levels[l].nrun > 0 ? 3 : 2; };
            if (lv.size() <= 5) {
                int combos = 1;
                for (int l : lv) combos *= modes(l);
                double tb = 1e300;
                int best = 0;
                for (int c = 0; c < combos; ++c) {
                    for (int i = 0, v = c; i < (int)lv.size(); v /= modes(lv[i]), ++i) d.tuneM2L[lv[i]] = 1 + v % modes(lv[i]);
                    const double t = timeStage(2, 3, 4);
                    if (t < tb) {
                        tb = t;
                        best = c;
                    }
                }
                for (int i = 0, v = best; i < (int)lv.size(); v /= modes(lv[i]), ++i) d.tuneM2L[lv[i]] = 1 + v % modes(lv[i]);
            } else {
                for (int l : lv) {
                    double tb = 1e300;
                    int mode = 0;
                    for (int m = 1; m <= modes(l); ++m) {
                        d.tuneM2L[l] = m;
                        const double t = timeStage(2, 3, 8);
                        if (t < tb) {
                            tb = t;
                            mode = m;
                        }
                    }
                    d.tuneM2L[l] = mode;
                }
            }
        }
        // near field beside the far-field chain, after it or (mode 0) in the
        // M2L launch, the finest-level M2L beside the M2M chain or in line,
        // and, beside the chain, the near field's blocks per CU (one per
        // item, or 1-4 resident blocks per CU that loop over the items):
        // every combination, since the in-order modes pay only when both
        // side streams are off
        // (and, for one right-hand side, whether the upward pass runs as the
        // persistent tree walk)
        const int NTW = nrhs == 1 ? 2 : 1;
        const int twKept = d.treeWalk;  // batched tuning keeps the one-RHS choice
        if (!(saved & MLFMA_FLAG_ONE_STREAM)) {
            constexpr int NNB = 5, NSL = 2;
            const int splits[NSL] = {100, 60};  // near items forked first (%), the rest beside the L2L
            double tm[4][2][NNB][2][NSL] = {};
            for (int rep = 0; rep < 2; ++rep)
                for (int m = 0; m < 4; ++m)
                    for (int sp = 0; sp < 2; ++sp)
                        for (int nbk = 0; nbk < (m ? NNB : 1); ++nbk)
                            for (int tw = 0; tw < NTW; ++tw)
                                for (int sl = 0; sl < (m ? NSL : 1); ++sl) {
                                    d.nearOverlap = m;
                                    d.m2lSplit = sp == 1;
                                    d.nearBlocks = nbk;
                                    d.treeWalk = NTW == 2 ? tw : twKept;
                                    d.nearSplit = splits[sl];
                                    tm[m][sp][nbk][tw][sl] += timeWhole(10);
                                }
            int bm = 2, bs = 1, bk = 0, bt = 0, bl = 0;
            for (int m = 0; m < 4; ++m)
                for (int sp = 0; sp < 2; ++sp)
                    for (int nbk = 0; nbk < (m ? NNB : 1); ++nbk)
                        for (int tw = 0; tw < NTW; ++tw)
                            for (int sl = 0; sl < (m ? NSL : 1); ++sl)
                                if (tm[m][sp][nbk][tw][sl] < tm[bm][bs][bk][bt][bl]) {
                                    bm = m;
                                    bs = sp;
                                    bk = nbk;
                                    bt = tw;
                                    bl = sl;
                                }
            d.nearOverlap = bm;
            d.m2lSplit = bs == 1;
            d.nearBlocks = bk;
            d.treeWalk = NTW == 2 ? bt : twKept;
            d.nearSplit = splits[bl];
            // how early the pair blocks of the shared M2L / near-field launch
            // end; that launch runs only for one right-hand side with the
            // stages in order (the in-order matvec mode, or solves run in order)
            auto tuneWindow = [&]() {
                const uint32_t f = d.flags;
                d.flags |= MLFMA_FLAG_ONE_STREAM;
                const int wins[4] = {50, 65, 80, 100};
                double tw[4] = {0, 0, 0, 0};
                for (int rep = 0; rep < 2; ++rep)
                    for (int w = 0; w < 4; ++w) {
                        d.nearWindow = wins[w];
                        tw[w] += timeWhole(10);
                    }
                d.flags = f;
                int bw = 3;
                for (int w = 0; w < 4; ++w)
                    if (tw[w] < tw[bw]) bw = w;
                d.nearWindow = wins[bw];
            };
            bool windowTuned = false;
            if (nrhs == 1 && bm == 0) {
                tuneWindow();
                windowTuned = true;
            }
            if (nrhs == 1 && d.Vin.size() == (size_t)n) {
                // a solve interleaves its own launches with the matvecs, so
                // the stage order of one-right-hand-side solves is chosen by
                // timing short GMRES runs both ways
                const uint32_t f = d.flags;
                d.flags &= ~MLFMA_FLAG_STAGE_TIMES;
                double tOverlap = 0, tSerial = 0;
                try {
                    timeSolveSteps(d, 8);
                    for (int rep = 0; rep < 3; ++rep) {
                        d.solveInOrder = false;
                        tOverlap += timeSolveSteps(d, 40);
                        d.solveInOrder = true;
                        tSerial += timeSolveSteps(d, 40);
                    }
                } catch (...) {
                    d.solveInOrder = false;
                    d.flags = f;
                    throw;
                }
                d.solveInOrder = false;
                d.flags = f;
                d.solveSerial = tSerial < tOverlap;
                if (d.solveSerial && !windowTuned) tuneWindow();
            }
        }
        const double after = timeWhole(10);
        if (ms_before) *ms_before = before;
        if (ms_after) *ms_after = after;
    } catch (...) {
        d.flags = saved;
        throw;
    }
    d.flags = saved;
    return MLFMA_OK;
    API_END
}

int mlfma_stage_times_detail(const mlfma_plan* plan, double* ms, int32_t cap, int32_t* count) {
    API_BEGIN
    const DevicePlan& d = P(plan);
    needDevice(d);
    if (!ms || !count) throw Error(MLFMA_ERR_ARG, "null argument");
    if (!d.timesValid) throw Error(MLFMA_ERR_STATE, "no matvec has run with MLFMA_FLAG_STAGE_TIMES set");
    *count = 7;
    for (int q = 0; q < 7 && q < cap; ++q) {
        float t = 0;
        hipCheck(hipEventElapsedTime(&t, d.ev[q], d.ev[q + 1]), "hipEventElapsedTime");
        ms[q] = t;
    }
    return MLFMA_OK;
    API_END
}

int mlfma_plan_set_param(mlfma_plan* plan, int32_t param, int32_t value) {
    API_BEGIN
    DevicePlan& d = P(plan);
    needDevice(d);
    switch (param) {
        case MLFMA_PARAM_NEAR_WINDOW:
            if (value < 1 || value > 100) throw Error(MLFMA_ERR_ARG, "near window must be 1..100 (%)");
            d.nearWindow = value;
            break;
        case MLFMA_PARAM_NEAR_OVERLAP:
            if (value < 0 || value > 3) throw Error(MLFMA_ERR_ARG, "near overlap mode must be 0..3");
            d.nearOverlap = value;
            break;
        case MLFMA_PARAM_M2L_SPLIT:
            if (value != 0 && value != 1) throw Error(MLFMA_ERR_ARG, "M2L split must be 0 or 1");
            d.m2lSplit = value == 1;
            break;
        case MLFMA_PARAM_SOLVE_IN_ORDER:
            if (value != 0 && value != 1) throw Error(MLFMA_ERR_ARG, "solve order must be 0 or 1");
            d.solveSerial = value == 1;
            break;
        case MLFMA_PARAM_NEAR_BLOCKS:
            if (value < 0 || value > 8) throw Error(MLFMA_ERR_ARG, "near blocks per CU must be 0..8");
            d.nearBlocks = value;
            break;
        case MLFMA_PARAM_TREE_WALK:
            if (value < 0 || value > 1) throw Error(MLFMA_ERR_ARG, "tree walk must be 0 or 1");
            d.treeWalk = value;
            break;
        case MLFMA_PARAM_NEAR_SPLIT:
            if (value < 0 || value > 100) throw Error(MLFMA_ERR_ARG, "near split must be 0..100 (%)");
            d.nearSplit = value;
            break;
        case MLFMA_PARAM_PROBE_GROUP:
            if (value < 0 || value > MLFMA_GROUP_LAST) throw Error(MLFMA_ERR_ARG, "unknown launch group");
            d.probeDrain(0);
            d.probeGroup = value;
            d.probeMs = 0;
            d.probeLaunches = 0;
            break;
        default:
            throw Error(MLFMA_ERR_ARG, "unknown parameter");
    }
    d.clearGraphs();
    return MLFMA_OK;
    API_END
}

int mlfma_plan_get_param(const mlfma_plan* plan, int32_t param, int32_t* value) {
    API_BEGIN
    const DevicePlan& d = P(plan);
    if (!value) throw Error(MLFMA_ERR_ARG, "null argument");
    switch (param) {
        case MLFMA_PARAM_NEAR_WINDOW: *value = d.nearWindow; break;
        case MLFMA_PARAM_NEAR_OVERLAP: *value = d.nearOverlap; break;
        case MLFMA_PARAM_M2L_SPLIT: *value = d.m2lSplit ? 1 : 0; break;
        case MLFMA_PARAM_SOLVE_IN_ORDER: *value = d.solveSerial ? 1 : 0; break;
        case MLFMA_PARAM_PROBE_GROUP: *value = d.probeGroup; break;
        case MLFMA_PARAM_NEAR_BLOCKS: *value = d.nearBlocks; break;
        case MLFMA_PARAM_TREE_WALK: *value = d.treeWalk; break;
        case MLFMA_PARAM_NEAR_SPLIT: *value = d.nearSplit; break;
        case MLFMA_PARAM_TREE_STALLS: {
            needDevice(d);
            if (!d.treeCtr) throw Error(MLFMA_ERR_STATE, "no tree-walk counters");
            hipCheck(hipSetDevice(d.device), "hipSetDevice");
            hipCheck(hipStreamSynchronize(d.stream), "tree stalls");
            int32_t v = 0;
            hipCheck(hipMemcpy(&v, d.treeCtr + d.treeCtrLen - 1, sizeof(v), hipMemcpyDeviceToHost), "tree stalls");
            *value = v;
            break;
        }
        default: throw Error(MLFMA_ERR_ARG, "unknown parameter");
    }
    return MLFMA_OK;
    API_END
}

// Tuning record: [version 4, nearOverlap, m2lSplit, nearWindow, solveSerial,
// levels L, tuneMinBlocks[0..L), tuneM2L[0..L), nearBlocks, treeWalk,
// nearSplit]; version 3 is the same without nearSplit (100), version 2 also
// without treeWalk (0), version 1 also without nearBlocks (0)
int mlfma_autotune_record(const mlfma_plan* plan, int32_t* rec, int32_t cap, int32_t* len) {
    API_BEGIN
    const DevicePlan& d = P(plan);
    if (!len) throw Error(MLFMA_ERR_ARG, "null argument");
    const int L = d.hp.tree.maxLevel + 1;
    std::vector<int32_t> r{4, d.nearOverlap, d.m2lSplit ? 1 : 0, d.nearWindow, d.solveSerial ? 1 : 0, L};
    for (int l = 0; l < L; ++l) r.push_back(l < (int)d.tuneMinBlocks.size() ? d.tuneMinBlocks[l] : 0);
    for (int l = 0; l < L; ++l) r.push_back(l < (int)d.tuneM2L.size() ? d.tuneM2L[l] : 0);
    r.push_back(d.nearBlocks);
    r.push_back(d.treeWalk);
    r.push_back(d.nearSplit);
    *len = (int32_t)r.size();
    if (rec)
        for (int32_t i = 0; i < std::min<int32_t>(cap, *len); ++i) rec[i] = r[i];
    return MLFMA_OK;
    API_END
}

int mlfma_autotune_apply(mlfma_plan* plan, const int32_t* rec, int32_t len) {
    API_BEGIN
    DevicePlan& d = P(plan);
    needDevice(d);
    const int L = d.hp.tree.maxLevel + 1;
    const bool v1 = rec && len == 6 + 2 * L && rec[0] == 1, v2 = rec && len == 7 + 2 * L && rec[0] == 2,
               v3 = rec && len == 8 + 2 * L && rec[0] == 3, v4 = rec && len == 9 + 2 * L && rec[0] == 4;
    if ((!v1 && !v2 && !v3 && !v4) || rec[5] != L)
        throw Error(MLFMA_ERR_ARG, "tuning record does not match this plan");
    const int nearBlocks = v2 || v3 || v4 ? rec[6 + 2 * L] : 0;
    const int treeWalk = v3 || v4 ? rec[7 + 2 * L] : 0;
    const int nearSplit = v4 ? rec[8 + 2 * L] : 100;
    if (rec[1] < 0 || rec[1] > 3 || rec[2] < 0 || rec[2] > 1 || rec[3] < 1 || rec[3] > 100 || rec[4] < 0 ||
        rec[4] > 1 || nearBlocks < 0 || nearBlocks > 8 || treeWalk < 0 || treeWalk > 1 || nearSplit < 0 ||
        nearSplit > 100)
        throw Error(MLFMA_ERR_ARG, "tuning record out of range");
    for (int l = 0; l < L; ++l)
        if (rec[6 + l] < 0 || rec[6 + L + l] < 0 || rec[6 + L + l] > 3)
            throw Error(MLFMA_ERR_ARG, "tuning record out of range");
    d.nearOverlap = rec[1];
    d.m2lSplit = rec[2] == 1;
    d.nearWindow = rec[3];
    d.solveSerial = rec[4] == 1;
    d.tuneMinBlocks.assign(rec + 6, rec + 6 + L);
    d.tuneM2L.assign(rec + 6 + L, rec + 6 + 2 * L);
    d.nearBlocks = nearBlocks;
    d.treeWalk = treeWalk;
    d.nearSplit = nearSplit;
    d.clearGraphs();
    return MLFMA_OK;
    API_END
}

int mlfma_precond_apply(mlfma_plan* plan, int32_t kind, const double* x, double* y, int32_t nrhs, int32_t adjoint,
                        int32_t ptr_mode, void* stream) {
    API_BEGIN
    DevicePlan& d = P(plan);
    needDevice(d);
    if (!x || !y || nrhs < 1) throw Error(MLFMA_ERR_ARG, "bad preconditioner arguments");
    if (kind != MLFMA_PRECOND_LEAF_BLOCKS) throw Error(MLFMA_ERR_ARG, "unknown preconditioner");
    hipCheck(hipSetDevice(d.device), "hipSetDevice");
    d.buildPrecond();
    const int64_t n = d.hp.src.n;
    hipStream_t st = stream ? (hipStream_t)stream : d.stream;
    if (ptr_mode == MLFMA_PTR_DEVICE) {
        d.applyPrecond((const cplx<double>*)x, (cplx<double>*)y, nrhs, n, adjoint != 0, st);
        hipCheck(hipGetLastError(), "preconditioner launch");
        if (!(d.flags & MLFMA_FLAG_ASYNC)) hipCheck(hipStreamSynchronize(st), "preconditioner");
        return MLFMA_OK;
    }
    if (ptr_mode != MLFMA_PTR_HOST) throw Error(MLFMA_ERR_ARG, "bad ptr_mode");
    d.ensureRhs(nrhs);
    const size_t bytes = (size_t)n * nrhs * sizeof(cplx<double>);
    hipCheck(hipMemcpyAsync(d.xs, x, bytes, hipMemcpyHostToDevice, st), "x upload");
    d.applyPrecond(d.xs, d.ys, nrhs, n, adjoint != 0, st);
    hipCheck(hipGetLastError(), "preconditioner launch");
    hipCheck(hipMemcpyAsync(y, d.ys, bytes, hipMemcpyDeviceToHost, st), "y download");
    hipCheck(hipStreamSynchronize(st), "preconditioner");
    return MLFMA_OK;
    API_END
}

int mlfma_probe_read(mlfma_plan* plan, double* ms, int64_t* launches) {
    API_BEGIN
    DevicePlan& d = P(plan);
    needDevice(d);
    if (!ms || !launches) throw Error(MLFMA_ERR_ARG, "null argument");
    hipCheck(hipSetDevice(d.device), "hipSetDevice");
    d.probeDrain(0);
    *ms = d.probeMs;
    *launches = d.probeLaunches;
    d.probeMs = 0;
    d.probeLaunches = 0;
    return MLFMA_OK;
    API_END
}

int mlfma_plan_set_flags(mlfma_plan* plan, uint32_t flags) {
    API_BEGIN
    DevicePlan& d = P(plan);
    if (flags & ~(MLFMA_FLAG_FRESH_LOCALS | MLFMA_FLAG_STAGE_TIMES | MLFMA_FLAG_M2L_GROUP | MLFMA_FLAG_M2L_TARGET |
                  MLFMA_FLAG_ASYNC | MLFMA_FLAG_PHYSICAL | MLFMA_FLAG_GRAPH | MLFMA_FLAG_ONE_STREAM))
        throw Error(MLFMA_ERR_ARG, "unknown flag bits");
    if ((flags ^ d.flags) & MLFMA_FLAG_PHYSICAL)
        throw Error(MLFMA_ERR_ARG, "MLFMA_FLAG_PHYSICAL is fixed when the plan is created");
    if (!d.hostOnly) {
        hipCheck(hipSetDevice(d.device), "hipSetDevice");
        hipCheck(hipStreamSynchronize(d.stream), "set flags");
        d.clearGraphs();  // captured launch sequences depend on the flags
    }
    d.flags = flags;
    return MLFMA_OK;
    API_END
}

int mlfma_level_boxes(const mlfma_plan* plan, int32_t level, int64_t* nbox, int32_t* coords, int32_t* is_leaf,
                      int32_t* nsrc, int64_t cap) {
    API_BEGIN
    const DevicePlan& d = P(plan);
    if (d.hp.forest) throw Error(MLFMA_ERR_STATE, "level introspection is not defined for forest plans");
    const Tree& t = d.hp.tree;
    if (level < 0 || level > t.maxLevel) throw Error(MLFMA_ERR_ARG, "level out of range");
    const int b = t.lvlBeg[level], e = t.lvlBeg[level + 1];
    *nbox = e - b;
    if (coords || is_leaf || nsrc) {
        if (cap < e - b) throw Error(MLFMA_ERR_ARG, "capacity too small");
        for (int i = b; i < e; ++i) {
            const Node& nd = t.nodes[i];
            if (coords) {
                coords[3 * (i - b)] = nd.ix;
                coords[3 * (i - b) + 1] = nd.iy;
                coords[3 * (i - b) + 2] = nd.iz;
            }
            if (is_leaf) is_leaf[i - b] = nd.leaf() ? 1 : 0;
            if (nsrc) nsrc[i - b] = (int32_t)nd.count();
        }
    }
    return MLFMA_OK;
    API_END
}

int mlfma_level_pairs(const mlfma_plan* plan, int32_t level, int32_t kind, int64_t* npairs, int32_t* pairs,
                      int32_t* src_level, int64_t cap) {
    API_BEGIN
    const DevicePlan& d = P(plan);
    if (d.hp.forest) throw Error(MLFMA_ERR_STATE, "level introspection is not defined for forest plans");
    const Tree& t = d.hp.tree;
    const Lists& L = d.hp.lists;
    if (level < 0 || level > t.maxLevel) throw Error(MLFMA_ERR_ARG, "level out of range");
    const int b = t.lvlBeg[level];
    std::vector<std::array<int, 3>> out;
    auto rel = [&](int node) { return node - t.lvlBeg[t.nodes[node].level]; };
    if (kind == 0) {
        for (int i = b; i < t.lvlBeg[level + 1]; ++i)
            for (int s : L.iList[i]) out.push_back({i - b, rel(s), t.nodes[s].level});
    } else if (kind == 1) {
        for (auto& p : L.nearPairs)
            if (t.nodes[p.first].level == level) out.push_back({rel(p.first), rel(p.second), t.nodes[p.second].level});
    } else if (kind == 2) {
        for (auto& p : L.nonNearPairs)
            if (t.nodes[p.first].level == level) out.push_back({rel(p.first), rel(p.second), t.nodes[p.second].level});
    } else
        throw Error(MLFMA_ERR_ARG, "bad pair kind");
    *npairs = (int64_t)out.size();
    if (pairs) {
        if (cap < (int64_t)out.size()) throw Error(MLFMA_ERR_ARG, "capacity too small");
        for (size_t q = 0; q < out.size(); ++q) {
            pairs[2 * q] = out[q][0];
            pairs[2 * q + 1] = out[q][1];
            if (src_level) src_level[q] = out[q][2];
        }
    }
    return MLFMA_OK;
    API_END
}

int mlfma_get_coeffs(const mlfma_plan* plan, int32_t level, int32_t which, double* out, int64_t cap) {
    API_BEGIN
    const DevicePlan& d = P(plan);
    if (d.hp.forest) throw Error(MLFMA_ERR_STATE, "level introspection is not defined for forest plans");
    needDevice(d);
    const Tree& t = d.hp.tree;
    if (level < 2 || level > t.maxLevel) throw Error(MLFMA_ERR_ARG, "coefficients exist for levels >= 2");
    const LevelBufs& L = d.levels[level];
    const int b = t.lvlBeg[level], e = t.lvlBeg[level + 1];
    const int64_t per = (int64_t)L.ndir * 2;
    if (cap < (int64_t)(e - b) * per * 2) throw Error(MLFMA_ERR_ARG, "capacity too small");
    std::vector<cplx<double>> buf((size_t)d.hp.boxes[level] * per);
    hipCheck(hipSetDevice(d.device), "hipSetDevice");
    hipCheck(hipStreamSynchronize(d.stream), "sync");
    if (d.useFp32) {
        std::vector<cplx<float>> f(buf.size());
        const cplx<float>* src = (which == 0 ? d.M32 : d.L32) + L.Moff;
        if (!f.empty()) hipCheck(hipMemcpy(f.data(), src, f.size() * sizeof(cplx<float>), hipMemcpyDeviceToHost), "copy");
        for (size_t i = 0; i < f.size(); ++i) buf[i] = cplx<double>{f[i].re, f[i].im};
    } else {
        const cplx<double>* src = (which == 0 ? d.M64 : d.L64) + L.Moff;
        if (!buf.empty()) hipCheck(hipMemcpy(buf.data(), src, buf.size() * sizeof(cplx<double>), hipMemcpyDeviceToHost), "copy");
    }
    std::memset(out, 0, sizeof(double) * (size_t)(e - b) * per * 2);
    for (int i = b; i < e; ++i) {
        const Node& nd = t.nodes[i];
        if (nd.slot < 0) continue;
        std::memcpy(out + (size_t)(i - b) * per * 2, buf.data() + (size_t)nd.slot * per, per * sizeof(cplx<double>));
    }
    return MLFMA_OK;
    API_END
}

int mlfma_get_translation(const mlfma_plan* plan, int32_t level, double* out, int64_t cap) {
    API_BEGIN
    const DevicePlan& d = P(plan);
    needDevice(d);
    if (level < 2 || level > d.hp.tree.maxLevel) throw Error(MLFMA_ERR_ARG, "translation tables exist for levels >= 2");
    const LevelBufs& L = d.levels[level];
    if (cap < (int64_t)316 * L.ndir * 2) throw Error(MLFMA_ERR_ARG, "capacity too small");
    std::vector<cplx<double>> T((size_t)316 * L.ndir);
    hipCheck(hipSetDevice(d.device), "hipSetDevice");
    hipCheck(hipMemcpy(T.data(), L.T64, T.size() * sizeof(cplx<double>), hipMemcpyDeviceToHost), "copy");
    for (int o = 0; o < 316; ++o)
        for (int q = 0; q < L.ndir; ++q) {
            const cplx<double> v = T[(size_t)o * L.ndir + q];
            out[2 * ((size_t)o * L.ndir + q)] = v.re / L.dirW[q];
            out[2 * ((size_t)o * L.ndir + q) + 1] = v.im / L.dirW[q];
        }
    return MLFMA_OK;
    API_END
}

}  // extern "C"
