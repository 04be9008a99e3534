// oracle.cpp — CPU FP64 restatement of the reference MLFMA (TEST INFRASTRUCTURE
// ONLY: the parity checker and CPU baseline; the product never uses it).
// Citations are file:line in the reference repository ellilu11/MLFMA.
#include "oracle.hpp"

#include <algorithm>
#include <cstdio>
#include <functional>
#include <queue>
#include <stdexcept>
#include <unordered_map>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

const double PI = std::acos(-1.0);  // src/math.h:10
static const double C0 = 299792458.0;   // src/phys.h:6
static const double MU0 = 1.256637E-6;  // src/phys.h:7
static const double P0 = 1.0E-8;        // src/phys.h:8
static const double FEPS = 1.0E-6;      // src/math.h:16
static const double EPS_NR = 1.0E-9;    // src/fmm/fmm.h:10
// Phys::C = -i c0 mu0 / 4pi (src/phys.h:10)
static cd physC() { return -IU * C0 * MU0 / (4.0 * PI); }
static bool approxZero(double x) { return std::fabs(x) < FEPS; }  // src/math.h:39-41

// ------------------------------------------------------------------ math
double legendreP(double x, int n, double* dP) {  // src/math.cpp:25-44
    double Pm = 1.0;
    if (!n) {
        if (dP) *dP = 0.0;
        return Pm;
    }
    double P = x;
    for (int i = 1; i < n; ++i) {
        double Pp = ((2.0 * i + 1) * x * P - i * Pm) / static_cast<double>(i + 1);
        Pm = P;
        P = Pp;
    }
    if (dP) *dP = n * (x * P - Pm) / (x * x - 1.0);
    return P;
}

cd sphericalHankel1(double x, int n) {  // src/math.cpp:68-84
    cd Hm = -IU * std::exp(IU * x) / x;
    if (!n) return Hm;
    cd H = -std::exp(IU * x) * (1.0 / x + IU / (x * x));
    for (int i = 1; i < n; ++i) {
        cd Hp = (2.0 * i + 1.0) / x * H - Hm;
        Hm = H;
        H = Hp;
    }
    return H;
}

void dyadicG(const V3& X, double k, cd G[3][3]) {  // src/math.h:132-142
    const double r = norm(X), kr = k * r, invkrsq = 1.0 / (kr * kr);
    const cd iinvkr = IU / kr;
    const V3 rhat = X / r;
    const cd a = 1.0 + iinvkr - invkrsq;
    const cd b = 1.0 + 3.0 * iinvkr - 3.0 * invkrsq;
    const cd e = std::exp(IU * kr) / r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            cd v = (i == j ? a : cd(0.0)) - (rhat[i] * rhat[j]) * b;
            G[i][j] = e * v;
        }
}

int flipIdxToRange(int i, int size) {  // src/math.h:144-154
    if (i < 0) return -i - 1;
    if (i >= size) return 2 * size - i - 1;
    return i;
}

int wrapIdxToRange(int i, int size) {  // src/math.h:156-166
    if (i < 0) return i + size;
    if (i >= size) return i - size;
    return i;
}

std::pair<cd, cd> givensRotation(cd z, cd w) {  // src/math.h:168-172 (Q2: no conjugates)
    const cd t = std::sqrt(z * z + w * w);
    return {z / t, w / t};
}

V3 idx2pm(int x) {  // src/math.cpp:8-17
    int xm4 = x % 4;
    V3 b(xm4 % 2, xm4 / 2, x / 4);
    for (int i = 0; i < 3; ++i) b[i] = (b[i] == 0 ? -1 : 1);
    return b;
}

cd powI(int m) {  // src/math.h:26-33
    switch (m % 4) {
        case 0: return 1.0;
        case 1: return IU;
        case 2: return -1.0;
        default: return -IU;
    }
}

static double infNorm(const V3& v) {
    return std::max(std::fabs(v[0]), std::max(std::fabs(v[1]), std::fabs(v[2])));
}

std::vector<double> iNodeDistances() {  // src/math.cpp:86-107
    std::vector<double> d;
    for (double dz = 0; dz <= 3; ++dz)
        for (double dy = 0; dy <= 3; ++dy)
            for (double dx = 0; dx <= 3; ++dx) {
                V3 v(dx, dy, dz);
                if (infNorm(v) > 1.0) d.push_back(norm(v));
            }
    std::sort(d.begin(), d.end());
    d.erase(std::unique(d.begin(), d.end()), d.end());
    return d;
}

std::vector<V3> iNodeDistVecs() {  // src/math.cpp:109-124
    std::vector<V3> v;
    for (double dz = -3; dz <= 3; ++dz)
        for (double dy = -3; dy <= 3; ++dy)
            for (double dx = -3; dx <= 3; ++dx) {
                V3 d(dx, dy, dz);
                if (infNorm(d) > 1.0) v.push_back(d);
            }
    return v;
}

static bool approxLess(double x, double y) {  // src/math.h:43-46
    if (std::fabs(x - y) < FEPS) return false;
    return x < y;
}
static bool vecEquals(const V3& X, const V3& Y) { return norm(X - Y) < FEPS; }  // src/math.h:48-50
static bool vecLessThan(const V3& X, const V3& Y) {  // src/math.h:52-60
    if (approxLess(X[0], Y[0])) return true;
    if (approxLess(Y[0], X[0])) return false;
    if (approxLess(X[1], Y[1])) return true;
    if (approxLess(Y[1], X[1])) return false;
    return X[2] < Y[2];
}

std::vector<V3> iNodeDirections() {  // src/math.cpp:126-147
    std::vector<V3> dirs;
    for (double dz = -3; dz <= 3; ++dz)
        for (double dy = -3; dy <= 3; ++dy)
            for (double dx = -3; dx <= 3; ++dx) {
                V3 d(dx, dy, dz);
                if (infNorm(d) > 1.0) dirs.push_back(d / norm(d));
            }
    std::sort(dirs.begin(), dirs.end(), vecLessThan);
    dirs.erase(std::unique(dirs.begin(), dirs.end(), vecEquals), dirs.end());
    return dirs;
}

// ------------------------------------------------------------------ interp
void gaussLegendre(int l, double eps, double a, double b, std::vector<double>& nodes,
                   std::vector<double>& weights) {  // src/interp.cpp:12-52
    const double leng = b - a;
    const double mid = (a + b) / 2.0;
    nodes.assign(l, 0.0);
    weights.assign(l, 0.0);
    const int kmax = l / 2;
    if (l % 2) {
        nodes[kmax] = mid;
        double dp;
        legendreP(0.0, l, &dp);
        weights[kmax] = leng / (dp * dp);
    }
    for (int k = 0; k < kmax; ++k) {
        double x = std::cos(PI * (4.0 * (kmax - k) - 1) / (4.0 * l + 2.0));
        double dpk;
        while (true) {
            double dp;
            double p = legendreP(x, l, &dp);
            x -= p / dp;  // Newton step; the weight below uses dp of the pre-step point
            if (std::fabs(p / dp) <= eps) {
                dpk = dp;
                break;
            }
        }
        const int kplus = l % 2 ? kmax + 1 + k : kmax + k;
        const int kminus = kmax - 1 - k;
        nodes[kplus] = leng / 2.0 * x + mid;
        nodes[kminus] = -leng / 2.0 * x + mid;
        weights[kplus] = leng / ((1.0 - x * x) * dpk * dpk);
        weights[kminus] = weights[kplus];
    }
}

int nearGLNodeIdx(double xi, int m, double a, double b) {  // src/interp.cpp:64-77
    const double leng = b - a;
    const double mid = (a + b) / 2.0;
    const double x = 2.0 * (xi - mid) / leng;
    const int idx = m - std::floor(((4.0 * m + 2.0) * std::acos(x) / PI + 1.0) / 4.0) - 1;
    return idx;
}

double lagrangeBasis(double x, const std::vector<double>& xs, int k) {  // src/interp.cpp:86-100
    double p = 1.0;
    for (int j = 0; j < (int)xs.size(); ++j) {
        if (j == k) continue;
        p *= (x - xs[j]) / (xs[k] - xs[j]);
    }
    return p;
}

// ------------------------------------------------------------------ sources
void buildQuads(Triangle& t, int prec) {  // src/sources/triangle.cpp:45-130
    auto bary = [&](double w0, double w1, double w2) {
        return (w0 * t.Xs[0] + w1 * t.Xs[1]) + w2 * t.Xs[2];
    };
    t.quads.clear();
    switch (prec) {
        case 0:
            t.quads.emplace_back(t.center, 1.0 / 2.0);
            break;
        case 1: {
            const double w = 1.0 / 6.0;
            t.quads.emplace_back(bary(2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0), w);
            t.quads.emplace_back(bary(1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0), w);
            t.quads.emplace_back(bary(1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0), w);
            break;
        }
        case 2: {
            t.quads.emplace_back(t.center, 0.1125);
            const double w1 = 0.066197076394253, al = 0.059715871789770, be = 0.470142064105115;
            t.quads.emplace_back(bary(al, be, be), w1);
            t.quads.emplace_back(bary(be, al, be), w1);
            t.quads.emplace_back(bary(be, be, al), w1);
            const double w2 = 0.0629695902724135, ga = 0.797426985353087, de = 0.101286507323456;
            t.quads.emplace_back(bary(ga, de, de), w2);
            t.quads.emplace_back(bary(de, ga, de), w2);
            t.quads.emplace_back(bary(de, de, ga), w2);
            break;
        }
        default: {
            t.quads.emplace_back(t.center, -0.074785022233841);
            const double w1 = 0.087807628716604, al = 0.479308067841920, be = 0.260345966079040;
            t.quads.emplace_back(bary(al, be, be), w1);
            t.quads.emplace_back(bary(be, al, be), w1);
            t.quads.emplace_back(bary(be, be, al), w1);
            const double w2 = 0.026673617804419, ga = 0.869739794195568, de = 0.065130102902216;
            t.quads.emplace_back(bary(ga, de, de), w2);
            t.quads.emplace_back(bary(de, ga, de), w2);
            t.quads.emplace_back(bary(de, de, ga), w2);
            const double w3 = 0.0385568804451285, e = 0.048690315425316, z = 0.312865496004874,
                         th = 0.638444188569810;
            t.quads.emplace_back(bary(e, z, th), w3);
            t.quads.emplace_back(bary(th, e, z), w3);
            t.quads.emplace_back(bary(z, th, e), w3);
            t.quads.emplace_back(bary(e, th, z), w3);
            t.quads.emplace_back(bary(z, e, th), w3);
            t.quads.emplace_back(bary(th, z, e), w3);
            break;
        }
    }
}

C3 Source::integratedPlaneWave(const V3& kvec) const {  // src/sources/rwg.cpp:46-86
    C3 rad{0.0, 0.0, 0.0};
    for (int t = 0; t < 2; ++t) {
        const Triangle& tri = *tris[t];
        const V3* Xs = tri.Xs;
        const V3* Ds = tri.Ds;
        const double alpha = dot(kvec, Ds[0]), beta = -dot(kvec, Ds[2]), gamma = alpha - beta;
        const double asq = alpha * alpha, bsq = beta * beta;
        const cd eA = std::exp(IU * alpha), eB = std::exp(IU * beta);
        const cd f0a = approxZero(alpha) ? -IU : (1.0 - eA) / alpha;
        const cd f0b = approxZero(beta) ? -IU : (1.0 - eB) / beta;
        const cd f1a = approxZero(alpha) ? cd(-0.5) : (1.0 - (1.0 - IU * alpha) * eA) / asq;
        const cd f1b = approxZero(beta) ? cd(-0.5) : (1.0 - (1.0 - IU * beta) * eB) / bsq;
        const V3 A = Xs[0] - Xpm[t];
        C3 rv;
        if (approxZero(gamma)) {
            const cd f2 = (eA * (asq + 2.0 * IU * alpha - 2.0) + 2.0) / (2.0 * alpha * asq);
            const V3 Dd = Ds[0] - Ds[2];
            for (int i = 0; i < 3; ++i) rv[i] = -f1a * A[i] - IU * f2 * Dd[i];
        } else {
            const cd I0 = (f0a - f0b) / gamma;
            const cd I1 = IU * (I0 + f1a);
            const cd I2 = -IU * (I0 + f1b);
            for (int i = 0; i < 3; ++i) rv[i] = I0 * A[i] + (I1 * Ds[0][i] - I2 * Ds[2][i]) / gamma;
        }
        const cd ph = std::exp(IU * dot(kvec, Xs[0]));
        const double s = sgn(t);
        for (int i = 0; i < 3; ++i) rad[i] += ph * rv[i] * s;
    }
    for (int i = 0; i < 3; ++i) rad[i] = leng * rad[i];
    return rad;
}

C3 Source::integratedPlaneWaveNumeric(const V3& kvec) const {  // src/sources/rwg.cpp:34-44
    C3 rad{0.0, 0.0, 0.0};
    for (int t = 0; t < 2; ++t)
        for (const auto& q : tris[t]->quads) {
            const cd e = q.second * std::exp(IU * dot(kvec, q.first));
            const V3 d = q.first - Xpm[t];
            for (int i = 0; i < 3; ++i) rad[i] += e * d[i] * sgn(t);
        }
    for (int i = 0; i < 3; ++i) rad[i] = leng * rad[i];
    return rad;
}

C3 Source::radAlongDir(const V3& X, const V3& kvec) const {
    if (kind == DIPOLE) {  // src/sources/dipole.h:47-51
        const cd e = std::exp(IU * dot(kvec, X - pos));
        return {e * phat[0], e * phat[1], e * phat[2]};
    }
    // src/sources/rwg.h:23-25
    const cd e = std::exp(IU * dot(kvec, X));
    C3 f = integratedPlaneWave(kvec);
    return {e * std::conj(f[0]), e * std::conj(f[1]), e * std::conj(f[2])};
}

C3 Source::farAlongDir(const V3& krhat) const {
    if (kind == DIPOLE) {  // src/sources/dipole.h:53-57
        const cd e = std::exp(-IU * dot(pos, krhat));
        return {e * phat[0], e * phat[1], e * phat[2]};
    }
    C3 f = integratedPlaneWave(krhat);  // src/sources/rwg.h:27-29
    return {std::conj(f[0]), std::conj(f[1]), std::conj(f[2])};
}

cd Source::integratedRad(const Source& src, double k) const {
    cd G[3][3];
    if (kind == DIPOLE) {  // src/sources/dipole.h:62-70
        if (pos == src.pos) return 0.0;
        dyadicG(pos - src.pos, k, G);
        cd r = 0.0;
        for (int i = 0; i < 3; ++i) {
            cd gi = G[i][0] * src.phat[0] + G[i][1] * src.phat[1] + G[i][2] * src.phat[2];
            r += gi * phat[i];  // conj(rad.dot(phat)) with phat real
        }
        return r;
    }
    // src/sources/rwg.cpp:92-139
    if (center == src.center) return 0.0;  // Q3
    cd intRad = 0.0;
    for (int ot = 0; ot < 2; ++ot) {
        const Triangle* obsTri = tris[ot];
        const V3& obsXpm = Xpm[ot];
        for (int st = 0; st < 2; ++st) {
            const Triangle* srcTri = src.tris[st];
            const V3& srcXpm = src.Xpm[st];
            if (obsTri->id == srcTri->id) continue;  // Q3: coincident triangles skipped
            for (const auto& oq : obsTri->quads) {
                const V3 vo = oq.first - obsXpm;
                for (const auto& sq : srcTri->quads) {
                    dyadicG(oq.first - sq.first, k, G);
                    const V3 vs = sq.first - srcXpm;
                    cd d = 0.0;
                    for (int i = 0; i < 3; ++i) {
                        cd gi = (sq.second * G[i][0]) * vs[0] + (sq.second * G[i][1]) * vs[1] +
                                (sq.second * G[i][2]) * vs[2];
                        d += (gi * sgn(st)) * vo[i];
                    }
                    intRad += oq.second * d * sgn(ot);
                }
            }
        }
    }
    return leng * src.leng * intRad;
}

void Source::buildVoltage(const PlaneWave& pw) {
    if (kind == DIPOLE) {
        // src/sources/dipole.h:29-32 run from the 3-argument constructor with the
        // default moment (p0, 0, 0) (src/sources/dipole.h:11-16); the 4-argument
        // constructor (src/sources/dipole.h:18-27) overwrites pol afterwards without
        // rebuilding the voltage, so the RHS always uses the default moment.
        const V3 polDefault(P0, 0.0, 0.0);
        voltage = -pw.amplitude * std::exp(IU * dot(pw.wavevec, pos)) * dot(polDefault, pw.pol);
        return;
    }
    // src/sources/rwg.h:18-21: -amp * conj(IPW(k).dot(pol)) = -amp * sum_i IPW_i pol_i
    C3 f = integratedPlaneWave(pw.wavevec);
    cd s = 0.0;
    for (int i = 0; i < 3; ++i) s += std::conj(f[i]) * pw.pol[i];
    voltage = -pw.amplitude * std::conj(s);
}

// ------------------------------------------------------------------ tree
Node* Model::newNode(const std::vector<int>& s, int branchIdx, Node* base, bool leaf) {
    // src/fmm/node.cpp:33-44
    auto n = std::make_unique<Node>();
    n->id = (int)nodes.size();
    n->srcs = s;
    n->branchIdx = branchIdx;
    n->base = base;
    n->leaf = leaf;
    n->nodeLeng = base == nullptr ? cfg.rootLeng : base->nodeLeng / 2.0;
    n->level = base == nullptr ? 0 : base->level + 1;
    if (base == nullptr)
        n->center = V3(0, 0, 0);
    else {
        V3 pm = idx2pm(branchIdx);
        for (int i = 0; i < 3; ++i) n->center[i] = base->center[i] + n->nodeLeng / 2.0 * pm[i];
    }
    if (leaf) maxLevel = std::max(maxLevel, n->level);  // src/fmm/leaf.cpp:11
    Node* raw = n.get();
    nodes.push_back(std::move(n));
    return raw;
}

void Model::splitStem(Node* n) {  // src/fmm/stem.cpp:3-26
    std::array<std::vector<int>, 8> bs;
    for (int s : n->srcs) {
        const V3 c = srcs[s].getCenter();
        int b = (c[0] > n->center[0]) + 2 * (c[1] > n->center[1]) + 4 * (c[2] > n->center[2]);
        bs[b].push_back(s);
    }
    for (int k = 0; k < 8; ++k) {
        const int childLevel = n->level + 1;
        bool stem = (int)bs[k].size() > cfg.maxNodeSrcs &&
                    (cfg.maxLevelCap < 0 || childLevel < cfg.maxLevelCap);
        Node* c = newNode(bs[k], k, n, !stem);
        n->branches.push_back(c);
        if (stem) splitStem(c);
    }
}

void Model::buildTree() {  // src/mlfma.cpp:26-30
    nodes.clear();
    maxLevel = 0;
    std::vector<int> all(srcs.size());
    for (size_t i = 0; i < srcs.size(); ++i) all[i] = (int)i;
    bool stem = direct ? false
                       : ((int)srcs.size() > cfg.maxNodeSrcs && (cfg.maxLevelCap < 0 || 0 < cfg.maxLevelCap));
    root = newNode(all, 0, nullptr, !stem);
    if (stem) splitStem(root);
}

// Direction table in the order of FMM::Dir (src/fmm/fmm.h:13-17):
// W E S N D U SW SE NW NE DW DE UW UE DS DN US UN DSW DSE DNW DNE USW USE UNW UNE
static const int DIRS[26][3] = {
    {-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1},
    {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 1, 0},
    {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1}, {1, 0, 1},
    {0, -1, -1}, {0, 1, -1}, {0, -1, 1}, {0, 1, 1},
    {-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {1, 1, -1},
    {-1, -1, 1}, {1, -1, 1}, {-1, 1, 1}, {1, 1, 1}};

static int dirIndex(int dx, int dy, int dz) {
    for (int i = 0; i < 26; ++i)
        if (DIRS[i][0] == dx && DIRS[i][1] == dy && DIRS[i][2] == dz) return i;
    return -1;
}

// src/fmm/nbor.cpp:10-483.  The reference spells out the 26 directions case by
// case; every case has the same structure, restated here once:
//   * the branch index toggles the bit of each axis the direction moves along;
//   * if no moving axis leaves the parent, the neighbour is that sibling;
//   * otherwise the parent is asked for its neighbour along the axes that leave
//     it (the "exit" direction).  A Stem answer yields its child with the toggled
//     index; a Leaf answer is returned only when the exit direction is the full
//     direction, and is dropped otherwise, because that leaf is found along the
//     lower-dimensional exit direction itself ("do not double count").
// The step of that rule for a node of branch index b: returns whether the
// neighbour lies outside the parent; *target is the branch index of the
// neighbour (the sibling, or the child of the parent's neighbour) and *exitDir
// the direction the parent is asked along.
bool neighborRule(int dir, int b, int* target, int* exitDir) {
    int t = b;
    int ex[3] = {0, 0, 0};
    bool exits = false;
    for (int a = 0; a < 3; ++a) {
        const int d = DIRS[dir][a];
        if (!d) continue;
        const int bit = (b >> a) & 1;
        t ^= (1 << a);
        if ((d > 0 && bit == 1) || (d < 0 && bit == 0)) {
            ex[a] = d;
            exits = true;
        }
    }
    *target = t;
    *exitDir = exits ? dirIndex(ex[0], ex[1], ex[2]) : -1;
    return exits;
}

// Whether child c of a stem neighbour along dir faces the node (bit 0 on the
// side a positive direction enters, bit 1 for a negative direction).
bool neighborChildFaces(int dir, int c) {
    for (int a = 0; a < 3; ++a) {
        const int d = DIRS[dir][a];
        if (d && ((c >> a) & 1) != (d > 0 ? 0 : 1)) return false;
    }
    return true;
}

Node* Model::neighborGeq(const Node* n, int dir) const {
    if (n->isRoot()) return nullptr;
    int target = 0, edir = -1;
    if (!neighborRule(dir, n->branchIdx, &target, &edir)) return n->base->branches[target];
    Node* nb = neighborGeq(n->base, edir);
    if (nb == nullptr) return nullptr;
    if (nb->leaf) return edir == dir ? nb : nullptr;
    return nb->branches[target];
}

// src/fmm/nbor.cpp:490-644: breadth-first descent into the children of the
// neighbour that face this node, in increasing branch order.
std::vector<Node*> Model::neighborsLeq(const Node* n, Node* nb, int dir) const {
    std::vector<Node*> out;
    if (n->isRoot()) return out;
    std::queue<Node*> q;
    q.push(nb);
    while (!q.empty()) {
        Node* x = q.front();
        if (x->leaf)
            out.push_back(x);
        else
            for (int c = 0; c < 8; ++c)
                if (neighborChildFaces(dir, c)) q.push(x->branches[c]);
        q.pop();
    }
    return out;
}

static bool contains(const std::vector<Node*>& v, const Node* x) {
    return std::find(v.begin(), v.end(), x) != v.end();
}

void Model::buildListsRec(Node* n) {
    if (n->leaf) leaves.push_back(n);  // src/fmm/leaf.cpp:51
    if (!n->isRoot()) {
        // buildNeighbors (src/fmm/stem.cpp:31-42, src/fmm/leaf.cpp:27-43)
        for (int d = 0; d < 26; ++d) {
            Node* nb = neighborGeq(n, d);
            if (nb) {
                n->nbors.push_back(nb);
                if (n->leaf) {
                    auto l = neighborsLeq(n, nb, d);
                    n->nearNbors.insert(n->nearNbors.end(), l.begin(), l.end());
                }
            }
        }
        // buildInteractionList (src/fmm/node.cpp:61-84)
        for (Node* bn : n->base->nbors) {
            if (bn->isSrcless()) continue;
            if (bn->leaf && !contains(n->nbors, bn)) {
                n->leafIlist.push_back(bn);
                continue;
            }
            for (Node* br : bn->branches)
                if (!contains(n->nbors, br) && !br->isSrcless()) n->iList.push_back(br);
        }
        // pushSelfToNearNonNbors (src/fmm/node.cpp:90-99)
        for (Node* l : n->leafIlist) {
            l->nearNonNbors.push_back(n);
            nonNearPairs.emplace_back(l, n);
        }
    }
    for (Node* br : n->branches) buildListsRec(br);
}

void Model::buildLists() {
    leaves.clear();
    nonNearPairs.clear();
    buildListsRec(root);
}

void Model::buildNearRads() {  // src/fmm/leaf.cpp:70-180
    const double k = wavenum;
    nearPairs.clear();
    // findNearNborPairs: one pair per unordered adjacent-leaf pair; orientation by
    // creation order (the reference compares heap addresses, src/fmm/leaf.cpp:76).
    for (Node* l : leaves)
        for (Node* nb : l->nearNbors)
            if (l->id < nb->id) nearPairs.emplace_back(l, nb);

    std::vector<std::vector<cd>> blocks(nearPairs.size());
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads)
    for (long p = 0; p < (long)nearPairs.size(); ++p) {
        const Node* o = nearPairs[p].first;
        const Node* s = nearPairs[p].second;
        std::vector<cd> b(o->srcs.size() * s->srcs.size());
        size_t q = 0;
        for (int io : o->srcs)
            for (int is : s->srcs) b[q++] = srcs[io].integratedRad(srcs[is], k);
        blocks[p] = std::move(b);
    }
    for (size_t p = 0; p < nearPairs.size(); ++p) nearPairs[p].first->nearRads.push_back(std::move(blocks[p]));

    std::vector<std::vector<cd>> nb(nonNearPairs.size());
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads)
    for (long p = 0; p < (long)nonNearPairs.size(); ++p) {
        const Node* o = nonNearPairs[p].first;
        const Node* s = nonNearPairs[p].second;
        std::vector<cd> b(o->srcs.size() * s->srcs.size());
        size_t q = 0;
        for (int io : o->srcs)
            for (int is : s->srcs) b[q++] = srcs[io].integratedRad(srcs[is], k);
        nb[p] = std::move(b);
    }
    for (size_t p = 0; p < nonNearPairs.size(); ++p) nonNearPairs[p].first->nonNearRads.push_back(std::move(nb[p]));

#pragma omp parallel for schedule(dynamic, 4) num_threads(threads)
    for (long li = 0; li < (long)leaves.size(); ++li) {
        Node* l = leaves[li];
        l->selfRads.clear();
        for (size_t io = 1; io < l->srcs.size(); ++io)
            for (size_t is = 0; is < io; ++is)
                l->selfRads.push_back(srcs[l->srcs[io]].integratedRad(srcs[l->srcs[is]], k));
    }
}

// ------------------------------------------------------------------ angles / tables
void Model::buildAngles(int level, Angles& a) const {  // src/fmm/angles.cpp:3-55
    const double k = wavenum;
    const double nodeLeng = cfg.rootLeng / std::pow(2.0, level);
    const int tau = (int)std::ceil((1.73 * k * nodeLeng +
                                    2.16 * std::pow((double)cfg.digits, 2.0 / 3.0) * std::pow(k * nodeLeng, 1.0 / 3.0)));
    a.L = (int)std::floor(0.50 * tau);
    const int nth = tau + 1;
    gaussLegendre(nth, EPS_NR, 0.0, PI, a.thetas, a.weights);
    for (int i = 0; i < nth; ++i) a.weights[i] = a.weights[i] * std::sin(a.thetas[i]);
    const int nph = 2 * nth;
    a.phis.resize(nph);
    for (int i = 0; i < nph; ++i) a.phis[i] = 2.0 * PI * i / static_cast<double>(nph);
    a.khat.resize((size_t)nth * nph);
    a.toThPh.resize((size_t)nth * nph);
    size_t d = 0;
    for (int it = 0; it < nth; ++it) {
        const double th = a.thetas[it];
        for (int ip = 0; ip < nph; ++ip) {
            const double ph = a.phis[ip];
            a.khat[d] = V3(std::sin(th) * std::cos(ph), std::sin(th) * std::sin(ph), std::cos(th));
            a.toThPh[d] = {std::cos(th) * std::cos(ph), std::cos(th) * std::sin(ph), -std::sin(th),
                           -std::sin(ph), std::cos(ph), 0.0};
            ++d;
        }
    }
}

struct PsiHash {  // src/maps.h:18-23
    size_t operator()(double x) const noexcept {
        long long q = static_cast<long long>(std::llround(x / 1.0E-9));
        return std::hash<long long>{}(q);
    }
};
struct PsiEq {  // src/maps.h:25-29
    bool operator()(double x, double y) const noexcept { return std::fabs(x - y) <= 1.0E-9; }
};

void Model::buildTablesLevel(int level, Tables& t) const {  // src/fmm/tables.cpp
    const int order = cfg.interpOrder;
    const double k = wavenum;
    // interpolation tables (src/fmm/tables.cpp:7-88), level < maxLevel
    if (level < maxLevel) {
        const Angles& src = angles[level + 1];
        const Angles& tgt = angles[level];
        const int mth = src.nth(), nth = tgt.nth();
        t.interpTheta.clear();
        for (int jth = 0; jth < nth; ++jth) {
            const double tt = tgt.thetas[jth];
            const int nearIdx = nearGLNodeIdx(tt, mth, 0.0, PI);
            std::vector<double> xs(2 * order);
            for (int ith = nearIdx + 1 - order, q = 0; ith <= nearIdx + order; ++ith, ++q) {
                const int f = flipIdxToRange(ith, mth);
                double st = src.thetas[f];
                if (ith < 0)
                    st *= -1.0;
                else if (ith >= mth)
                    st = 2.0 * PI - st;
                xs[q] = st;
            }
            InterpPair ip;
            ip.nearIdx = nearIdx;
            ip.coeffs.resize(2 * order);
            for (int q = 0; q < 2 * order; ++q) ip.coeffs[q] = lagrangeBasis(tt, xs, q);
            t.interpTheta.push_back(ip);
        }
        const int mph = src.nph(), nph = tgt.nph();
        t.interpPhi.clear();
        for (int jph = 0; jph < nph; ++jph) {
            const double tp = tgt.phis[jph];
            const int nearIdx = (int)std::floor(mph * tp / (2.0 * PI));
            std::vector<double> xs(2 * order);
            for (int iph = nearIdx + 1 - order, q = 0; iph <= nearIdx + order; ++iph, ++q)
                xs[q] = 2.0 * PI * iph / static_cast<double>(mph);
            InterpPair ip;
            ip.nearIdx = nearIdx;
            ip.coeffs.resize(2 * order);
            for (int q = 0; q < 2 * order; ++q) ip.coeffs[q] = lagrangeBasis(tp, xs, q);
            t.interpPhi.push_back(ip);
        }
    }
    // Translation tables are only read by M2L, which never runs below level 2
    // (a level-1 node's parent is the root, whose neighbour list is empty).
    if (level < 2) return;
    const Angles& ang = angles[level];
    const int nth = ang.nth(), nph = ang.nph(), nDir = nth * nph;
    const int L = ang.L;
    const int nps = (int)std::floor(cfg.overInterp * (nth - 1));
    t.nps = nps;
    const double nodeLeng = cfg.rootLeng / std::pow(2.0, level);
    // getAlpha (src/fmm/tables.cpp:90-127)
    t.alpha.assign(dists.size(), std::vector<cd>(nps));
    for (size_t di = 0; di < dists.size(); ++di) {
        const double kr = k * dists[di] * nodeLeng;
        std::vector<cd> h(L + 1);
        for (int l = 0; l <= L; ++l) h[l] = sphericalHankel1(kr, l);
        for (int ips = 0; ips < nps; ++ips) {
            const double xi = std::cos(PI * ips / static_cast<double>(nps - 1));
            cd coeff = 0.0;
            for (int l = 0; l <= L; ++l) coeff += powI(l) * (2.0 * l + 1.0) * h[l] * legendreP(xi, l, nullptr);
            t.alpha[di][ips] = IU * k / (4.0 * PI) * coeff;
        }
    }
    // getInterpPsi (src/fmm/tables.cpp:129-176), with the reference's hash map
    std::vector<double> psis;
    psis.reserve((size_t)nDir * rhats.size());
    for (int d = 0; d < nDir; ++d)
        for (const V3& rh : rhats) psis.push_back(std::acos(dot(ang.khat[d], rh)));
    std::sort(psis.begin(), psis.end());
    psis.erase(std::unique(psis.begin(), psis.end()), psis.end());
    std::unordered_map<double, InterpPair, PsiHash, PsiEq> ipsi;
    long merged = 0;
    for (double psi : psis) {
        const int nearIdx = (int)std::floor((nps - 1) * psi / PI);
        std::vector<double> xs(2 * order);
        for (int ips = nearIdx + 1 - order, q = 0; ips <= nearIdx + order; ++ips, ++q)
            xs[q] = PI * ips / static_cast<double>(nps - 1);
        InterpPair ip;
        ip.nearIdx = nearIdx;
        ip.coeffs.resize(2 * order);
        for (int q = 0; q < 2 * order; ++q) ip.coeffs[q] = lagrangeBasis(psi, xs, q);
        if (!ipsi.emplace(psi, ip).second) ++merged;
    }
    t.psiKeys = (long)ipsi.size();
    t.psiMerged = merged;
    // buildTranslationTable (src/fmm/tables.cpp:178-220)
    t.transl.assign(dXs.size(), std::vector<cd>(nDir));
    for (size_t o = 0; o < dXs.size(); ++o) {
        const V3& dX = dXs[o];
        const double r = norm(dX);
        const V3 rhat = dX / r;
        int di = -1;
        for (size_t q = 0; q < dists.size(); ++q)
            if (!(r + 1.0E-1 < dists[q]) && !(dists[q] + 1.0E-1 < r)) di = (int)q;  // Maps::DoubleComp
        const std::vector<cd>& al = t.alpha[di];
        for (int d = 0; d < nDir; ++d) {
            const double psi = std::acos(dot(ang.khat[d], rhat));
            auto it = ipsi.find(psi);
            if (it == ipsi.end()) throw std::runtime_error("psi key not found (unordered_map::at)");
            const InterpPair& ip = it->second;
            cd T = 0.0;
            for (int ips = ip.nearIdx + 1 - order, q = 0; q < 2 * order; ++ips, ++q)
                T += al[flipIdxToRange(ips, nps)] * ip.coeffs[q];  // Q1
            t.transl[o][d] = T;
        }
    }
}

void Model::buildTables() {  // src/fmm/node.cpp:46-56
    angles.assign(maxLevel + 1, Angles());
    for (int l = 0; l <= maxLevel; ++l) buildAngles(l, angles[l]);
    dists = iNodeDistances();
    rhats = iNodeDirections();
    dXs = iNodeDistVecs();
    tables.assign(maxLevel + 1, Tables());
    for (int l = 0; l <= maxLevel; ++l) buildTablesLevel(l, tables[l]);
}

void Model::resizeCoeffs() {  // src/fmm/stem.cpp:59-67, src/fmm/leaf.cpp:60-65
    for (auto& n : nodes) {
        const size_t nd = (size_t)angles[n->level].nth() * angles[n->level].nph();
        n->coeffs.assign(nd, C2{0.0, 0.0});
        n->localCoeffs.assign(nd, C2{0.0, 0.0});
    }
}

void Model::buildRadPats() {  // src/fmm/leaf.cpp:185-207
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (long li = 0; li < (long)leaves.size(); ++li) {
        Node* l = leaves[li];
        const Angles& a = angles[l->level];
        const int nd = a.nth() * a.nph();
        l->radPats.assign(nd, std::vector<C2>());
        for (int d = 0; d < nd; ++d) {
            const V3 kvec(a.khat[d][0] * wavenum, a.khat[d][1] * wavenum, a.khat[d][2] * wavenum);
            const auto& M = a.toThPh[d];
            std::vector<C2> rp(l->srcs.size());
            for (size_t s = 0; s < l->srcs.size(); ++s) {
                C3 r = srcs[l->srcs[s]].radAlongDir(l->center, kvec);
                rp[s][0] = M[0] * r[0] + M[1] * r[1] + M[2] * r[2];
                rp[s][1] = M[3] * r[0] + M[4] * r[1] + M[5] * r[2];
            }
            l->radPats[d] = std::move(rp);
        }
    }
}

void Model::setupAll() {
    buildTree();
    buildLists();
    buildNearRads();
    buildTables();
    resizeCoeffs();
    buildRadPats();
    lvec.assign(srcs.size(), 0.0);
    rvec.assign(srcs.size(), 0.0);
}

// ------------------------------------------------------------------ matvec passes
void Model::addInterp(const std::vector<C2>& in, std::vector<C2>& out, int srcLvl, int tgtLvl) const {
    // src/fmm/stem.cpp:133-184
    const int order = cfg.interpOrder;
    const int mth = angles[srcLvl].nth(), mph = angles[srcLvl].nph();
    const int nth = angles[tgtLvl].nth(), nph = angles[tgtLvl].nph();
    const Tables& tb = tables[std::min(srcLvl, tgtLvl)];
    std::vector<C2> inner((size_t)nth * mph, C2{0.0, 0.0});
    for (int jth = 0; jth < nth; ++jth) {
        const InterpPair& ip = tb.interpTheta[jth];
        for (int ith = ip.nearIdx + 1 - order, q = 0; ith <= ip.nearIdx + order; ++ith, ++q) {
            const int f = flipIdxToRange(ith, mth);
            const bool oor = ith != f;
            for (int iph = 0; iph < mph; ++iph) {
                int is = iph;
                if (oor) is += (iph < mph / 2) ? mph / 2 : -mph / 2;
                const C2& v = in[(size_t)f * mph + is];
                C2& o = inner[(size_t)jth * mph + iph];
                for (int c = 0; c < 2; ++c) o[c] += ip.coeffs[q] * v[c] * sgn(oor);
            }
        }
    }
    for (int jph = 0; jph < nph; ++jph) {
        const InterpPair& ip = tb.interpPhi[jph];
        for (int iph = ip.nearIdx + 1 - order, q = 0; iph <= ip.nearIdx + order; ++iph, ++q) {
            const int w = wrapIdxToRange(iph, mph);
            for (int jth = 0; jth < nth; ++jth) {
                const C2& v = inner[(size_t)jth * mph + w];
                C2& o = out[(size_t)jth * nph + jph];
                for (int c = 0; c < 2; ++c) o[c] += ip.coeffs[q] * v[c];
            }
        }
    }
}

void Model::addAnterp(const std::vector<C2>& in, std::vector<C2>& out, int srcLvl, int tgtLvl) const {
    // src/fmm/stem.cpp:186-244
    const int order = cfg.interpOrder;
    const int mth = angles[srcLvl].nth(), mph = angles[srcLvl].nph();
    const int nth = angles[tgtLvl].nth(), nph = angles[tgtLvl].nph();
    const Tables& tb = tables[std::min(srcLvl, tgtLvl)];
    std::vector<C2> inner((size_t)mth * nph, C2{0.0, 0.0});
    for (int iph = 0; iph < mph; ++iph) {
        const InterpPair& ip = tb.interpPhi[iph];
        for (int jph = -order; jph < nph + order; ++jph) {
            const int q = jph - (ip.nearIdx + 1 - order);
            if (q < 0 || q >= 2 * order) continue;
            const int w = wrapIdxToRange(jph, nph);
            for (int ith = 0; ith < mth; ++ith) {
                const C2& v = in[(size_t)ith * mph + iph];
                C2& o = inner[(size_t)ith * nph + w];
                for (int c = 0; c < 2; ++c) o[c] += ip.coeffs[q] * v[c];
            }
        }
    }
    for (int ith = 0; ith < mth; ++ith) {
        const InterpPair& ip = tb.interpTheta[ith];
        for (int jth = -order; jth < nth + order; ++jth) {
            const int q = jth - (ip.nearIdx + 1 - order);
            if (q < 0 || q >= 2 * order) continue;
            const int f = flipIdxToRange(jth, nth);
            const bool oor = jth != f;
            for (int jph = 0; jph < nph; ++jph) {
                int js = jph;
                if (oor) js += (jph < nph / 2) ? nph / 2 : -nph / 2;
                const C2& v = inner[(size_t)ith * nph + jph];
                C2& o = out[(size_t)f * nph + js];
                for (int c = 0; c < 2; ++c) o[c] += ip.coeffs[q] * v[c] * sgn(oor);
            }
        }
    }
}

void Model::buildMpoleRec(Node* n) {
    const Angles& a = angles[n->level];
    const size_t nd = (size_t)a.nth() * a.nph();
    if (n->leaf) {  // S2M, src/fmm/leaf.cpp:212-230
        if (n->isSrcless() || n->isRoot()) return;
        for (size_t d = 0; d < nd; ++d) {
            C2 c{0.0, 0.0};
            const auto& rp = n->radPats[d];
            for (size_t s = 0; s < n->srcs.size(); ++s) {
                const cd x = lvec[srcs[n->srcs[s]].idx];
                c[0] += x * rp[s][0];
                c[1] += x * rp[s][1];
            }
            n->coeffs[d] = c;
        }
        return;
    }
    // M2M, src/fmm/stem.cpp:72-103
    const Angles& ac = angles[n->level + 1];
    const size_t md = (size_t)ac.nth() * ac.nph();
    std::fill(n->coeffs.begin(), n->coeffs.end(), C2{0.0, 0.0});
    for (Node* br : n->branches) {
        if (br->isSrcless()) continue;
        buildMpoleRec(br);
        const V3 dX = n->center - br->center;
        std::vector<C2> sh(md);
        for (size_t d = 0; d < md; ++d) {
            const V3 kv(ac.khat[d][0] * wavenum, ac.khat[d][1] * wavenum, ac.khat[d][2] * wavenum);
            const cd e = std::exp(IU * dot(kv, dX));
            sh[d] = C2{e * br->coeffs[d][0], e * br->coeffs[d][1]};
        }
        addInterp(sh, n->coeffs, n->level + 1, n->level);
    }
}

int Model::offsetIndex(const V3& dXn) const {
    int ix = (int)std::lround(dXn[0]), iy = (int)std::lround(dXn[1]), iz = (int)std::lround(dXn[2]);
    int idx = 0;
    for (int dz = -3; dz <= 3; ++dz)
        for (int dy = -3; dy <= 3; ++dy)
            for (int dx = -3; dx <= 3; ++dx) {
                if (std::max(std::abs(dx), std::max(std::abs(dy), std::abs(dz))) <= 1) continue;
                if (dx == ix && dy == iy && dz == iz) return idx;
                ++idx;
            }
    return -1;
}

void Model::m2l(Node* n) {  // src/fmm/node.cpp:105-134
    if (n->iList.empty()) return;  // Q4: locals are not zeroed
    std::fill(n->localCoeffs.begin(), n->localCoeffs.end(), C2{0.0, 0.0});
    const Tables& tb = tables[n->level];
    const size_t nd = n->localCoeffs.size();
    for (Node* s : n->iList) {
        const V3 dX = n->center - s->center;
        const int o = offsetIndex(dX / n->nodeLeng);
        if (o < 0) throw std::runtime_error("M2L offset outside the 316-entry table");
        const std::vector<cd>& T = tb.transl[o];
        for (size_t d = 0; d < nd; ++d) {
            n->localCoeffs[d][0] += T[d] * s->coeffs[d][0];
            n->localCoeffs[d][1] += T[d] * s->coeffs[d][1];
        }
    }
    const Angles& a = angles[n->level];
    const double phw = 2.0 * PI / static_cast<double>(a.nph());
    size_t d = 0;
    for (int it = 0; it < a.nth(); ++it) {
        const double w = a.weights[it] * phw;
        for (int ip = 0; ip < a.nph(); ++ip, ++d) {
            n->localCoeffs[d][0] *= w;
            n->localCoeffs[d][1] *= w;
        }
    }
}

std::vector<C2> Model::shiftedLocal(const Node* p, int branchIdx) const {  // src/fmm/stem.cpp:109-131
    const Angles& ap = angles[p->level];
    const Angles& ac = angles[p->level + 1];
    std::vector<C2> out((size_t)ac.nth() * ac.nph(), C2{0.0, 0.0});
    if (p->iList.empty()) return out;  // Q5
    const V3 dX = p->branches[branchIdx]->center - p->center;
    const size_t md = (size_t)ap.nth() * ap.nph();
    std::vector<C2> sh(md);
    for (size_t d = 0; d < md; ++d) {
        const V3 kv(ap.khat[d][0] * wavenum, ap.khat[d][1] * wavenum, ap.khat[d][2] * wavenum);
        const cd e = std::exp(IU * dot(kv, dX));
        sh[d] = C2{e * p->localCoeffs[d][0], e * p->localCoeffs[d][1]};
    }
    addAnterp(sh, out, p->level, p->level + 1);
    return out;
}

void Model::buildLocalRec(Node* n) {  // src/fmm/stem.cpp:250-270, src/fmm/leaf.cpp:236-251
    if (!n->isRoot()) {
        m2l(n);
        if (!n->base->isRoot()) {
            std::vector<C2> s = shiftedLocal(n->base, n->branchIdx);
            for (size_t d = 0; d < s.size(); ++d) {
                n->localCoeffs[d][0] = n->localCoeffs[d][0] + s[d][0];
                n->localCoeffs[d][1] = n->localCoeffs[d][1] + s[d][1];
            }
        }
    }
    if (!n->leaf)
        for (Node* br : n->branches) buildLocalRec(br);
}

void Model::evalFarSols(Node* l) {  // src/fmm/leaf.cpp:257-280
    if (l->isSrcless() || l->level <= 1) return;
    const Angles& a = angles[l->level];
    const int nd = a.nth() * a.nph();
    const cd Ck = physC() * wavenum;
    for (size_t io = 0; io < l->srcs.size(); ++io) {
        cd acc = 0.0;
        for (int d = 0; d < nd; ++d) {
            const C2& r = l->radPats[d][io];
            const C2& c = l->localCoeffs[d];
            acc += std::conj(r[0]) * c[0] + std::conj(r[1]) * c[1];  // Hermitian dot
        }
        rvec[srcs[l->srcs[io]].idx] += Ck * acc;
    }
}

void Model::evalPairSols(Node* o, const Node* s, const std::vector<cd>& rads) {  // src/fmm/leaf.cpp:329-355
    const size_t no = o->srcs.size(), ns = s->srcs.size();
    std::vector<cd> so(no, 0.0), ss(ns, 0.0);
    size_t q = 0;
    for (size_t io = 0; io < no; ++io)
        for (size_t is = 0; is < ns; ++is) {
            const cd r = rads[q++];
            so[io] += lvec[srcs[s->srcs[is]].idx] * r;
            ss[is] += lvec[srcs[o->srcs[io]].idx] * r;
        }
    const cd Ck = physC() * wavenum;
    for (size_t i = 0; i < no; ++i) rvec[srcs[o->srcs[i]].idx] += Ck * so[i];
    for (size_t i = 0; i < ns; ++i) rvec[srcs[s->srcs[i]].idx] += Ck * ss[i];
}

void Model::evalSelfSols(Node* l) {  // src/fmm/leaf.cpp:361-383
    const size_t n = l->srcs.size();
    std::vector<cd> so(n, 0.0);
    size_t q = 0;
    for (size_t io = 1; io < n; ++io)
        for (size_t is = 0; is < io; ++is) {
            const cd r = l->selfRads[q++];
            so[io] += lvec[srcs[l->srcs[is]].idx] * r;
            so[is] += lvec[srcs[l->srcs[io]].idx] * r;
        }
    const cd Ck = physC() * wavenum;
    for (size_t i = 0; i < n; ++i) rvec[srcs[l->srcs[i]].idx] += Ck * so[i];
}

void Model::evaluateSols() {  // src/fmm/leaf.cpp:388-407
    for (Node* l : leaves) evalFarSols(l);
    for (auto& pr : nearPairs) {
        const size_t pi = pr.first->leafPairIdx++;
        evalPairSols(pr.first, pr.second, pr.first->nearRads[pi]);
    }
    for (Node* l : leaves) {
        for (Node* nn : l->nearNonNbors) evalPairSols(l, nn, l->nonNearRads[l->nonNearPairIdx++]);
        evalSelfSols(l);
        l->leafPairIdx = 0;
        l->nonNearPairIdx = 0;
    }
}

void Model::matvec(const cd* x, cd* y) {  // src/solver.cpp:41-57 with resetRvec
    for (size_t i = 0; i < srcs.size(); ++i) {
        lvec[i] = x[i];
        rvec[i] = 0.0;
    }
    if (!root->leaf) {
        buildMpoleRec(root);
        buildLocalRec(root);
    }
    evaluateSols();
    for (size_t i = 0; i < srcs.size(); ++i) y[i] = rvec[i];
}

void Model::denseMatvec(const cd* x, cd* y) const {
    // evalDirect (src/mlfma.cpp:84-107): a single root Leaf, so the matvec is
    // evalSelfSols over all sources with the strict lower triangle used both ways.
    const size_t n = srcs.size();
    std::vector<cd> acc(n, 0.0);
    const double k = wavenum;
#pragma omp parallel for schedule(dynamic, 8) num_threads(threads)
    for (long io = 0; io < (long)n; ++io) {
        cd a = 0.0;
        for (size_t is = 0; is < n; ++is) {
            if ((size_t)io == is) continue;
            // entry (max, min) as computed by the reference, used symmetrically
            const size_t hi = std::max((size_t)io, is), lo = std::min((size_t)io, is);
            a += x[is] * srcs[hi].integratedRad(srcs[lo], k);
        }
        acc[io] = a;
    }
    const cd Ck = physC() * wavenum;
    for (size_t i = 0; i < n; ++i) y[i] = Ck * acc[i];
}

void Model::farfield(const cd* I, int ndir, const double* dirs, cd* E) const {  // src/fmm/node.cpp:147-178
    const cd Ck = physC() * wavenum;
    for (int d = 0; d < ndir; ++d) {
        const V3 rh(dirs[3 * d], dirs[3 * d + 1], dirs[3 * d + 2]);
        const V3 kr = wavenum * rh;
        C3 acc{0.0, 0.0, 0.0};
        for (size_t n = 0; n < srcs.size(); ++n) {
            C3 f = srcs[n].farAlongDir(kr);
            for (int i = 0; i < 3; ++i) acc[i] += I[srcs[n].idx] * f[i];
        }
        for (int i = 0; i < 3; ++i) {
            cd v = 0.0;
            for (int j = 0; j < 3; ++j) v += ((i == j ? 1.0 : 0.0) - rh[i] * rh[j]) * acc[j];
            E[3 * d + i] = Ck * v;
        }
    }
}

// ------------------------------------------------------------------ GMRES (src/solver.cpp)
SolveResult gmresCompat(Model& m, int maxIter, double eps, cd* currents, bool record) {
    const size_t N = m.srcs.size();
    SolveResult res;
    res.record = record;
    std::vector<cd> Q(N * (size_t)(maxIter + 1));
    std::vector<std::vector<cd>> H;  // columns, each of length k+2
    std::vector<cd> g(maxIter + 1, 0.0), vc(maxIter, 0.0), vs(maxIter, 0.0);
    std::vector<cd> l(N), r(N);
    // src/solver.cpp:29-38: lvec = -V, g0 = |lvec|, lvec normalised
    for (size_t i = 0; i < N; ++i) l[i] = -m.srcs[i].voltage;
    double nrm2 = 0.0;
    for (size_t i = 0; i < N; ++i) nrm2 += std::norm(l[i]);
    const double g0 = std::sqrt(nrm2);
    g[0] = g0;
    if (!(g0 > 0.0) || !std::isfinite(g0)) {
        res.status = -4;
        for (size_t i = 0; i < N; ++i) currents[i] = 0.0;
        return res;
    }
    res.g0 = g0;
    for (size_t i = 0; i < N; ++i) {
        l[i] /= g0;
        Q[i] = l[i];
    }
    int iter = 0;
    do {
        m.matvec(l.data(), r.data());  // evalRvec (src/solver.cpp:41-57)
        const int k = iter;
        // iterateArnoldi (src/solver.cpp:59-83): modified Gram-Schmidt
        std::vector<cd> h(k + 2);
        for (int i = 0; i <= k; ++i) {
            const cd* qi = &Q[(size_t)i * N];
            cd hi = 0.0;
            for (size_t j = 0; j < N; ++j) hi += std::conj(qi[j]) * r[j];
            for (size_t j = 0; j < N; ++j) r[j] -= hi * qi[j];
            h[i] = hi;
        }
        double rn = 0.0;
        for (size_t j = 0; j < N; ++j) rn += std::norm(r[j]);
        h[k + 1] = std::sqrt(rn);
        for (size_t j = 0; j < N; ++j) {
            l[j] = r[j] / h[k + 1];
            Q[(size_t)(k + 1) * N + j] = l[j];
        }
        if (record) res.Hraw.push_back(h);
        // updateGvec / applyGivensRotation (src/solver.cpp:85-113)
        for (int i = 0; i <= k - 1; ++i) {
            const cd hp = vc[i] * h[i] + vs[i] * h[i + 1];
            h[i + 1] = -vs[i] * h[i] + vc[i] * h[i + 1];
            h[i] = hp;
        }
        auto cs = givensRotation(h[k], h[k + 1]);
        h[k] = cs.first * h[k] + cs.second * h[k + 1];
        h[k + 1] = 0.0;
        vc[k] = cs.first;
        vs[k] = cs.second;
        H.push_back(h);
        g[k + 1] = -vs[k] * g[k];
        g[k] = vc[k] * g[k];
        ++iter;
        const double rr = std::abs(g[iter]) / g0;
        res.resid.push_back(rr);
        if (!std::isfinite(rr)) {
            res.status = -4;
            break;
        }
        if (!(rr > eps && iter < maxIter)) break;
    } while (true);
    res.iters = iter;
    // y = R \ g[0:iter] (R upper triangular; partial-pivot LU does no swaps), x = Q y
    std::vector<cd> y(iter);
    for (int i = iter - 1; i >= 0; --i) {
        cd s = g[i];
        for (int j = i + 1; j < iter; ++j) s -= H[j][i] * y[j];
        y[i] = s / H[i][i];
    }
    for (size_t j = 0; j < N; ++j) {
        cd s = 0.0;
        for (int i = 0; i < iter; ++i) s += Q[(size_t)i * N + j] * y[i];
        currents[j] = s;
    }
    if (record) {
        res.y = y;
        res.Q.assign(Q.begin(), Q.begin() + (size_t)(iter + 1) * N);
    }
    return res;
}

}  // namespace orc
