// host.cpp — host-side plan construction (see host.hpp).
#include "host.hpp"

#include <algorithm>
#include <numeric>

namespace mlfma {

static const double kPi = std::acos(-1.0);

// ---------------------------------------------------------------- sources

// Triangle rules as barycentric (w0,w1,w2) x weight (weights sum to 1/2):
// centroid, 3-point, 7-point and 13-point rules with the abscissae of
// src/sources/triangle.cpp:45-130.
struct Rule {
    int n;
    double b[13][4];
};
static Rule triangleRule(int prec) {
    Rule r{};
    auto perm3 = [&](double a, double b, double w, int at) {
        r.b[at][0] = a; r.b[at][1] = b; r.b[at][2] = b; r.b[at][3] = w;
        r.b[at + 1][0] = b; r.b[at + 1][1] = a; r.b[at + 1][2] = b; r.b[at + 1][3] = w;
        r.b[at + 2][0] = b; r.b[at + 2][1] = b; r.b[at + 2][2] = a; r.b[at + 2][3] = w;
    };
    const double third = 1.0 / 3.0;
    switch (prec) {
        case 0:
            r.n = 1;
            r.b[0][0] = r.b[0][1] = r.b[0][2] = third;
            r.b[0][3] = 0.5;
            break;
        case 1:
            r.n = 3;
            perm3(2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0, 0);
            break;
        case 2:
            r.n = 7;
            r.b[0][0] = r.b[0][1] = r.b[0][2] = third;
            r.b[0][3] = 0.1125;
            perm3(0.059715871789770, 0.470142064105115, 0.066197076394253, 1);
            perm3(0.797426985353087, 0.101286507323456, 0.0629695902724135, 4);
            break;
        default: {
            r.n = 13;
            r.b[0][0] = r.b[0][1] = r.b[0][2] = third;
            r.b[0][3] = -0.074785022233841;
            perm3(0.479308067841920, 0.260345966079040, 0.087807628716604, 1);
            perm3(0.869739794195568, 0.065130102902216, 0.026673617804419, 4);
            const double e = 0.048690315425316, z = 0.312865496004874, t = 0.638444188569810,
                         w = 0.0385568804451285;
            const double s[6][3] = {{e, z, t}, {t, e, z}, {z, t, e}, {e, t, z}, {z, e, t}, {t, z, e}};
            for (int i = 0; i < 6; ++i) {
                r.b[7 + i][0] = s[i][0]; r.b[7 + i][1] = s[i][1]; r.b[7 + i][2] = s[i][2];
                r.b[7 + i][3] = w;
            }
            break;
        }
    }
    return r;
}

static bool finite3(const double* p) { return std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]); }

Sources buildSources(const mlfma_geometry& g, int quadPrec) {
    Sources s;
    s.kind = g.kind;
    s.n = g.n;
    if (g.n <= 0) throw Error(MLFMA_ERR_ARG, "geometry has no unknowns");
    s.ctr.resize(g.n);
    if (g.kind == MLFMA_SRC_DIPOLE) {
        if (!g.dip_pos || !g.dip_mom) throw Error(MLFMA_ERR_ARG, "dipole arrays missing");
        s.pos.resize(g.n);
        s.phat.resize(g.n);
        for (int64_t i = 0; i < g.n; ++i) {
            const double* p = g.dip_pos + 3 * i;
            const double* m = g.dip_mom + 3 * i;
            if (!finite3(p) || !finite3(m)) throw Error(MLFMA_ERR_NONFINITE, "non-finite dipole input");
            s.pos[i] = V3(p[0], p[1], p[2]);
            V3 mv(m[0], m[1], m[2]);
            const double mn = norm(mv);
            if (!(mn > 0)) throw Error(MLFMA_ERR_GEOMETRY, "dipole with zero moment");
            s.phat[i] = V3(mv.x / mn, mv.y / mn, mv.z / mn);
            s.ctr[i] = s.pos[i];
        }
        return s;
    }
    if (g.kind != MLFMA_SRC_RWG) throw Error(MLFMA_ERR_ARG, "unknown source kind");
    if (!g.verts || !g.tris || !g.rwgs || g.ntris <= 0 || g.nverts <= 0)
        throw Error(MLFMA_ERR_ARG, "RWG arrays missing");
    const Rule rule = triangleRule(quadPrec);
    s.tris.resize(g.ntris);
    for (int64_t t = 0; t < g.ntris; ++t) {
        TriQuad& T = s.tris[t];
        for (int j = 0; j < 3; ++j) {
            const int vi = g.tris[3 * t + j];
            if (vi < 0 || vi >= g.nverts) throw Error(MLFMA_ERR_GEOMETRY, "triangle vertex index out of range");
            const double* v = g.verts + 3 * vi;
            if (!finite3(v)) throw Error(MLFMA_ERR_NONFINITE, "non-finite vertex");
            for (int a = 0; a < 3; ++a) T.X[j][a] = v[a];
        }
        for (int a = 0; a < 3; ++a) {
            T.D0[a] = T.X[1][a] - T.X[0][a];
            T.D2[a] = T.X[0][a] - T.X[2][a];
        }
        T.nq = rule.n;
        for (int q = 0; q < rule.n; ++q) {
            for (int a = 0; a < 3; ++a)
                T.q[q][a] = (rule.b[q][0] * T.X[0][a] + rule.b[q][1] * T.X[1][a]) + rule.b[q][2] * T.X[2][a];
            T.q[q][3] = rule.b[q][3];
        }
        if (rule.n == 1 || (quadPrec >= 2)) {
            // the centroid node is (X0 + X1 + X2) / 3 (src/sources/triangle.cpp:11)
            if (quadPrec != 1)
                for (int a = 0; a < 3; ++a) T.q[0][a] = ((T.X[0][a] + T.X[1][a]) + T.X[2][a]) / 3.0;
        }
    }
    for (int t = 0; t < 2; ++t) {
        s.tri[t].resize(g.n);
        s.xpm[t].resize(g.n);
    }
    s.leng.resize(g.n);
    for (int64_t i = 0; i < g.n; ++i) {
        const int32_t* r = g.rwgs + 4 * i;
        for (int j = 0; j < 2; ++j)
            if (r[j] < 0 || r[j] >= g.nverts) throw Error(MLFMA_ERR_GEOMETRY, "RWG vertex index out of range");
        for (int j = 2; j < 4; ++j)
            if (r[j] < 0 || r[j] >= g.ntris) throw Error(MLFMA_ERR_GEOMETRY, "RWG triangle index out of range");
        const double* a = g.verts + 3 * r[0];
        const double* b = g.verts + 3 * r[1];
        V3 A(a[0], a[1], a[2]), B(b[0], b[1], b[2]);
        s.ctr[i] = V3((A.x + B.x) / 2.0, (A.y + B.y) / 2.0, (A.z + B.z) / 2.0);
        s.leng[i] = norm(A - B);
        for (int t = 0; t < 2; ++t) {
            const int tid = r[2 + t];
            s.tri[t][i] = tid;
            bool found = false;
            for (int j = 0; j < 3; ++j) {
                const int vi = g.tris[3 * tid + j];
                if (vi != r[0] && vi != r[1]) {
                    const double* v = g.verts + 3 * vi;
                    s.xpm[t][i] = V3(v[0], v[1], v[2]);
                    found = true;
                }
            }
            if (!found) throw Error(MLFMA_ERR_GEOMETRY, "RWG triangle has no free vertex");
        }
    }
    return s;
}

// ---------------------------------------------------------------- tree

int Tree::find(int level, int ix, int iy, int iz) const {
    auto it = lookup.find(boxKey(level, ix, iy, iz));
    return it == lookup.end() ? -1 : it->second;
}

Tree buildTree(const Sources& s, double rootLeng, int maxNodeSrcs, int maxLevelCap) {
    Tree t;
    t.perm.resize(s.n);
    std::iota(t.perm.begin(), t.perm.end(), 0);
    Node root;
    root.leng = rootLeng;
    root.center = V3(0, 0, 0);
    root.sbeg = 0;
    root.send = s.n;
    t.nodes.push_back(root);
    t.lvlBeg.push_back(0);
    std::vector<int64_t> tmp;
    for (int level = 0;; ++level) {
        const int beg = t.lvlBeg[level];
        const int end = (int)t.nodes.size();
        bool any = false;
        for (int id = beg; id < end; ++id) {
            const int64_t cnt = t.nodes[id].count();
            const bool split = cnt > maxNodeSrcs && (maxLevelCap < 0 || level < maxLevelCap);
            if (!split) continue;
            if (level >= 19) throw Error(MLFMA_ERR_GEOMETRY, "octree deeper than 20 levels (coincident sources?)");
            any = true;
            const Node P = t.nodes[id];
            // stable partition of the node's sources by octant: bit0 x, bit1 y, bit2 z,
            // a source goes to the upper half only when strictly above the centre
            // (src/fmm/stem.cpp:11-14)
            std::array<std::vector<int64_t>, 8> b;
            for (int64_t q = P.sbeg; q < P.send; ++q) {
                const V3& c = s.ctr[t.perm[q]];
                const int o = (c.x > P.center.x) + 2 * (c.y > P.center.y) + 4 * (c.z > P.center.z);
                b[o].push_back(t.perm[q]);
            }
            int64_t at = P.sbeg;
            const int first = (int)t.nodes.size();
            t.nodes[id].child0 = first;
            const double cl = P.leng / 2.0;
            for (int o = 0; o < 8; ++o) {
                Node c;
                c.level = level + 1;
                c.parent = id;
                c.branch = o;
                c.leng = cl;
                const double pmx = (o & 1) ? 1.0 : -1.0, pmy = (o & 2) ? 1.0 : -1.0, pmz = (o & 4) ? 1.0 : -1.0;
                c.center = V3(P.center.x + cl / 2.0 * pmx, P.center.y + cl / 2.0 * pmy, P.center.z + cl / 2.0 * pmz);
                c.ix = 2 * P.ix + (o & 1);
                c.iy = 2 * P.iy + ((o >> 1) & 1);
                c.iz = 2 * P.iz + ((o >> 2) & 1);
                c.sbeg = at;
                for (int64_t v : b[o]) t.perm[at++] = v;
                c.send = at;
                t.nodes.push_back(c);
            }
        }
        if (!any) break;
        t.lvlBeg.push_back(end);
    }
    t.lvlBeg.push_back((int)t.nodes.size());
    t.maxLevel = 0;
    for (const Node& n : t.nodes)
        if (n.leaf()) t.maxLevel = std::max(t.maxLevel, n.level);
    for (int id = 0; id < (int)t.nodes.size(); ++id) {
        const Node& n = t.nodes[id];
        t.lookup[boxKey(n.level, n.ix, n.iy, n.iz)] = id;
    }
    return t;
}

// ---------------------------------------------------------------- lists

// The 26 neighbour directions in the order of FMM::Dir (src/fmm/fmm.h:13-17).
static const int kDirs[26][3] = {
    {-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1},
    {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 1, 0},
    {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1}, {1, 0, 1},
    {0, -1, -1}, {0, 1, -1}, {0, -1, 1}, {0, 1, 1},
    {-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {1, 1, -1},
    {-1, -1, 1}, {1, -1, 1}, {-1, 1, 1}, {1, 1, 1}};

int offsetIndex(int dx, int dy, int dz) {
    static int table[7][7][7];
    static bool init = false;
    if (!init) {
        int idx = 0;
        for (int z = -3; z <= 3; ++z)
            for (int y = -3; y <= 3; ++y)
                for (int x = -3; x <= 3; ++x) {
                    const int m = std::max(std::abs(x), std::max(std::abs(y), std::abs(z)));
                    table[z + 3][y + 3][x + 3] = m > 1 ? idx++ : -1;
                }
        init = true;
    }
    if (dx < -3 || dx > 3 || dy < -3 || dy > 3 || dz < -3 || dz > 3) return -1;
    return table[dz + 3][dy + 3][dx + 3];
}

void offsetVector(int idx, int& dx, int& dy, int& dz) {
    int k = 0;
    for (int z = -3; z <= 3; ++z)
        for (int y = -3; y <= 3; ++y)
            for (int x = -3; x <= 3; ++x) {
                if (std::max(std::abs(x), std::max(std::abs(y), std::abs(z))) <= 1) continue;
                if (k++ == idx) {
                    dx = x; dy = y; dz = z;
                    return;
                }
            }
    dx = dy = dz = 0;
}

// Neighbours of equal or greater size (src/fmm/nbor.cpp:10-483) in geometric
// form: for each of the 26 directions, the deepest existing node at this
// node's level or above that contains the adjacent box of this node's size.
// Each such node is listed once; the reference reaches the same set through
// its case analysis, which drops a coarse leaf along every direction except
// the one pointing from the ancestor that borders it.
static void geqNeighbors(const Tree& t, int id, std::vector<int>& out, std::vector<int>& dirs) {
    const Node& n = t.nodes[id];
    const int side = 1 << n.level;
    for (int d = 0; d < 26; ++d) {
        const int x = n.ix + kDirs[d][0], y = n.iy + kDirs[d][1], z = n.iz + kDirs[d][2];
        if (x < 0 || y < 0 || z < 0 || x >= side || y >= side || z >= side) continue;
        int found = -1;
        for (int lv = n.level; lv >= 0; --lv) {
            const int sh = n.level - lv;
            found = t.find(lv, x >> sh, y >> sh, z >> sh);
            if (found >= 0) break;
        }
        if (found < 0) continue;
        if (std::find(out.begin(), out.end(), found) != out.end()) continue;
        out.push_back(found);
        dirs.push_back(d);
    }
}

Lists buildLists(const Tree& t) {
    const int nn = (int)t.nodes.size();
    Lists L;
    L.nbors.resize(nn);
    L.nborDir.resize(nn);
    L.iList.resize(nn);
    L.leafIlist.resize(nn);
    L.nearNbors.resize(nn);
    for (int id = 1; id < nn; ++id) geqNeighbors(t, id, L.nbors[id], L.nborDir[id]);
    // list 1: leaves touching a leaf (src/fmm/nbor.cpp:490-644): coarse leaf
    // neighbours as they are, same-level stems descended towards this leaf
    for (int id = 1; id < nn; ++id) {
        if (!t.nodes[id].leaf()) continue;
        for (size_t q = 0; q < L.nbors[id].size(); ++q) {
            const int d = L.nborDir[id][q];
            std::vector<int> stack{L.nbors[id][q]};
            size_t head = 0;
            while (head < stack.size()) {
                const int x = stack[head++];
                const Node& nx = t.nodes[x];
                if (nx.leaf()) {
                    L.nearNbors[id].push_back(x);
                    continue;
                }
                for (int c = 0; c < 8; ++c) {
                    bool ok = true;
                    for (int a = 0; a < 3; ++a) {
                        const int dd = kDirs[d][a];
                        if (dd && (((c >> a) & 1) != (dd > 0 ? 0 : 1))) ok = false;
                    }
                    if (ok) stack.push_back(nx.child0 + c);
                }
            }
        }
    }
    // lists 2 and 4 (src/fmm/node.cpp:61-84) and list-3/4 pairs (src/fmm/node.cpp:90-99)
    for (int id = 1; id < nn; ++id) {
        const Node& n = t.nodes[id];
        std::vector<int> mine = L.nbors[id];
        std::sort(mine.begin(), mine.end());
        auto inMine = [&](int x) { return std::binary_search(mine.begin(), mine.end(), x); };
        for (int bn : L.nbors[n.parent]) {
            const Node& b = t.nodes[bn];
            if (b.count() == 0) continue;
            if (b.leaf()) {
                if (!inMine(bn)) L.leafIlist[id].push_back(bn);
                continue;
            }
            for (int c = 0; c < 8; ++c) {
                const int ch = b.child0 + c;
                if (!inMine(ch) && t.nodes[ch].count() > 0) L.iList[id].push_back(ch);
            }
        }
        for (int l4 : L.leafIlist[id]) L.nonNearPairs.emplace_back(l4, id);
    }
    for (int id = 1; id < nn; ++id) {
        if (!t.nodes[id].leaf() || t.nodes[id].count() == 0) continue;
        for (int nb : L.nearNbors[id])
            if (id < nb && t.nodes[nb].count() > 0) L.nearPairs.emplace_back(id, nb);
    }
    return L;
}

// ---------------------------------------------------------------- grids

static double legendreWithDeriv(double x, int n, double* dp) {
    // three-term recurrence; derivative from (x^2-1) P_n' = n (x P_n - P_{n-1})
    double pm = 1.0, p = x;
    if (n == 0) {
        *dp = 0.0;
        return 1.0;
    }
    for (int i = 1; i < n; ++i) {
        const double pn = ((2.0 * i + 1) * x * p - i * pm) / static_cast<double>(i + 1);
        pm = p;
        p = pn;
    }
    *dp = n * (x * p - pm) / (x * x - 1.0);
    return p;
}

// Gauss-Legendre rule on [a,b] by Newton iteration from the asymptotic
// initial guess, stopping once the Newton step is at most 1e-9; the weight
// uses the derivative evaluated before the final step (src/interp.cpp:12-52).
void gaussLegendre(int n, double a, double b, std::vector<double>& x, std::vector<double>& w) {
    const double len = b - a, mid = (a + b) / 2.0;
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    const int half = n / 2;
    if (n % 2) {
        double dp;
        legendreWithDeriv(0.0, n, &dp);
        x[half] = mid;
        w[half] = len / (dp * dp);
    }
    for (int k = 0; k < half; ++k) {
        double z = std::cos(kPi * (4.0 * (half - k) - 1) / (4.0 * n + 2.0));
        double dpz;
        for (;;) {
            double dp;
            const double p = legendreWithDeriv(z, n, &dp);
            const double step = p / dp;
            z -= step;
            if (std::fabs(step) <= 1.0e-9) {
                dpz = dp;
                break;
            }
        }
        const int up = (n % 2) ? half + 1 + k : half + k;
        const int dn = half - 1 - k;
        x[up] = len / 2.0 * z + mid;
        x[dn] = -len / 2.0 * z + mid;
        w[up] = len / ((1.0 - z * z) * dpz * dpz);
        w[dn] = w[up];
    }
}

Grid buildGrid(int level, double k, double rootLeng, int digits) {
    // excess-bandwidth rule and sampling of src/fmm/angles.cpp:3-31
    Grid g;
    const double a = rootLeng / std::pow(2.0, level);
    const int tau = (int)std::ceil(1.73 * k * a + 2.16 * std::pow((double)digits, 2.0 / 3.0) * std::pow(k * a, 1.0 / 3.0));
    g.L = (int)std::floor(0.50 * tau);
    g.nth = tau + 1;
    g.nph = 2 * g.nth;
    gaussLegendre(g.nth, 0.0, kPi, g.theta, g.weight);
    for (int i = 0; i < g.nth; ++i) g.weight[i] = g.weight[i] * std::sin(g.theta[i]);
    g.phi.resize(g.nph);
    for (int i = 0; i < g.nph; ++i) g.phi[i] = 2.0 * kPi * i / static_cast<double>(g.nph);
    return g;
}

static double lagrange(double x, const double* xs, int n, int k) {
    double p = 1.0;
    for (int j = 0; j < n; ++j) {
        if (j == k) continue;
        p *= (x - xs[j]) / (xs[k] - xs[j]);
    }
    return p;
}

static inline int reflectIndex(int i, int n) { return i < 0 ? -i - 1 : (i >= n ? 2 * n - i - 1 : i); }
static inline int wrapIndex(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

InterpTable buildInterp(const Grid& fine, const Grid& coarse, int p) {
    // Lagrange stencils of src/fmm/tables.cpp:7-88 (fine level l+1 -> coarse level l)
    InterpTable T;
    T.p = p;
    const int mth = fine.nth, mph = fine.nph, nth = coarse.nth, nph = coarse.nph;
    if (p > mth || p > mph) throw Error(MLFMA_ERR_ARG, "interp_order exceeds the finest angular grid");
    T.thNear.resize(nth);
    T.thW.resize((size_t)nth * 2 * p);
    std::vector<double> xs(2 * p);
    for (int j = 0; j < nth; ++j) {
        const double th = coarse.theta[j];
        // nearest lower fine node from the asymptotic node formula (src/interp.cpp:64-77)
        const double xm = 2.0 * (th - kPi / 2.0) / kPi;
        const int near = mth - (int)std::floor(((4.0 * mth + 2.0) * std::acos(xm) / kPi + 1.0) / 4.0) - 1;
        T.thNear[j] = near;
        for (int i = near + 1 - p, q = 0; i <= near + p; ++i, ++q) {
            double v = fine.theta[reflectIndex(i, mth)];
            if (i < 0) v = -v;
            else if (i >= mth) v = 2.0 * kPi - v;
            xs[q] = v;
        }
        for (int q = 0; q < 2 * p; ++q) T.thW[(size_t)j * 2 * p + q] = lagrange(th, xs.data(), 2 * p, q);
    }
    T.phNear.resize(nph);
    T.phW.resize((size_t)nph * 2 * p);
    for (int j = 0; j < nph; ++j) {
        const double ph = coarse.phi[j];
        const int near = (int)std::floor(mph * ph / (2.0 * kPi));
        T.phNear[j] = near;
        for (int i = near + 1 - p, q = 0; i <= near + p; ++i, ++q) xs[q] = 2.0 * kPi * i / static_cast<double>(mph);
        for (int q = 0; q < 2 * p; ++q) T.phW[(size_t)j * 2 * p + q] = lagrange(ph, xs.data(), 2 * p, q);
    }
    // transposes in gather form: for each fine index, the coarse entries that
    // scatter into it (src/fmm/stem.cpp:186-244)
    {
        std::vector<std::vector<std::pair<int, double>>> g(mph);
        for (int j = 0; j < nph; ++j)
            for (int q = 0; q < 2 * p; ++q) {
                const int f = wrapIndex(T.phNear[j] + 1 - p + q, mph);
                g[f].emplace_back(j, T.phW[(size_t)j * 2 * p + q]);
            }
        T.atPhPtr.assign(1, 0);
        for (int f = 0; f < mph; ++f) {
            for (auto& e : g[f]) {
                T.atPhSrc.push_back(e.first);
                T.atPhW.push_back(e.second);
            }
            T.atPhPtr.push_back((int)T.atPhSrc.size());
        }
    }
    {
        struct E { int j; double w; int flip; };
        std::vector<std::vector<E>> g(mth);
        for (int j = 0; j < nth; ++j)
            for (int q = 0; q < 2 * p; ++q) {
                const int ie = T.thNear[j] + 1 - p + q;
                const int f = reflectIndex(ie, mth);
                g[f].push_back(E{j, T.thW[(size_t)j * 2 * p + q], ie != f ? 1 : 0});
            }
        T.atThPtr.assign(1, 0);
        for (int f = 0; f < mth; ++f) {
            for (auto& e : g[f]) {
                T.atThSrc.push_back(e.j);
                T.atThW.push_back(e.w);
                T.atThFlip.push_back(e.flip);
            }
            T.atThPtr.push_back((int)T.atThSrc.size());
        }
    }
    return T;
}

// ---------------------------------------------------------------- plan

// Creation order of the reference's nodes: a stem creates its eight children
// in octant order and each child builds its subtree before the next sibling
// (src/fmm/stem.cpp:3-26), i.e. depth-first preorder.  The reference heads a
// near pair with the leaf that compares lower (src/fmm/leaf.cpp:76); creation
// order stands in for that comparison.
static std::vector<int> creationOrder(const Tree& t) {
    std::vector<int> ord(t.nodes.size(), -1), stack{0};
    int next = 0;
    while (!stack.empty()) {
        const int id = stack.back();
        stack.pop_back();
        ord[id] = next++;
        const Node& n = t.nodes[id];
        if (!n.leaf())
            for (int c = 7; c >= 0; --c) stack.push_back(n.child0 + c);
    }
    return ord;
}

static void buildNearItems(HostPlan& hp) {
    const Tree& t = hp.tree;
    const Lists& L = hp.lists;
    const std::vector<int> ord = creationOrder(t);
    // per head leaf: the column nodes of the pairs it heads
    std::vector<std::vector<int>> heads(t.nodes.size());
    for (auto& pr : L.nearPairs) {
        const bool firstHeads = ord[pr.first] < ord[pr.second];
        heads[firstHeads ? pr.first : pr.second].push_back(firstHeads ? pr.second : pr.first);
    }
    for (auto& pr : L.nonNearPairs) {  // (leaf, list-3 node): the leaf is the observer
        if (t.nodes[pr.first].count() == 0 || t.nodes[pr.second].count() == 0) continue;
        heads[pr.first].push_back(pr.second);
    }
    int64_t nnz = 0;
    for (int id = 0; id < (int)t.nodes.size(); ++id) {
        const Node& n = t.nodes[id];
        if (n.leaf() && n.count() > 0) {
            int64_t nc = n.count();
            for (int cn : heads[id]) nc += t.nodes[cn].count();
            nnz += n.count() * nc;
        }
    }
    // entries per item: enough items to spread over the device, none too long
#ifndef MLFMA_NEAR_ITEM_CAP
#define MLFMA_NEAR_ITEM_CAP 8192
#endif
    const int64_t cap = std::max<int64_t>(MLFMA_NEAR_ITEM_CAP, (nnz + 8191) / 8192);
    int64_t voff = 0, parts = 0;
    std::vector<std::vector<int64_t>> sums(t.perm.size());
    for (int id = 0; id < (int)t.nodes.size(); ++id) {
        const Node& n = t.nodes[id];
        if (!n.leaf() || n.count() == 0) continue;
        auto& hc = heads[id];
        std::sort(hc.begin(), hc.end(), [&](int a, int b) { return t.nodes[a].sbeg < t.nodes[b].sbeg; });
        const int64_t base = (int64_t)hp.nearCols.size();
        for (int64_t q = n.sbeg; q < n.send; ++q) hp.nearCols.push_back(q);
        for (int cn : hc)
            for (int64_t q = t.nodes[cn].sbeg; q < t.nodes[cn].send; ++q) hp.nearCols.push_back(q);
        const int64_t ncol = (int64_t)hp.nearCols.size() - base;
        const int64_t nself = n.count();
        for (int64_t r0 = 0; r0 < n.count(); r0 += 128) {
            const int64_t nr = std::min<int64_t>(128, n.count() - r0);
            const int64_t cw = std::max<int64_t>(1, cap / nr);
            for (int64_t c0 = 0; c0 < ncol; c0 += cw) {
                const int64_t c1 = std::min(ncol, c0 + cw);
                HostPlan::NearItem it;
                it.rbeg = n.sbeg + r0;
                it.nr = (int)nr;
                it.nc = (int)(c1 - c0);
                it.nself = (int)std::max<int64_t>(0, std::min(c1, nself) - c0);
                it.cptr = base + c0;
                it.voff = voff;
                it.rpo = parts;
                it.cpo = parts + nr;
                for (int64_t r = 0; r < nr; ++r) sums[it.rbeg + r].push_back(it.rpo + r);
                for (int64_t c = it.nself; c < it.nc; ++c) sums[hp.nearCols[it.cptr + c]].push_back(it.cpo + c - it.nself);
                parts += nr + (it.nc - it.nself);
                voff += nr * it.nc;
                hp.nearItems.push_back(it);
            }
        }
    }
    if (parts >= ((int64_t)1 << 31)) throw Error(MLFMA_ERR_ARG, "near field too large for one plan");
    hp.nearNnz = voff;
    hp.nearParts = parts;
    hp.nearSumPtr.assign(1, 0);
    for (auto& v : sums) {
        for (int64_t q : v) hp.nearSumIdx.push_back((int32_t)q);
        hp.nearSumPtr.push_back((int64_t)hp.nearSumIdx.size());
    }
}

void buildHostPlan(HostPlan& hp, const mlfma_geometry& g, const mlfma_planewave& pw, const mlfma_config& cfg) {
    if (cfg.interp_order < 1) throw Error(MLFMA_ERR_ARG, "interp_order must be >= 1");
    if (!(cfg.root_leng > 0) || !(pw.wavenum > 0)) throw Error(MLFMA_ERR_ARG, "root_leng and wavenum must be > 0");
    if (cfg.digits < 1) throw Error(MLFMA_ERR_ARG, "digits must be >= 1");
    if (!(cfg.over_interp > 0)) throw Error(MLFMA_ERR_ARG, "over_interp must be > 0");
    if (cfg.quad_prec < 0 || cfg.quad_prec > 3) throw Error(MLFMA_ERR_ARG, "quad_prec must be 0..3");
    hp.cfg = cfg;
    hp.pw = pw;
    hp.k = pw.wavenum;
    hp.src = buildSources(g, cfg.quad_prec);
    hp.tree = buildTree(hp.src, cfg.root_leng, cfg.max_node_srcs, cfg.max_level);
    const Tree& t = hp.tree;
    if (t.maxLevel > 15) throw Error(MLFMA_ERR_GEOMETRY, "tree deeper than 15 levels");
    hp.lists = buildLists(t);
    const Lists& L = hp.lists;
    const int ML = t.maxLevel;
    hp.grids.resize(ML + 1);
    for (int l = 0; l <= ML; ++l) hp.grids[l] = buildGrid(l, hp.k, cfg.root_leng, cfg.digits);
    hp.interp.assign(ML + 1, InterpTable());
    for (int l = 2; l < ML; ++l) hp.interp[l] = buildInterp(hp.grids[l + 1], hp.grids[l], cfg.interp_order);

    // slots: non-empty boxes per level in node order
    hp.boxes.assign(ML + 1, 0);
    hp.slotNode.assign(ML + 1, {});
    for (int id = 0; id < (int)t.nodes.size(); ++id) {
        Node& n = hp.tree.nodes[id];
        if (n.count() == 0) continue;
        n.slot = hp.boxes[n.level]++;
        hp.slotNode[n.level].push_back(id);
    }
    // leaves with far-field work (src/fmm/leaf.cpp:212-230, 257-280: level >= 2)
    int64_t rp = 0;
    for (int l = 2; l <= ML; ++l)
        for (int id : hp.slotNode[l]) {
            const Node& n = t.nodes[id];
            if (!n.leaf()) continue;
            hp.farLeaves.push_back({id, l, n.slot, n.sbeg, n.count(), rp});
            rp += n.count() * hp.grids[l].ndir();
        }
    hp.radpatEntries = rp;
    // M2L worklists
    hp.m2lPtr.assign(ML + 1, {});
    hp.m2lSrc.assign(ML + 1, {});
    hp.m2lOff.assign(ML + 1, {});
    hp.m2lPairs.assign(ML + 1, 0);
    for (int l = 2; l <= ML; ++l) {
        auto& ptr = hp.m2lPtr[l];
        ptr.assign(1, 0);
        for (int id : hp.slotNode[l]) {
            const Node& n = t.nodes[id];
            for (int s : L.iList[id]) {
                const Node& m = t.nodes[s];
                const int o = offsetIndex(n.ix - m.ix, n.iy - m.iy, n.iz - m.iz);
                if (o < 0) throw Error(MLFMA_ERR_GEOMETRY, "interaction-list offset outside the 316-entry table");
                hp.m2lSrc[l].push_back(m.slot);
                hp.m2lOff[l].push_back(o);
            }
            ptr.push_back((int)hp.m2lSrc[l].size());
            if (L.iList[id].empty()) ++hp.staleBoxes;
        }
        hp.m2lPairs[l] = (int64_t)hp.m2lSrc[l].size();
    }
    // M2M (into levels >= 2) and L2L (from levels >= 2)
    hp.m2mPtr.assign(ML + 1, {});
    hp.m2mChild.assign(ML + 1, {});
    hp.m2mOct.assign(ML + 1, {});
    hp.l2lChild.assign(ML + 1, {});
    hp.l2lParent.assign(ML + 1, {});
    hp.l2lOct.assign(ML + 1, {});
    for (int l = 2; l < ML; ++l) {
        auto& ptr = hp.m2mPtr[l];
        ptr.assign(1, 0);
        for (int id : hp.slotNode[l]) {
            const Node& n = t.nodes[id];
            if (!n.leaf())
                for (int c = 0; c < 8; ++c) {
                    const Node& ch = t.nodes[n.child0 + c];
                    if (ch.count() == 0) continue;
                    hp.m2mChild[l].push_back(ch.slot);
                    hp.m2mOct[l].push_back(c);
                }
            ptr.push_back((int)hp.m2mChild[l].size());
        }
    }
    for (int l = 3; l <= ML; ++l)
        for (int id : hp.slotNode[l]) {
            const Node& n = t.nodes[id];
            const int p = n.parent;
            if (L.iList[p].empty()) {  // src/fmm/stem.cpp:115
                ++hp.truncatedL2L;
                continue;
            }
            hp.l2lChild[l].push_back(n.slot);
            hp.l2lParent[l].push_back(t.nodes[p].slot);
            hp.l2lOct[l].push_back(n.branch);
        }
    buildNearItems(hp);
}

// ---------------------------------------------------------------- forests
void mergeHostPlans(HostPlan& out, std::vector<HostPlan>& parts) {
    if (parts.empty()) throw Error(MLFMA_ERR_ARG, "a forest needs at least one scatterer");
    const HostPlan& p0 = parts[0];
    double rootMax = 0;
    for (const HostPlan& p : parts) {
        const mlfma_config &a = p.cfg, &b = p0.cfg;
        if (p.src.kind != p0.src.kind || a.digits != b.digits || a.interp_order != b.interp_order ||
            a.over_interp != b.over_interp || a.quad_prec != b.quad_prec || a.precision != b.precision ||
            a.flags != b.flags || p.k != p0.k)
            throw Error(MLFMA_ERR_ARG, "the scatterers of a forest must share the source kind, wavenumber, digits, "
                                       "interpolation order, oversampling, quadrature, precision and flags");
        rootMax = std::max(rootMax, p.cfg.root_leng);
    }
    // level shift: root_leng = rootMax / 2^shift
    int D = 0;
    out.levelShift.clear();
    for (const HostPlan& p : parts) {
        const double r = std::log2(rootMax / p.cfg.root_leng);
        const int sh = (int)std::lround(r);
        if (std::fabs(r - sh) > 1e-9)
            throw Error(MLFMA_ERR_ARG, "forest root sides must be the largest one over powers of two");
        out.levelShift.push_back(sh);
        D = std::max(D, p.tree.maxLevel + sh);
    }
    if (D > 15) throw Error(MLFMA_ERR_GEOMETRY, "forest deeper than 15 levels");
    out.forest = true;
    out.cfg = p0.cfg;
    out.cfg.root_leng = rootMax;
    out.cfg.max_level = D;
    out.pw = p0.pw;
    out.k = p0.k;
    out.grids.resize(D + 1);
    for (int l = 0; l <= D; ++l) out.grids[l] = buildGrid(l, out.k, rootMax, out.cfg.digits);
    out.interp.assign(D + 1, InterpTable());
    for (int l = 2; l < D; ++l) out.interp[l] = buildInterp(out.grids[l + 1], out.grids[l], out.cfg.interp_order);
    // sources, tree nodes, permutation
    Sources& S = out.src;
    S.kind = p0.src.kind;
    Tree& T = out.tree;
    T.maxLevel = D;
    out.segOff.assign(1, 0);
    std::vector<int> nodeBase, triBase;
    std::vector<int64_t> srcBase;
    for (const HostPlan& p : parts) {
        const int64_t sb = S.n;
        const int nb = (int)T.nodes.size(), tb = (int)S.tris.size();
        srcBase.push_back(sb);
        nodeBase.push_back(nb);
        triBase.push_back(tb);
        const Sources& s = p.src;
        S.n += s.n;
        S.ctr.insert(S.ctr.end(), s.ctr.begin(), s.ctr.end());
        S.pos.insert(S.pos.end(), s.pos.begin(), s.pos.end());
        S.phat.insert(S.phat.end(), s.phat.begin(), s.phat.end());
        S.tris.insert(S.tris.end(), s.tris.begin(), s.tris.end());
        for (int q = 0; q < 2; ++q) {
            for (int32_t t : s.tri[q]) S.tri[q].push_back(t + tb);
            S.xpm[q].insert(S.xpm[q].end(), s.xpm[q].begin(), s.xpm[q].end());
        }
        S.leng.insert(S.leng.end(), s.leng.begin(), s.leng.end());
        for (Node nd : p.tree.nodes) {
            if (nd.parent >= 0) nd.parent += nb;
            if (nd.child0 >= 0) nd.child0 += nb;
            nd.level += out.levelShift.back();
            nd.sbeg += sb;
            nd.send += sb;
            nd.slot = -1;
            T.nodes.push_back(nd);
        }
        for (int64_t q : p.tree.perm) T.perm.push_back(q + sb);
        out.segOff.push_back(S.n);
    }
    // slots per forest level: non-empty boxes at their own level >= 2, in
    // scatterer order (boxes above a scatterer's level 2 take no work)
    out.boxes.assign(D + 1, 0);
    out.slotNode.assign(D + 1, {});
    for (size_t s = 0; s < parts.size(); ++s)
        for (int l = 0; l <= D; ++l) {
            const int own = l - out.levelShift[s];
            if (own < 2 || own > parts[s].tree.maxLevel) continue;
            for (int id : parts[s].slotNode[own]) {
                Node& nd = T.nodes[nodeBase[s] + id];
                nd.slot = out.boxes[l]++;
                out.slotNode[l].push_back(nodeBase[s] + id);
            }
        }
    auto slotOf = [&](size_t s, int own, int ownSlot) {
        return T.nodes[nodeBase[s] + parts[s].slotNode[own][ownSlot]].slot;
    };
    // far-field leaves
    out.radpatEntries = 0;
    for (size_t s = 0; s < parts.size(); ++s)
        for (const auto& f : parts[s].farLeaves) {
            HostPlan::LeafItem it = f;
            it.node = f.node + nodeBase[s];
            it.level = f.level + out.levelShift[s];
            it.slot = T.nodes[it.node].slot;
            it.sbeg = f.sbeg + srcBase[s];
            it.rpOff = out.radpatEntries;
            out.farLeaves.push_back(it);
            out.radpatEntries += it.cnt * out.grids[it.level].ndir();
        }
    // worklists per forest level
    out.m2lPtr.assign(D + 1, {});
    out.m2lSrc.assign(D + 1, {});
    out.m2lOff.assign(D + 1, {});
    out.m2lPairs.assign(D + 1, 0);
    out.m2mPtr.assign(D + 1, {});
    out.m2mChild.assign(D + 1, {});
    out.m2mOct.assign(D + 1, {});
    out.l2lChild.assign(D + 1, {});
    out.l2lParent.assign(D + 1, {});
    out.l2lOct.assign(D + 1, {});
    for (int l = 2; l <= D; ++l) {
        out.m2lPtr[l].assign(1, 0);
        if (l < D) out.m2mPtr[l].assign(1, 0);
        for (size_t s = 0; s < parts.size(); ++s) {
            const HostPlan& p = parts[s];
            const int own = l - out.levelShift[s];
            if (own < 2 || own > p.tree.maxLevel) continue;
            const int nslot = (int)p.slotNode[own].size();
            // M2L: sources are slots of the same level
            for (int q = 0; q < nslot; ++q) {
                for (int e = p.m2lPtr[own][q]; e < p.m2lPtr[own][q + 1]; ++e) {
                    out.m2lSrc[l].push_back(slotOf(s, own, p.m2lSrc[own][e]));
                    out.m2lOff[l].push_back(p.m2lOff[own][e]);
                }
                out.m2lPtr[l].push_back((int)out.m2lSrc[l].size());
            }
            // M2M into this level: children at own + 1
            if (l < D)
                for (int q = 0; q < nslot; ++q) {
                    if (own < (int)p.m2mPtr.size() && !p.m2mPtr[own].empty())
                        for (int e = p.m2mPtr[own][q]; e < p.m2mPtr[own][q + 1]; ++e) {
                            out.m2mChild[l].push_back(slotOf(s, own + 1, p.m2mChild[own][e]));
                            out.m2mOct[l].push_back(p.m2mOct[own][e]);
                        }
                    out.m2mPtr[l].push_back((int)out.m2mChild[l].size());
                }
            // L2L into this level: parents at own - 1
            if (own >= 3)
                for (size_t e = 0; e < p.l2lChild[own].size(); ++e) {
                    out.l2lChild[l].push_back(slotOf(s, own, p.l2lChild[own][e]));
                    out.l2lParent[l].push_back(slotOf(s, own - 1, p.l2lParent[own][e]));
                    out.l2lOct[l].push_back(p.l2lOct[own][e]);
                }
        }
        out.m2lPairs[l] = (int64_t)out.m2lSrc[l].size();
    }
    // near field: items, columns and partial sums side by side
    out.nearSumPtr.assign(1, 0);
    for (size_t s = 0; s < parts.size(); ++s) {
        const HostPlan& p = parts[s];
        const int64_t cb = (int64_t)out.nearCols.size(), vb = out.nearNnz, pb = out.nearParts;
        for (auto it : p.nearItems) {
            it.rbeg += srcBase[s];
            it.cptr += cb;
            it.voff += vb;
            it.rpo += pb;
            it.cpo += pb;
            out.nearItems.push_back(it);
        }
        for (int64_t c : p.nearCols) out.nearCols.push_back(c + srcBase[s]);
        const int64_t ib = (int64_t)out.nearSumIdx.size();
        for (int32_t q : p.nearSumIdx) out.nearSumIdx.push_back((int32_t)(q + pb));
        for (size_t q = 1; q < p.nearSumPtr.size(); ++q) out.nearSumPtr.push_back(p.nearSumPtr[q] + ib);
        out.nearNnz += p.nearNnz;
        out.nearParts += p.nearParts;
        out.staleBoxes += p.staleBoxes;
        out.truncatedL2L += p.truncatedL2L;
    }
    if (out.nearParts >= ((int64_t)1 << 31)) throw Error(MLFMA_ERR_ARG, "near field too large for one plan");
}

}  // namespace mlfma
