// oracle_capi.cpp — ctypes-facing C API of the CPU oracle (TEST INFRASTRUCTURE
// ONLY).  Input structs share the layout of include/mlfma_c.h so that a test
// builds one geometry description and hands it to both the oracle and the
// library under test.
#include <cstring>
#include <stdexcept>
#include <string>

#include "../include/mlfma_c.h"
#include "oracle.hpp"

using namespace orc;

namespace {
thread_local std::string g_err;

struct Handle {
    Model m;
};

int fail(const std::exception& e) {
    g_err = e.what();
    return -1;
}

void buildSources(Model& m, const mlfma_geometry* g) {
    m.srcs.clear();
    m.tris.clear();
    if (g->kind == MLFMA_SRC_DIPOLE) {
        for (int64_t i = 0; i < g->n; ++i) {
            Source s;
            s.kind = DIPOLE;
            s.idx = (size_t)i;
            s.pos = V3(g->dip_pos[3 * i], g->dip_pos[3 * i + 1], g->dip_pos[3 * i + 2]);
            s.pol = V3(g->dip_mom[3 * i], g->dip_mom[3 * i + 1], g->dip_mom[3 * i + 2]);
            s.pmag = norm(s.pol);
            s.phat = s.pol / s.pmag;
            s.buildVoltage(m.pw);
            m.srcs.push_back(s);
        }
        return;
    }
    // importTriangles / importRWG (src/fileio.h:150-202), Triangle ctor
    // (src/sources/triangle.cpp:5-21), RWG ctor (src/sources/rwg.cpp:3-26)
    m.tris.resize(g->ntris);
    for (int64_t t = 0; t < g->ntris; ++t) {
        Triangle& tr = m.tris[t];
        tr.id = (int)t;
        for (int j = 0; j < 3; ++j) {
            tr.vIdx[j] = g->tris[3 * t + j];
            const double* v = g->verts + 3 * tr.vIdx[j];
            tr.Xs[j] = V3(v[0], v[1], v[2]);
        }
        tr.center = ((tr.Xs[0] + tr.Xs[1]) + tr.Xs[2]) / 3.0;
        for (int i = 0; i < 3; ++i) tr.Ds[i] = tr.Xs[wrapIdxToRange(i + 1, 3)] - tr.Xs[i];
    }
    for (auto& tr : m.tris) {
        (void)tr;
    }
    for (int64_t i = 0; i < g->n; ++i) {
        const int32_t* r = g->rwgs + 4 * i;
        Source s;
        s.kind = RWG;
        s.idx = (size_t)i;
        s.tris[0] = &m.tris[r[2]];
        s.tris[1] = &m.tris[r[3]];
        const double* a = g->verts + 3 * r[0];
        const double* b = g->verts + 3 * r[1];
        s.X0 = V3(a[0], a[1], a[2]);
        s.X1 = V3(b[0], b[1], b[2]);
        s.center = (s.X0 + s.X1) / 2.0;
        s.leng = norm(s.X0 - s.X1);
        for (int t = 0; t < 2; ++t)
            for (int j = 0; j < 3; ++j) {
                const int vi = s.tris[t]->vIdx[j];
                if (vi != r[0] && vi != r[1]) {
                    const double* v = g->verts + 3 * vi;
                    s.Xpm[t] = V3(v[0], v[1], v[2]);
                }
            }
        m.srcs.push_back(s);
    }
}

void toCd(const double* in, std::vector<cd>& out, size_t n) {
    out.resize(n);
    for (size_t i = 0; i < n; ++i) out[i] = cd(in[2 * i], in[2 * i + 1]);
}
void fromCd(const std::vector<cd>& in, double* out) {
    for (size_t i = 0; i < in.size(); ++i) {
        out[2 * i] = in[i].real();
        out[2 * i + 1] = in[i].imag();
    }
}
}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

// mode: 0 = FMM tree (reference main), 1 = evalDirect single-leaf tree
int orc_create(const mlfma_geometry* g, const mlfma_planewave* pw, const mlfma_config* c, int mode,
               int threads, void** out) {
    try {
        auto* h = new Handle();
        Model& m = h->m;
        m.cfg.digits = c->digits;
        m.cfg.interpOrder = c->interp_order;
        m.cfg.overInterp = c->over_interp;
        m.cfg.rootLeng = c->root_leng;
        m.cfg.maxNodeSrcs = c->max_node_srcs;
        m.cfg.maxLevelCap = c->max_level;
        m.cfg.quadPrec = c->quad_prec;
        m.pw.amplitude = pw->amplitude;
        m.pw.pol = V3(pw->pol[0], pw->pol[1], pw->pol[2]);
        m.pw.wavevec = pw->wavenum * V3(pw->khat[0], pw->khat[1], pw->khat[2]);
        m.pw.wavenum = pw->wavenum;
        m.wavenum = pw->wavenum;
        m.direct = (mode == 1);
        m.threads = threads > 0 ? threads : 1;
        if (g->kind == MLFMA_SRC_RWG) {
            // Triangle quadrature is built once the triangles exist.
        }
        buildSources(m, g);
        for (auto& t : m.tris) buildQuads(t, m.cfg.quadPrec);
        if (g->kind == MLFMA_SRC_RWG)
            for (auto& s : m.srcs) s.buildVoltage(m.pw);
        if (m.direct) {
            m.buildTree();
            m.buildLists();
            m.lvec.assign(m.srcs.size(), 0.0);
            m.rvec.assign(m.srcs.size(), 0.0);
        } else {
            m.setupAll();
        }
        *out = h;
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

void orc_destroy(void* h) { delete static_cast<Handle*>(h); }

// info: [n, maxLevel, nNodes, nLeaves, nNearPairs, nNonNearPairs]
int orc_info(void* hp, int64_t* out) {
    Model& m = static_cast<Handle*>(hp)->m;
    out[0] = (int64_t)m.srcs.size();
    out[1] = m.maxLevel;
    out[2] = (int64_t)m.nodes.size();
    out[3] = (int64_t)m.leaves.size();
    out[4] = (int64_t)m.nearPairs.size();
    out[5] = (int64_t)m.nonNearPairs.size();
    return 0;
}

int orc_voltage(void* hp, double* v) {
    Model& m = static_cast<Handle*>(hp)->m;
    for (size_t i = 0; i < m.srcs.size(); ++i) {
        v[2 * i] = m.srcs[i].voltage.real();
        v[2 * i + 1] = m.srcs[i].voltage.imag();
    }
    return 0;
}

int orc_matvec(void* hp, const double* x, double* y) {
    try {
        Model& m = static_cast<Handle*>(hp)->m;
        std::vector<cd> xi, yo(m.srcs.size());
        toCd(x, xi, m.srcs.size());
        if (m.direct)
            m.denseMatvec(xi.data(), yo.data());
        else
            m.matvec(xi.data(), yo.data());
        fromCd(yo, y);
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int orc_dense_matvec(void* hp, const double* x, double* y) {
    Model& m = static_cast<Handle*>(hp)->m;
    std::vector<cd> xi, yo(m.srcs.size());
    toCd(x, xi, m.srcs.size());
    m.denseMatvec(xi.data(), yo.data());
    fromCd(yo, y);
    return 0;
}

int orc_solve(void* hp, int maxIter, double eps, double* x, double* resid, int* iters) {
    try {
        Model& m = static_cast<Handle*>(hp)->m;
        std::vector<cd> cur(m.srcs.size());
        SolveResult r = gmresCompat(m, maxIter, eps, cur.data());
        fromCd(cur, x);
        for (size_t i = 0; i < r.resid.size(); ++i) resid[i] = r.resid[i];
        *iters = r.iters;
        return r.status;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// Solve with the Arnoldi record: H (maxIter+1) x maxIter column-major
// (unrotated Hessenberg), Q N x (maxIter+1) column-major, y maxIter, g0.
int orc_solve_record(void* hp, int maxIter, double eps, double* x, double* resid, int* iters, double* H, double* Q,
                     double* y, double* g0) {
    try {
        Model& m = static_cast<Handle*>(hp)->m;
        const size_t N = m.srcs.size();
        std::vector<cd> cur(N);
        SolveResult r = gmresCompat(m, maxIter, eps, cur.data(), true);
        fromCd(cur, x);
        for (size_t i = 0; i < r.resid.size(); ++i) resid[i] = r.resid[i];
        *iters = r.iters;
        *g0 = r.g0;
        for (size_t j = 0; j < r.Hraw.size(); ++j)
            for (size_t i = 0; i < r.Hraw[j].size(); ++i) {
                H[2 * (j * (maxIter + 1) + i)] = r.Hraw[j][i].real();
                H[2 * (j * (maxIter + 1) + i) + 1] = r.Hraw[j][i].imag();
            }
        for (size_t q = 0; q < r.Q.size(); ++q) {
            Q[2 * q] = r.Q[q].real();
            Q[2 * q + 1] = r.Q[q].imag();
        }
        for (size_t q = 0; q < r.y.size(); ++q) {
            y[2 * q] = r.y[q].real();
            y[2 * q + 1] = r.y[q].imag();
        }
        return r.status;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int orc_farfield(void* hp, const double* I, int ndir, const double* dirs, double* E) {
    Model& m = static_cast<Handle*>(hp)->m;
    std::vector<cd> ci;
    toCd(I, ci, m.srcs.size());
    std::vector<cd> e(3 * (size_t)ndir);
    m.farfield(ci.data(), ndir, dirs, e.data());
    fromCd(e, E);
    return 0;
}

// angles of one level: sizes [nth, nph, L]; arrays optional
int orc_angles(void* hp, int level, int* sizes, double* thetas, double* weights, double* phis) {
    Model& m = static_cast<Handle*>(hp)->m;
    if (level < 0 || level > m.maxLevel || m.angles.empty()) return -1;
    const Angles& a = m.angles[level];
    sizes[0] = a.nth();
    sizes[1] = a.nph();
    sizes[2] = a.L;
    if (thetas) std::memcpy(thetas, a.thetas.data(), a.thetas.size() * sizeof(double));
    if (weights) std::memcpy(weights, a.weights.data(), a.weights.size() * sizeof(double));
    if (phis) std::memcpy(phis, a.phis.data(), a.phis.size() * sizeof(double));
    return 0;
}

// translation operators of one level: 316 x nDir complex; extra: [nps, psiKeys, psiMerged]
int orc_translation(void* hp, int level, double* T, int64_t* extra) {
    Model& m = static_cast<Handle*>(hp)->m;
    if (level < 2 || level > m.maxLevel) return -1;
    const Tables& t = m.tables[level];
    size_t q = 0;
    for (const auto& row : t.transl)
        for (const cd& v : row) {
            T[q++] = v.real();
            T[q++] = v.imag();
        }
    if (extra) {
        extra[0] = t.nps;
        extra[1] = t.psiKeys;
        extra[2] = t.psiMerged;
    }
    return 0;
}

// alpha table of one level: 15 x nps complex
int orc_alpha(void* hp, int level, double* A) {
    Model& m = static_cast<Handle*>(hp)->m;
    if (level < 2 || level > m.maxLevel) return -1;
    size_t q = 0;
    for (const auto& row : m.tables[level].alpha)
        for (const cd& v : row) {
            A[q++] = v.real();
            A[q++] = v.imag();
        }
    return 0;
}

// interp tables of level l (from level l+1): which 0 = theta, 1 = phi.
// coeffs: count x 2p, near: count
int orc_interp(void* hp, int level, int which, double* coeffs, int* nearIdx) {
    Model& m = static_cast<Handle*>(hp)->m;
    if (level < 0 || level >= m.maxLevel) return -1;
    const auto& v = which == 0 ? m.tables[level].interpTheta : m.tables[level].interpPhi;
    size_t q = 0;
    for (size_t i = 0; i < v.size(); ++i) {
        for (double c : v[i].coeffs) coeffs[q++] = c;
        nearIdx[i] = v[i].nearIdx;
    }
    return (int)v.size();
}

// Apply interpolation (dir = 0, level l+1 -> l) or anterpolation (dir = 1,
// level l -> l+1) to a 2-component field; out is accumulated into.
int orc_apply_interp(void* hp, int level, int dir, const double* in, double* out) {
    Model& m = static_cast<Handle*>(hp)->m;
    if (level < 0 || level >= m.maxLevel) return -1;
    const int src = dir == 0 ? level + 1 : level;
    const int tgt = dir == 0 ? level : level + 1;
    const size_t ns = (size_t)m.angles[src].nth() * m.angles[src].nph();
    const size_t nt = (size_t)m.angles[tgt].nth() * m.angles[tgt].nph();
    std::vector<C2> vi(ns), vo(nt);
    for (size_t d = 0; d < ns; ++d)
        for (int c = 0; c < 2; ++c) vi[d][c] = cd(in[4 * d + 2 * c], in[4 * d + 2 * c + 1]);
    for (size_t d = 0; d < nt; ++d)
        for (int c = 0; c < 2; ++c) vo[d][c] = cd(out[4 * d + 2 * c], out[4 * d + 2 * c + 1]);
    if (dir == 0)
        m.addInterp(vi, vo, src, tgt);
    else
        m.addAnterp(vi, vo, src, tgt);
    for (size_t d = 0; d < nt; ++d)
        for (int c = 0; c < 2; ++c) {
            out[4 * d + 2 * c] = vo[d][c].real();
            out[4 * d + 2 * c + 1] = vo[d][c].imag();
        }
    return 0;
}

// The neighbour rule (src/fmm/nbor.cpp) for direction dir (FMM::Dir order) and
// branch index b: out = [exits the parent, target branch, exit direction or
// -1, a Leaf answer is kept (1) or dropped (0)]; and the children of a stem
// neighbour that face the node (out[0] = count, out[1..] = branches).
int orc_nbor_rule(int dir, int b, int32_t* out) {
    if (dir < 0 || dir >= 26 || b < 0 || b >= 8) return -1;
    int t = 0, e = -1;
    const bool ex = neighborRule(dir, b, &t, &e);
    out[0] = ex ? 1 : 0;
    out[1] = t;
    out[2] = e;
    out[3] = (ex && e == dir) ? 1 : 0;
    return 0;
}
int orc_nbor_leq_children(int dir, int32_t* out) {
    if (dir < 0 || dir >= 26) return -1;
    int k = 0;
    for (int c = 0; c < 8; ++c)
        if (neighborChildFaces(dir, c)) out[1 + k++] = c;
    out[0] = k;
    return 0;
}

// Nodes in creation order: per node [level, ix, iy, iz, leaf, nsrcs, id]
int orc_nodes(void* hp, int32_t* out, int64_t cap) {
    Model& m = static_cast<Handle*>(hp)->m;
    if ((int64_t)m.nodes.size() * 7 > cap) return -1;
    for (size_t i = 0; i < m.nodes.size(); ++i) {
        const Node* n = m.nodes[i].get();
        const double h = n->nodeLeng;
        int32_t* o = out + 7 * i;
        o[0] = n->level;
        for (int a = 0; a < 3; ++a)
            o[1 + a] = (int32_t)std::lround((n->center[a] + m.cfg.rootLeng / 2.0) / h - 0.5);
        o[4] = n->leaf ? 1 : 0;
        o[5] = (int32_t)n->srcs.size();
        o[6] = n->id;
    }
    return (int)m.nodes.size();
}

// Lists as node-id pairs: kind 0 = iList (node, src), 1 = near pairs,
// 2 = nonNear pairs (leaf, node), 3 = nbors (node, nbor), 4 = nearNbors (leaf, leaf)
int64_t orc_pairs(void* hp, int kind, int32_t* out, int64_t cap) {
    Model& m = static_cast<Handle*>(hp)->m;
    int64_t q = 0;
    auto put = [&](int a, int b) {
        if (out && 2 * q + 1 < cap) {
            out[2 * q] = a;
            out[2 * q + 1] = b;
        }
        ++q;
    };
    if (kind == 1) {
        for (auto& p : m.nearPairs) put(p.first->id, p.second->id);
    } else if (kind == 2) {
        for (auto& p : m.nonNearPairs) put(p.first->id, p.second->id);
    } else {
        for (auto& n : m.nodes) {
            const auto& v = kind == 0 ? n->iList : (kind == 3 ? n->nbors : n->nearNbors);
            for (Node* x : v) put(n->id, x->id);
        }
    }
    return q;
}

// coefficients of every node at one level (creation order), which 0 = mpole, 1 = local
int orc_coeffs(void* hp, int level, int which, double* out, int64_t cap) {
    Model& m = static_cast<Handle*>(hp)->m;
    int64_t q = 0;
    for (auto& n : m.nodes) {
        if (n->level != level) continue;
        const auto& v = which == 0 ? n->coeffs : n->localCoeffs;
        for (const C2& c : v)
            for (int j = 0; j < 2; ++j) {
                if (q + 1 >= cap) return -1;
                out[q++] = c[j].real();
                out[q++] = c[j].imag();
            }
    }
    return (int)(q / 4);
}

// ---- math KAT entry points
double orc_legendre(double x, int n, double* dp) { return legendreP(x, n, dp); }
void orc_hankel(double x, int n, double* out) {
    cd h = sphericalHankel1(x, n);
    out[0] = h.real();
    out[1] = h.imag();
}
void orc_gauss_legendre(int l, double eps, double a, double b, double* nodes, double* weights) {
    std::vector<double> n, w;
    gaussLegendre(l, eps, a, b, n, w);
    std::memcpy(nodes, n.data(), l * sizeof(double));
    std::memcpy(weights, w.data(), l * sizeof(double));
}
int orc_near_gl_idx(double x, int m, double a, double b) { return nearGLNodeIdx(x, m, a, b); }
double orc_lagrange(double x, const double* xs, int n, int k) {
    std::vector<double> v(xs, xs + n);
    return lagrangeBasis(x, v, k);
}
int orc_inode_distances(double* out) {
    auto d = iNodeDistances();
    for (size_t i = 0; i < d.size(); ++i) out[i] = d[i];
    return (int)d.size();
}
int orc_inode_directions(double* out) {
    auto d = iNodeDirections();
    for (size_t i = 0; i < d.size(); ++i)
        for (int a = 0; a < 3; ++a) out[3 * i + a] = d[i][a];
    return (int)d.size();
}
int orc_inode_distvecs(double* out) {
    auto d = iNodeDistVecs();
    for (size_t i = 0; i < d.size(); ++i)
        for (int a = 0; a < 3; ++a) out[3 * i + a] = d[i][a];
    return (int)d.size();
}
void orc_dyadic(const double* X, double k, double* G) {
    cd g[3][3];
    dyadicG(V3(X[0], X[1], X[2]), k, g);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            G[2 * (3 * i + j)] = g[i][j].real();
            G[2 * (3 * i + j) + 1] = g[i][j].imag();
        }
}
void orc_givens(const double* z, const double* w, double* cs) {
    auto r = givensRotation(cd(z[0], z[1]), cd(w[0], w[1]));
    cs[0] = r.first.real();
    cs[1] = r.first.imag();
    cs[2] = r.second.real();
    cs[3] = r.second.imag();
}
// RWG plane-wave integral of source i: analytic (numeric=0) or quadrature (numeric=1)
int orc_rwg_ipw(void* hp, int i, const double* kvec, int numeric, double* out) {
    Model& m = static_cast<Handle*>(hp)->m;
    const Source& s = m.srcs[i];
    if (s.kind != RWG) return -1;
    V3 k(kvec[0], kvec[1], kvec[2]);
    C3 f = numeric ? s.integratedPlaneWaveNumeric(k) : s.integratedPlaneWave(k);
    for (int a = 0; a < 3; ++a) {
        out[2 * a] = f[a].real();
        out[2 * a + 1] = f[a].imag();
    }
    return 0;
}
// Near-field entry (obs i, src j) as the reference computes it
int orc_entry(void* hp, int i, int j, double* out) {
    Model& m = static_cast<Handle*>(hp)->m;
    cd v = m.srcs[i].integratedRad(m.srcs[j], m.wavenum);
    out[0] = v.real();
    out[1] = v.imag();
    return 0;
}
// Radiation pattern of source i about point X along kvec, projected (theta, phi)
int orc_radpat(void* hp, int i, const double* X, const double* khat, double* out) {
    Model& m = static_cast<Handle*>(hp)->m;
    const double th = std::acos(khat[2]);
    const double ph = std::atan2(khat[1], khat[0]);
    const V3 kv = m.wavenum * V3(khat[0], khat[1], khat[2]);
    C3 r = m.srcs[i].radAlongDir(V3(X[0], X[1], X[2]), kv);
    const double M[6] = {std::cos(th) * std::cos(ph), std::cos(th) * std::sin(ph), -std::sin(th),
                         -std::sin(ph), std::cos(ph), 0.0};
    cd a = M[0] * r[0] + M[1] * r[1] + M[2] * r[2];
    cd b = M[3] * r[0] + M[4] * r[1] + M[5] * r[2];
    out[0] = a.real();
    out[1] = a.imag();
    out[2] = b.real();
    out[3] = b.imag();
    return 0;
}

}  // extern "C"
