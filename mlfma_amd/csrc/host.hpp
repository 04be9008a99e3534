// host.hpp — host-side plan construction for the MI355X MLFMA library:
// sources, octree, interaction lists, angular grids, interpolation tables and
// the flat worklists the HIP kernels walk.
//
// Reference counterparts (ellilu11/MLFMA): src/fmm/stem.cpp:3-26 (tree),
// src/fmm/nbor.cpp + src/fmm/node.cpp:61-99 (lists), src/fmm/angles.cpp
// (grids), src/fmm/tables.cpp:7-88 (interpolation tables),
// src/fmm/leaf.cpp:70-180 (near-field pair structure).
#pragma once

#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/mlfma_c.h"

namespace mlfma {

using cd = std::complex<double>;

struct V3 {
    double x = 0, y = 0, z = 0;
    V3() = default;
    V3(double a, double b, double c) : x(a), y(b), z(c) {}
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// ---------------------------------------------------------------- sources
constexpr int MAXQ = 13;

struct TriQuad {            // one triangle with its quadrature rule
    double X[3][3];         // vertices in face order
    double D0[3], D2[3];    // X1-X0, X0-X2 (the edges the closed-form integral uses)
    int nq = 0;
    double q[MAXQ][4];      // x, y, z, weight
};

struct Sources {
    int kind = MLFMA_SRC_DIPOLE;
    int64_t n = 0;
    std::vector<V3> ctr;    // classification point (Source::getCenter)
    // dipoles
    std::vector<V3> pos, phat;
    // RWG
    std::vector<TriQuad> tris;
    std::vector<int32_t> tri[2];
    std::vector<V3> xpm[2];
    std::vector<double> leng;
};

Sources buildSources(const mlfma_geometry& g, int quadPrec);

// ---------------------------------------------------------------- tree
struct Node {
    int level = 0;
    int parent = -1;
    int child0 = -1;        // first of 8 contiguous children, -1 for a leaf
    int branch = 0;
    int ix = 0, iy = 0, iz = 0;
    double leng = 0;
    V3 center;
    int64_t sbeg = 0, send = 0;  // range in the tree order of the sources
    int slot = -1;               // index among non-empty boxes of its level
    bool leaf() const { return child0 < 0; }
    int64_t count() const { return send - sbeg; }
};

struct Tree {
    std::vector<Node> nodes;          // level order; children contiguous
    std::vector<int> lvlBeg;          // nodes of level l: [lvlBeg[l], lvlBeg[l+1])
    int maxLevel = 0;                 // deepest leaf
    std::vector<int64_t> perm;        // perm[tree position] = input index
    std::unordered_map<uint64_t, int> lookup;  // (level, ix, iy, iz) -> node
    int find(int level, int ix, int iy, int iz) const;
};

Tree buildTree(const Sources& s, double rootLeng, int maxNodeSrcs, int maxLevelCap);

// ---------------------------------------------------------------- lists
struct Lists {
    std::vector<std::vector<int>> nbors;      // per node, distinct neighbours (ids)
    std::vector<std::vector<int>> nborDir;    // direction index of each neighbour
    std::vector<std::vector<int>> iList;      // list 2
    std::vector<std::vector<int>> leafIlist;  // list 4
    std::vector<std::vector<int>> nearNbors;  // list 1 (leaves only)
    std::vector<std::pair<int, int>> nearPairs;     // unordered adjacent leaves, both non-empty
    std::vector<std::pair<int, int>> nonNearPairs;  // (list-4 leaf, node)
};

Lists buildLists(const Tree& t);

// 316 M2L offsets in the order of Math::getINodeDistVecs (src/math.cpp:109-124)
int offsetIndex(int dx, int dy, int dz);
void offsetVector(int idx, int& dx, int& dy, int& dz);

// ---------------------------------------------------------------- grids & tables
struct Grid {
    int nth = 0, nph = 0, L = 0;
    std::vector<double> theta, weight, phi;  // weight includes sin(theta)
    int ndir() const { return nth * nph; }
};

Grid buildGrid(int level, double wavenum, double rootLeng, int digits);
void gaussLegendre(int n, double a, double b, std::vector<double>& x, std::vector<double>& w);

// Lagrange tables for one level transition (source level l+1 -> target level l)
struct InterpTable {
    int p = 0;                               // stencil half-width
    std::vector<int> thNear, phNear;         // per target theta / phi
    std::vector<double> thW, phW;            // per target x 2p
    // gather form of the transpose (anterpolation onto the finer grid)
    std::vector<int> atPhPtr, atPhSrc;       // per fine phi: coarse phi indices
    std::vector<double> atPhW;
    std::vector<int> atThPtr, atThSrc, atThFlip;  // per fine theta: coarse theta, pole flag
    std::vector<double> atThW;
};

InterpTable buildInterp(const Grid& fine, const Grid& coarse, int p);

// Host-side plan data (everything the device plan uploads)
struct HostPlan {
    mlfma_config cfg{};
    mlfma_planewave pw{};
    double k = 0;
    Sources src;
    Tree tree;
    Lists lists;
    std::vector<Grid> grids;                   // per level 0..maxLevel
    std::vector<InterpTable> interp;           // interp[l]: l+1 -> l, for 2 <= l < maxLevel
    std::vector<int> boxes;                    // non-empty boxes per level
    std::vector<std::vector<int>> slotNode;    // per level: slot -> node
    // worklists
    struct LeafItem { int node, level, slot; int64_t sbeg, cnt, rpOff; };
    std::vector<LeafItem> farLeaves;           // non-empty leaves at level >= 2
    int64_t radpatEntries = 0;
    // M2L CSR per level over target slots
    std::vector<std::vector<int>> m2lPtr, m2lSrc, m2lOff;
    std::vector<int64_t> m2lPairs;
    // M2M per level l (target): per non-empty stem slot, children (slot, octant)
    std::vector<std::vector<int>> m2mPtr, m2mChild, m2mOct;
    // L2L per level l+1 (target): (child slot, parent slot, octant)
    std::vector<std::vector<int>> l2lChild, l2lParent, l2lOct;
    // near field, stored once per unordered pair as the reference stores it
    // (src/fmm/leaf.cpp:108-180) and applied in both directions
    // (src/fmm/leaf.cpp:329-383).  Every non-empty leaf heads one row block:
    // its self block (full square, zero diagonal, applied to the rows only)
    // followed by the columns of the pairs it heads.  Row blocks are cut into
    // work items of at most 128 rows and a bounded number of entries; values
    // are column-major per item.  Each item writes partial sums (nr for its
    // rows, nc - nself for its pair columns); nearSumPtr/nearSumIdx list the
    // partials of every tree position in a fixed order.
    struct NearItem {
        int64_t rbeg;  // first row (tree position)
        int nr, nc, nself;
        int64_t cptr, voff, rpo, cpo;
    };
    std::vector<NearItem> nearItems;           // in value order
    std::vector<int64_t> nearCols;             // tree positions of the columns
    std::vector<int64_t> nearSumPtr;           // per tree position (n + 1)
    std::vector<int32_t> nearSumIdx;
    int64_t nearNnz = 0, nearParts = 0;
    int64_t staleBoxes = 0, truncatedL2L = 0;
    // a forest of several scatterers (mergeHostPlans): segOff[s] .. segOff[s+1]
    // are scatterer s's unknowns (input order), levelShift[s] maps its own
    // level l to the forest level l + levelShift[s]
    bool forest = false;
    std::vector<int64_t> segOff;
    std::vector<int> levelShift;
};

void buildHostPlan(HostPlan& hp, const mlfma_geometry& g, const mlfma_planewave& pw,
                   const mlfma_config& cfg);

// One plan for several independent scatterers: their trees side by side, the
// levels aligned by box size (root sides must be the largest one over powers
// of two), the worklists of every level concatenated so that each stage runs
// once for all of them; unknowns of scatterer s follow those of s - 1.
void mergeHostPlans(HostPlan& out, std::vector<HostPlan>& parts);

// Integer box coordinates and M2L offsets
inline uint64_t boxKey(int level, int ix, int iy, int iz) {
    return ((uint64_t)level << 60) | ((uint64_t)(uint32_t)ix << 40) | ((uint64_t)(uint32_t)iy << 20) |
           (uint64_t)(uint32_t)iz;
}

}  // namespace mlfma
