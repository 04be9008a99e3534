// oracle.hpp — CPU FP64 restatement of the reference MLFMA matvec and GMRES.
//
// TEST INFRASTRUCTURE ONLY.  This code is the parity checker for the MI355X
// library in mlfma_amd/: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only to check or time against it.  The
// product never links, loads or calls anything under oracle/.
//
// It restates, function by function, the reference's algorithm (ellilu11/MLFMA)
// in plain C++17 with std::complex and no Eigen.  Each function cites the
// reference file:line it follows.  The reference's object structure (recursive
// Stem/Leaf tree, pointer lists, per-node coefficient vectors, libstdc++ hash
// maps for the psi table) is kept on purpose, so that the summation orders and
// the quirks listed below are those of the reference:
//   Q1 psi-stencil half-sample flip          src/fmm/tables.cpp:200-213
//   Q2 non-unitary complex Givens             src/math.h:168-172
//   Q3 self / coincident-triangle terms = 0   src/sources/rwg.cpp:97,113-116
//   Q4 stale locals when the iList is empty   src/fmm/node.cpp:105-108
//   Q5 L2L returns zero when parent iList = 0 src/fmm/stem.cpp:115
//   psi hash-map key merging (EPS_psi=1e-9)   src/maps.h:11-31, src/fmm/tables.cpp:129-176
// Deviations (each documented where it occurs):
//   * near-pair orientation uses node creation order instead of heap address
//     order (src/fmm/leaf.cpp:76); this changes summation order only.
//   * optional level cap on the adaptive split rule, to express the fixed
//     leaf sizes of the synthetic workloads (the reference splits only on
//     maxNodeSrcs, src/fmm/stem.cpp:19).
//   * the Krylov basis is preallocated instead of conservativeResize
//     (src/solver.cpp:76); arithmetic unchanged.
#pragma once

#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace orc {

using cd = std::complex<double>;
const cd IU(0.0, 1.0);
extern const double PI;

struct V3 {
    double c[3];
    V3() : c{0, 0, 0} {}
    V3(double x, double y, double z) : c{x, y, z} {}
    double& operator[](int i) { return c[i]; }
    double operator[](int i) const { return c[i]; }
};
inline V3 operator+(const V3& a, const V3& b) { return V3(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
inline V3 operator-(const V3& a, const V3& b) { return V3(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
inline V3 operator*(double s, const V3& a) { return V3(s * a[0], s * a[1], s * a[2]); }
inline V3 operator/(const V3& a, double s) { return V3(a[0] / s, a[1] / s, a[2] / s); }
inline double dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline double norm(const V3& a) { return std::sqrt(dot(a, a)); }
inline bool operator==(const V3& a, const V3& b) { return a[0] == b[0] && a[1] == b[1] && a[2] == b[2]; }

using C3 = std::array<cd, 3>;
using C2 = std::array<cd, 2>;

// ---------------------------------------------------------------- math (src/math.{h,cpp})
double legendreP(double x, int n, double* dP);          // src/math.cpp:25-44
cd sphericalHankel1(double x, int n);                   // src/math.cpp:68-84
void dyadicG(const V3& X, double k, cd G[3][3]);        // src/math.h:132-142
int flipIdxToRange(int i, int size);                    // src/math.h:144-154
int wrapIdxToRange(int i, int size);                    // src/math.h:156-166
std::pair<cd, cd> givensRotation(cd z, cd w);           // src/math.h:168-172
V3 idx2pm(int x);                                       // src/math.cpp:8-17
std::vector<double> iNodeDistances();                   // src/math.cpp:86-107
std::vector<V3> iNodeDistVecs();                        // src/math.cpp:109-124
std::vector<V3> iNodeDirections();                      // src/math.cpp:126-147
inline double sgn(int k) { return (k % 2) ? -1.0 : 1.0; }  // src/math.h:22-24
cd powI(int m);                                          // src/math.h:26-33

// ---------------------------------------------------------------- interp (src/interp.cpp)
void gaussLegendre(int l, double eps, double a, double b,
                   std::vector<double>& nodes, std::vector<double>& weights);  // :12-52
int nearGLNodeIdx(double xi, int m, double a, double b);                        // :64-77
double lagrangeBasis(double x, const std::vector<double>& xs, int k);           // :86-100

// ---------------------------------------------------------------- inputs
struct PlaneWave {            // src/excite.h:6-28
    double amplitude = 1.0;
    V3 pol{1, 0, 0};
    V3 wavevec{0, 0, 1};
    double wavenum = 1.0;
};

struct Triangle {             // src/sources/triangle.{h,cpp}
    int id = 0;
    int vIdx[3] = {0, 0, 0};
    V3 Xs[3];
    V3 Ds[3];
    V3 center;
    std::vector<std::pair<V3, double>> quads;
};

void buildQuads(Triangle& t, int prec);  // src/sources/triangle.cpp:45-130

enum SrcKind { DIPOLE = 0, RWG = 1 };

struct Source {               // src/sources/source.h:10-40 (+ dipole.h, rwg.h)
    SrcKind kind = DIPOLE;
    size_t idx = 0;
    cd voltage = 0.0;
    // dipole
    V3 pos, pol, phat;
    double pmag = 0.0;
    // RWG
    const Triangle* tris[2] = {nullptr, nullptr};
    V3 Xpm[2];
    V3 X0, X1, center;
    double leng = 0.0;

    V3 getCenter() const { return kind == DIPOLE ? pos : center; }
    C3 integratedPlaneWave(const V3& kvec) const;              // src/sources/rwg.cpp:28-87
    C3 integratedPlaneWaveNumeric(const V3& kvec) const;       // src/sources/rwg.cpp:34-44
    C3 radAlongDir(const V3& X, const V3& kvec) const;        // dipole.h:47-51, rwg.h:23-25
    C3 farAlongDir(const V3& krhat) const;                    // dipole.h:53-57, rwg.h:27-29
    cd integratedRad(const Source& src, double k) const;      // dipole.h:62-70, rwg.cpp:92-139
    void buildVoltage(const PlaneWave& pw);                   // dipole.h:29-32, rwg.h:18-21
};

struct Config {               // src/config.h:50-73 (+ level cap, see header)
    int digits = 3;
    int interpOrder = 3;
    double overInterp = 4.0;
    double rootLeng = 4.0;
    int maxNodeSrcs = 16;
    int maxLevelCap = -1;
    int quadPrec = 2;
};

// ---------------------------------------------------------------- FMM
struct Angles {               // src/fmm/angles.{h,cpp}
    std::vector<double> thetas, weights, phis;
    std::vector<V3> khat;
    std::vector<std::array<double, 6>> toThPh;  // 2x3 row-major
    int L = 0;
    int nth() const { return (int)thetas.size(); }
    int nph() const { return (int)phis.size(); }
};

struct InterpPair {
    std::vector<double> coeffs;
    int nearIdx = 0;
};

struct Tables {               // src/fmm/tables.{h,cpp}
    std::vector<InterpPair> interpTheta, interpPhi;
    std::vector<std::vector<cd>> transl;       // indexed like iNodeDistVecs()
    std::vector<std::vector<cd>> alpha;        // [dist][ips]
    int nps = 0;
    long psiKeys = 0;                          // number of distinct psi keys kept
    long psiMerged = 0;                        // psis merged into an existing key
};

struct Node {
    int id = 0;
    int level = 0;
    int branchIdx = 0;
    Node* base = nullptr;
    bool leaf = false;
    double nodeLeng = 0.0;
    V3 center;
    std::vector<int> srcs;                 // indices into Model::srcs (reference order)
    std::vector<Node*> branches;
    std::vector<Node*> nbors, iList, leafIlist, nearNbors, nearNonNbors;
    std::vector<C2> coeffs, localCoeffs;
    // leaf data (src/fmm/leaf.h:59-81)
    std::vector<std::vector<cd>> nearRads, nonNearRads;
    std::vector<cd> selfRads;
    std::vector<std::vector<C2>> radPats;  // [dir][src]
    size_t leafPairIdx = 0, nonNearPairIdx = 0;
    bool isSrcless() const { return srcs.empty(); }
    bool isRoot() const { return base == nullptr; }
};

// The neighbour rule of src/fmm/nbor.cpp for one (direction, branch index):
// see oracle.cpp.  Directions in the order of FMM::Dir.
bool neighborRule(int dir, int b, int* target, int* exitDir);
bool neighborChildFaces(int dir, int c);

struct Model {
    Config cfg;
    PlaneWave pw;
    double wavenum = 1.0;
    std::vector<Triangle> tris;
    std::vector<Source> srcs;
    std::vector<std::unique_ptr<Node>> nodes;     // creation order
    Node* root = nullptr;
    int maxLevel = 0;
    std::vector<Node*> leaves;                    // DFS order (src/fmm/leaf.cpp:48)
    std::vector<std::pair<Node*, Node*>> nearPairs;
    std::vector<std::pair<Node*, Node*>> nonNearPairs;
    std::vector<Angles> angles;
    std::vector<Tables> tables;
    std::vector<double> dists;
    std::vector<V3> rhats;
    std::vector<V3> dXs;
    std::vector<cd> lvec, rvec;
    bool direct = false;                          // evalDirect single-leaf tree
    int threads = 1;                              // setup parallelism only

    // setup, in the order of src/mlfma.cpp:21-64
    void buildTree();                 // src/fmm/stem.cpp:3-26, src/mlfma.cpp:26-30
    void buildLists();                // src/fmm/stem.cpp:48-57, src/fmm/leaf.cpp:50-58
    void buildNearRads();             // src/fmm/leaf.cpp:108-180
    void buildTables();               // src/fmm/node.cpp:46-56
    void resizeCoeffs();              // src/fmm/stem.cpp:59-67
    void buildRadPats();              // src/fmm/leaf.cpp:185-207
    void setupAll();

    // matvec: rvec = 0; rvec += Z lvec  (src/solver.cpp:41-57)
    void matvec(const cd* x, cd* y);
    // direct O(N^2) matvec with the same entries (evalDirect semantics, src/mlfma.cpp:84-107)
    void denseMatvec(const cd* x, cd* y) const;
    // far field E(rhat) = C k (I - rr) sum_n I_n f_n(k rhat) (src/fmm/node.cpp:147-178)
    void farfield(const cd* I, int ndir, const double* dirs, cd* E) const;

    // internals
    Node* newNode(const std::vector<int>& s, int branchIdx, Node* base, bool leaf);
    void splitStem(Node* n);
    Node* neighborGeq(const Node* n, int dir) const;
    std::vector<Node*> neighborsLeq(const Node* n, Node* nb, int dir) const;
    void buildListsRec(Node* n);
    void buildMpoleRec(Node* n);
    void buildLocalRec(Node* n);
    void m2l(Node* n);
    std::vector<C2> shiftedLocal(const Node* parent, int branchIdx) const;
    void addInterp(const std::vector<C2>& in, std::vector<C2>& out, int srcLvl, int tgtLvl) const;
    void addAnterp(const std::vector<C2>& in, std::vector<C2>& out, int srcLvl, int tgtLvl) const;
    void evalFarSols(Node* leaf);
    void evalPairSols(Node* obsLeaf, const Node* srcNode, const std::vector<cd>& rads);
    void evalSelfSols(Node* leaf);
    void evaluateSols();
    int offsetIndex(const V3& dXn) const;
    void buildAngles(int level, Angles& a) const;
    void buildTablesLevel(int level, Tables& t) const;
};

// compat GMRES of src/solver.cpp:5-146
struct SolveResult {
    int iters = 0;
    std::vector<double> resid;   // |g_{k}|/g0 after each iteration
    int status = 0;              // 0 ok, -4 breakdown/non-finite
    // optional record: Arnoldi basis (columns), unrotated Hessenberg columns, y
    bool record = false;
    std::vector<cd> Q;
    std::vector<std::vector<cd>> Hraw;
    std::vector<cd> y;
    double g0 = 0;
};
SolveResult gmresCompat(Model& m, int maxIter, double eps, cd* currents, bool record = false);

}  // namespace orc
