// dense.hpp — small dense complex linear algebra on the host for the
// deflated restarts of GMRES(m) (GMRES-DR, R. B. Morgan, "GMRES with
// deflated restarting", SIAM J. Sci. Comput. 24 (2002) 20-37): LU solves,
// Householder QR and Hessenberg reduction, eigenvalues by shifted QR
// iterations and eigenvectors by inverse iteration (the textbook algorithms,
// Golub & Van Loan, "Matrix Computations", 4th ed., 7.4-7.6), and the
// restart itself.
// Matrices are at most a few hundred rows (the restart length), so
// everything here is plain O(m^3) loops.  The reference has no counterpart:
// its GMRES never restarts (src/solver.cpp:116-146).
#pragma once
#include <algorithm>
#include <cmath>
#include <complex>
#include <limits>
#include <numeric>
#include <vector>

namespace mlfma {
namespace dense {

using zc = std::complex<double>;

// Complex products written out: std::complex's operator* calls the
// runtime's NaN-checking routine per product unless limited-range arithmetic
// is enabled for the whole translation unit.
inline zc mulz(zc a, zc b) {
    return zc(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
}
inline zc mulcz(zc a, zc b) {  // conj(a) b
    return zc(a.real() * b.real() + a.imag() * b.imag(), a.real() * b.imag() - a.imag() * b.real());
}
inline zc divz(zc a, zc b) {  // Smith's algorithm (no overflow in |b|^2)
    const double br = b.real(), bi = b.imag();
    if (std::abs(br) >= std::abs(bi)) {
        const double r = bi / br, t = 1.0 / (br + bi * r);
        return zc((a.real() + a.imag() * r) * t, (a.imag() - a.real() * r) * t);
    }
    const double r = br / bi, t = 1.0 / (br * r + bi);
    return zc((a.real() * r + a.imag()) * t, (a.imag() * r - a.real()) * t);
}

// column-major r x c matrix
struct Mat {
    int r = 0, c = 0;
    std::vector<zc> a;
    Mat() = default;
    Mat(int rows, int cols) : r(rows), c(cols), a((size_t)rows * cols, zc(0.0)) {}
    zc& operator()(int i, int j) { return a[(size_t)j * r + i]; }
    const zc& operator()(int i, int j) const { return a[(size_t)j * r + i]; }
    zc* col(int j) { return a.data() + (size_t)j * r; }
    const zc* col(int j) const { return a.data() + (size_t)j * r; }
};

inline Mat eye(int n) {
    Mat I(n, n);
    for (int i = 0; i < n; ++i) I(i, i) = 1.0;
    return I;
}

// A B
inline Mat mul(const Mat& A, const Mat& B) {
    Mat C(A.r, B.c);
    for (int j = 0; j < B.c; ++j) {
        zc* cj = C.col(j);
        for (int l = 0; l < A.c; ++l) {
            const zc b = B(l, j);
            if (b == zc(0.0)) continue;
            const zc* al = A.col(l);
            for (int i = 0; i < A.r; ++i) cj[i] += mulz(al[i], b);
        }
    }
    return C;
}

// A^H B
inline Mat mulH(const Mat& A, const Mat& B) {
    Mat C(A.c, B.c);
    for (int j = 0; j < B.c; ++j)
        for (int i = 0; i < A.c; ++i) {
            const zc *ai = A.col(i), *bj = B.col(j);
            zc s = 0.0;
            for (int l = 0; l < A.r; ++l) s += mulcz(ai[l], bj[l]);
            C(i, j) = s;
        }
    return C;
}

inline double normF(const Mat& A) {
    double s = 0;
    for (const zc& v : A.a) s += std::norm(v);
    return std::sqrt(s);
}

// Solves A X = B in place of B by LU with partial pivoting; false when a
// pivot is zero or not finite.
inline bool luSolve(Mat A, Mat& B) {
    const int n = A.r;
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (std::abs(A(i, k)) > std::abs(A(p, k))) p = i;
        const double piv = std::abs(A(p, k));
        if (!(piv > 0) || !std::isfinite(piv)) return false;
        if (p != k) {
            for (int j = 0; j < n; ++j) std::swap(A(k, j), A(p, j));
            for (int j = 0; j < B.c; ++j) std::swap(B(k, j), B(p, j));
        }
        const zc inv = divz(zc(1.0), A(k, k));
        for (int i = k + 1; i < n; ++i) A(i, k) = mulz(A(i, k), inv);
        for (int j = k + 1; j < n; ++j) {  // column-wise rank-1 update
            const zc akj = A(k, j);
            if (akj == zc(0.0)) continue;
            zc* aj = A.col(j);
            const zc* ak = A.col(k);
            for (int i = k + 1; i < n; ++i) aj[i] -= mulz(ak[i], akj);
        }
        for (int j = 0; j < B.c; ++j) {
            const zc bk = B(k, j);
            if (bk == zc(0.0)) continue;
            zc* bj = B.col(j);
            const zc* ak = A.col(k);
            for (int i = k + 1; i < n; ++i) bj[i] -= mulz(ak[i], bk);
        }
    }
    for (int j = 0; j < B.c; ++j) {
        zc* bj = B.col(j);
        for (int i = n - 1; i >= 0; --i) {
            bj[i] = divz(bj[i], A(i, i));
            const zc v = bj[i];
            const zc* ai = A.col(i);
            for (int l = 0; l < i; ++l) bj[l] -= mulz(ai[l], v);
        }
    }
    return true;
}

// Householder reflector of x (length len, stride 1): v with v^H v = 1 and
// beta such that (I - 2 v v^H) x = beta e_1; false when x is zero.
inline bool householder(const zc* x, int len, std::vector<zc>& v, zc& beta) {
    double nx = 0;
    for (int i = 0; i < len; ++i) nx += std::norm(x[i]);
    nx = std::sqrt(nx);
    v.assign(x, x + len);
    if (nx == 0) return false;
    const double a0 = std::abs(x[0]);
    const zc ph = a0 > 0 ? x[0] / a0 : zc(1.0);
    beta = -ph * nx;  // opposite phase to x_0: no cancellation in v_0
    v[0] -= beta;
    double nv = 0;
    for (int i = 0; i < len; ++i) nv += std::norm(v[i]);
    nv = std::sqrt(nv);
    for (int i = 0; i < len; ++i) v[i] /= nv;
    return true;
}

// rows r0..r0+len of columns c0..c1 of A: A = (I - 2 v v^H) A
inline void reflectRows(Mat& A, const std::vector<zc>& v, int r0, int len, int c0, int c1) {
    for (int j = c0; j < c1; ++j) {
        zc* aj = A.col(j) + r0;
        zc s = 0.0;
        for (int i = 0; i < len; ++i) s += mulcz(v[i], aj[i]);
        s *= 2.0;
        for (int i = 0; i < len; ++i) aj[i] -= mulz(s, v[i]);
    }
}

// columns c0..c0+len of rows 0..nr of A: A = A (I - 2 v v^H)
inline void reflectCols(Mat& A, const std::vector<zc>& v, int c0, int len, int nr) {
    std::vector<zc> s(nr, zc(0.0));
    for (int l = 0; l < len; ++l) {
        const zc* al = A.col(c0 + l);
        for (int i = 0; i < nr; ++i) s[i] += mulz(al[i], v[l]);
    }
    for (int l = 0; l < len; ++l) {
        zc* al = A.col(c0 + l);
        const zc f = 2.0 * std::conj(v[l]);
        for (int i = 0; i < nr; ++i) al[i] -= mulz(s[i], f);
    }
}

// Full QR of A (r x c, r >= c): Q (r x r, unitary) and R (r x c, upper
// triangular) with A = Q R.
inline void qrFull(const Mat& A, Mat& Q, Mat& R) {
    const int r = A.r, c = A.c;
    R = A;
    Q = eye(r);
    std::vector<zc> v;
    for (int j = 0; j < std::min(r - 1, c); ++j) {
        zc beta;
        if (!householder(R.col(j) + j, r - j, v, beta)) continue;
        reflectRows(R, v, j, r - j, j, c);
        for (int i = j + 1; i < r; ++i) R(i, j) = 0.0;
        reflectCols(Q, v, j, r - j, r);
    }
}

// Householder reduction to upper Hessenberg form: A = Q H Q^H.  A column
// whose entries below the subdiagonal are already zero is skipped, and a
// reflector spans only the rows down to the last nonzero entry, so a matrix
// that is Hessenberg except in a leading block costs what that block costs
// (the restart's matrices are such).
inline void hessenberg(const Mat& A, Mat& H, Mat& Q) {
    const int n = A.r;
    H = A;
    Q = eye(n);
    std::vector<zc> v;
    for (int j = 0; j + 2 < n; ++j) {
        int last = n - 1;
        while (last > j + 1 && H(last, j) == zc(0.0)) --last;
        if (last == j + 1) continue;
        const int len = last - j;
        zc beta;
        if (!householder(H.col(j) + j + 1, len, v, beta)) continue;
        reflectRows(H, v, j + 1, len, j, n);
        reflectCols(H, v, j + 1, len, n);
        reflectCols(Q, v, j + 1, len, n);
        for (int i = j + 2; i <= last; ++i) H(i, j) = 0.0;
    }
}

// Eigenvalues of an upper Hessenberg H: single-shift QR sweeps on the
// active block (Wilkinson shift of its trailing 2 x 2 block, an exceptional
// shift every tenth sweep without deflation), Givens rotations
// G = [[c, s], [-conj(s), c]] (c real), deflation of negligible subdiagonal
// entries.  False when a block does not converge in 30 sweeps per row.
inline bool hessEigvals(Mat T, std::vector<zc>& w) {
    const int n = T.r;
    w.assign(n, zc(0.0));
    const double eps = std::numeric_limits<double>::epsilon();
    const double tnorm = std::max(normF(T), std::numeric_limits<double>::min());
    int hi = n - 1, sweeps = 0, total = 0;
    while (hi >= 0) {
        if (hi == 0) {
            w[0] = T(0, 0);
            break;
        }
        // deflation test in the 1-norm of the real and imaginary parts
        // (|re| + |im|, as LAPACK's complex QR does: no square roots)
        auto cabs1 = [](zc z) { return std::abs(z.real()) + std::abs(z.imag()); };
        int l = hi;
        for (; l > 0; --l) {
            double s = cabs1(T(l - 1, l - 1)) + cabs1(T(l, l));
            if (s == 0) s = tnorm;
            if (cabs1(T(l, l - 1)) <= eps * s) {
                T(l, l - 1) = 0.0;
                break;
            }
        }
        if (l == hi) {  // 1 x 1 block converged
            w[hi] = T(hi, hi);
            --hi;
            sweeps = 0;
            continue;
        }
        if (++total > 30 * n) return false;
        ++sweeps;
        zc mu;
        if (sweeps % 10 == 0) {
            mu = T(hi, hi) + std::abs(T(hi, hi - 1).real()) + std::abs(T(hi, hi - 1).imag());
        } else {
            const zc a = T(hi - 1, hi - 1), b = T(hi - 1, hi), c = T(hi, hi - 1), d = T(hi, hi);
            const zc half = 0.5 * (a + d);
            const zc disc = std::sqrt(0.25 * mulz(a - d, a - d) + mulz(b, c));
            const zc m1 = half + disc, m2 = half - disc;
            mu = std::abs(m1 - d) < std::abs(m2 - d) ? m1 : m2;
        }
        for (int k = l; k < hi; ++k) {
            const zc x = k == l ? T(l, l) - mu : T(k, k - 1);
            const zc y = k == l ? T(l + 1, l) : T(k + 1, k - 1);
            const double ax = std::abs(x), rho = std::hypot(ax, std::abs(y));
            if (rho == 0) continue;
            const double c = ax / rho;
            const zc s = mulz(ax > 0 ? x / ax : zc(1.0), std::conj(y)) / rho;
            for (int j = (k == l ? l : k - 1); j <= hi; ++j) {
                const zc p = T(k, j), q = T(k + 1, j);
                T(k, j) = c * p + mulz(s, q);
                T(k + 1, j) = c * q - mulcz(s, p);
            }
            if (k > l) T(k + 1, k - 1) = 0.0;
            const int iend = std::min(k + 2, hi);
            for (int i = l; i <= iend; ++i) {
                const zc p = T(i, k), q = T(i, k + 1);
                T(i, k) = c * p + mulcz(s, q);
                T(i, k + 1) = c * q - mulz(s, p);
            }
        }
    }
    return true;
}

// An eigenvector of the upper Hessenberg H for the eigenvalue theta by
// inverse iteration (three solves with (H - theta I), LU with the row
// exchanges Hessenberg form allows; a zero pivot becomes eps |H|), started
// from a vector that depends on `seed` so that equal eigenvalues of a
// non-defective H give independent vectors.  Unit 2-norm.
inline void hessEigvec(const Mat& H, zc theta, int seed, std::vector<zc>& x) {
    const int n = H.r;
    const double tiny = std::max(normF(H), 1.0) * std::numeric_limits<double>::epsilon();
    // LU of (H - theta I): L unit lower bidiagonal (multipliers lm, row
    // exchange flags sw), U upper triangular in U
    Mat U = H;
    for (int i = 0; i < n; ++i) U(i, i) -= theta;
    std::vector<zc> lm(n, zc(0.0));
    std::vector<char> sw(n, 0);
    for (int k = 0; k + 1 < n; ++k) {
        if (std::abs(U(k + 1, k)) > std::abs(U(k, k))) {
            for (int j = k; j < n; ++j) std::swap(U(k, j), U(k + 1, j));
            sw[k] = 1;
        }
        if (std::abs(U(k, k)) < tiny) U(k, k) = tiny;
        lm[k] = divz(U(k + 1, k), U(k, k));
        U(k + 1, k) = 0.0;
        for (int j = k + 1; j < n; ++j) U(k + 1, j) -= mulz(lm[k], U(k, j));
    }
    if (std::abs(U(n - 1, n - 1)) < tiny) U(n - 1, n - 1) = tiny;
    x.assign(n, zc(0.0));
    for (int i = 0; i < n; ++i) x[i] = zc(1.0 + 0.37 * ((i * 7 + seed * 13) % 11), 0.21 * ((i * 5 + seed * 3) % 7));
    for (int it = 0; it < 3; ++it) {
        for (int k = 0; k + 1 < n; ++k) {  // forward: L y = P x
            if (sw[k]) std::swap(x[k], x[k + 1]);
            x[k + 1] -= mulz(lm[k], x[k]);
        }
        for (int i = n - 1; i >= 0; --i) {  // back: U x = y
            zc s = x[i];
            for (int j = i + 1; j < n; ++j) s -= mulz(U(i, j), x[j]);
            x[i] = divz(s, U(i, i));
        }
        double nr = 0;
        for (const zc& v : x) nr += std::norm(v);
        nr = std::sqrt(nr);
        for (zc& v : x) v /= nr;
    }
}

// Eigenvalues w of A and unit eigenvectors (columns of X) for the indices
// `want` of w (all when empty); false when the QR iteration does not
// converge.
inline bool eig(const Mat& A, std::vector<zc>& w, Mat& X, const std::vector<int>& want = {}) {
    const int n = A.r;
    Mat H, Q;
    hessenberg(A, H, Q);
    if (!hessEigvals(H, w)) return false;
    std::vector<int> idx = want;
    if (idx.empty()) {
        idx.resize(n);
        std::iota(idx.begin(), idx.end(), 0);
    }
    X = Mat(n, (int)idx.size());
    std::vector<zc> y;
    for (size_t c = 0; c < idx.size(); ++c) {
        hessEigvec(H, w[idx[c]], (int)c, y);
        zc* xc = X.col((int)c);
        for (int l = 0; l < n; ++l) {
            const zc* ql = Q.col(l);
            for (int r = 0; r < n; ++r) xc[r] += mulz(ql[r], y[l]);
        }
    }
    return true;
}

// ---- the deflated restart
//
// A cycle of the device GMRES(m) ends with the Arnoldi-type relation
//   A Z_m = V_{m+1} H,   H = Omega^H [R; 0],
// where V_{m+1} (the cycle's orthonormal basis) and Z_m (its inputs: the k0
// deflated vectors kept from the previous restart, then V's columns k0..m-1)
// are on the device; R (m x m upper triangular), the rotations
// G_i = [[conj c_i, conj s_i], [-s_i, c_i]] on rows (i, i+1) (the identity
// for i < k0) and the rotated right-hand side g are the device's Givens
// state.  Z_m = V_m T with T = I except its first k0 columns, which are the
// first k0 columns of Qprev^H (rows 0..k0; Qprev from the previous restart),
// so A V_m = V_{m+1} Hh with Hh = H T^{-1}.  The restart keeps the k
// harmonic Ritz vectors of smallest modulus (eigenvectors g of
// Hh_m + f h, h the last row of Hh and Hh_m^H f = h^H) and the residual
// direction:
//   P_{k+1} = orth[g_1 .. g_k (a zero row appended), c],  c = Omega^H g_m e_m
//   (the residual b - A x of the cycle's solution is V_{m+1} c),
//   Hbar = P_{k+1}^H Hh P_k = Qhat [Rk; 0]  (full QR),
// and hands back
//   Zk' = V_m Pk (the next cycle's first k inputs),
//   V'  = V_{m+1} U with U = P_{k+1} Qhat (its first k+1 basis vectors),
//   A Zk' = V' [Rk; 0]: the next cycle's first k Hessenberg columns, already
//   upper triangular (identity rotations), and g' = Qhat^H P_{k+1}^H c its
//   rotated right-hand side (b - A x = V' g').
struct DrState {
    int k = 0;           // deflated vectors of the next cycle
    Mat Pk;              // m x k
    Mat U;               // (m + 1) x (k + 1)
    Mat Rk;              // k x k, upper triangular
    std::vector<zc> g;   // k + 1
    Mat Qhat;            // (k + 1) x (k + 1): the next restart's Qprev
    double defect = 0;   // |(I - P P^H) Hh Pk| / |Hh|: how far the kept
                         // vectors are from an exact harmonic Ritz space
};

// R: m x m (upper triangle used); cs, sn: m rotations; gm: the rotated
// right-hand side's last entry; Qprev: (k0 + 1) x (k0 + 1) when k0 > 0.
// False when an intermediate step fails (singular matrix, no Schur
// convergence, a kept vector in the span of the others): the caller then
// restarts plainly from the residual.
inline bool drRestart(const Mat& R, const std::vector<zc>& cs, const std::vector<zc>& sn, zc gm, int k0,
                      const Mat& Qprev, int k, DrState& out) {
    const int m = R.c;
    if (k < 1 || k + 1 >= m || k0 < 0 || k0 >= m) return false;
    // H = Omega^H [R; 0] and c = Omega^H g_m e_m (G_i^H = [[c, -conj s], [s, conj c]])
    Mat H(m + 1, m);
    for (int j = 0; j < m; ++j)
        for (int i = 0; i <= j; ++i) H(i, j) = R(i, j);
    std::vector<zc> cv(m + 1, zc(0.0));
    cv[m] = gm;
    for (int i = m - 1; i >= 0; --i) {
        const zc c = cs[i], s = sn[i];
        for (int j = 0; j < m; ++j) {
            const zc p = H(i, j), q = H(i + 1, j);
            H(i, j) = c * p - std::conj(s) * q;
            H(i + 1, j) = s * p + std::conj(c) * q;
        }
        const zc p = cv[i], q = cv[i + 1];
        cv[i] = c * p - std::conj(s) * q;
        cv[i + 1] = s * p + std::conj(c) * q;
    }
    // Hh = H T^{-1}: T = blockdiag(B, I) with B = [first k0 columns of
    // Qprev^H, e_k0] ((k0 + 1) x (k0 + 1)), so only H's first k0 + 1 columns change
    Mat Hh = H;
    if (k0 > 0) {
        if (Qprev.r != k0 + 1 || Qprev.c != k0 + 1) return false;
        Mat B(k0 + 1, k0 + 1);
        for (int j = 0; j < k0; ++j)
            for (int i = 0; i <= k0; ++i) B(i, j) = std::conj(Qprev(j, i));
        B(k0, k0) = 1.0;
        // X B = H1 for X: B^T X^T = H1^T
        Mat Bt(k0 + 1, k0 + 1), Xt(k0 + 1, m + 1);
        for (int j = 0; j <= k0; ++j)
            for (int i = 0; i <= k0; ++i) Bt(i, j) = B(j, i);
        for (int i = 0; i <= m; ++i)
            for (int j = 0; j <= k0; ++j) Xt(j, i) = H(i, j);
        if (!luSolve(Bt, Xt)) return false;
        for (int i = 0; i <= m; ++i)
            for (int j = 0; j <= k0; ++j) Hh(i, j) = Xt(j, i);
    }
    // harmonic Ritz pairs: (Hm + f h) g = theta g, Hm^H f = h^H (h: row m of Hh)
    Mat Hm(m, m), f(m, 1);
    for (int j = 0; j < m; ++j) {
        for (int i = 0; i < m; ++i) Hm(i, j) = Hh(i, j);
        f(j, 0) = std::conj(Hh(m, j));
    }
    Mat HmH(m, m);
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) HmH(i, j) = std::conj(Hm(j, i));
    if (!luSolve(HmH, f)) return false;
    Mat F = Hm;
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) F(i, j) += f(i, 0) * Hh(m, j);
    Mat Hf, Qf;
    hessenberg(F, Hf, Qf);
    std::vector<zc> th;
    if (!hessEigvals(Hf, th)) return false;
    std::vector<int> ord(m);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return std::abs(th[a]) < std::abs(th[b]); });
    // P_{k+1}: the k eigenvectors, then c, orthonormalized (two Gram-Schmidt passes)
    Mat P(m + 1, k + 1);
    std::vector<zc> y;
    for (int j = 0; j < k; ++j) {
        hessEigvec(Hf, th[ord[j]], j, y);
        for (int l = 0; l < m; ++l) {
            const zc* ql = Qf.col(l);
            for (int i = 0; i < m; ++i) P(i, j) += mulz(ql[i], y[l]);
        }
    }
    for (int i = 0; i <= m; ++i) P(i, k) = cv[i];
    for (int j = 0; j <= k; ++j) {
        double n0 = 0;
        for (int i = 0; i <= m; ++i) n0 += std::norm(P(i, j));
        n0 = std::sqrt(n0);
        if (!(n0 > 0) || !std::isfinite(n0)) return false;
        for (int pass = 0; pass < 2; ++pass)
            for (int q = 0; q < j; ++q) {
                zc s = 0.0;
                for (int i = 0; i <= m; ++i) s += std::conj(P(i, q)) * P(i, j);
                for (int i = 0; i <= m; ++i) P(i, j) -= s * P(i, q);
            }
        double n1 = 0;
        for (int i = 0; i <= m; ++i) n1 += std::norm(P(i, j));
        n1 = std::sqrt(n1);
        if (!(n1 > 1e-10 * n0)) return false;
        for (int i = 0; i <= m; ++i) P(i, j) /= n1;
    }
    Mat Pk(m, k);
    for (int j = 0; j < k; ++j)
        for (int i = 0; i < m; ++i) Pk(i, j) = P(i, j);
    const Mat HP = mul(Hh, Pk);   // (m + 1) x k
    const Mat Hbar = mulH(P, HP); // (k + 1) x k
    Mat D = HP;                   // the part of Hh Pk outside span(P)
    const Mat PHbar = mul(P, Hbar);
    for (size_t i = 0; i < D.a.size(); ++i) D.a[i] -= PHbar.a[i];
    const double hn = normF(Hh);
    out.defect = hn > 0 ? normF(D) / hn : 0;
    Mat Qh, Rf;
    qrFull(Hbar, Qh, Rf);
    out.k = k;
    out.Pk = Pk;
    out.U = mul(P, Qh);
    out.Rk = Mat(k, k);
    for (int j = 0; j < k; ++j)
        for (int i = 0; i <= j; ++i) out.Rk(i, j) = Rf(i, j);
    Mat cm(m + 1, 1);
    for (int i = 0; i <= m; ++i) cm(i, 0) = cv[i];
    const Mat gp = mulH(out.U, cm);
    out.g.resize(k + 1);
    for (int i = 0; i <= k; ++i) out.g[i] = gp(i, 0);
    out.Qhat = Qh;
    for (int i = 0; i < k; ++i)
        if (!(std::abs(out.Rk(i, i)) > 0) || !std::isfinite(std::abs(out.Rk(i, i)))) return false;
    return true;
}

}  // namespace dense
}  // namespace mlfma
