// C entry points over mlfma_amd/csrc/dense.hpp for the CPU tests (built by
// tests/test_dense_host.py with g++; complex values interleaved re/im,
// matrices column-major).
#include <cstring>

#include "../mlfma_amd/csrc/dense.hpp"

using mlfma::dense::Mat;
using mlfma::dense::zc;

static Mat load(int r, int c, const double* p) {
    Mat A(r, c);
    for (size_t i = 0; i < A.a.size(); ++i) A.a[i] = zc(p[2 * i], p[2 * i + 1]);
    return A;
}
static void store(const Mat& A, double* p) {
    for (size_t i = 0; i < A.a.size(); ++i) {
        p[2 * i] = A.a[i].real();
        p[2 * i + 1] = A.a[i].imag();
    }
}

extern "C" {

int dense_eig(int n, const double* A, double* w, double* X) {
    std::vector<zc> ww;
    Mat XX;
    if (!mlfma::dense::eig(load(n, n, A), ww, XX)) return 0;
    for (int i = 0; i < n; ++i) {
        w[2 * i] = ww[i].real();
        w[2 * i + 1] = ww[i].imag();
    }
    store(XX, X);
    return 1;
}

void dense_qr(int r, int c, const double* A, double* Q, double* R) {
    Mat QQ, RR;
    mlfma::dense::qrFull(load(r, c, A), QQ, RR);
    store(QQ, Q);
    store(RR, R);
}

int dense_lu_solve(int n, int nrhs, const double* A, double* B) {
    Mat BB = load(n, nrhs, B);
    if (!mlfma::dense::luSolve(load(n, n, A), BB)) return 0;
    store(BB, B);
    return 1;
}

// outputs sized for k: Pk m x k, U (m+1) x (k+1), Rk k x k, g k+1, Qhat (k+1)^2
int dense_dr(int m, const double* R, const double* cs, const double* sn, const double* gm, int k0,
             const double* Qprev, int k, double* Pk, double* U, double* Rk, double* g, double* Qhat,
             double* defect) {
    std::vector<zc> c(m), s(m);
    for (int i = 0; i < m; ++i) {
        c[i] = zc(cs[2 * i], cs[2 * i + 1]);
        s[i] = zc(sn[2 * i], sn[2 * i + 1]);
    }
    const Mat Qp = k0 > 0 ? load(k0 + 1, k0 + 1, Qprev) : Mat();
    mlfma::dense::DrState st;
    if (!mlfma::dense::drRestart(load(m, m, R), c, s, zc(gm[0], gm[1]), k0, Qp, k, st)) return 0;
    store(st.Pk, Pk);
    store(st.U, U);
    store(st.Rk, Rk);
    for (int i = 0; i <= k; ++i) {
        g[2 * i] = st.g[i].real();
        g[2 * i + 1] = st.g[i].imag();
    }
    store(st.Qhat, Qhat);
    *defect = st.defect;
    return 1;
}
}
