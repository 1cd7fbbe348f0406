// expm.h -- small dense complex matrices on the host in FP64 and their
// exponential, for the Krylov driver's Hessenberg matrix.
//
// The reference evaluates exp((-beta - i t) Hfull) with Eigen's
// MatrixFunctions in the run's FpType (inc/algo.h:331-332).  This build keeps
// the small matrices in FP64 for both state precisions and uses the scaling
// and squaring Pade method of N. J. Higham, SIAM J. Matrix Anal. Appl. 26
// (2005) 1179-1193: Pade degree m in {3,5,7,9,13} chosen from the 1-norm
// against Higham's theta_m bounds for double precision, 2^-s scaling so that
// ||A||_1 <= theta_13, and s squarings.
#pragma once

#ifndef _REAPERS_H_
#error "Do not include this file directly. Include the master header file reapers.h"
#endif

namespace dense {

using cd = std::complex<double>;

// column-major n x n complex matrix
struct CMat {
    int n = 0;
    std::vector<cd> a;
    CMat() = default;
    explicit CMat(int n_) : n(n_), a((size_t)n_ * n_) {}
    cd &operator()(int i, int j) { return a[(size_t)j * n + i]; }
    const cd &operator()(int i, int j) const { return a[(size_t)j * n + i]; }
    static CMat identity(int n) {
        CMat m(n);
        for (int i = 0; i < n; i++) m(i, i) = 1.0;
        return m;
    }
    // resize keeping the top-left block, new entries zero
    void grow(int nn) {
        CMat m(nn);
        for (int j = 0; j < std::min(n, nn); j++)
            for (int i = 0; i < std::min(n, nn); i++) m(i, j) = (*this)(i, j);
        *this = std::move(m);
    }
};

inline CMat matmul(const CMat &A, const CMat &B)
{
    const int n = A.n;
    CMat C(n);
    for (int j = 0; j < n; j++)
        for (int k = 0; k < n; k++) {
            const cd b = B(k, j);
            if (b == cd(0)) continue;
            const cd *ak = &A.a[(size_t)k * n];
            cd *cj = &C.a[(size_t)j * n];
            for (int i = 0; i < n; i++) cj[i] += ak[i] * b;
        }
    return C;
}

inline double norm1(const CMat &A)
{
    double best = 0;
    for (int j = 0; j < A.n; j++) {
        double s = 0;
        for (int i = 0; i < A.n; i++) s += std::abs(A(i, j));
        best = std::max(best, s);
    }
    return best;
}

// solve D X = B with partial pivoting; B is overwritten by X
inline void lu_solve(CMat D, CMat &B)
{
    const int n = D.n;
    for (int k = 0; k < n; k++) {
        int p = k;
        double best = -1;
        for (int i = k; i < n; i++) {
            double v = std::abs(D(i, k));
            if (v > best) { best = v; p = i; }
        }
        if (p != k)
            for (int j = 0; j < n; j++) {
                std::swap(D(k, j), D(p, j));
                std::swap(B(k, j), B(p, j));
            }
        const cd inv = 1.0 / D(k, k);
        for (int i = k + 1; i < n; i++) {
            const cd l = D(i, k) * inv;
            if (l == cd(0)) continue;
            for (int j = k + 1; j < n; j++) D(i, j) -= l * D(k, j);
            for (int j = 0; j < n; j++) B(i, j) -= l * B(k, j);
        }
    }
    for (int j = 0; j < n; j++)
        for (int i = n - 1; i >= 0; i--) {
            cd s = B(i, j);
            for (int k = i + 1; k < n; k++) s -= D(i, k) * B(k, j);
            B(i, j) = s / D(i, i);
        }
}

// sum_k w[k] M[k] + w0 I
inline CMat combo(const std::vector<const CMat *> &M, const std::vector<double> &w, double w0, int n)
{
    CMat R(n);
    for (size_t k = 0; k < M.size(); k++)
        for (size_t q = 0; q < R.a.size(); q++) R.a[q] += w[k] * M[k]->a[q];
    for (int i = 0; i < n; i++) R(i, i) += w0;
    return R;
}

inline CMat expm(const CMat &A0)
{
    const int n = A0.n;
    if (n == 0) return CMat();
    static const double theta[] = {1.495585217958292e-2, 2.539398330063230e-1,
                                   9.504178996162932e-1, 2.097847961257068e0,
                                   5.371920351148152e0};
    double nrm = norm1(A0);
    CMat A = A0;
    int s = 0, deg;
    if (nrm <= theta[0]) deg = 3;
    else if (nrm <= theta[1]) deg = 5;
    else if (nrm <= theta[2]) deg = 7;
    else if (nrm <= theta[3]) deg = 9;
    else {
        deg = 13;
        s = std::max(0, (int)std::ceil(std::log2(nrm / theta[4])));
        const double sc = std::ldexp(1.0, -s);
        for (auto &v : A.a) v *= sc;
    }
    CMat U, V;
    CMat A2 = matmul(A, A);
    if (deg == 3) {
        U = matmul(A, combo({&A2}, {1.}, 60., n));
        V = combo({&A2}, {12.}, 120., n);
    } else if (deg == 5) {
        CMat A4 = matmul(A2, A2);
        U = matmul(A, combo({&A4, &A2}, {1., 420.}, 15120., n));
        V = combo({&A4, &A2}, {30., 3360.}, 30240., n);
    } else if (deg == 7) {
        CMat A4 = matmul(A2, A2), A6 = matmul(A4, A2);
        U = matmul(A, combo({&A6, &A4, &A2}, {1., 1512., 277200.}, 8648640., n));
        V = combo({&A6, &A4, &A2}, {56., 25200., 1995840.}, 17297280., n);
    } else if (deg == 9) {
        CMat A4 = matmul(A2, A2), A6 = matmul(A4, A2), A8 = matmul(A6, A2);
        U = matmul(A, combo({&A8, &A6, &A4, &A2}, {1., 3960., 2162160., 302702400.},
                            8821612800., n));
        V = combo({&A8, &A6, &A4, &A2}, {90., 110880., 30270240., 2075673600.}, 17643225600., n);
    } else {
        const double b[] = {64764752532480000., 32382376266240000., 7771770303897600.,
                            1187353796428800., 129060195264000., 10559470521600.,
                            670442572800., 33522128640., 1323241920., 40840800.,
                            960960., 16380., 182., 1.};
        CMat A4 = matmul(A2, A2), A6 = matmul(A4, A2);
        CMat Ui = matmul(A6, combo({&A6, &A4, &A2}, {b[13], b[11], b[9]}, 0., n));
        CMat Ul = combo({&A6, &A4, &A2}, {b[7], b[5], b[3]}, b[1], n);
        for (size_t q = 0; q < Ui.a.size(); q++) Ui.a[q] += Ul.a[q];
        U = matmul(A, Ui);
        V = matmul(A6, combo({&A6, &A4, &A2}, {b[12], b[10], b[8]}, 0., n));
        CMat Vl = combo({&A6, &A4, &A2}, {b[6], b[4], b[2]}, b[0], n);
        for (size_t q = 0; q < V.a.size(); q++) V.a[q] += Vl.a[q];
    }
    CMat P(n), Q(n);
    for (size_t q = 0; q < P.a.size(); q++) {
        P.a[q] = V.a[q] + U.a[q];
        Q.a[q] = V.a[q] - U.a[q];
    }
    lu_solve(Q, P);
    for (int k = 0; k < s; k++) P = matmul(P, P);
    return P;
}

// eigen-decomposition of the symmetric tridiagonal matrix with diagonal a
// and off-diagonal b, by cyclic Jacobi rotations (the Lanczos matrices are
// small); eigenvector k is column k of Q (column-major)
inline void tridiag_eigen(const std::vector<double> &a, const std::vector<double> &b,
                          std::vector<double> &w, std::vector<double> &Q)
{
    const int n = (int)a.size();
    std::vector<double> A((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) A[(size_t)i * n + i] = a[i];
    for (int i = 0; i + 1 < n; i++) {
        A[(size_t)i * n + i + 1] = b[i];
        A[(size_t)(i + 1) * n + i] = b[i];
    }
    Q.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) Q[(size_t)i * n + i] = 1.0;
    auto at = [&](int i, int j) -> double & { return A[(size_t)j * n + i]; };
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) off += at(p, q) * at(p, q);
        if (off < 1e-300) break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = at(p, q);
                if (apq == 0) continue;
                double th = (at(q, q) - at(p, p)) / (2 * apq);
                double t = (th >= 0 ? 1.0 : -1.0) / (std::abs(th) + std::sqrt(th * th + 1));
                double c = 1 / std::sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < n; k++) {
                    double akp = at(k, p), akq = at(k, q);
                    at(k, p) = c * akp - s * akq;
                    at(k, q) = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    double apk = at(p, k), aqk = at(q, k);
                    at(p, k) = c * apk - s * aqk;
                    at(q, k) = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double qkp = Q[(size_t)p * n + k], qkq = Q[(size_t)q * n + k];
                    Q[(size_t)p * n + k] = c * qkp - s * qkq;
                    Q[(size_t)q * n + k] = s * qkp + c * qkq;
                }
            }
    }
    w.resize(n);
    for (int i = 0; i < n; i++) w[i] = at(i, i);
}

}  // namespace dense
