// hostmat.h -- dense complex matrices in host memory and their Hermitian
// eigensystems, for the CPU backend's dense forms of an operator sum (the
// reference's HostSumOps::MatrixType, get_eigensystem and matexp,
// inc/cpuimpl.h:336-420, which use Eigen).
//
// HostMatrix<FpType> is column-major like DevMatrix and offers the same
// operations (sums, scalar and matrix products, adjoint, trace, Frobenius
// norm, element access), so that user code such as
//     Matrix A = syk.fermion_ops(N - 1);
//     v = (rq * A * rq * B).trace();
// reads the same on either backend.
//
// heigh() diagonalizes a Hermitian matrix in FP64: Householder reduction to
// a Hermitian tridiagonal matrix (the reflectors chosen so that each one
// zeroes a column below the subdiagonal), a diagonal phase similarity that
// makes the off-diagonal real and non-negative, and the implicit QL method
// with Wilkinson shifts on the real symmetric tridiagonal matrix (the tql2
// procedure of EISPACK; Bowdler, Martin, Reinsch and Wilkinson, Numer.
// Math. 11 (1968) 293-306), whose rotations are accumulated into the
// eigenvectors.  Eigenvalues come back in ascending order.
#pragma once

#ifndef _REAPERS_H_
#error "Do not include this file directly. Include the master header file reapers.h"
#endif

template<RealScalar FpType>
class HostMatrix {
    size_t rowdim = 0, coldim = 0;
    std::vector<complex<FpType>> a;

    void check_same_shape(const HostMatrix &o) const {
        if (rowdim != o.rowdim || coldim != o.coldim)
            ThrowException(InvalidArgument, "matrix dimensions", "must match");
    }

public:
    HostMatrix() = default;
    // a zero matrix of rows x cols
    explicit HostMatrix(size_t rows, size_t cols = 1)
        : rowdim(rows), coldim(cols), a(rows * cols) {}

    static HostMatrix Identity(size_t rows, size_t cols) {
        if (rows != cols) ThrowException(InvalidArgument, "identity", "must be square");
        HostMatrix m(rows, cols);
        for (size_t i = 0; i < rows; i++) m.a[i * rows + i] = 1;
        return m;
    }
    static HostMatrix Zero(size_t rows, size_t cols) { return HostMatrix(rows, cols); }

    template<RealScalar F1>
    explicit HostMatrix(const HostMatrix<F1> &m) : rowdim(m.rows()), coldim(m.cols()),
                                                    a(m.size()) {
        const complex<F1> *p = m.data();
        for (size_t i = 0; i < a.size(); i++) a[i] = complex<FpType>(p[i]);
    }

    size_t rows() const { return rowdim; }
    size_t cols() const { return coldim; }
    size_t size() const { return rowdim * coldim; }
    size_t bytes() const { return size() * sizeof(complex<FpType>); }
    complex<FpType> *data() { return a.data(); }
    const complex<FpType> *data() const { return a.data(); }

    complex<FpType> &operator()(size_t r, size_t c) { return a[c * rowdim + r]; }
    const complex<FpType> &operator()(size_t r, size_t c) const { return a[c * rowdim + r]; }

    bool operator==(const HostMatrix &o) const {
        return rowdim == o.rowdim && coldim == o.coldim && a == o.a;
    }

    HostMatrix &operator+=(const HostMatrix &o) {
        check_same_shape(o);
        for (size_t i = 0; i < a.size(); i++) a[i] += o.a[i];
        return *this;
    }
    HostMatrix &operator-=(const HostMatrix &o) {
        check_same_shape(o);
        for (size_t i = 0; i < a.size(); i++) a[i] -= o.a[i];
        return *this;
    }
    friend HostMatrix operator+(HostMatrix x, const HostMatrix &y) { return x += y; }
    friend HostMatrix operator-(HostMatrix x, const HostMatrix &y) { return x -= y; }

    HostMatrix &operator*=(complex<FpType> c) {
        for (auto &v : a) v *= c;
        return *this;
    }
    friend HostMatrix operator*(complex<FpType> c, HostMatrix m) { return m *= c; }
    friend HostMatrix operator*(HostMatrix m, complex<FpType> c) { return m *= c; }

    // matrix product, columns of the result in parallel; the inner sums run
    // over k with the left matrix's columns contiguous
    HostMatrix operator*(const HostMatrix &o) const {
        if (coldim != o.rowdim) {
            std::stringstream ss;
            ss << "must match the row dimension of the second matrix (" << coldim
               << " != " << o.rowdim << ")";
            ThrowException(InvalidArgument, "column dimension of first matrix", ss.str().c_str());
        }
        HostMatrix r(rowdim, o.coldim);
        const long nc = (long)o.coldim;
#pragma omp parallel for schedule(static)
        for (long j = 0; j < nc; j++) {
            complex<FpType> *rj = &r.a[(size_t)j * rowdim];
            for (size_t k = 0; k < coldim; k++) {
                const complex<FpType> b = o.a[(size_t)j * o.rowdim + k];
                if (b == complex<FpType>(0)) continue;
                const complex<FpType> *ak = &a[k * rowdim];
                for (size_t i = 0; i < rowdim; i++) rj[i] += ak[i] * b;
            }
        }
        return r;
    }

    HostMatrix adjoint() const {
        HostMatrix r(coldim, rowdim);
        for (size_t j = 0; j < coldim; j++)
            for (size_t i = 0; i < rowdim; i++) r(j, i) = std::conj((*this)(i, j));
        return r;
    }

    complex<FpType> trace() const {
        if (rowdim != coldim)
            ThrowException(InvalidArgument, "column dimension of matrix",
                           "must equal the row dimension");
        std::complex<double> s = 0;
        for (size_t i = 0; i < rowdim; i++) s += std::complex<double>((*this)(i, i));
        return complex<FpType>((FpType)s.real(), (FpType)s.imag());
    }

    FpType mat_norm() const {
        double s = 0;
        for (auto &v : a) s += std::norm(std::complex<double>(v));
        return (FpType)std::sqrt(s);
    }
};

namespace internal {

// Eigensystem of the n x n Hermitian matrix A (column-major, FP64; the
// lower triangle is read): w ascending, V column eigenvectors, A = V w V^H.
inline void heigh(size_t n, std::vector<std::complex<double>> A, std::vector<double> &w,
                  std::vector<std::complex<double>> &V)
{
    using cd = std::complex<double>;
    auto at = [&](size_t i, size_t j) -> cd & { return A[j * n + i]; };
    // make A exactly Hermitian from its lower triangle
    for (size_t j = 0; j < n; j++) {
        at(j, j) = at(j, j).real();
        for (size_t i = j + 1; i < n; i++) at(j, i) = std::conj(at(i, j));
    }
    // Householder reduction: step k zeroes column k below row k + 1 with
    // P = I - 2 v v^H (|v| = 1, v supported on rows k+1..n-1); the trailing
    // block becomes P A P = A - 2 (v u^H + u v^H), u = A v - (v^H A v) v
    std::vector<std::vector<cd>> refl;
    std::vector<cd> offd(n > 0 ? n - 1 : 0);
    std::vector<cd> v(n), u(n);
    for (size_t k = 0; k + 2 < n; k++) {
        double xn2 = 0;
        for (size_t i = k + 1; i < n; i++) xn2 += std::norm(at(i, k));
        const double xn = std::sqrt(xn2);
        cd x0 = at(k + 1, k);
        double tail = xn2 - std::norm(x0);
        if (xn == 0 || tail <= 1e-300 * xn2) {
            offd[k] = x0;
            refl.emplace_back();
            continue;
        }
        const cd ph = std::abs(x0) > 0 ? x0 / std::abs(x0) : cd(1);
        const cd alpha = -ph * xn;
        std::fill(v.begin(), v.end(), cd(0));
        v[k + 1] = x0 - alpha;
        for (size_t i = k + 2; i < n; i++) v[i] = at(i, k);
        double vn = 0;
        for (size_t i = k + 1; i < n; i++) vn += std::norm(v[i]);
        vn = std::sqrt(vn);
        for (size_t i = k + 1; i < n; i++) v[i] /= vn;
        // u = A v on the trailing block (columns k+1..n-1)
        const long lo = (long)k + 1, hi = (long)n;
#pragma omp parallel for schedule(static)
        for (long i = lo; i < hi; i++) {
            cd s = 0;
            for (size_t j = k + 1; j < n; j++) s += at((size_t)i, j) * v[j];
            u[i] = s;
        }
        cd K = 0;
        for (size_t i = k + 1; i < n; i++) K += std::conj(v[i]) * u[i];
        for (size_t i = k + 1; i < n; i++) u[i] -= K.real() * v[i];
#pragma omp parallel for schedule(static)
        for (long j = lo; j < hi; j++) {
            const cd vj = std::conj(v[j]), uj = std::conj(u[j]);
            for (size_t i = k + 1; i < n; i++) at(i, (size_t)j) -= 2.0 * (v[i] * uj + u[i] * vj);
        }
        offd[k] = alpha;
        for (size_t i = k + 2; i < n; i++) at(i, k) = 0;
        refl.emplace_back(v.begin() + k + 1, v.end());
    }
    if (n >= 2) offd[n - 2] = at(n - 1, n - 2);
    // diagonal phases D with D^H T D real: d_0 = 1, d_{k+1} = d_k offd_k / |offd_k|
    std::vector<double> d(n), e(n, 0.0);
    std::vector<cd> phase(n, cd(1));
    for (size_t i = 0; i < n; i++) d[i] = at(i, i).real();
    for (size_t k = 0; k + 1 < n; k++) {
        const double r = std::abs(offd[k]);
        e[k] = r;
        phase[k + 1] = r > 0 ? phase[k] * (offd[k] / r) : phase[k];
    }
    // implicit QL with Wilkinson shifts on (d, e); Z accumulates rotations
    std::vector<double> Z(n * n, 0.0);
    for (size_t i = 0; i < n; i++) Z[i * n + i] = 1;
    const double eps = std::numeric_limits<double>::epsilon();
    for (size_t l = 0; l < n; l++) {
        int iter = 0;
        size_t m;
        while (true) {
            for (m = l; m + 1 < n; m++) {
                const double dd = std::abs(d[m]) + std::abs(d[m + 1]);
                if (std::abs(e[m]) <= eps * dd) break;
            }
            if (m == l) break;
            if (++iter > 200) ThrowException(RuntimeError, "heigh", 200);
            double g = (d[l + 1] - d[l]) / (2 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
            double s = 1, c = 1, p = 0;
            bool deflated = false;
            for (size_t ii = m; ii-- > l;) {
                double f = s * e[ii], b = c * e[ii];
                r = std::hypot(f, g);
                e[ii + 1] = r;
                if (r == 0) {
                    d[ii + 1] -= p;
                    e[m] = 0;
                    deflated = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[ii + 1] - p;
                r = (d[ii] - g) * s + 2 * c * b;
                p = s * r;
                d[ii + 1] = g + p;
                g = c * r - b;
                double *z0 = &Z[ii * n], *z1 = &Z[(ii + 1) * n];
                for (size_t q = 0; q < n; q++) {
                    const double zq1 = z1[q];
                    z1[q] = s * z0[q] + c * zq1;
                    z0[q] = c * z0[q] - s * zq1;
                }
            }
            if (deflated) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0;
        }
    }
    // ascending order
    std::vector<size_t> ord(n);
    for (size_t i = 0; i < n; i++) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](size_t x, size_t y) { return d[x] < d[y]; });
    w.resize(n);
    V.assign(n * n, cd(0));
    // eigenvectors of A: Q D Z, Q = P_0 P_1 ... applied right to left
    const long nl = (long)n;
#pragma omp parallel for schedule(static)
    for (long jj = 0; jj < nl; jj++) {
        const size_t src = ord[(size_t)jj];
        cd *col = &V[(size_t)jj * n];
        for (size_t i = 0; i < n; i++) col[i] = phase[i] * Z[src * n + i];
        for (size_t k = refl.size(); k-- > 0;) {
            const auto &r = refl[k];
            if (r.empty()) continue;
            cd dot = 0;
            for (size_t i = 0; i < r.size(); i++) dot += std::conj(r[i]) * col[k + 1 + i];
            for (size_t i = 0; i < r.size(); i++) col[k + 1 + i] -= 2.0 * r[i] * dot;
        }
    }
    for (size_t i = 0; i < n; i++) w[i] = d[ord[i]];
}

// the byte budget of the CPU backend's matexp caches, shared by all
// operators (CPUImpl's cache_reserve, inc/cpuimpl.h:570-600: 16 GB)
struct HostCacheBudget {
    size_t max = (size_t)16 << 30;
    size_t used = 0;
    static HostCacheBudget &get() {
        static HostCacheBudget b;
        return b;
    }
};

}  // namespace internal
