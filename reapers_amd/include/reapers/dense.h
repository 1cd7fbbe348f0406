// dense.h -- device matrices and the exact-diagonalization path.
//
// DevMatrix<FpType> is the reference's DevSumOps::MatrixType
// (inc/gpuimpl.h:139-300): a column-major complex matrix in device memory
// with sums, scalar and matrix products, trace and Frobenius norm.
// DevSumOps gains the reference's get_matrix(len), get_eigensystem(len) and
// matexp(c, len) (inc/gpuimpl.h:370-388, 1179-1263): the dense form is built
// on the device from the term list, the eigensystem comes from rocSOLVER's
// heevd, and exp(c H) = V diag(exp(c w)) V^H is formed by one column-scaling
// kernel and one GEMM.  Results of matexp are kept in a cache per operator
// keyed by c, under the byte budget GPUImpl::cache_reserve keeps for all
// operators (half of the free device memory by default,
// GPUImpl::set_max_cache_size changes it; inc/gpuimpl.h:560-597); when an
// insertion exceeds it, the least-used entries of that cache are dropped
// (inc/cpuimpl.h:29-88).  Every cached form is dropped when the term list
// changes.
//
// This is the small-N path (n = 2^len <= 2^15): it serves as an on-device
// check of the Krylov path and computes trace-definition correlators.
#pragma once

#ifndef _REAPERS_H_
#error "Do not include this file directly. Include the master header file reapers.h"
#endif

template<RealScalar FpType>
class DevMatrix {
    void *ptr = nullptr;
    size_t rowdim = 0, coldim = 0;

    static constexpr int prec() { return internal::rp_prec<FpType>(); }

    void alloc(size_t r, size_t c) {
        rowdim = r;
        coldim = c;
        if (!r || !c) return;
        GPUImpl::ensure_context();
        int st = rp_malloc(&ptr, r * c * sizeof(complex<FpType>));
        if (st != RP_OK)
            internal::throw_status(__FILE__, __LINE__, __func__, st, r * c * sizeof(complex<FpType>));
    }

    void free_mem() {
        if (ptr) rp_free(ptr);
        ptr = nullptr;
        rowdim = coldim = 0;
    }

    void check_same_shape(const DevMatrix &o) const {
        if (rowdim != o.rowdim || coldim != o.coldim)
            ThrowException(InvalidArgument, "matrix dimensions", "must match");
    }

public:
    DevMatrix() = default;
    // a zero matrix of rows x cols
    explicit DevMatrix(size_t rows, size_t cols = 1) {
        alloc(rows, cols);
        if (ptr) RP_CALL(rp_zero_vec(prec(), ptr, rows * cols));
    }
    DevMatrix(const DevMatrix &m) {
        alloc(m.rowdim, m.coldim);
        if (ptr) RP_CALL(rp_memcpy(ptr, m.ptr, bytes(), RP_D2D));
    }
    DevMatrix(DevMatrix &&m) : ptr(m.ptr), rowdim(m.rowdim), coldim(m.coldim) {
        m.ptr = nullptr;
        m.rowdim = m.coldim = 0;
    }
    DevMatrix &operator=(const DevMatrix &m) {
        if (this == &m) return *this;
        if (rowdim * coldim != m.rowdim * m.coldim) {
            free_mem();
            alloc(m.rowdim, m.coldim);
        }
        rowdim = m.rowdim;
        coldim = m.coldim;
        if (ptr) RP_CALL(rp_memcpy(ptr, m.ptr, bytes(), RP_D2D));
        return *this;
    }
    DevMatrix &operator=(DevMatrix &&m) {
        if (this == &m) return *this;
        free_mem();
        std::swap(ptr, m.ptr);
        std::swap(rowdim, m.rowdim);
        std::swap(coldim, m.coldim);
        return *this;
    }
    ~DevMatrix() { free_mem(); }

    static DevMatrix Identity(size_t rows, size_t cols) {
        if (rows != cols) ThrowException(InvalidArgument, "identity", "must be square");
        DevMatrix m;
        m.alloc(rows, cols);
        RP_CALL(rp_dense_eye(prec(), rows, m.ptr));
        return m;
    }

    // a host matrix given column-major
    static DevMatrix from_host(size_t rows, size_t cols, const complex<FpType> *host) {
        DevMatrix m;
        m.alloc(rows, cols);
        RP_CALL(rp_memcpy(m.ptr, host, m.bytes(), RP_H2D));
        return m;
    }

    size_t rows() const { return rowdim; }
    size_t cols() const { return coldim; }
    size_t size() const { return rowdim * coldim; }
    size_t bytes() const { return size() * sizeof(complex<FpType>); }
    void *data() { return ptr; }
    const void *data() const { return ptr; }

    void download(complex<FpType> *host) const {
        if (ptr) RP_CALL(rp_memcpy(host, ptr, bytes(), RP_D2H));
    }

    std::vector<complex<FpType>> to_host() const {
        std::vector<complex<FpType>> v(size());
        download(v.data());
        return v;
    }

    complex<FpType> operator()(size_t r, size_t c) const {
        if (r >= rowdim || c >= coldim) ThrowException(InvalidArgument, "index", "out of range");
        complex<FpType> v;
        RP_CALL(rp_memcpy(&v, (const complex<FpType> *)ptr + (r + c * rowdim), sizeof(v), RP_D2H));
        return v;
    }

    bool operator==(const DevMatrix &o) const {
        if (rowdim != o.rowdim || coldim != o.coldim) return false;
        return to_host() == o.to_host();
    }

    DevMatrix &operator+=(const DevMatrix &o) {
        check_same_shape(o);
        if (ptr) RP_CALL(rp_add_vec(prec(), ptr, o.ptr, size()));
        return *this;
    }
    DevMatrix &operator-=(const DevMatrix &o) {
        check_same_shape(o);
        if (ptr) RP_CALL(rp_axpy(prec(), ptr, -1.0, 0.0, o.ptr, size()));
        return *this;
    }
    friend DevMatrix operator+(const DevMatrix &a, const DevMatrix &b) {
        DevMatrix r(a);
        r += b;
        return r;
    }
    friend DevMatrix operator-(const DevMatrix &a, const DevMatrix &b) {
        DevMatrix r(a);
        r -= b;
        return r;
    }

    DevMatrix &operator*=(complex<FpType> c) {
        if (ptr) RP_CALL(rp_scale_vec(prec(), ptr, (double)c.real(), (double)c.imag(), size()));
        return *this;
    }
    friend DevMatrix operator*(complex<FpType> c, const DevMatrix &m) {
        DevMatrix r(m);
        r *= c;
        return r;
    }
    friend DevMatrix operator*(const DevMatrix &m, complex<FpType> c) { return c * m; }

    DevMatrix operator*(const DevMatrix &o) const {
        if (coldim != o.rowdim) {
            std::stringstream ss;
            ss << "must match the row dimension of the second matrix (" << coldim
               << " != " << o.rowdim << ")";
            ThrowException(InvalidArgument, "column dimension of first matrix", ss.str().c_str());
        }
        DevMatrix r;
        r.alloc(rowdim, o.coldim);
        RP_CALL(rp_dense_gemm(prec(), rowdim, o.coldim, coldim, 0, 0, 1.0, 0.0, ptr, rowdim,
                              o.ptr, o.rowdim, 0.0, 0.0, r.ptr, rowdim));
        return r;
    }

    // the conjugate transpose
    DevMatrix adjoint() const {
        DevMatrix id = Identity(rowdim, rowdim);
        DevMatrix r;
        r.alloc(coldim, rowdim);
        RP_CALL(rp_dense_gemm(prec(), coldim, rowdim, rowdim, 1, 0, 1.0, 0.0, ptr, rowdim,
                              id.ptr, rowdim, 0.0, 0.0, r.ptr, coldim));
        return r;
    }

    complex<FpType> trace() const {
        if (rowdim != coldim)
            ThrowException(InvalidArgument, "column dimension of matrix", "must equal the row dimension");
        double re, im;
        RP_CALL(rp_dense_trace(prec(), rowdim, ptr, &re, &im));
        return {(FpType)re, (FpType)im};
    }

    FpType mat_norm() const {
        if (!ptr) return 0;
        double n;
        RP_CALL(rp_vec_norm(prec(), ptr, size(), &n));
        return (FpType)n;
    }
};

namespace internal {

// matexp results of one operator keyed by c, with their use counts
template<RealScalar FpType>
class MatexpCache {
    struct Entry {
        DevMatrix<FpType> m;
        u64 hits;
    };
    std::map<std::pair<FpType, FpType>, Entry> cache;

    void drop(typename std::map<std::pair<FpType, FpType>, Entry>::iterator it) {
        GPUImpl::cache_release(it->second.m.bytes());
        cache.erase(it);
    }

public:
    MatexpCache() = default;
    MatexpCache(const MatexpCache &) = delete;
    MatexpCache &operator=(const MatexpCache &) = delete;
    ~MatexpCache() { clear(); }

    const DevMatrix<FpType> *find(complex<FpType> c) {
        auto it = cache.find({c.real(), c.imag()});
        if (it == cache.end()) return nullptr;
        it->second.hits++;
        return &it->second.m;
    }

    // the least-used entries are dropped until the result fits the budget
    // (or the cache is empty: a result larger than the budget is kept alone)
    const DevMatrix<FpType> &insert(complex<FpType> c, DevMatrix<FpType> &&m) {
        size_t sz = m.bytes();
        while (!GPUImpl::cache_reserve(sz)) {
            if (cache.empty()) {
                GPUImpl::set_max_cache_size(std::max(GPUImpl::get_max_cache_size(),
                                                     GPUImpl::get_current_cache_size() + sz));
                continue;
            }
            auto victim = cache.begin();
            for (auto it = cache.begin(); it != cache.end(); ++it)
                if (it->second.hits < victim->second.hits) victim = it;
            drop(victim);
        }
        auto it = cache.insert_or_assign({c.real(), c.imag()}, Entry{std::move(m), 0}).first;
        return it->second.m;
    }

    size_t entries() const { return cache.size(); }

    void clear() {
        while (!cache.empty()) drop(cache.begin());
    }
};

// the dense forms of one operator at one chain length
template<RealScalar FpType>
struct DenseCache {
    int len = -1;
    std::unique_ptr<DevMatrix<FpType>> mat;
    std::unique_ptr<DevMatrix<FpType>> vecs;   // column eigenvectors
    std::vector<FpType> vals;                  // eigenvalues, ascending
    DevMatrix<FpType> dvals;                   // the same on the device (packed reals)
    MatexpCache<FpType> expm;
};

}  // namespace internal

template<RealScalar FpType>
internal::DenseCache<FpType> &DevSumOps<FpType>::dense_at(int len) const {
    if (len < 1 || len > 15) ThrowException(InvalidArgument, "len", "must be 1..15 for dense forms");
    if (!dense || dense->len != len) {
        dense = std::make_unique<internal::DenseCache<FpType>>();
        dense->len = len;
    }
    return *dense;
}

template<RealScalar FpType>
const DevMatrix<FpType> &DevSumOps<FpType>::get_matrix(int len) const {
    auto &d = dense_at(len);
    if (!d.mat) {
        size_t n = (size_t)1 << len;
        auto m = std::make_unique<DevMatrix<FpType>>(n, n);
        std::vector<u64> bits;
        std::vector<double> coeff;
        this->export_terms(bits, coeff);
        RP_CALL(rp_dense_build(internal::rp_prec<FpType>(), len, bits.data(), coeff.data(),
                               bits.size(), m->data()));
        d.mat = std::move(m);
    }
    return *d.mat;
}

template<RealScalar FpType>
typename DevSumOps<FpType>::EigenSystem DevSumOps<FpType>::get_eigensystem(int len) const {
    auto &d = dense_at(len);
    if (!d.vecs) {
        size_t n = (size_t)1 << len;
        auto v = std::make_unique<DevMatrix<FpType>>(get_matrix(len));
        // n reals fit in the first half of an n x 1 complex buffer
        DevMatrix<FpType> w(n, 1);
        RP_CALL(rp_dense_eigh(internal::rp_prec<FpType>(), n, v->data(), w.data()));
        d.vals.resize(n);
        RP_CALL(rp_memcpy(d.vals.data(), w.data(), n * sizeof(FpType), RP_D2H));
        d.vecs = std::move(v);
        d.dvals = std::move(w);
    }
    return {d.vals, *d.vecs};
}

// exp(c H) from the eigensystem (_matexp, inc/gpuimpl.h:1235-1252)
template<RealScalar FpType>
DevMatrix<FpType> DevSumOps<FpType>::_matexp(complex<FpType> c, int len) const {
    get_eigensystem(len);
    auto &d = *dense;
    size_t n = (size_t)1 << len;
    DevMatrix<FpType> res(n, n), work(n, n);
    RP_CALL(rp_dense_expm_eig(internal::rp_prec<FpType>(), n, d.vecs->data(), d.dvals.data(),
                              (double)c.real(), (double)c.imag(), res.data(), work.data()));
    return res;
}

template<RealScalar FpType>
const DevMatrix<FpType> &DevSumOps<FpType>::matexp(complex<FpType> c, int len) const {
    auto &d = dense_at(len);
    if (auto *m = d.expm.find(c)) return *m;
    DevMatrix<FpType> r = _matexp(c, len);
    return dense_at(len).expm.insert(c, std::move(r));
}

template<RealScalar FpType>
size_t DevSumOps<FpType>::matexp_cache_entries() const {
    return dense ? dense->expm.entries() : 0;
}
