// sumops.h -- HostSumOps: an ordered sum of Pauli strings held on the host.
//
// Semantics as the reference's HostSumOps (inc/cpuimpl.h:98-507):
//   +=  merges a string already present (linear search on the bits), erases a
//       term whose coefficient becomes exactly zero, otherwise appends;
//   *   forms all products op0*op1 in (left, right) order and merges them;
//   a scalar product multiplies every coefficient (merging again);
//   only a scalar zero (bits 0, coefficient 0) is ever skipped on insert.
// Term order is insertion order, which fixes the order of the coefficients a
// generator produces.  The device copy is a derived class (DevSumOps in
// gpuimpl.h) that overrides release() to drop its uploaded layout.
#pragma once

#ifndef _REAPERS_H_
#error "Do not include this file directly. Include the master header file reapers.h"
#endif

template<RealScalar FpType = DefFpType>
class HostSumOps {
    template<RealScalar F1> friend class HostSumOps;

public:
    using RealScalarType = FpType;
    using ComplexScalarType = complex<FpType>;

protected:
    std::vector<SpinOp<FpType>> ops;

    // the dense forms of the CPU backend, built on request
    struct HostDense;
    mutable std::unique_ptr<HostDense> hdense;

    // called on every mutation: drops the cached dense forms; derived
    // classes drop their device state too
    virtual void release() const { hdense.reset(); }

public:
    HostSumOps(const SpinOp<FpType> &op) : ops{op} {}
    HostSumOps(const HostSumOps &op) : ops{op.ops} {}
    HostSumOps(HostSumOps &&op) : ops(std::move(op.ops)) { op.release(); }

    template<RealScalar F1>
    explicit HostSumOps(const HostSumOps<F1> &op) {
        for (auto o : op) {
            if (!(o == 0)) ops.emplace_back(o);
        }
    }

    HostSumOps(const std::initializer_list<SpinOp<FpType>> &l = {}) {
        for (auto &&o : l) {
            if (!(o == 0)) *this += o;
        }
    }

    virtual ~HostSumOps() {}

    HostSumOps &operator=(const HostSumOps &rhs) {
        if (this == &rhs) return *this;
        release();
        ops = rhs.ops;
        return *this;
    }

    HostSumOps &operator=(HostSumOps &&rhs) {
        if (this == &rhs) return *this;
        release();
        ops = std::move(rhs.ops);
        rhs.release();
        return *this;
    }

    template<RealScalar F1>
    HostSumOps &operator=(const HostSumOps<F1> &rhs) {
        release();
        ops.clear();
        for (auto op : rhs) {
            if (!(op == 0)) ops.emplace_back(SpinOp<FpType>(op));
        }
        return *this;
    }

    HostSumOps &operator=(const SpinOp<FpType> &rhs) {
        release();
        ops.clear();
        if (!(rhs == 0)) ops.emplace_back(rhs);
        return *this;
    }

    bool operator==(const HostSumOps &rhs) const {
        HostSumOps res{*this - rhs};
        return res.ops.empty();
    }

    template<ScalarType S>
    bool operator==(S s) const {
        if (s == S{}) return ops.empty();
        return ops.size() == 1 && ops[0] == s;
    }

    auto begin() const { return std::cbegin(ops); }
    auto end() const { return std::cend(ops); }
    auto empty() const { return ops.empty(); }
    auto size() const { return ops.size(); }

    // merge-on-insert (inc/cpuimpl.h:233-253)
    HostSumOps &operator+=(const SpinOp<FpType> &op) {
        release();
        if (op == 0) return *this;
        for (auto it = ops.begin(); it != ops.end(); it++) {
            if (it->bits == op.bits) {
                it->coeff += op.coeff;
                if (it->coeff == complex<FpType>(0)) ops.erase(it);
                return *this;
            }
        }
        ops.push_back(op);
        return *this;
    }

    HostSumOps &operator-=(SpinOp<FpType> op) {
        release();
        op *= -1;
        return *this += op;
    }

    HostSumOps &operator+=(const HostSumOps &rhs) {
        release();
        for (auto &&o : rhs) *this += o;
        return *this;
    }

    HostSumOps &operator-=(HostSumOps rhs) {
        release();
        rhs *= -1;
        return *this += rhs;
    }

    HostSumOps operator+(const SpinOp<FpType> &op) const { HostSumOps r(*this); r += op; return r; }
    HostSumOps operator+(const HostSumOps &op) const { HostSumOps r(*this); r += op; return r; }
    HostSumOps operator-(const SpinOp<FpType> &op) const { HostSumOps r(*this); r -= op; return r; }
    HostSumOps operator-(const HostSumOps &op) const { HostSumOps r(*this); r -= op; return r; }

    HostSumOps &operator*=(const HostSumOps &rhs) {
        release();
        *this = operator*(rhs);
        return *this;
    }

    HostSumOps &operator*=(const complex<FpType> &rhs) {
        release();
        *this = operator*(rhs);
        return *this;
    }

    HostSumOps &operator/=(const complex<FpType> &s) {
        release();
        *this = operator*(complex<FpType>{1.0, 0.0} / s);
        return *this;
    }

    // all products, left factor outer (inc/cpuimpl.h:325-333)
    HostSumOps operator*(const HostSumOps &rhs) const {
        HostSumOps res;
        for (auto &&op0 : *this)
            for (auto &&op1 : rhs) res += op0 * op1;
        return res;
    }

    HostSumOps operator*(complex<FpType> rhs) const {
        HostSumOps res;
        for (auto &&op0 : *this) res += rhs * op0;
        return res;
    }

    friend HostSumOps operator*(complex<FpType> s, const HostSumOps &o) { return o * s; }

    HostSumOps operator<<(typename SpinOp<FpType>::IndexType len) const {
        HostSumOps res;
        for (auto &&op : *this) res += op << len;
        return res;
    }

    bool is_hermitian() const {
        for (auto &op : *this)
            if (!op.is_hermitian()) return false;
        return true;
    }

    // the term list as flat arrays: SpinOp bits and (re, im) coefficients,
    // the form the C ABI's rp_terms_upload takes
    void export_terms(std::vector<u64> &bits, std::vector<double> &coeff) const {
        bits.resize(ops.size());
        coeff.resize(2 * ops.size());
        for (size_t i = 0; i < ops.size(); i++) {
            bits[i] = ops[i].bits;
            coeff[2 * i] = (double)ops[i].coeff.real();
            coeff[2 * i + 1] = (double)ops[i].coeff.imag();
        }
    }

    // build a term list from flat arrays without merging (the arrays are
    // taken to hold distinct strings, e.g. a list exported above)
    static HostSumOps from_terms(const u64 *bits, const double *coeff, size_t n) {
        HostSumOps res;
        res.ops.reserve(n);
        for (size_t i = 0; i < n; i++)
            res.ops.push_back(SpinOp<FpType>(
                complex<FpType>{(FpType)coeff[2 * i], (FpType)coeff[2 * i + 1]}, bits[i]));
        return res;
    }

    // Dense forms on len sites for the CPU backend (get_matrix,
    // get_eigensystem, matexp: inc/cpuimpl.h:336-420), cached until the
    // term list changes.  Matrix element <i|H|i ^ x> = sum over the terms
    // with flip mask x of coeff (-1)^popcount((i ^ x) & z) (SpinOp::apply).
    using MatrixType = HostMatrix<FpType>;
    using EigenVals = std::vector<FpType>;
    using EigenVecs = HostMatrix<FpType>;
    using EigenSystem = std::tuple<const EigenVals &, const EigenVecs &>;
    const MatrixType &get_matrix(int len) const;
    EigenSystem get_eigensystem(int len) const;
    // exp(c H) = V diag(exp(c w)) V^H, kept per c under the CPU backend's
    // cache budget, least-used entries dropped first (inc/cpuimpl.h:29-88)
    const MatrixType &matexp(complex<FpType> c, int len) const;
    size_t matexp_cache_entries() const;

    friend std::ostream &operator<<(std::ostream &os, const HostSumOps &s) {
        bool first = true;
        if (s.ops.empty()) os << "0";
        for (auto &&op : s) {
            std::stringstream ss;
            ss << op;
            std::string str(ss.str());
            if (str == "0") continue;
            if (!first) {
                if (str[0] == '-') {
                    str.insert(str.begin() + 1, ' ');
                    str.insert(str.begin(), ' ');
                } else {
                    str.insert(0, " + ");
                }
            }
            os << str;
            first = false;
        }
        return os;
    }
};

template<RealScalar FpType>
struct HostSumOps<FpType>::HostDense {
    int len = -1;
    std::unique_ptr<HostMatrix<FpType>> mat, vecs;
    std::vector<FpType> vals;
    std::vector<double> vals64;
    std::vector<std::complex<double>> vecs64;
    struct Entry {
        HostMatrix<FpType> m;
        u64 hits;
    };
    std::map<std::pair<FpType, FpType>, Entry> expm;
    ~HostDense() {
        auto &b = internal::HostCacheBudget::get();
        for (auto &e : expm) b.used -= std::min(b.used, e.second.m.bytes());
    }
};

template<RealScalar FpType>
const HostMatrix<FpType> &HostSumOps<FpType>::get_matrix(int len) const {
    if (len < 1 || len > 15) ThrowException(InvalidArgument, "len", "must be 1..15 for dense forms");
    if (!hdense || hdense->len != len) {
        hdense = std::make_unique<HostDense>();
        hdense->len = len;
    }
    if (!hdense->mat) {
        const size_t n = (size_t)1 << len;
        auto m = std::make_unique<HostMatrix<FpType>>(n, n);
        for (auto &op : ops) {
            const u64 b = op.get_bit_string();
            const u32 x = internal::flip_mask(b), z = internal::phase_mask(b);
            const complex<FpType> c = op.get_coeff();
            for (size_t i = 0; i < n; i++) {
                const size_t col = i ^ x;
                (*m)(i, col) += (__builtin_popcountll(col & z) & 1) ? -c : c;
            }
        }
        hdense->mat = std::move(m);
    }
    return *hdense->mat;
}

template<RealScalar FpType>
typename HostSumOps<FpType>::EigenSystem HostSumOps<FpType>::get_eigensystem(int len) const {
    const auto &m = get_matrix(len);
    auto &d = *hdense;
    if (!d.vecs) {
        const size_t n = m.rows();
        std::vector<std::complex<double>> a(n * n);
        for (size_t i = 0; i < n * n; i++) a[i] = std::complex<double>(m.data()[i]);
        internal::heigh(n, std::move(a), d.vals64, d.vecs64);
        d.vals.resize(n);
        for (size_t i = 0; i < n; i++) d.vals[i] = (FpType)d.vals64[i];
        auto v = std::make_unique<HostMatrix<FpType>>(n, n);
        for (size_t i = 0; i < n * n; i++)
            v->data()[i] = complex<FpType>((FpType)d.vecs64[i].real(), (FpType)d.vecs64[i].imag());
        d.vecs = std::move(v);
    }
    return {d.vals, *d.vecs};
}

template<RealScalar FpType>
const HostMatrix<FpType> &HostSumOps<FpType>::matexp(complex<FpType> c, int len) const {
    get_eigensystem(len);
    auto &d = *hdense;
    auto key = std::make_pair(c.real(), c.imag());
    if (auto it = d.expm.find(key); it != d.expm.end()) {
        it->second.hits++;
        return it->second.m;
    }
    const size_t n = (size_t)1 << len;
    // V diag(e) V^H in FP64 from the FP64 eigensystem
    std::vector<std::complex<double>> e(n);
    const std::complex<double> cc(c.real(), c.imag());
    for (size_t j = 0; j < n; j++) e[j] = std::exp(cc * d.vals64[j]);
    HostMatrix<FpType> r(n, n);
    const long nl = (long)n;
#pragma omp parallel for schedule(static)
    for (long col = 0; col < nl; col++) {
        std::vector<std::complex<double>> acc(n);
        for (size_t j = 0; j < n; j++) {
            const std::complex<double> f = e[j] * std::conj(d.vecs64[j * n + (size_t)col]);
            const std::complex<double> *vj = &d.vecs64[j * n];
            for (size_t i = 0; i < n; i++) acc[i] += vj[i] * f;
        }
        for (size_t i = 0; i < n; i++)
            r((size_t)i, (size_t)col) = complex<FpType>((FpType)acc[i].real(), (FpType)acc[i].imag());
    }
    auto &b = internal::HostCacheBudget::get();
    const size_t sz = r.bytes();
    while (b.used + sz > b.max && !d.expm.empty()) {
        auto victim = d.expm.begin();
        for (auto it = d.expm.begin(); it != d.expm.end(); ++it)
            if (it->second.hits < victim->second.hits) victim = it;
        b.used -= std::min(b.used, victim->second.m.bytes());
        d.expm.erase(victim);
    }
    b.used += sz;
    auto it = d.expm.insert_or_assign(key, typename HostDense::Entry{std::move(r), 0}).first;
    return it->second.m;
}

template<RealScalar FpType>
size_t HostSumOps<FpType>::matexp_cache_entries() const {
    return hdense ? hdense->expm.size() : 0;
}
