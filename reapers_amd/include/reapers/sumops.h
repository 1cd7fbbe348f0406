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

    // called on every mutation; derived classes drop cached device state
    virtual void release() const {}

public:
    HostSumOps(const SpinOp<FpType> &op) : ops{op} {}
    HostSumOps(const HostSumOps &op) : ops{op.ops} {}
    HostSumOps(HostSumOps &&op) = default;

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
