// ops.h -- SpinOp: one Pauli string with a complex coefficient.
//
// Same encoding and semantics as the reference's SpinOp (inc/ops.h:118-624):
// two bits per site in a u64, the X (flip) bit at position 2s and the Z
// (phase) bit at 2s+1, so id=00, sx=01, sz=10, sy=11 = sx.sz, and the
// coefficient of sigma_y carries its i (inc/ops.h:285-288).  The bit work
// here uses whole-mask arithmetic instead of the reference's per-site loops:
//   product sign  (inc/ops.h:151-161, 319-327): a minus for every site where
//                 the left factor has a Z bit and the right factor an X bit,
//                 i.e. parity(((b0 >> 1) & b1) & 0x5555...);
//   apply         (inc/ops.h:421-440): res = n ^ x, sign parity(res & z).
#pragma once

#ifndef _REAPERS_H_
#error "Do not include this file directly. Include the master header file reapers.h"
#endif

#define NUM_SPIN_SITES (32)
#define MAX_SPIN_INDEX (NUM_SPIN_SITES - 1)
#define NUM_FERMION_FIELDS (2 * NUM_SPIN_SITES)
#define MAX_FIELD_INDEX (NUM_FERMION_FIELDS - 1)

using u8 = std::uint8_t;
using u32 = std::uint32_t;
using u64 = std::uint64_t;

// imaginary-unit literals (inc/ops.h:47-72): 2_i is i*2 in DefFpType,
// _if in float, _id in double
constexpr complex<DefFpType> operator""_i(unsigned long long d)
{
    return complex<DefFpType>{0, static_cast<DefFpType>(d)};
}
constexpr complex<DefFpType> operator""_i(long double d)
{
    return complex<DefFpType>{0, static_cast<DefFpType>(d)};
}
constexpr complex<float> operator""_if(unsigned long long d)
{
    return complex<float>{0, static_cast<float>(d)};
}
constexpr complex<float> operator""_if(long double d)
{
    return complex<float>{0, static_cast<float>(d)};
}
constexpr complex<double> operator""_id(unsigned long long d)
{
    return complex<double>{0, static_cast<double>(d)};
}
constexpr complex<double> operator""_id(long double d)
{
    return complex<double>{0, static_cast<double>(d)};
}

namespace internal {
constexpr u64 kEvenBits = 0x5555555555555555ull;
// gather bits 2s (X) or 2s+1 (Z) of a SpinOp bit string into bit s
inline u32 flip_mask(u64 bits)
{
    u32 r = 0;
    for (u64 b = bits & kEvenBits; b; b &= b - 1) {
        int p = __builtin_ctzll(b);
        r |= 1u << (p >> 1);
    }
    return r;
}
inline u32 phase_mask(u64 bits) { return flip_mask(bits >> 1); }
}  // namespace internal

template<RealScalar FpType>
class HostSumOps;

template<RealScalar FpType = DefFpType>
class SpinOp {
public:
    using IndexType = u8;
    using StateType = u32;
    using ReprType = u64;

private:
    enum class Sigma : u8 { id = 0b00, x = 0b01, y = 0b11, z = 0b10 };

    complex<FpType> coeff;
    ReprType bits;

    template<RealScalar F1> friend class SpinOp;
    template<RealScalar F1> friend class HostSumOps;

    SpinOp(complex<FpType> c, ReprType b) : coeff(c), bits(b) {}

    static void check_spin_index(IndexType n) {
        if (n > MAX_SPIN_INDEX) ThrowException(SpinIndexTooLarge, n, MAX_SPIN_INDEX);
    }

    int count_sy() const {
        return __builtin_popcountll((bits >> 1) & bits & internal::kEvenBits) & 3;
    }

public:
    template<RealScalar F1>
    explicit SpinOp(const SpinOp<F1> &op) : coeff(op.coeff), bits(op.bits) {}

    static SpinOp zero() { return SpinOp(0, 0); }
    static SpinOp identity() { return SpinOp(1, 0); }
    static SpinOp sigma_x(IndexType n) {
        check_spin_index(n);
        return SpinOp(1, ReprType(Sigma::x) << (n * 2));
    }
    static SpinOp sigma_y(IndexType n) {
        check_spin_index(n);
        return SpinOp(complex<FpType>{0, 1}, ReprType(Sigma::y) << (n * 2));
    }
    static SpinOp sigma_z(IndexType n) {
        check_spin_index(n);
        return SpinOp(1, ReprType(Sigma::z) << (n * 2));
    }

    SpinOp &operator*=(complex<FpType> s) { coeff *= s; return *this; }
    SpinOp &operator/=(complex<FpType> s) { coeff *= complex<FpType>{1.0, 0.0} / s; return *this; }
    SpinOp &operator*=(const SpinOp &rhs) { *this = *this * rhs; return *this; }
    friend SpinOp operator*(complex<FpType> s, SpinOp op) { return SpinOp(op.coeff * s, op.bits); }
    SpinOp operator*(complex<FpType> s) const { return SpinOp(coeff * s, bits); }

    // right-multiply by op1; a minus for each site (left Z bit, right X bit)
    SpinOp operator*(const SpinOp &op1) const {
        bool minus = __builtin_popcountll((bits >> 1) & op1.bits & internal::kEvenBits) & 1;
        return SpinOp((minus ? -coeff : coeff) * op1.coeff, bits ^ op1.bits);
    }

    SpinOp operator<<(IndexType n) const { return SpinOp(coeff, bits << (n * 2)); }

    bool operator==(const SpinOp &o) const { return bits == o.bits && coeff == o.coeff; }
    template<ScalarType S>
    bool operator==(S s) const { return is_scalar() && (coeff == complex<FpType>(s)); }

    ReprType get_bit_string() const { return bits; }
    complex<FpType> get_coeff() const { return coeff; }

    // coefficient in the standard-sigma_y convention (inc/ops.h:366-377)
    complex<FpType> coefficient() const {
        switch (count_sy()) {
        case 0: return coeff;
        case 1: return complex<FpType>{0, -1} * coeff;
        case 2: return -coeff;
        default: return complex<FpType>{0, 1} * coeff;
        }
    }

    bool is_hermitian() const { return coefficient().imag() == 0; }
    bool is_scalar() const { return !bits; }

    // res = n with the X bits flipped; returns +/- coeff (inc/ops.h:421-440)
    complex<FpType> apply(StateType n, StateType &res) const {
        res = n ^ internal::flip_mask(bits);
        bool minus = __builtin_popcount(res & internal::phase_mask(bits)) & 1;
        return minus ? -coeff : coeff;
    }

    // the site operator at site n, as 0..3 (id, sx, sz, sy codes)
    int site(IndexType n) const {
        check_spin_index(n);
        return (int)((bits >> (2 * n)) & 3u);
    }

    friend std::ostream &operator<<(std::ostream &os, const SpinOp &op) {
        static const char *names[4] = {"id", "sx", "sz", "sy"};
        if (op.coeff == complex<FpType>(0)) { os << "0"; return os; }
        complex<FpType> c = op.coefficient();
        if (c != complex<FpType>(1) && c != complex<FpType>(-1)) {
            if (c == complex<FpType>(0, 1)) os << "i";
            else if (c == complex<FpType>(0, -1)) os << "-i";
            else if (c.real() * c.imag() != 0) os << "(" << c.real() << "+" << c.imag() << "i" << ")";
            else if (c.imag() == 0) os << c.real();
            else os << c.imag() << "i";
            os << "*";
        } else if (c == complex<FpType>(-1)) {
            os << "-";
        }
        if (op.is_scalar()) { os << "Id"; return os; }
        bool dot = false;
        for (int s = 0; s < NUM_SPIN_SITES; s++) {
            int code = (int)((op.bits >> (2 * s)) & 3u);
            if (!code) continue;
            if (dot) os << ".";
            os << names[code] << "(" << s << ")";
            dot = true;
        }
        return os;
    }
};
