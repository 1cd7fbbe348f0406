// blkops.h -- operators and vectors in parity-block form.
//
// A parity-conserving problem on 2^(L+1) amplitudes is kept as two halves of
// 2^L amplitudes each, the "L" and the "R" parity sector, and an operator as
// a 2x2 matrix of blocks acting on them:
//
//     BlockDiag      ( LL  0  )    products of an even number of Majoranas,
//                    ( 0   RR )    Hamiltonians
//     BlockAntiDiag  ( 0   LR )    a single Majorana
//                    ( RL  0  )
//     BlockOp        ( LL  LR )    anything else
//                    ( RL  RR )
//     BlockVec       ( L )         a vector; BlockState is a BlockVec of States
//                    ( R )
//
// Each block carries a null flag: a null block is the zero block, holds a
// default-constructed T and takes no part in sums or products.  Sums build
// their result's flags from the blocks' sizes, products from the operands'
// flags, and scaling by zero nulls every block -- the reference's rules
// (inc/blkops.h:90-457 for the types, 465-728 for the arithmetic).  The
// member names LL, LR, RL, RR, L, R and nullXX are the reference's, so that
// code written against it (ham.LL, ops[n].LR, st.nullR) reads the same.
//
// A block type T needs size() (zero when empty), +=, *= (scalar) and the
// products used; T is a SubspaceView<SumOps> for operators and a State for
// vectors.  get_matrix(), matexp() and trace() are forwarded to the blocks
// when T has them.
#pragma once

#ifndef _REAPERS_H_
#error "Do not include this file directly. Include the master header file reapers.h"
#endif

template<typename T> struct BlockDiag;
template<typename T> struct BlockAntiDiag;
template<typename T> struct BlockOp;
template<typename T> struct BlockVec;

namespace blk {
// overload probes: each accepts its template and every class derived from
// one of its specializations (FermionOpBlockForm derives from BlockAntiDiag)
template<typename T> void probe_diag(const BlockDiag<T> &);
template<typename T> void probe_anti(const BlockAntiDiag<T> &);
template<typename T> void probe_op(const BlockOp<T> &);
template<typename T> void probe_vec(const BlockVec<T> &);

template<typename A, typename B>
using prod_t = std::remove_cvref_t<decltype(std::declval<const A &>() * std::declval<const B &>())>;

// the product of two blocks, or an empty block when either is null
template<typename A, typename B>
prod_t<A, B> mul(bool null_a, const A &a, bool null_b, const B &b) {
    if (null_a || null_b) return {};
    return a * b;
}

// res += term, where both may be null; res stays null only if both are
template<typename T>
void accumulate(bool &null_res, T &res, bool null_term, T &&term) {
    if (null_term) return;
    if (null_res) res = std::move(term);
    else res += term;
    null_res = false;
}

// a += b for one pair of blocks, copying or moving b into a null a
template<typename T, typename U>
void add_block(bool &null_a, T &a, bool null_b, U &&b) {
    if (null_b) return;
    if (null_a) a = std::forward<U>(b);
    else a += b;
    null_a = false;
}

// a = b for one block, resetting a to empty when b is null
template<typename T, typename U>
void assign_block(bool &null_a, T &a, bool null_b, U &&b) {
    if (null_b) a = T{};
    else a = std::forward<U>(b);
    null_a = null_b;
}

template<typename T, typename S>
void scale_block(bool null_a, T &a, S c) {
    if (!null_a) a *= c;
}

template<typename T>
concept HasMatrix = requires(const T &t) { t.get_matrix(); };
}  // namespace blk

template<typename B>
concept BlockDiagType = requires(const std::remove_cvref_t<B> &b) { blk::probe_diag(b); };
template<typename B>
concept BlockAntiDiagType = requires(const std::remove_cvref_t<B> &b) { blk::probe_anti(b); };
template<typename B>
concept BlockOpType = requires(const std::remove_cvref_t<B> &b) { blk::probe_op(b); };
template<typename B>
concept BlockVecType = requires(const std::remove_cvref_t<B> &b) { blk::probe_vec(b); };
template<typename B>
concept BlockMatrix = BlockDiagType<B> || BlockAntiDiagType<B>;
template<typename B>
concept BlockForm = BlockMatrix<B> || BlockVecType<B>;

template<typename T>
struct BlockDiag {
    bool nullLL, nullRR;
    T LL, RR;

    BlockDiag() : nullLL(true), nullRR(true), LL{}, RR{} {}
    BlockDiag(T B) : nullLL(!B.size()), nullRR(!B.size()), LL(B), RR(B) {}
    BlockDiag(T LL, T RR) : nullLL(!LL.size()), nullRR(!RR.size()), LL(std::move(LL)),
                            RR(std::move(RR)) {}
    BlockDiag(const BlockDiag &) = default;
    BlockDiag(BlockDiag &&) = default;
    // blocks of another type, e.g. another precision (ham64 = ham32 in the
    // production driver, examples/syk.cc:457-462)
    template<typename T1>
        requires(!std::is_same_v<T1, T>)
    BlockDiag(const BlockDiag<T1> &b)
        : nullLL(b.nullLL), nullRR(b.nullRR), LL(b.nullLL ? T{} : T(b.LL)),
          RR(b.nullRR ? T{} : T(b.RR)) {}
    template<typename T1>
        requires(!std::is_same_v<T1, T>)
    BlockDiag &operator=(const BlockDiag<T1> &b) {
        blk::assign_block(nullLL, LL, b.nullLL, b.LL);
        blk::assign_block(nullRR, RR, b.nullRR, b.RR);
        return *this;
    }
    BlockDiag &operator=(const BlockDiag &b) {
        blk::assign_block(nullLL, LL, b.nullLL, b.LL);
        blk::assign_block(nullRR, RR, b.nullRR, b.RR);
        return *this;
    }
    BlockDiag &operator=(BlockDiag &&b) {
        blk::assign_block(nullLL, LL, b.nullLL, std::move(b.LL));
        blk::assign_block(nullRR, RR, b.nullRR, std::move(b.RR));
        return *this;
    }
    bool operator==(const BlockDiag &) const = default;

    BlockDiag &operator+=(const BlockDiag &rhs) {
        blk::add_block(nullLL, LL, rhs.nullLL, rhs.LL);
        blk::add_block(nullRR, RR, rhs.nullRR, rhs.RR);
        return *this;
    }

    template<ScalarType S>
    BlockDiag &operator*=(S c) {
        if (c == S{}) {
            *this = BlockDiag{};
            return *this;
        }
        blk::scale_block(nullLL, LL, c);
        blk::scale_block(nullRR, RR, c);
        return *this;
    }

    template<typename... Args>
    auto get_matrix(Args &&...args) const {
        using M = std::remove_cvref_t<decltype(LL.get_matrix(args...))>;
        BlockDiag<M> res(nullLL ? M{} : LL.get_matrix(args...),
                         nullRR ? M{} : RR.get_matrix(args...));
        res.nullLL = nullLL;
        res.nullRR = nullRR;
        return res;
    }

    template<typename... Args>
    auto matexp(Args &&...args) const {
        using M = std::remove_cvref_t<decltype(LL.matexp(args...))>;
        BlockDiag<M> res(nullLL ? M{} : LL.matexp(args...), nullRR ? M{} : RR.matexp(args...));
        res.nullLL = nullLL;
        res.nullRR = nullRR;
        return res;
    }

    template<typename... Args>
    auto trace(Args &&...args) const {
        using R = std::remove_cvref_t<decltype(LL.trace(args...))>;
        return (nullLL ? R{} : LL.trace(args...)) + (nullRR ? R{} : RR.trace(args...));
    }

    template<typename... Args>
    static BlockDiag identity(Args &&...args) {
        return BlockDiag(T::identity(std::forward<Args>(args)...));
    }
};

template<typename T>
struct BlockAntiDiag {
    bool nullLR, nullRL;
    T LR, RL;

    BlockAntiDiag() : nullLR(true), nullRL(true), LR{}, RL{} {}
    BlockAntiDiag(T LR, T RL) : nullLR(!LR.size()), nullRL(!RL.size()), LR(std::move(LR)),
                                RL(std::move(RL)) {}
    BlockAntiDiag(const BlockAntiDiag &) = default;
    BlockAntiDiag(BlockAntiDiag &&) = default;
    BlockAntiDiag &operator=(const BlockAntiDiag &b) {
        blk::assign_block(nullLR, LR, b.nullLR, b.LR);
        blk::assign_block(nullRL, RL, b.nullRL, b.RL);
        return *this;
    }
    BlockAntiDiag &operator=(BlockAntiDiag &&b) {
        blk::assign_block(nullLR, LR, b.nullLR, std::move(b.LR));
        blk::assign_block(nullRL, RL, b.nullRL, std::move(b.RL));
        return *this;
    }
    bool operator==(const BlockAntiDiag &) const = default;

    BlockAntiDiag &operator+=(const BlockAntiDiag &rhs) {
        blk::add_block(nullLR, LR, rhs.nullLR, rhs.LR);
        blk::add_block(nullRL, RL, rhs.nullRL, rhs.RL);
        return *this;
    }

    template<ScalarType S>
    BlockAntiDiag &operator*=(S c) {
        if (c == S{}) {
            *this = BlockAntiDiag{};
            return *this;
        }
        blk::scale_block(nullLR, LR, c);
        blk::scale_block(nullRL, RL, c);
        return *this;
    }

    template<typename... Args>
    auto get_matrix(Args &&...args) const {
        using M = std::remove_cvref_t<decltype(LR.get_matrix(args...))>;
        BlockAntiDiag<M> res(nullLR ? M{} : LR.get_matrix(args...),
                             nullRL ? M{} : RL.get_matrix(args...));
        res.nullLR = nullLR;
        res.nullRL = nullRL;
        return res;
    }

    // an anti-diagonal operator has no diagonal
    double trace() const { return 0.0; }
};

template<typename T>
struct BlockOp : BlockDiag<T>, BlockAntiDiag<T> {
    BlockOp(T LL = {}, T LR = {}, T RL = {}, T RR = {})
        : BlockDiag<T>(std::move(LL), std::move(RR)), BlockAntiDiag<T>(std::move(LR), std::move(RL)) {}
    BlockOp(const BlockOp &) = default;
    BlockOp(BlockOp &&) = default;
    BlockOp &operator=(const BlockOp &) = default;
    BlockOp &operator=(BlockOp &&) = default;
    bool operator==(const BlockOp &) const = default;

    explicit BlockOp(const BlockDiag<T> &b) : BlockDiag<T>(b) {}
    explicit BlockOp(const BlockAntiDiag<T> &b) : BlockAntiDiag<T>(b) {}

    BlockOp &operator=(const BlockDiag<T> &b) {
        BlockDiag<T>::operator=(b);
        BlockAntiDiag<T>::operator=(BlockAntiDiag<T>{});
        return *this;
    }
    BlockOp &operator=(const BlockAntiDiag<T> &b) {
        BlockDiag<T>::operator=(BlockDiag<T>{});
        BlockAntiDiag<T>::operator=(b);
        return *this;
    }

    BlockOp &operator+=(const BlockOp &rhs) {
        BlockDiag<T>::operator+=(rhs);
        BlockAntiDiag<T>::operator+=(rhs);
        return *this;
    }
    BlockOp &operator+=(const BlockDiag<T> &rhs) {
        BlockDiag<T>::operator+=(rhs);
        return *this;
    }
    BlockOp &operator+=(const BlockAntiDiag<T> &rhs) {
        BlockAntiDiag<T>::operator+=(rhs);
        return *this;
    }

    template<ScalarType S>
    BlockOp &operator*=(S c) {
        BlockDiag<T>::operator*=(c);
        BlockAntiDiag<T>::operator*=(c);
        return *this;
    }

    template<typename... Args>
    auto get_matrix(Args &&...args) const {
        using M = std::remove_cvref_t<decltype(this->LL.get_matrix(args...))>;
        auto m = [&](bool null, const T &b) { return null ? M{} : b.get_matrix(args...); };
        BlockOp<M> res(m(this->nullLL, this->LL), m(this->nullLR, this->LR),
                       m(this->nullRL, this->RL), m(this->nullRR, this->RR));
        res.nullLL = this->nullLL;
        res.nullLR = this->nullLR;
        res.nullRL = this->nullRL;
        res.nullRR = this->nullRR;
        return res;
    }

    template<typename... Args>
    auto trace(Args &&...args) const {
        return BlockDiag<T>::trace(std::forward<Args>(args)...);
    }
};

template<typename T>
struct BlockVec {
    bool nullL, nullR;
    T L, R;

    BlockVec(T L = {}, T R = {}) : nullL(!L.size()), nullR(!R.size()), L(std::move(L)),
                                   R(std::move(R)) {}
    BlockVec(const BlockVec &) = default;
    BlockVec(BlockVec &&) = default;
    BlockVec &operator=(const BlockVec &b) {
        blk::assign_block(nullL, L, b.nullL, b.L);
        blk::assign_block(nullR, R, b.nullR, b.R);
        return *this;
    }
    BlockVec &operator=(BlockVec &&b) {
        blk::assign_block(nullL, L, b.nullL, std::move(b.L));
        blk::assign_block(nullR, R, b.nullR, std::move(b.R));
        return *this;
    }
    bool operator==(const BlockVec &) const = default;

    BlockVec &operator+=(const BlockVec &rhs) {
        blk::add_block(nullL, L, rhs.nullL, rhs.L);
        blk::add_block(nullR, R, rhs.nullR, rhs.R);
        return *this;
    }

    template<ScalarType S>
    BlockVec &operator*=(S c) {
        if (c == S{}) {
            *this = BlockVec{};
            return *this;
        }
        blk::scale_block(nullL, L, c);
        blk::scale_block(nullR, R, c);
        return *this;
    }

    template<BlockMatrix B>
    BlockVec &operator*=(const B &op) {
        *this = op * *this;
        return *this;
    }
};

// ---- sums (inc/blkops.h:465-528) ----

template<typename T0, typename T1>
auto operator+(const BlockDiag<T0> &a, const BlockDiag<T1> &b) {
    auto pick = [](bool na, const auto &x, bool nb, const auto &y) {
        return na ? y : (nb ? x : x + y);
    };
    return BlockDiag(pick(a.nullLL, a.LL, b.nullLL, b.LL), pick(a.nullRR, a.RR, b.nullRR, b.RR));
}

template<typename T0, typename T1>
auto operator+(const BlockAntiDiag<T0> &a, const BlockAntiDiag<T1> &b) {
    auto pick = [](bool na, const auto &x, bool nb, const auto &y) {
        return na ? y : (nb ? x : x + y);
    };
    return BlockAntiDiag(pick(a.nullLR, a.LR, b.nullLR, b.LR), pick(a.nullRL, a.RL, b.nullRL, b.RL));
}

template<typename T>
BlockOp<T> operator+(const BlockOp<T> &a, const BlockOp<T> &b) {
    BlockOp<T> r(a);
    r += b;
    return r;
}

template<typename T>
BlockVec<T> operator+(const BlockVec<T> &a, const BlockVec<T> &b) {
    BlockVec<T> r(a);
    r += b;
    return r;
}

// a - b = (-1) b + a, in that order
template<typename T0, typename T1>
BlockDiag<T1> operator-(const BlockDiag<T0> &a, const BlockDiag<T1> &b) {
    BlockDiag<T1> r(b);
    r *= -1;
    r += a;
    return r;
}

template<typename T0, typename T1>
BlockAntiDiag<T1> operator-(const BlockAntiDiag<T0> &a, const BlockAntiDiag<T1> &b) {
    BlockAntiDiag<T1> r(b);
    r *= -1;
    r += a;
    return r;
}

template<typename T>
BlockOp<T> operator-(const BlockOp<T> &a, const BlockOp<T> &b) {
    BlockOp<T> r(b);
    r *= -1;
    r += a;
    return r;
}

template<typename T>
BlockVec<T> operator-(const BlockVec<T> &a, const BlockVec<T> &b) {
    BlockVec<T> r(b);
    r *= -1;
    r += a;
    return r;
}

// ---- scalars (inc/blkops.h:530-550) ----

template<ScalarType S, BlockForm B>
B operator*(S c, const B &b) {
    B r(b);
    r *= c;
    return r;
}

template<ScalarType S, BlockForm B>
B operator*(const B &b, S c) {
    return c * b;
}

template<ScalarType S, BlockForm B>
B &operator/=(B &b, S c) {
    b *= S{1} / c;
    return b;
}

// ---- products of block matrices (inc/blkops.h:552-698) ----
//
// Diagonal and anti-diagonal combine by the parity rule:
//     diag * diag = diag,  diag * anti = anti,  anti * anti = diag;
// a BlockOp operand makes the result a BlockOp.  A result block is null
// when every product contributing to it has a null factor.

template<typename T0, typename T1>
auto operator*(const BlockDiag<T0> &a, const BlockDiag<T1> &b) {
    using P = blk::prod_t<T0, T1>;
    BlockDiag<P> r(blk::mul(a.nullLL, a.LL, b.nullLL, b.LL), blk::mul(a.nullRR, a.RR, b.nullRR, b.RR));
    r.nullLL = a.nullLL || b.nullLL;
    r.nullRR = a.nullRR || b.nullRR;
    return r;
}

template<typename T0, typename T1>
auto operator*(const BlockDiag<T0> &a, const BlockAntiDiag<T1> &b) {
    using P = blk::prod_t<T0, T1>;
    BlockAntiDiag<P> r(blk::mul(a.nullLL, a.LL, b.nullLR, b.LR),
                       blk::mul(a.nullRR, a.RR, b.nullRL, b.RL));
    r.nullLR = a.nullLL || b.nullLR;
    r.nullRL = a.nullRR || b.nullRL;
    return r;
}

template<typename T0, typename T1>
auto operator*(const BlockAntiDiag<T0> &a, const BlockDiag<T1> &b) {
    using P = blk::prod_t<T0, T1>;
    BlockAntiDiag<P> r(blk::mul(a.nullLR, a.LR, b.nullRR, b.RR),
                       blk::mul(a.nullRL, a.RL, b.nullLL, b.LL));
    r.nullLR = a.nullLR || b.nullRR;
    r.nullRL = a.nullRL || b.nullLL;
    return r;
}

template<typename T0, typename T1>
auto operator*(const BlockAntiDiag<T0> &a, const BlockAntiDiag<T1> &b) {
    using P = blk::prod_t<T0, T1>;
    BlockDiag<P> r(blk::mul(a.nullLR, a.LR, b.nullRL, b.RL), blk::mul(a.nullRL, a.RL, b.nullLR, b.LR));
    r.nullLL = a.nullLR || b.nullRL;
    r.nullRR = a.nullRL || b.nullLR;
    return r;
}

namespace blk {
// the four blocks of a general product, each the sum of two block products
// (the first term of each sum is the one the reference evaluates first)
template<typename P, typename A, typename B>
BlockOp<P> full_product(bool aLL, const A &LL0, bool aLR, const A &LR0, bool aRL, const A &RL0,
                        bool aRR, const A &RR0, bool bLL, const B &LL1, bool bLR, const B &LR1,
                        bool bRL, const B &RL1, bool bRR, const B &RR1) {
    BlockOp<P> r;
    auto block = [](bool &nul, P &dst, bool n0, const A &x0, bool n1, const B &y0, bool n2,
                    const A &x1, bool n3, const B &y1) {
        nul = true;
        dst = P{};
        accumulate(nul, dst, n0 || n1, mul(n0, x0, n1, y0));
        accumulate(nul, dst, n2 || n3, mul(n2, x1, n3, y1));
    };
    block(r.nullLL, r.LL, aLL, LL0, bLL, LL1, aLR, LR0, bRL, RL1);
    block(r.nullLR, r.LR, aLL, LL0, bLR, LR1, aLR, LR0, bRR, RR1);
    block(r.nullRL, r.RL, aRL, RL0, bLL, LL1, aRR, RR0, bRL, RL1);
    block(r.nullRR, r.RR, aRL, RL0, bLR, LR1, aRR, RR0, bRR, RR1);
    return r;
}
}  // namespace blk

// every product with a BlockOp operand treats the missing blocks of a
// BlockDiag or BlockAntiDiag operand as null
template<typename T0, typename T1>
auto operator*(const BlockOp<T0> &a, const BlockOp<T1> &b) {
    using P = blk::prod_t<T0, T1>;
    return blk::full_product<P>(a.nullLL, a.LL, a.nullLR, a.LR, a.nullRL, a.RL, a.nullRR, a.RR,
                                b.nullLL, b.LL, b.nullLR, b.LR, b.nullRL, b.RL, b.nullRR, b.RR);
}

template<typename T0, typename T1>
auto operator*(const BlockOp<T0> &a, const BlockDiag<T1> &b) {
    using P = blk::prod_t<T0, T1>;
    const T1 z{};
    return blk::full_product<P>(a.nullLL, a.LL, a.nullLR, a.LR, a.nullRL, a.RL, a.nullRR, a.RR,
                                b.nullLL, b.LL, true, z, true, z, b.nullRR, b.RR);
}

template<typename T0, typename T1>
auto operator*(const BlockOp<T0> &a, const BlockAntiDiag<T1> &b) {
    using P = blk::prod_t<T0, T1>;
    const T1 z{};
    return blk::full_product<P>(a.nullLL, a.LL, a.nullLR, a.LR, a.nullRL, a.RL, a.nullRR, a.RR,
                                true, z, b.nullLR, b.LR, b.nullRL, b.RL, true, z);
}

template<typename T0, typename T1>
auto operator*(const BlockDiag<T0> &a, const BlockOp<T1> &b) {
    using P = blk::prod_t<T0, T1>;
    const T0 z{};
    return blk::full_product<P>(a.nullLL, a.LL, true, z, true, z, a.nullRR, a.RR,
                                b.nullLL, b.LL, b.nullLR, b.LR, b.nullRL, b.RL, b.nullRR, b.RR);
}

template<typename T0, typename T1>
auto operator*(const BlockAntiDiag<T0> &a, const BlockOp<T1> &b) {
    using P = blk::prod_t<T0, T1>;
    const T0 z{};
    return blk::full_product<P>(true, z, a.nullLR, a.LR, a.nullRL, a.RL, true, z,
                                b.nullLL, b.LL, b.nullLR, b.LR, b.nullRL, b.RL, b.nullRR, b.RR);
}

// ---- block matrices on block vectors (inc/blkops.h:700-735) ----

template<typename T0, typename T1>
auto operator*(const BlockDiag<T0> &op, const BlockVec<T1> &v) {
    using P = blk::prod_t<T0, T1>;
    BlockVec<P> r(blk::mul(op.nullLL, op.LL, v.nullL, v.L), blk::mul(op.nullRR, op.RR, v.nullR, v.R));
    r.nullL = op.nullLL || v.nullL;
    r.nullR = op.nullRR || v.nullR;
    return r;
}

template<typename T0, typename T1>
auto operator*(const BlockAntiDiag<T0> &op, const BlockVec<T1> &v) {
    using P = blk::prod_t<T0, T1>;
    BlockVec<P> r(blk::mul(op.nullLR, op.LR, v.nullR, v.R), blk::mul(op.nullRL, op.RL, v.nullL, v.L));
    r.nullL = op.nullLR || v.nullR;
    r.nullR = op.nullRL || v.nullL;
    return r;
}

template<typename T0, typename T1>
auto operator*(const BlockOp<T0> &op, const BlockVec<T1> &v) {
    return static_cast<const BlockDiag<T0> &>(op) * v + static_cast<const BlockAntiDiag<T0> &>(op) * v;
}

// inner product <a|b>: the sum of the two sectors' inner products
template<typename T0, typename T1>
auto operator*(const BlockVec<T0> &a, const BlockVec<T1> &b) {
    using P = blk::prod_t<T0, T1>;
    P l = (a.nullL || b.nullL) ? P{} : a.L * b.L;
    P r = (a.nullR || b.nullR) ? P{} : a.R * b.R;
    return l + r;
}

template<typename T>
std::ostream &operator<<(std::ostream &os, const BlockDiag<T> &b) {
    return os << "BlockDiag<" << b.LL << ", " << b.RR << ">";
}

template<typename T>
std::ostream &operator<<(std::ostream &os, const BlockAntiDiag<T> &b) {
    return os << "BlockAntiDiag<" << b.LR << ", " << b.RL << ">";
}

template<typename T>
std::ostream &operator<<(std::ostream &os, const BlockOp<T> &b) {
    return os << "BlockOp<" << b.LL << ", " << b.LR << ", " << b.RL << ", " << b.RR << ">";
}

template<typename T>
std::ostream &operator<<(std::ostream &os, const BlockVec<T> &b) {
    return os << "BlockVec<" << b.L << ", " << b.R << ">";
}
