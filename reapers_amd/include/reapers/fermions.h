// fermions.h -- Majorana operators as Jordan-Wigner Pauli strings.
//
// SubspaceView is a SumOps with a fixed spin chain length (the reference's
// SubspaceView, inc/fermions.h:39-205).  FermionOpNonBlockForm(N, n) is the
// reference's chi_n on L = N/2 sites (inc/fermions.h:254-306):
//
//   chi_n = {sigma_y if n odd, sigma_z if n even}((n-1)/2)
//           . prod_{(n+1)/2 <= i < N/2} sigma_x(i)       (/ sqrt(2) if standard)
//
// so its flip mask covers the sites [n/2, N/2) and its phase mask is the
// single site (n-1)/2 (none for n = 0).  It is built by the same sequence of
// products as the reference, so coefficients agree bit for bit.
#pragma once

#ifndef _REAPERS_H_
#error "Do not include this file directly. Include the master header file reapers.h"
#endif

template<typename SumOpsTy>
class SubspaceView : public SumOpsTy {
    template<typename S1> friend class SubspaceView;
    int len;

    void ensure_equal_len(const SubspaceView &rhs) const {
        if (this->empty() || rhs.empty()) return;
        if (len != rhs.len) ThrowException(InvalidArgument, "spin chain length", "must be equal");
    }

public:
    SubspaceView() : len{} {}
    template<typename S1>
    SubspaceView(int len, S1 &&ops) : SumOpsTy(std::forward<S1>(ops)), len(len) {}
    SubspaceView(const SubspaceView &) = default;
    SubspaceView(SubspaceView &&) = default;
    SubspaceView &operator=(const SubspaceView &) = default;
    SubspaceView &operator=(SubspaceView &&) = default;

    // a view over term lists of another precision (inc/fermions.h:62-72):
    // the production driver promotes its FP32 Hamiltonian with ham64 = ham32
    template<typename S1>
        requires(!std::is_same_v<S1, SumOpsTy>)
    SubspaceView(const SubspaceView<S1> &view)
        : SumOpsTy(static_cast<const S1 &>(view)), len(view.len) {}

    template<typename S1>
        requires(!std::is_same_v<S1, SumOpsTy>)
    SubspaceView &operator=(const SubspaceView<S1> &view) {
        SumOpsTy::operator=(static_cast<const S1 &>(view));
        len = view.len;
        return *this;
    }

    static auto identity(int len) {
        return SubspaceView<SumOpsTy>(len, SpinOp<typename SumOpsTy::RealScalarType>::identity());
    }

    using SumOpsTy::operator+=;
    using SumOpsTy::operator-=;
    using SumOpsTy::operator*=;
    using SumOpsTy::operator/=;
    using SumOpsTy::operator==;

    // operand orders as in the reference (inc/fermions.h:87-175); every
    // product is HostSumOps::operator* on the underlying term lists
    SubspaceView operator+(const SumOpsTy &rhs) const {
        return SubspaceView(len, SumOpsTy::operator+(rhs));
    }

    template<typename OpTy>
    friend SubspaceView operator+(OpTy &&lhs, const SubspaceView &rhs) {
        return SubspaceView{rhs.len, std::forward<OpTy>(lhs) + static_cast<const SumOpsTy &>(rhs)};
    }

    SubspaceView operator+(const SubspaceView &rhs) const {
        ensure_equal_len(rhs);
        if (len) return *this + static_cast<const SumOpsTy &>(rhs);
        return rhs;
    }

    SubspaceView &operator+=(const SubspaceView &rhs) {
        ensure_equal_len(rhs);
        if (len) SumOpsTy::operator+=(static_cast<const SumOpsTy &>(rhs));
        else *this = rhs;
        return *this;
    }

    SubspaceView operator-(const SumOpsTy &rhs) const {
        return SubspaceView(len, SumOpsTy::operator-(rhs));
    }

    SubspaceView operator-(const SubspaceView &rhs) const {
        ensure_equal_len(rhs);
        return SubspaceView(len ? len : rhs.len, SumOpsTy::operator-(static_cast<const SumOpsTy &>(rhs)));
    }

    SubspaceView &operator-=(const SubspaceView &rhs) {
        ensure_equal_len(rhs);
        if (!len) len = rhs.len;
        SumOpsTy::operator-=(static_cast<const SumOpsTy &>(rhs));
        return *this;
    }

    SubspaceView operator*(const SumOpsTy &rhs) const {
        return SubspaceView(len, SumOpsTy::operator*(rhs));
    }

    template<typename OpTy>
    friend SubspaceView operator*(OpTy &&lhs, const SubspaceView &rhs) {
        return SubspaceView{rhs.len, std::forward<OpTy>(lhs) * static_cast<const SumOpsTy &>(rhs)};
    }

    SubspaceView operator*(const SubspaceView &rhs) const {
        ensure_equal_len(rhs);
        return static_cast<const SumOpsTy &>(*this) * rhs;
    }

    SubspaceView &operator*=(const SubspaceView &rhs) {
        ensure_equal_len(rhs);
        SumOpsTy::operator*=(static_cast<const SumOpsTy &>(rhs));
        return *this;
    }

    bool operator==(const SubspaceView &rhs) const {
        if (len != rhs.len) return false;
        return SumOpsTy::operator==(static_cast<const SumOpsTy &>(rhs));
    }

    int spin_chain_length() const { return len; }

    // dense forms on this view's chain length (inc/fermions.h:183-203)
    using MatrixType = typename SumOpsTy::MatrixType;
    decltype(auto) get_matrix() const { return SumOpsTy::get_matrix(len); }
    // the view reads as its dense matrix (Matrix A = syk.fermion_ops(n))
    operator const MatrixType &() const { return SumOpsTy::get_matrix(len); }
    MatrixType operator*(const MatrixType &m) const { return SumOpsTy::get_matrix(len) * m; }
    decltype(auto) get_eigensystem() const { return SumOpsTy::get_eigensystem(len); }
    decltype(auto) matexp(typename SumOpsTy::RealScalarType c) const {
        return SumOpsTy::matexp(typename SumOpsTy::ComplexScalarType(c), len);
    }
    decltype(auto) matexp(typename SumOpsTy::ComplexScalarType c) const {
        return SumOpsTy::matexp(c, len);
    }
    using SumOpsTy::matexp;
    using SumOpsTy::get_matrix;
    using SumOpsTy::get_eigensystem;
};

template<RealScalar FpType = DefFpType, typename Impl = DefImpl>
class FermionOpNonBlockForm : public SubspaceView<typename Impl::template SumOps<FpType>> {
    using BaseType = SubspaceView<typename Impl::template SumOps<FpType>>;
    using IndexType = u8;
    IndexType N, n;

    static void check_field_index(IndexType N, IndexType n) {
        if (N < 2) ThrowException(InvalidArgument, "N", "must be at least two");
        if (N & 1) ThrowException(InvalidArgument, "N", "must be even");
        if (n > N) ThrowException(FieldIndexTooLarge, n, N);
    }

public:
    using ProductType = BaseType;

    FermionOpNonBlockForm(IndexType N, IndexType n, bool standard_gamma = true)
        : BaseType(N / 2, SpinOp<FpType>::identity()), N(N), n(n) {
        check_field_index(N, n);
        for (int i = (n + 1) / 2; i < N / 2; i++) *this *= SpinOp<FpType>::sigma_x(i);
        if (n != 0)
            *this *= (n & 1) ? SpinOp<FpType>::sigma_y((n - 1) / 2)
                             : SpinOp<FpType>::sigma_z((n - 1) / 2);
        if (standard_gamma) *this /= sqrt(2.0);
    }

    using BaseType::operator*;
    using BaseType::operator+=;
    using BaseType::operator*=;
    using BaseType::operator/=;
    using BaseType::operator==;
};

// Majorana chi_n in parity-block form (the reference's FermionOpBlockForm,
// inc/fermions.h:308-368).  Every chi_n anticommutes with the parity of the
// N/2 sites, so it is block anti-diagonal between the two parity sectors of
// 2^(N/2-1) amplitudes each, and its blocks are operators on N/2-1 sites:
//     n <= N-2:  LR = RL = chi'_n, the Majorana of N-2 fermions without the
//                1/sqrt(2) (for n = N-2 that is sigma_z(N/2-2));
//     n == N-1:  LR = -i 1, RL = +i 1;
// all divided by sqrt(2) for the standard normalization, as a product with
// 1/sqrt(2.0) rounded to FpType.  A product of two is block diagonal.
template<RealScalar FpType = DefFpType, typename Impl = DefImpl>
class FermionOpBlockForm
    : public BlockAntiDiag<SubspaceView<typename Impl::template SumOps<FpType>>> {
    using BlockType = SubspaceView<typename Impl::template SumOps<FpType>>;
    using BaseType = BlockAntiDiag<BlockType>;
    using IndexType = u8;
    IndexType N, n;

public:
    using ProductType = BlockDiag<BlockType>;

    FermionOpBlockForm(IndexType N, IndexType n, bool standard_gamma = true)
        : BaseType(BlockType(N / 2 - 1, SpinOp<FpType>::identity()),
                   BlockType(N / 2 - 1, SpinOp<FpType>::identity())),
          N(N), n(n) {
        // same checks as the non-block form; n == N and N == 2 are then
        // rejected by the (N-2)-fermion operator below
        if (N < 2) ThrowException(InvalidArgument, "N", "must be at least two");
        if (N & 1) ThrowException(InvalidArgument, "N", "must be even");
        if (n > N) ThrowException(FieldIndexTooLarge, n, N);
        if (n == N - 1) {
            this->LR *= complex<FpType>{0, -1};
            this->RL *= complex<FpType>{0, 1};
        } else {
            this->LR = FermionOpNonBlockForm<FpType, Impl>(N - 2, n, false);
            this->RL = this->LR;
        }
        if (standard_gamma) *this /= sqrt(2.0);
    }

    int num_fermions() const { return N; }
    int index() const { return n; }
};

#ifdef REAPERS_USE_PARITY_BLOCKS
template<RealScalar FpType = DefFpType, typename Impl = DefImpl>
using FermionOp = FermionOpBlockForm<FpType, Impl>;
#else
template<RealScalar FpType = DefFpType, typename Impl = DefImpl>
using FermionOp = FermionOpNonBlockForm<FpType, Impl>;
#endif
