// state.h -- State: a vector in the 2^len Hilbert space of a spin chain,
// owning a ring of backend buffers (the reference's State,
// inc/state.h:43-443).  buf(i) is bufs[(curbuf + i) % nbufs]; evolution
// algorithms enlarge the ring and rotate curbuf instead of copying vectors.
#pragma once

#ifndef _REAPERS_H_
#error "Do not include this file directly. Include the master header file reapers.h"
#endif

template<RealScalar FpType, typename Impl>
class State;

// statistics and options of a restarted Krylov evolution
struct KrylovStats {
    long matvecs = 0;
    long restarts = 0;
    long gs_passes = 0;
    int breakdown = 0;
    int substeps = 0;
    int rejected = 0;         // schedule proposals the error estimate rejected
    double last_nrm = 0;
    double host_ms = 0;       // time in the host-side small-matrix work
};

enum class Reorth { Selective, Always };

// What a schedule sees when it decides the number of sub-steps (before the
// evolution) and the Krylov dimension of a restart (after the restart's first
// Arnoldi step, so that alpha0 and beta0 are this restart's).
struct RestartInfo {
    int len = 0;              // spin chain length
    long nterms = 0;          // terms of H
    double h_rms = 0;         // sqrt(sum |c|^2) = RMS eigenvalue of H
    double h_l1 = 0;          // sum |c| >= spectral radius
    double apply_passes = 1;  // HBM traffic of one H*v in state-vector sizes
    double t = 0, beta = 0;   // the whole evolution
    double tau = 0, taub = 0; // this sub-step
    double tol = 0;
    int substep = 0, nsub = 1;
    int restart = 0;          // restarts done in this sub-step
    double last_update = 0;   // ||Hexp0|| of the previous restart (0 before)
    double prev_update = 0;   // ||Hexp0|| of the one before that (0 before)
    int prev_krydim = 0;      // Krylov dimension of the previous restart
    double alpha0 = 0;        // H(0,0) = <v0|H|v0> of this restart
    double beta0 = 0;         // H(1,0) = ||H v0 - alpha0 v0|| of this restart
};

// Chooses the number of sub-steps of an evolution and the Krylov dimension
// of every restart.  The default schedule is the reference's: one sub-step,
// the same dimension for every restart.
class KrylovSchedule {
public:
    virtual ~KrylovSchedule() {}
    virtual int substeps(const RestartInfo &) { return 1; }
    virtual int krydim(const RestartInfo &) = 0;
};

struct KrylovOptions {
    int krydim = 5;
    double tol = 0;           // 0: epsilon of the state's precision
    int max_iters = 0;
    Reorth reorth = Reorth::Selective;
    double substep = 0;       // 0: whole t in one evolution
    KrylovSchedule *schedule = nullptr;   // null: fixed krydim / substep
    int max_krydim = 64;      // bound on a schedule's proposals
};

namespace EvolutionAlgorithm {
#include "algo.h"
}

template<typename Impl>
using DefEvolutionAlgorithm = EvolutionAlgorithm::Krylov<Impl>;

template<RealScalar FpType = DefFpType, typename Impl = DefImpl>
class State {
    template<RealScalar F1, typename I1> friend class State;

    using IndexType = size_t;
    typename SpinOp<FpType>::IndexType len;
    int curbuf;
    std::vector<typename Impl::template VecType<FpType>> bufs;

public:
    State() : len(0), curbuf(0) {}

    State(typename SpinOp<FpType>::IndexType len) : len(len), curbuf(0) {
        if (len == 0) ThrowException(InvalidArgument, "len", "cannot be zero");
        if (len > NUM_SPIN_SITES)
            ThrowException(InvalidArgument, "len", "cannot be larger than NUM_SPIN_SITES");
        bufs.emplace_back(dim());
    }

    State(const State &st) : len(st.len), curbuf(0) {
        if (!len) return;
        bufs.emplace_back(dim());
        Impl::copy_vec(dim(), buf(), st.buf());
    }

    State(State &&st) : State() {
        std::swap(len, st.len);
        std::swap(bufs, st.bufs);
        std::swap(curbuf, st.curbuf);
    }

    ~State() { len = 0; curbuf = 0; }

    // precision conversion (inc/state.h:104-109); double -> float rounds to
    // nearest unless another mode is given
    template<RealScalar F1>
    explicit State(const State<F1, Impl> &st, int rounding = RP_ROUND_NEAREST)
        : len(st.len), curbuf(0) {
        if (!len) return;
        bufs.emplace_back(dim());
        Impl::copy_vec(dim(), buf(), st.buf(), rounding);
    }

    // conversion between backends (inc/state.h:114-118): the current vector
    // is copied into the other backend's memory; a precision change, if
    // any, is made on the source backend first
    template<RealScalar F1, typename I1>
        requires(!std::is_same_v<I1, Impl>)
    explicit State(const State<F1, I1> &st) : len(st.len), curbuf(0) {
        if (!len) return;
        if constexpr (std::is_same_v<F1, FpType>) {
            bufs.emplace_back(static_cast<typename Impl::template VecType<FpType>>(
                st.bufs[st.curbuf % st.num_bufs()]));
        } else {
            State<FpType, I1> conv(st);
            bufs.emplace_back(static_cast<typename Impl::template VecType<FpType>>(
                conv.bufs[conv.curbuf % conv.num_bufs()]));
        }
    }

    void random_state(u64 seed = 0) {
        if (!len) return;
        Impl::init_random(dim(), buf(), seed);
    }

    void zero_state() {
        if (!len) return;
        Impl::zero_vec(dim(), buf());
    }

    int num_bufs() const { return (int)bufs.size(); }
    auto spin_chain_length() const { return len; }
    IndexType dim() const { return (IndexType)1 << len; }
    size_t size() const { return len ? dim() : 0; }

    typename Impl::template BufType<FpType> buf(int i = 0) {
        assert(len != 0 && num_bufs() >= i + 1);
        return bufs[(curbuf + i) % num_bufs()].get();
    }
    typename Impl::template ConstBufType<FpType> buf(int i = 0) const {
        assert(len != 0 && num_bufs() >= i + 1);
        return bufs[(curbuf + i) % num_bufs()].get();
    }

    // the backend vector object of buf(i) (the fused Krylov sweeps take a
    // list of them)
    const typename Impl::template VecType<FpType> &vec(int i = 0) const {
        assert(len != 0 && num_bufs() >= i + 1);
        return bufs[(curbuf + i) % num_bufs()];
    }

    void inc_curbuf(int i = 1) {
        if (!len) return;
        curbuf = (curbuf + i) % num_bufs();
    }

    void enlarge(int n) {
        if (!len) return;
        int nbufs = num_bufs();
        for (int i = nbufs; i < n; i++) bufs.emplace_back(dim());
    }

    void swap_bufs(int i, int j) {
        if (!len) return;
        auto n = num_bufs();
        i = ((i + curbuf) % n + n) % n;
        j = ((j + curbuf) % n + n) % n;
        if (i != j) swap(bufs[i], bufs[j]);
    }

    complex<FpType> operator[](IndexType i) const {
        if (i >= dim()) ThrowException(StateIndexTooLarge, i, dim() - 1);
        return buf()[i];
    }

    // a writable amplitude (inc/state.h:225-237): st[i] = c copies one value
    // to the device
    typename Impl::template ElemRefType<FpType> operator[](IndexType i) {
        if (i >= dim()) ThrowException(StateIndexTooLarge, i, dim() - 1);
        return buf()[i];
    }

    // whole-vector host transfers (interleaved complex<FpType>)
    void load(const complex<FpType> *host) {
        if constexpr (std::is_pointer_v<typename Impl::template BufType<FpType>>)
            std::memcpy(buf(), host, dim() * sizeof(complex<FpType>));
        else
            buf().upload(host);
    }
    void store(complex<FpType> *host) const {
        if constexpr (std::is_pointer_v<typename Impl::template BufType<FpType>>)
            std::memcpy(host, buf(), dim() * sizeof(complex<FpType>));
        else
            buf().download(host);
    }

    State &operator=(const State &st) {
        if (this == &st) return *this;
        if (dim() != st.dim() || !len) {
            bufs.clear();
            curbuf = 0;
            len = st.len;
            if (len) bufs.emplace_back(dim());
        }
        if (len) Impl::copy_vec(dim(), buf(), st.buf());
        return *this;
    }

    State &operator=(State &&st) {
        if (this == &st) return *this;
        std::swap(len, st.len);
        std::swap(curbuf, st.curbuf);
        std::swap(bufs, st.bufs);
        return *this;
    }

    template<RealScalar F1>
    State &operator=(const State<F1, Impl> &st) {
        if (dim() != st.dim() || !len) {
            bufs.clear();
            curbuf = 0;
            len = st.len;
            if (len) bufs.emplace_back(dim());
        }
        if (len) Impl::copy_vec(dim(), buf(), st.buf());
        return *this;
    }

    State &operator+=(const State &rhs) {
        if (len != rhs.len) ThrowException(InvalidArgument, "dimensions of states must match");
        if (len) Impl::add_vec(dim(), buf(), rhs.buf());
        return *this;
    }

    State operator+(const State &rhs) const {
        if (len != rhs.len) ThrowException(InvalidArgument, "dimensions of states must match");
        if (!len) return {};
        State res(rhs.len);
        Impl::add_vec(dim(), res.buf(), buf(), rhs.buf());
        return res;
    }

    State &operator-=(const State &rhs) {
        if (len != rhs.len) ThrowException(InvalidArgument, "dimensions of states must match");
        if (len) Impl::add_vec(dim(), buf(), FpType(-1.0), rhs.buf());
        return *this;
    }

    State operator-(const State &rhs) const {
        if (!len) return {};
        State res{*this};
        res -= rhs;
        return res;
    }

    // s = ops * s, through a second ring buffer (inc/state.h:291-302)
    State &operator*=(const typename Impl::template SumOps<FpType> &ops) {
        if (!len) return *this;
        enlarge(2);
        auto &&v = buf();
        inc_curbuf();
        Impl::apply_ops_overwrite(len, buf(), ops, v);
        return *this;
    }

    State &operator*=(complex<FpType> s) {
        if (len) Impl::scale_vec(dim(), buf(), s);
        return *this;
    }

    FpType norm() const {
        if (!len) return {};
        return Impl::vec_norm(dim(), buf());
    }

    double norm64() const {
        if (!len) return 0;
        return Impl::vec_norm64(dim(), buf());
    }

    void normalize() {
        if (!len) return;
        Impl::scale_vec(dim(), buf(), (FpType)(1.0 / norm64()));
    }

    bool operator==(const State &st) const {
        if (len != st.len) return false;
        if (!len) return true;
        State diff{*this - st};
        return diff.norm() == 0.0;
    }

    friend State operator*(const typename Impl::template SumOps<FpType> &ops, const State &s) {
        if (!s.len) return {};
        State res(s.len);
        Impl::apply_ops_overwrite(s.len, res.buf(), ops, s.buf());
        return res;
    }

    // <s0|s1> (inc/state.h:368-375)
    friend complex<FpType> operator*(const State &s0, const State &s1) {
        if (s0.dim() != s1.dim()) ThrowException(InvalidArgument, "Dimensions of states", "must match");
        if (!s0.len) return {};
        return Impl::vec_prod(s0.dim(), s0.buf(), s1.buf());
    }

    std::complex<double> inner64(const State &s1) const {
        if (dim() != s1.dim()) ThrowException(InvalidArgument, "Dimensions of states", "must match");
        return Impl::vec_prod64(dim(), buf(), s1.buf());
    }

    template<template<typename> typename Algo = DefEvolutionAlgorithm, typename... Args>
    FpType ground_state(const typename Impl::template SumOps<FpType> &ham, Args &&...args) {
        if (!len) return {};
        return Algo<Impl>::ground_state(*this, ham, std::forward<Args>(args)...);
    }

    // |psi> = exp(-(i t + beta) H) |psi> (inc/state.h:413-424)
    template<template<typename> typename Algo = DefEvolutionAlgorithm, RealScalar F1,
             RealScalar F2, typename... Args>
    void evolve(const typename Impl::template SumOps<FpType> &ham, F1 t, F2 beta = 0.0,
                Args &&...args) {
        if ((t == 0.0) && (beta == 0.0)) return;
        if (!len) return;
        Algo<Impl>::evolve(*this, ham, static_cast<FpType>(t), static_cast<FpType>(beta),
                           std::forward<Args>(args)...);
    }

    // evolve with options and statistics (FP64 times)
    KrylovStats evolve_ex(const typename Impl::template SumOps<FpType> &ham, double t,
                          double beta, const KrylovOptions &opt) {
        KrylovStats st;
        if ((t == 0.0) && (beta == 0.0)) return st;
        if (!len) return st;
        EvolutionAlgorithm::Krylov<Impl>::evolve_ex(*this, ham, t, beta, opt, st);
        return st;
    }

    void gc() {
        if (!len) return;
        if (curbuf != 0) swap_bufs(0, -curbuf);
        while (num_bufs() > 1) bufs.pop_back();
        curbuf = 0;
    }
};

// A state in parity-block form (the reference's BlockState, inc/state.h:
// 447-584): the L and the R parity sector as two States of len sites each.
// Amplitude i of the full 2^(len+1) vector is L[i] below 2^len and
// R[i - 2^len] from there on, L first; a null sector reads as zeros.  The
// sectors are independent problems on the device: evolve() runs each one's
// restarted Krylov with its own block of the Hamiltonian, and a single
// Majorana (BlockAntiDiag) moves amplitudes between them.
template<RealScalar FpType = DefFpType, typename Impl = DefImpl>
struct BlockState : BlockVec<State<FpType, Impl>> {
    using StateType = State<FpType, Impl>;
    using Base = BlockVec<StateType>;

private:
    using LenType = typename SpinOp<FpType>::IndexType;
    LenType slen;

    LenType len_of(const Base &b) const {
        return b.L.size() ? b.L.spin_chain_length() : b.R.spin_chain_length();
    }

    // a null sector holds an empty State; give it storage again
    void define_sectors() {
        if (!this->L.size()) this->L = StateType(slen);
        if (!this->R.size()) this->R = StateType(slen);
        this->nullL = this->nullR = false;
    }

public:
    // len is the length of each sector's spin chain (the full space has
    // 2^(len+1) amplitudes)
    BlockState(LenType len) : Base(StateType(len), StateType(len)), slen(len) {}
    BlockState(const Base &b) : Base(b), slen(len_of(b)) {}
    BlockState(Base &&b) : Base(std::move(b)), slen(len_of(*this)) {}
    BlockState(const BlockState &) = default;
    BlockState(BlockState &&) = default;
    // another precision or backend, sector by sector (null sectors stay null)
    template<RealScalar F1, typename I1>
        requires(!std::is_same_v<F1, FpType> || !std::is_same_v<I1, Impl>)
    explicit BlockState(const BlockState<F1, I1> &b)
        : Base(b.nullL ? StateType{} : StateType(b.L), b.nullR ? StateType{} : StateType(b.R)),
          slen(b.spin_chain_length()) {}
    template<RealScalar F1>
        requires(!std::is_same_v<F1, FpType>)
    BlockState &operator=(const BlockState<F1, Impl> &b) {
        if (b.nullL) { this->L = StateType{}; this->nullL = true; }
        else { this->L = b.L; this->nullL = false; }
        if (b.nullR) { this->R = StateType{}; this->nullR = true; }
        else { this->R = b.R; this->nullR = false; }
        slen = b.spin_chain_length();
        return *this;
    }
    BlockState &operator=(const Base &b) {
        Base::operator=(b);
        if (len_of(b)) slen = len_of(b);
        return *this;
    }
    BlockState &operator=(Base &&b) {
        LenType l = len_of(b);
        Base::operator=(std::move(b));
        if (l) slen = l;
        return *this;
    }
    BlockState &operator=(const BlockState &b) {
        Base::operator=(b);
        slen = b.slen;
        return *this;
    }
    BlockState &operator=(BlockState &&b) {
        Base::operator=(std::move(b));
        slen = b.slen;
        return *this;
    }

    // each sector's own random vector, then the pair normalized together;
    // null sectors become defined again
    void random_state(u64 seed = 0) {
        define_sectors();
        this->L.random_state(seed);
        this->R.random_state(seed ? seed ^ 0x9e3779b97f4a7c15ull : 0);
        normalize();
    }

    // zero vectors in both sectors, which are then defined (not null)
    void zero_state() {
        define_sectors();
        this->L.zero_state();
        this->R.zero_state();
    }

    LenType spin_chain_length() const { return slen; }

    size_t dim() const { return (size_t)1 << (spin_chain_length() + 1); }

    complex<FpType> operator[](size_t i) const {
        size_t half = dim() / 2;
        if (i < half) return this->nullL ? complex<FpType>{} : this->L[i];
        return this->nullR ? complex<FpType>{} : this->R[i - half];
    }

    // a writable amplitude; both sectors must be defined (inc/state.h:542-550)
    typename Impl::template ElemRefType<FpType> operator[](size_t i) {
        if (this->nullL || this->nullR)
            ThrowException(InvalidArgument, "Both parity blocks must be defined.");
        size_t half = dim() / 2;
        return i < half ? this->L[i] : this->R[i - half];
    }

    double norm64() const {
        double l = this->nullL ? 0.0 : this->L.norm64();
        double r = this->nullR ? 0.0 : this->R.norm64();
        return std::sqrt(l * l + r * r);
    }

    FpType norm() const { return (FpType)norm64(); }

    void normalize() {
        if (this->nullL && this->nullR) return;
        double n = norm64();
        if (!this->nullL) this->L *= complex<FpType>((FpType)(1.0 / n));
        if (!this->nullR) this->R *= complex<FpType>((FpType)(1.0 / n));
    }

    // <this|other>, accumulated in FP64 per sector
    std::complex<double> inner64(const BlockState &other) const {
        std::complex<double> s{};
        if (!this->nullL && !other.nullL) s += this->L.inner64(other.L);
        if (!this->nullR && !other.nullR) s += this->R.inner64(other.R);
        return s;
    }

    // Lanczos in each sector; the state keeps the sector with the lower
    // energy and the other becomes null (inc/state.h:487-505)
    template<template<typename> typename Algo = DefEvolutionAlgorithm, typename HamOpTy,
             typename... Args>
    FpType ground_state(const BlockDiag<HamOpTy> &ham, Args &&...args) {
        FpType e0 = this->L.template ground_state<Algo>(ham.LL, args...);
        FpType e1 = this->R.template ground_state<Algo>(ham.RR, args...);
        if (e0 <= e1) {
            this->R = StateType{};
            this->nullR = true;
            this->nullL = false;
            return e0;
        }
        this->L = StateType{};
        this->nullL = true;
        this->nullR = false;
        return e1;
    }

    // |psi> = exp(-(it + beta) H) |psi>, each defined sector with its block
    template<template<typename> typename Algo = DefEvolutionAlgorithm, RealScalar F1,
             RealScalar F2, typename HamOpTy, typename... Args>
    void evolve(const BlockDiag<HamOpTy> &ham, F1 t, F2 beta = 0.0, Args &&...args) {
        if ((t == 0.0) && (beta == 0.0)) return;
        if (!this->nullL) this->L.template evolve<Algo>(ham.LL, t, beta, args...);
        if (!this->nullR) this->R.template evolve<Algo>(ham.RR, t, beta, args...);
    }

    // the same with options; statistics summed over the sectors
    template<typename HamOpTy>
    KrylovStats evolve_ex(const BlockDiag<HamOpTy> &ham, double t, double beta,
                          const KrylovOptions &opt) {
        KrylovStats st;
        for (int s = 0; s < 2; s++) {
            bool null = s ? this->nullR : this->nullL;
            if (null) continue;
            StateType &v = s ? this->R : this->L;
            const HamOpTy &h = s ? ham.RR : ham.LL;
            KrylovStats k = v.evolve_ex(h, t, beta, opt);
            st.matvecs += k.matvecs;
            st.restarts += k.restarts;
            st.gs_passes += k.gs_passes;
            st.breakdown |= k.breakdown;
            st.substeps += k.substeps;
            st.rejected += k.rejected;
            st.host_ms += k.host_ms;
            st.last_nrm = k.last_nrm;
        }
        return st;
    }

    void gc() {
        if (!this->nullL) this->L.gc();
        if (!this->nullR) this->R.gc();
    }
};
