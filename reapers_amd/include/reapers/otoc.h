// otoc.h -- out-of-time-order correlator F(t) = <i| A(t) B A(t) B |i> with
// A(t) = exp(iHt) A exp(-iHt) and |i> = exp(-beta H / 8) |psi>, by four
// Krylov evolutions and four operator applications per time point: the
// sequence of the reference's otoc_krylov (examples/simple-syk-krylov.cc:
// 19-66) and OTOC::evolve (examples/syk.cc:270-284):
//     psi = B |i>;  psi = exp(-(it + beta/4) H) psi;  psi = A psi;
//     psi = exp(-(-it + beta/4) H) psi;  psi = B psi;
//     psi = exp(-(it + beta/4) H) psi;  psi = A psi;  psi = exp(iHt) psi;
//     F(t) = <i|psi>.
// Differences that are deliberate: the inner product accumulates in FP64,
// the time grid is the caller's (the reference uses t_k = k tmax / nsteps),
// the values are returned unnormalized (the reference divides by F(0)), and
// the initial state's Krylov buffers are released after its imaginary-time
// evolution so that only one state carries a Krylov basis (at 30 qubits one
// FP32 buffer is 8.6 GB).
#pragma once

#ifndef _REAPERS_H_
#error "Do not include this file directly. Include the master header file reapers.h"
#endif

namespace internal {
inline void add_stats(KrylovStats &st, const KrylovStats &s) {
    st.matvecs += s.matvecs;
    st.restarts += s.restarts;
    st.gs_passes += s.gs_passes;
    st.breakdown |= s.breakdown;
    st.substeps += s.substeps;
    st.rejected += s.rejected;
    st.host_ms += s.host_ms;
    st.last_nrm = s.last_nrm;
}
}  // namespace internal

// The four-evolution OTOC with the forward (+t) evolutions under fwd and the
// backward (-t) ones under bwd; the initial state is pre-evolved under bwd.
// fwd == bwd is the single-Hamiltonian form; in parity-block form fwd is
// the R block, bwd the L block, and A = B is a Majorana's LR block, which
// carries the state from one sector to the other (examples/syk.cc:270-284).
template<RealScalar FpType, typename Impl>
std::vector<std::complex<double>> otoc2(const typename Impl::template SumOps<FpType> &fwd,
                                        const typename Impl::template SumOps<FpType> &bwd,
                                        const typename Impl::template SumOps<FpType> &A,
                                        const typename Impl::template SumOps<FpType> &B,
                                        State<FpType, Impl> &init,
                                        const std::vector<double> &times, double beta,
                                        const KrylovOptions &opt, KrylovStats &st)
{
    if (beta != 0.0) {
        internal::add_stats(st, init.evolve_ex(bwd, 0.0, beta / 8, opt));
        init.gc();
    }
    std::vector<std::complex<double>> v;
    v.reserve(times.size());
    State<FpType, Impl> psi;
    for (double t : times) {
        psi = init;
        psi *= B;
        internal::add_stats(st, psi.evolve_ex(fwd, t, beta / 4, opt));
        psi *= A;
        internal::add_stats(st, psi.evolve_ex(bwd, -t, beta / 4, opt));
        psi *= B;
        internal::add_stats(st, psi.evolve_ex(fwd, t, beta / 4, opt));
        psi *= A;
        internal::add_stats(st, psi.evolve_ex(bwd, -t, 0.0, opt));
        v.push_back(init.inner64(psi));
        st.matvecs += 4;   // the four operator applications
    }
    return v;
}

template<RealScalar FpType, typename Impl>
std::vector<std::complex<double>> otoc(const typename Impl::template SumOps<FpType> &ham,
                                       const typename Impl::template SumOps<FpType> &A,
                                       const typename Impl::template SumOps<FpType> &B,
                                       State<FpType, Impl> &init,
                                       const std::vector<double> &times, double beta,
                                       const KrylovOptions &opt, KrylovStats &st)
{
    return otoc2<FpType, Impl>(ham, ham, A, B, init, times, beta, opt, st);
}

// The reference's finite-temperature n-point functions of one disorder
// realization in parity-block form (BaseEval::eval with the OTOC and Green
// evaluators, examples/syk.cc:120-284), starting from a state s0 of the L
// sector:
//   1. E0 = the Lanczos ground-state energy of ham.LL (krylov_dim vectors,
//      eps 0.1, from a random state; examples/syk.cc:137-140);
//   2. ham.LL and ham.RR are shifted by -E0 times the identity, so that the
//      imaginary-time factors exp(-beta (H - E0) / k) stay of order one;
//   3. s0 is pre-evolved: exp(-beta/8 H_L) for the OTOC, exp(-beta/2 H_L)
//      for the Green's function;
//   4. for t = k tmax / nsteps, k = 0..nsteps, s = s0 and
//        OTOC:  s *= op; s = e^{-(it+b/4) H_R} s; s = e^{-(-it+b/4) H_L} s;
//               s *= op; s = e^{-(it+b/4) H_R} s; s = e^{it H_L} s
//        Green: s = e^{-it H_R} s; s = e^{it H_L} s
//      and v_k = <s0|s>;
//   5. every v_k is divided by v_0.
// op is ops[otoc_idx].LR (ops[N-2].LR when otoc_idx is outside [0, N-2]);
// the Green's function is that of the last Majorana, whose blocks are
// multiples of the identity.  The shifted Hamiltonian is left in ham, as in
// the reference; e0 receives E0.
enum class NPoint { Otoc, Green };

template<RealScalar FpType, typename Impl, typename HamOpTy>
std::vector<std::complex<double>> npoint_block(NPoint kind, BlockDiag<HamOpTy> &ham,
                                               const typename Impl::template SumOps<FpType> &op,
                                               State<FpType, Impl> &s0, double tmax, int nsteps,
                                               double beta, int krylov_dim,
                                               const KrylovOptions &opt, KrylovStats &st,
                                               double *e0 = nullptr)
{
    if (nsteps < 0) ThrowException(InvalidArgument, "nsteps", "must not be negative");
    if (ham.nullLL || ham.nullRR)
        ThrowException(InvalidArgument, "Hamiltonian", "must have both parity blocks");
    FpType gnd;
    {
        State<FpType, Impl> g(s0.spin_chain_length());
        gnd = g.ground_state(ham.LL, krylov_dim, (FpType)0.1);
    }
    if (e0) *e0 = (double)gnd;
    ham.LL += (-gnd) * SpinOp<FpType>::identity();
    ham.RR += (-gnd) * SpinOp<FpType>::identity();
    double pre = kind == NPoint::Otoc ? beta / 8 : beta / 2;
    if (pre != 0.0) internal::add_stats(st, s0.evolve_ex(ham.LL, 0.0, pre, opt));
    s0.gc();
    std::vector<std::complex<double>> v;
    v.reserve(nsteps + 1);
    double dt = nsteps ? tmax / nsteps : 0.0;
    State<FpType, Impl> s;
    for (int k = 0; k <= nsteps; k++) {
        double t = k * dt;
        s = s0;
        if (kind == NPoint::Otoc) {
            s *= op;
            internal::add_stats(st, s.evolve_ex(ham.RR, t, beta / 4, opt));
            internal::add_stats(st, s.evolve_ex(ham.LL, -t, beta / 4, opt));
            s *= op;
            internal::add_stats(st, s.evolve_ex(ham.RR, t, beta / 4, opt));
            internal::add_stats(st, s.evolve_ex(ham.LL, -t, 0.0, opt));
            st.matvecs += 2;
        } else {
            internal::add_stats(st, s.evolve_ex(ham.RR, t, 0.0, opt));
            internal::add_stats(st, s.evolve_ex(ham.LL, -t, 0.0, opt));
        }
        v.push_back(s0.inner64(s));
    }
    std::complex<double> v0 = v[0];
    for (auto &x : v) x /= v0;
    return v;
}

// The trace definitions of the same functions (BaseEval::eval with trace
// set, and evolve_trace; examples/syk.cc:171-185 and 286-298), by exact
// diagonalization: with expmbH = exp(-beta H_L), m0 = exp(-it H_R),
// m1 = exp(it H_L) and m2 = op as a dense matrix,
//     OTOC:  v_k = Tr(m1 m0 m2 m1 m0 m2 expmbH)
//     Green: v_k = Tr(m1 m0 expmbH)
// each divided by v_0 (no E0 shift on this path, as in the reference).  The
// matrix exponentials come from the operators' matexp caches.  For sectors
// of up to 15 sites.
template<RealScalar FpType, typename HamOpTy, typename OpTy>
std::vector<std::complex<double>> npoint_block_trace(NPoint kind, const BlockDiag<HamOpTy> &ham,
                                                     const OpTy &op, double tmax, int nsteps,
                                                     double beta)
{
    if (nsteps < 0) ThrowException(InvalidArgument, "nsteps", "must not be negative");
    if (ham.nullLL || ham.nullRR)
        ThrowException(InvalidArgument, "Hamiltonian", "must have both parity blocks");
    const auto &expmbH = ham.LL.matexp((FpType)(-beta));
    const auto &m2 = op.get_matrix();
    std::vector<std::complex<double>> v;
    double dt = nsteps ? tmax / nsteps : 0.0;
    for (int k = 0; k <= nsteps; k++) {
        FpType t = (FpType)(k * dt);
        const auto &m0 = ham.RR.matexp(complex<FpType>{0, -t});
        const auto &m1 = ham.LL.matexp(complex<FpType>{0, t});
        complex<FpType> tr = kind == NPoint::Otoc ? (m1 * m0 * m2 * m1 * m0 * m2 * expmbH).trace()
                                                  : (m1 * m0 * expmbH).trace();
        v.push_back(std::complex<double>(tr));
    }
    std::complex<double> v0 = v[0];
    for (auto &x : v) x /= v0;
    return v;
}

// The reference's text outputs of the SYK driver: the Hamiltonian dump
// ("H_LL = <LL>" and "H_RR = <RR>", one line each, in SumOps' operator<<
// form; examples/syk.cc:429-438) and the result file (one "t Re Im" line
// per time point t = k dt, each number in an ostream's default format;
// examples/syk.cc:230-235).
template<typename HamOpTy>
void dump_block_ham(std::ostream &os, const BlockDiag<HamOpTy> &ham)
{
    os << "H_LL = " << ham.LL << std::endl;
    os << "H_RR = " << ham.RR << std::endl;
}

template<typename C>
void write_results(std::ostream &os, double dt, const std::vector<C> &v)
{
    for (size_t k = 0; k < v.size(); k++)
        os << dt * k << " " << v[k].real() << " " << v[k].imag() << std::endl;
}
