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

template<RealScalar FpType, typename Impl>
std::vector<std::complex<double>> otoc(const typename Impl::template SumOps<FpType> &ham,
                                       const typename Impl::template SumOps<FpType> &A,
                                       const typename Impl::template SumOps<FpType> &B,
                                       State<FpType, Impl> &init,
                                       const std::vector<double> &times, double beta,
                                       const KrylovOptions &opt, KrylovStats &st)
{
    auto add = [&st](const KrylovStats &s) {
        st.matvecs += s.matvecs;
        st.restarts += s.restarts;
        st.gs_passes += s.gs_passes;
        st.breakdown |= s.breakdown;
        st.substeps += s.substeps;
        st.rejected += s.rejected;
        st.host_ms += s.host_ms;
        st.last_nrm = s.last_nrm;
    };
    if (beta != 0.0) {
        add(init.evolve_ex(ham, 0.0, beta / 8, opt));
        init.gc();
    }
    std::vector<std::complex<double>> v;
    v.reserve(times.size());
    State<FpType, Impl> psi;
    for (double t : times) {
        psi = init;
        psi *= B;
        add(psi.evolve_ex(ham, t, beta / 4, opt));
        psi *= A;
        add(psi.evolve_ex(ham, -t, beta / 4, opt));
        psi *= B;
        add(psi.evolve_ex(ham, t, beta / 4, opt));
        psi *= A;
        add(psi.evolve_ex(ham, -t, 0.0, opt));
        v.push_back(init.inner64(psi));
        st.matvecs += 4;   // the four operator applications
    }
    return v;
}
