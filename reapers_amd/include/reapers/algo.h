// algo.h -- evolution algorithms on State (included inside
// namespace EvolutionAlgorithm by state.h, as in the reference).
//
// Krylov<Impl>::evolve computes s = exp(-(beta + i t) H) s by the restarted
// Arnoldi method of Eiermann and Ernst (SIAM J. Numer. Anal. 44 (2006)
// 2481-2504, Alg. 2), following the reference's Krylov::evolve
// (inc/algo.h:266-365) step by step: Arnoldi of dimension m from the
// normalized vector, H appended block-bidiagonally to Hfull with the previous
// restart's beta at (r*m, r*m-1), Hexp = exp((-beta - i t) Hfull), the answer
// updated with beta0 * Hexp[(r*m + i), 0] * V_i, convergence when the 2-norm
// of that column block drops below tol, restart from V_m.
//
// Differences that are deliberate:
//   - Gram-Schmidt: the reference subtracts one projection at a time with a
//     blocking dot product each (inc/algo.h:86-96); here each pass is one
//     fused sweep that forms all j+1 coefficients from the same vector
//     (classical Gram-Schmidt) in FP64 and applies them in a second sweep.
//     The reference's selective criterion decides whether to repeat a pass
//     (at most 3, eta = 1/sqrt(2): inc/algo.h:79-101), or, with
//     reorth = Reorth::Always, every step makes exactly two passes.
//   - Hfull, H, beta and the norms are FP64 for FP32 states as well.
//   - the pass that ends the Gram-Schmidt iteration also normalizes the new
//     vector when its norm follows from the coefficients without
//     cancellation (||w - V h||^2 = ||w||^2 - ||h||^2 >= ||w||^2 / 2), which
//     saves the separate scaling sweep.
//   - the optional sub-step splits t into equal pieces, each a separate
//     restarted evolution (the reference evolves the whole t at once).
#pragma once

template<typename Impl>
class Krylov {
    template<RealScalar FpType>
    using SumOps = typename Impl::template SumOps<FpType>;
    using cd = std::complex<double>;

    // iterated fused Gram-Schmidt of buf(j+1) against buf(0..j).  A pass
    // that ends the iteration (its norm at least eta times the previous one)
    // may normalize buf(j+1) in the same sweep; normalized reports it.
    // first(V, r, fuse_min, normalized) makes the first pass (by default the
    // plain sweep; the Arnoldi step passes the apply fused with it).
    template<RealScalar FpType, typename FirstPass>
    static bool orthogonalize(State<FpType, Impl> &s, dense::CMat &H, int j, double &beta,
                              Reorth mode, KrylovStats &st, bool &normalized,
                              FirstPass &&first) {
        const double eta = 1 / std::sqrt(2.0);
        const size_t dim = s.dim();
        std::vector<const typename Impl::template VecType<FpType> *> V(j + 1);
        for (int i = 0; i <= j; i++) V[i] = &s.vec(i);
        std::vector<cd> h(j + 1, cd(0)), r;
        double old_norm = 1.0, norm = eta / 2;
        int niter = 0;
        const int max_passes = mode == Reorth::Always ? 2 : 3;
        normalized = false;
        while (niter < max_passes && norm != 0 &&
               (mode == Reorth::Always || norm < eta * old_norm)) {
            niter++;
            // the norm this pass must reach to end the iteration
            double fuse_min = 0;
            if (mode == Reorth::Always) fuse_min = niter == max_passes ? 1e-300 : 0;
            else fuse_min = eta * norm;
            double nrm = niter == 1 ? first(V, r, fuse_min, normalized)
                                    : Impl::gs_pass(dim, V, s.buf(j + 1), r, fuse_min, normalized);
            for (int i = 0; i <= j; i++) h[i] += r[i];
            old_norm = norm;
            norm = nrm;
            st.gs_passes++;
            if (normalized) break;
        }
        beta = norm;
        for (int i = 0; i <= j; i++) H(i, j) = h[i];
        return norm != 0 && (normalized || norm >= eta * old_norm);
    }

    // Arnoldi steps j0 .. m-1 from the orthonormal buf(0..j0) (inc/algo.h:144-173;
    // j0 = 0 is the reference's m-step Arnoldi from the normalized buf(0))
    template<RealScalar FpType>
    static bool arnoldi(State<FpType, Impl> &s, const SumOps<FpType> &A, dense::CMat &H, int j0,
                        int &m, double &beta, Reorth mode, KrylovStats &st) {
        const size_t dim = s.dim();
        const auto len = s.spin_chain_length();
        for (int j = j0; j < m; j++) {
            bool normalized, ok;
            using VecList = std::vector<const typename Impl::template VecType<FpType> *>;
            bool fused = false;
            if constexpr (requires(VecList &V, std::vector<cd> &r, bool &nz) {
                              Impl::apply_gs(len, s.buf(j + 1), A, s.buf(j), V, r, 0.0, nz);
                              Impl::fused_first_pass();
                          })
                fused = Impl::fused_first_pass();
            if (fused) {
                // the apply and the first Gram-Schmidt pass in one sweep
                ok = orthogonalize(s, H, j, beta, mode, st, normalized,
                                   [&](const VecList &V, std::vector<cd> &r, double fuse_min,
                                       bool &nz) {
                                       if constexpr (requires {
                                                         Impl::apply_gs(len, s.buf(j + 1), A,
                                                                        s.buf(j), V, r, fuse_min,
                                                                        nz);
                                                     })
                                           return Impl::apply_gs(len, s.buf(j + 1), A, s.buf(j),
                                                                 V, r, fuse_min, nz);
                                       else
                                           return 0.0;
                                   });
            } else {
                Impl::apply_ops_overwrite(len, s.buf(j + 1), A, s.buf(j));
                ok = orthogonalize(s, H, j, beta, mode, st, normalized,
                                   [&](const VecList &V, std::vector<cd> &r, double fuse_min,
                                       bool &nz) {
                                       return Impl::gs_pass(dim, V, s.buf(j + 1), r, fuse_min, nz);
                                   });
            }
            st.matvecs++;
            if (!ok) {
                m = j + 1;
                return false;
            }
            if (j != m - 1) H(j + 1, j) = beta;
            if (!normalized) Impl::scale_vec(dim, s.buf(j + 1), (FpType)(1.0 / beta));
        }
        return true;
    }

    // One restarted evolution over (t_real, t_imag).  Restart r uses Krylov
    // dimension m_r: the options' krydim, or the schedule's proposal, made
    // after the restart's first Arnoldi step (which every dimension shares).
    // Hfull grows by blocks of those sizes.  A proposal is rejected -- and
    // the options' krydim used for the rest of the evolution -- when a
    // restart's update norm fails to shrink, the restarted method's
    // a-posteriori sign that the block was too small for the remaining time.
    template<RealScalar FpType>
    static void evolve_once(State<FpType, Impl> &s, const SumOps<FpType> &ops, double t_real,
                            double t_imag, const KrylovOptions &opt, RestartInfo &ri,
                            KrylovStats &st) {
        const double tol = ri.tol;
        const size_t dim = s.dim();
        int mmax = opt.krydim;
        if (opt.schedule) mmax = std::max(mmax, opt.max_krydim);
        s.enlarge(mmax + 2);
        const int ans = mmax + 1;          // buffer index of the answer
        Impl::zero_vec(dim, s.buf(ans));
        const double beta0 = Impl::vec_norm64(dim, s.buf());
        Impl::scale_vec(dim, s.buf(), (FpType)(1.0 / beta0));
        dense::CMat Hfull;
        int niters = 0, off = 0;
        double oldbeta = 0.0, lastnrm = 0.0;
        bool use_schedule = opt.schedule != nullptr;
        ri.restart = 0;
        ri.last_update = ri.prev_update = 0;
        ri.prev_krydim = 0;
        while (true) {
            niters++;
            dense::CMat H(mmax);
            double beta = 0;
            int m = 1;
            bool brokedown = !arnoldi(s, ops, H, 0, m, beta, opt.reorth, st);
            if (!brokedown) {
                int krydim = opt.krydim;
                if (use_schedule) {
                    ri.alpha0 = H(0, 0).real();
                    ri.beta0 = beta;
                    krydim = std::max(2, std::min(opt.schedule->krydim(ri), mmax));
                }
                m = krydim;
                H(1, 0) = beta;
                brokedown = !arnoldi(s, ops, H, 1, m, beta, opt.reorth, st);
            }
            H.grow(m);
            auto t0 = std::chrono::steady_clock::now();
            if (niters == 1) {
                Hfull = H;
            } else {
                Hfull.grow(off + m);
                Hfull(off, off - 1) = oldbeta;
                for (int i = 0; i < m; i++)
                    for (int j = 0; j < m; j++) Hfull(off + i, off + j) = H(i, j);
            }
            dense::CMat tH = Hfull;
            const cd z{-t_imag, -t_real};
            for (auto &v : tH.a) v *= z;
            dense::CMat E = dense::expm(tH);
            std::vector<cd> c(m);
            double nrm2 = 0;
            for (int i = 0; i < m; i++) {
                cd e = E(off + i, 0);
                c[i] = beta0 * e;
                nrm2 += std::norm(e);
            }
            st.host_ms += std::chrono::duration<double, std::milli>(
                              std::chrono::steady_clock::now() - t0).count();
            std::vector<const typename Impl::template VecType<FpType> *> V(m);
            for (int i = 0; i < m; i++) V[i] = &s.vec(i);
            Impl::lincomb(dim, s.buf(ans), V, c);
            const double nrm = std::sqrt(nrm2);
            st.last_nrm = nrm;
            if ((opt.max_iters && niters >= opt.max_iters) || brokedown || nrm < tol) {
                st.breakdown |= brokedown ? 1 : 0;
                // the answer becomes the current vector
                s.swap_bufs(0, ans);
                break;
            }
            if (use_schedule && niters > 1 && nrm >= lastnrm) {
                use_schedule = false;
                st.rejected++;
            }
            lastnrm = nrm;
            // restart from v_m
            s.swap_bufs(0, m);
            off += m;
            oldbeta = beta;
            ri.restart = niters;
            ri.prev_update = ri.last_update;
            ri.last_update = nrm;
            ri.prev_krydim = m;
        }
        st.restarts += niters;
    }

public:
    // exp(-(t_imag + i t_real) H) s with the reference's argument list
    // (inc/algo.h:266-269)
    template<RealScalar FpType>
    static void evolve(State<FpType, Impl> &s, const SumOps<FpType> &ops, FpType t_real,
                       FpType t_imag = 0.0, int krydim = 5, FpType tol = epsilon<FpType>(),
                       int max_iters = 0) {
        KrylovOptions opt;
        opt.krydim = krydim;
        opt.tol = (double)tol;
        opt.max_iters = max_iters;
        KrylovStats st;
        evolve_ex(s, ops, (double)t_real, (double)t_imag, opt, st);
    }

    template<RealScalar FpType>
    static void evolve_ex(State<FpType, Impl> &s, const SumOps<FpType> &ops, double t_real,
                          double t_imag, const KrylovOptions &opt, KrylovStats &st) {
        if (opt.krydim < 2) ThrowException(InvalidArgument, "krydim", "must be at least two");
        // the last Arnoldi step's Gram-Schmidt sweep and the answer update
        // each take krydim basis vectors in one fused sweep
        if (opt.krydim > RP_MAX_SWEEP_VECS)
            ThrowException(InvalidArgument, "krydim", "must be at most 64");
        if (opt.schedule && opt.max_krydim > RP_MAX_SWEEP_VECS)
            ThrowException(InvalidArgument, "max_krydim", "must be at most 64");
        RestartInfo ri;
        ri.len = s.spin_chain_length();
        ri.nterms = (long)ops.size();
        double s2 = 0, s1 = 0;
        for (auto &op : ops) {
            double a = std::abs(std::complex<double>(op.get_coeff()));
            s1 += a;
            s2 += a * a;
        }
        ri.h_rms = std::sqrt(s2);
        ri.h_l1 = s1;
        if (opt.schedule) ri.apply_passes = Impl::apply_passes(ri.len, ops);
        ri.t = t_real;
        ri.beta = t_imag;
        ri.tol = opt.tol > 0 ? opt.tol : (double)epsilon<FpType>();
        int nsub = 1;
        if (opt.schedule) {
            ri.tau = t_real;
            ri.taub = t_imag;
            nsub = std::max(1, std::min(64, opt.schedule->substeps(ri)));
        } else if (opt.substep > 0) {
            double span = std::max(std::abs(t_real), std::abs(t_imag));
            nsub = std::max(1, (int)std::ceil(span / opt.substep - 1e-12));
        }
        ri.nsub = nsub;
        ri.tau = t_real / nsub;
        ri.taub = t_imag / nsub;
        for (int k = 0; k < nsub; k++) {
            ri.substep = k;
            evolve_once(s, ops, t_real / nsub, t_imag / nsub, opt, ri, st);
        }
        st.substeps += nsub;
    }

    // Lanczos ground state (inc/algo.h:180-257), starting from the state's
    // current vector (the reference starts from a Haar random state)
    template<RealScalar FpType>
    static FpType ground_state(State<FpType, Impl> &s, const SumOps<FpType> &ham,
                               int krydim = 5, FpType eps = 10 * epsilon<FpType>(),
                               int maxiter = 0, bool randomize = true) {
        const size_t dim = s.dim();
        const auto len = s.spin_chain_length();
        if (krydim < 2) ThrowException(InvalidArgument, "krydim", "must be at least two");
        s.enlarge(krydim + 1);
        if (randomize) s.random_state();
        std::vector<double> alpha(krydim), beta(krydim > 1 ? krydim - 1 : 1);
        int numiter = 0;
        double energy = 0;
        while (true) {
            for (int i = 0; i < krydim; i++) {
                Impl::apply_ops_overwrite(len, s.buf(i + 1), ham, s.buf(i));
                alpha[i] = Impl::vec_prod64(dim, s.buf(i), s.buf(i + 1)).real();
                if (i == krydim - 1) break;
                Impl::add_vec(dim, s.buf(i + 1), (FpType)(-alpha[i]), s.buf(i));
                if (i != 0) Impl::add_vec(dim, s.buf(i + 1), (FpType)(-beta[i - 1]), s.buf(i - 1));
                beta[i] = Impl::vec_norm64(dim, s.buf(i + 1));
                Impl::scale_vec(dim, s.buf(i + 1), (FpType)(1.0 / beta[i]));
            }
            std::vector<double> w, Q;
            dense::tridiag_eigen(alpha, beta, w, Q);
            int idx = 0;
            for (int i = 1; i < krydim; i++) if (w[i] < w[idx]) idx = i;
            energy = w[idx];
            std::vector<cd> c(krydim);
            for (int i = 0; i < krydim; i++) c[i] = Q[(size_t)idx * krydim + i];
            std::vector<const typename Impl::template VecType<FpType> *> V(krydim);
            for (int i = 0; i < krydim; i++) V[i] = &s.vec(i);
            Impl::zero_vec(dim, s.buf(krydim));
            Impl::lincomb(dim, s.buf(krydim), V, c);
            s.inc_curbuf(krydim);
            s.normalize();
            numiter++;
            if (!(beta[0] > (double)eps && (!maxiter || numiter < maxiter))) break;
        }
        return (FpType)energy;
    }
};

// Evolution and ground state by exact diagonalization (the reference's
// ExactDiagonalization, inc/algo.h:28-54): s = exp(-(beta + i t) H) s as one
// product with the cached matrix exponential, and the ground state as the
// eigenvector of the lowest eigenvalue (the first of the ascending list, as
// minCoeff returns the first minimum).  For chains of up to 15 sites.
template<typename Impl>
class ExactDiagonalization {
    template<RealScalar FpType>
    using SumOps = typename Impl::template SumOps<FpType>;

public:
    template<RealScalar FpType>
    static FpType ground_state(State<FpType, Impl> &s, const SumOps<FpType> &ham) {
        auto len = s.spin_chain_length();
        auto &&es = ham.get_eigensystem(len);
        const auto &vals = std::get<0>(es);
        size_t idx = std::min_element(vals.begin(), vals.end()) - vals.begin();
        Impl::copy_vec(s.dim(), s.buf(), std::get<1>(es), idx);
        return vals[idx];
    }

    template<RealScalar FpType, typename... Args>
    static void evolve(State<FpType, Impl> &s, const SumOps<FpType> &ops, FpType t,
                       FpType beta = 0.0, Args &&...) {
        s.enlarge(2);
        const auto &m = ops.matexp(complex<FpType>{-beta, -t}, s.spin_chain_length());
        Impl::mat_mul(s.dim(), s.buf(1), m, s.buf());
        s.inc_curbuf(1);
    }
};
