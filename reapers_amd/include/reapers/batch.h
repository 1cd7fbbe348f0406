// batch.h -- many right-hand sides evolved together: nprob Hamiltonians of
// uneven size (disorder realizations), nrhs states each, in the block-state
// layout of libreapers_hip.so's batch entry points.  No counterpart in the
// reference, which evolves one State at a time (inc/state.h:413-424); per
// right-hand side the arithmetic is the reference's restarted Krylov
// (inc/algo.h:266-365): Arnoldi from the normalized state, the Hessenberg
// block appended to Hfull with the previous restart's beta at
// (off, off - 1), Hexp = exp(-(beta + i t) Hfull), the answer updated with
// beta0 Hexp[off + i, 0] V_i, convergence when that column block's 2-norm
// drops below tol, restart from V_m.
//
// Differences from the single-state evolution: every Arnoldi step makes two
// Gram-Schmidt passes (a fixed kernel sequence, so that a restart's m steps
// are one captured HIP graph replayed for every restart); the small-matrix
// exponentials of all right-hand sides run on the host in parallel between
// restarts; a right-hand side that has converged stops contributing, and a
// problem whose states have all converged is skipped by every kernel.
#pragma once

#ifndef _REAPERS_H_
#error "Do not include this file directly. Include the master header file reapers.h"
#endif

template<RealScalar FpType>
class BatchedStates {
    rp_batch_t b = nullptr;
    rp_batch_info_t info{};
    std::vector<uint64_t> off;
    std::vector<int> lens;
    int nprob = 0, ns = 1, kmax = 0;
    std::vector<void *> bufs;   // bufs[0]: the states; then Krylov vectors, answer
    double *H_dev = nullptr, *nrm_dev = nullptr, *c_dev = nullptr;

    static constexpr int prec = std::is_same_v<FpType, float> ? RP_C64 : RP_C128;

    void *alloc_vec() {
        void *p = nullptr;
        RP_CALL(rp_malloc(&p, info.vec_bytes));
        return p;
    }
    void enlarge(int n) {
        while ((int)bufs.size() < n) bufs.push_back(alloc_vec());
    }

public:
    // hams[p] acts on lens[p] qubits; kmax bounds the Krylov dimension; flags
    // as rp_batch_create (RP_BATCH_AUTOTUNE, RP_BATCH_ACC_GROUP, RP_BATCH_WIDE,
    // RP_BATCH_PLANES, RP_BATCH_TILE_BITS(b))
    BatchedStates(const std::vector<const HostSumOps<FpType> *> &hams, std::vector<int> lens_,
                  int nrhs, int kmax_, int flags = RP_BATCH_AUTOTUNE)
        : lens(std::move(lens_)), nprob((int)hams.size()), ns(nrhs), kmax(kmax_) {
        if ((int)lens.size() != nprob)
            ThrowException(InvalidArgument, "lens", "must have one entry per Hamiltonian");
        std::vector<u64> bits;
        std::vector<double> coeff;
        std::vector<int64_t> toff(1, 0);
        for (auto *h : hams) {
            std::vector<u64> b1;
            std::vector<double> c1;
            h->export_terms(b1, c1);
            double s1 = 0, s2 = 0;
            for (size_t q = 0; q < b1.size(); q++) {
                double a = std::abs(std::complex<double>(c1[2 * q], c1[2 * q + 1]));
                s1 += a;
                s2 += a * a;
            }
            pterms.push_back((long)b1.size());
            ph_l1.push_back(s1);
            ph_rms.push_back(std::sqrt(s2));
            bits.insert(bits.end(), b1.begin(), b1.end());
            coeff.insert(coeff.end(), c1.begin(), c1.end());
            toff.push_back((int64_t)bits.size());
        }
        RP_CALL(rp_batch_create(&b, prec, nprob, lens.data(), bits.data(), coeff.data(),
                                toff.data(), nrhs, kmax, flags));
        RP_CALL(rp_batch_info(b, &info));
        off.resize(nprob + 1);
        RP_CALL(rp_batch_offsets(b, off.data()));
        std::vector<uint64_t> groups(nprob);
        RP_CALL(rp_batch_groups(b, groups.data()));
        for (int p = 0; p < nprob; p++) ppasses.push_back((double)groups[p] + 1.0);
        const size_t nr = (size_t)nprob * ns;
        void *p;
        RP_CALL(rp_malloc(&p, nr * (kmax + 1) * kmax * 2 * sizeof(double)));
        H_dev = (double *)p;
        RP_CALL(rp_malloc(&p, nr * sizeof(double)));
        nrm_dev = (double *)p;
        RP_CALL(rp_malloc(&p, nr * 2 * kmax * sizeof(double)));
        c_dev = (double *)p;
        bufs.push_back(alloc_vec());
        RP_CALL(rp_zero_vec(prec, bufs[0], info.nbasis * ns));
    }
    BatchedStates(const BatchedStates &) = delete;
    BatchedStates &operator=(const BatchedStates &) = delete;
    ~BatchedStates() {
        for (auto &g : graphs) rp_graph_free(g.second);
        for (void *p : bufs) rp_free(p);
        for (void *p : {(void *)H_dev, (void *)nrm_dev, (void *)c_dev})
            if (p) rp_free(p);
        if (b) rp_batch_free(b);
    }

    const rp_batch_info_t &layout() const { return info; }
    const std::vector<uint64_t> &offsets() const { return off; }
    size_t size() const { return info.nbasis * ns; }   // complex elements

    void load(const complex<FpType> *host) { RP_CALL(rp_memcpy(bufs[0], host, info.vec_bytes, RP_H2D)); }
    void store(complex<FpType> *host) const { RP_CALL(rp_memcpy(host, bufs[0], info.vec_bytes, RP_D2H)); }

    // states = H states (every problem)
    void apply() {
        enlarge(2);
        std::vector<int> act(nprob, 1);
        RP_CALL(rp_batch_set_active(b, act.data()));
        RP_CALL(rp_batch_apply(b, bufs[1], bufs[0], 0));
        std::swap(bufs[0], bufs[1]);
    }

    // states = exp(-(beta + i t) H_p) states by restarted Krylov.  Without a
    // schedule every restart has dimension m; with one, the sub-step count is
    // the largest any state asks for, and each restart's dimension the
    // largest the active states propose after the restart's first Arnoldi
    // step (bounded by kmax); a restart whose update norm fails to shrink for
    // some state rejects the schedule for the rest of the sub-step, which
    // then continues with m.  restarts[rhs] (optional) receives the restarts
    // each right-hand side needed.
    KrylovStats evolve(double t, double beta, int m, double tol, int max_iters = 0,
                       bool use_graph = true, std::vector<int> *restarts = nullptr,
                       KrylovSchedule *sched = nullptr) {
        if (m < 2 || m > kmax) ThrowException(InvalidArgument, "m", "must be in [2, kmax]");
        const int nr = nprob * ns;
        KrylovStats st;
        std::vector<int> rs(nr, 0);
        int nsub = 1;
        if (sched) {
            for (int p = 0; p < nprob; p++) {
                RestartInfo ri = base_info(p, t, beta, tol);
                ri.tau = t;
                ri.taub = beta;
                nsub = std::max(nsub, std::min(64, sched->substeps(ri)));
            }
        }
        for (int k = 0; k < nsub; k++)
            evolve_once(t / nsub, beta / nsub, t, beta, m, tol, max_iters, use_graph, sched, k,
                        nsub, st, rs);
        st.substeps = nsub;
        if (restarts) *restarts = rs;
        return st;
    }

private:
    std::map<std::pair<int, int>, rp_graph_t> graphs;   // Arnoldi steps [j0, j1)
    std::vector<long> pterms;
    std::vector<double> ph_rms, ph_l1, ppasses;

    RestartInfo base_info(int p, double t, double beta, double tol) const {
        RestartInfo ri;
        ri.len = lens[p];
        ri.nterms = pterms[p];
        ri.h_rms = ph_rms[p];
        ri.h_l1 = ph_l1[p];
        ri.apply_passes = ppasses[p];
        ri.t = t;
        ri.beta = beta;
        ri.tol = tol;
        return ri;
    }

    void run_steps(int j0, int j1, bool use_graph) {
        if (!use_graph) {
            for (int j = j0; j < j1; j++) RP_CALL(rp_batch_arnoldi_step(b, j, H_dev));
            return;
        }
        auto key = std::make_pair(j0, j1);
        auto it = graphs.find(key);
        if (it == graphs.end()) {
            rp_graph_t g = nullptr;
            RP_CALL(rp_batch_capture_arnoldi(b, j0, j1, H_dev, &g));
            it = graphs.emplace(key, g).first;
        }
        RP_CALL(rp_graph_launch(it->second));
    }

    void evolve_once(double t, double beta, double t_all, double beta_all, int m_fixed,
                     double tol, int max_iters, bool use_graph, KrylovSchedule *sched, int substep,
                     int nsub, KrylovStats &st, std::vector<int> &rs) {
        using cd = std::complex<double>;
        const int nr = nprob * ns;
        const int mmax = sched ? kmax : m_fixed;
        // buffers: V_0 .. V_mmax at bufs[1 .. mmax+1], the answer at bufs[mmax+2]
        enlarge(mmax + 3);
        std::vector<void *> V(bufs.begin() + 1, bufs.begin() + mmax + 2);
        void *ans = bufs[mmax + 2];
        RP_CALL(rp_memcpy(V[0], bufs[0], info.vec_bytes, RP_D2D));
        RP_CALL(rp_batch_set_basis(b, V.data(), mmax + 1));
        std::vector<int> act(nprob, 1);
        RP_CALL(rp_batch_set_active(b, act.data()));
        RP_CALL(rp_batch_normalize(b, 0, nrm_dev));
        std::vector<double> beta0(nr);
        RP_CALL(rp_memcpy(beta0.data(), nrm_dev, nr * sizeof(double), RP_D2H));
        RP_CALL(rp_zero_vec(prec, ans, info.nbasis * ns));
        std::vector<char> done(nr, 0);
        for (int r = 0; r < nr; r++) done[r] = beta0[r] == 0;
        std::vector<dense::CMat> Hfull(nr);
        std::vector<double> oldbeta(nr, 0.0), last(nr, 0.0), prev(nr, 0.0);
        const size_t hstride = (size_t)2 * (kmax + 1) * kmax;
        std::vector<double> Hh(hstride * nr), ch((size_t)2 * kmax * nr);
        auto Hget = [&](int r, int i, int j) {
            const double *q = Hh.data() + hstride * r + 2 * ((size_t)j * (kmax + 1) + i);
            return cd(q[0], q[1]);
        };
        int niters = 0, off_blk = 0, mprev = 0;
        bool use_sched = sched != nullptr;
        const cd z{-beta, -t};
        while (true) {
            niters++;
            for (int p = 0; p < nprob; p++) {
                act[p] = 0;
                for (int s = 0; s < ns; s++) act[p] |= !done[p * ns + s];
            }
            RP_CALL(rp_batch_set_active(b, act.data()));
            int m = m_fixed;
            if (use_sched) {
                // first Arnoldi step, then the proposals from alpha0, beta0
                run_steps(0, 1, use_graph);
                RP_CALL(rp_memcpy(Hh.data(), H_dev, Hh.size() * sizeof(double), RP_D2H));
                m = 2;
                for (int r = 0; r < nr; r++) {
                    if (done[r]) continue;
                    RestartInfo ri = base_info(r / ns, t_all, beta_all, tol);
                    ri.tau = t;
                    ri.taub = beta;
                    ri.substep = substep;
                    ri.nsub = nsub;
                    ri.restart = niters - 1;
                    ri.last_update = last[r];
                    ri.prev_update = prev[r];
                    ri.prev_krydim = mprev;
                    ri.alpha0 = Hget(r, 0, 0).real();
                    ri.beta0 = Hget(r, 1, 0).real();
                    m = std::max(m, sched->krydim(ri));
                }
                m = std::min(m, kmax);
                run_steps(1, m, use_graph);
            } else {
                run_steps(0, m, use_graph);
            }
            for (int p = 0; p < nprob; p++)
                if (act[p]) st.matvecs += (long)m * ns;
            st.gs_passes += 2L * m;
            RP_CALL(rp_memcpy(Hh.data(), H_dev, Hh.size() * sizeof(double), RP_D2H));
            auto t0 = std::chrono::steady_clock::now();
            std::fill(ch.begin(), ch.end(), 0.0);
            std::vector<char> now_done(done);
            std::vector<double> nrm(nr, 0.0);
            const double eta = 1 / std::sqrt(2.0);
#pragma omp parallel for schedule(dynamic)
            for (int r = 0; r < nr; r++) {
                if (done[r]) continue;
                // Arnoldi breakdown at step j (the reference's test after its
                // passes, inc/algo.h:79-106): the second pass left less than
                // eta of the first pass's norm, or nothing.  The Krylov space
                // is then invariant, the block is cut to j + 1 columns and
                // the answer is exact (inc/algo.h:356).
                int mr = m;
                bool brk = false;
                for (int j = 0; j < m && !brk; j++) {
                    cd s = Hget(r, j + 1, j);
                    if (s.real() == 0.0 || s.real() < eta * s.imag()) {
                        brk = true;
                        mr = j + 1;
                    }
                }
                dense::CMat &F = Hfull[r];
                F.grow(off_blk + mr);
                if (off_blk) F(off_blk, off_blk - 1) = oldbeta[r];
                for (int i = 0; i < mr; i++)
                    for (int j = 0; j < mr; j++)
                        F(off_blk + i, off_blk + j) =
                            i == j + 1 ? cd(Hget(r, i, j).real()) : i <= j ? Hget(r, i, j) : 0.0;
                dense::CMat tH = F;
                for (auto &v : tH.a) v *= z;
                dense::CMat E = dense::expm(tH);
                double nrm2 = 0;
                for (int i = 0; i < mr; i++) {
                    cd e = E(off_blk + i, 0);
                    cd c = beta0[r] * e;
                    ch[(size_t)2 * kmax * r + 2 * i] = c.real();
                    ch[(size_t)2 * kmax * r + 2 * i + 1] = c.imag();
                    nrm2 += std::norm(e);
                }
                nrm[r] = std::sqrt(nrm2);
                oldbeta[r] = Hget(r, m, m - 1).real();
                if (brk || nrm[r] < tol || (max_iters && niters >= max_iters)) now_done[r] = 1;
            }
            st.host_ms += std::chrono::duration<double, std::milli>(
                              std::chrono::steady_clock::now() - t0).count();
            RP_CALL(rp_memcpy(c_dev, ch.data(), ch.size() * sizeof(double), RP_H2D));
            RP_CALL(rp_batch_lincomb(b, ans, m, c_dev, 2 * kmax));
            bool all = true, grew = false;
            for (int r = 0; r < nr; r++) {
                if (done[r]) continue;
                rs[r]++;
                if (niters > 1 && nrm[r] >= last[r]) grew = true;
                prev[r] = last[r];
                last[r] = nrm[r];
            }
            done = now_done;
            for (char d : done) all = all && d;
            if (all) break;
            if (use_sched && grew) {
                use_sched = false;
                st.rejected++;
            }
            off_blk += m;
            mprev = m;
            st.restarts++;
            // restart from V_m: the graphs' pointers stay fixed, so copy
            RP_CALL(rp_memcpy(V[0], V[m], info.vec_bytes, RP_D2D));
        }
        st.restarts++;
        std::swap(bufs[0], bufs[mmax + 2]);
        RP_CALL(rp_synchronize());
    }

public:
};
