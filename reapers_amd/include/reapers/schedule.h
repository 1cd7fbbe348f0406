// schedule.h -- Krylov schedules beyond the reference's fixed one (the
// reference's Krylov::evolve uses one dimension for every restart and the
// whole t at once, inc/algo.h:266-365).
//
// ListSchedule:  a given number of sub-steps and a list of Krylov dimensions,
//                one per restart (the last one repeats).
// MLPSchedule:   a trained 12 -> 64 -> 64 -> 2 network proposes log2 of the
//                Krylov dimension of each restart and log2 of the number of
//                sub-steps, evaluated on the GPU by rp_mlp_infer.  The
//                evolution's a-posteriori check (an update norm that fails to
//                shrink) rejects its proposals and falls back to the fixed
//                dimension.
#pragma once

#ifndef _REAPERS_H_
#error "Do not include this file directly. Include the master header file reapers.h"
#endif

class ListSchedule : public KrylovSchedule {
    std::vector<int> dims;
    int nsub;

public:
    ListSchedule(std::vector<int> dims, int nsub = 1) : dims(std::move(dims)), nsub(nsub) {
        if (this->dims.empty()) ThrowException(InvalidArgument, "dims", "cannot be empty");
    }
    int substeps(const RestartInfo &) override { return nsub; }
    int krydim(const RestartInfo &ri) override {
        return dims[std::min<size_t>(ri.restart, this->dims.size() - 1)];
    }
};

// The 12 inputs of the proposer network.  Scale-free where possible so that a
// network trained on small chains transfers to large ones: products of the
// norm estimates with the (sub-)step length, logs of tolerances and update
// norms, the apply's cost relative to one vector pass.  for_substeps selects
// the evolution-level decision (no restart information yet).
inline void proposer_features(const RestartInfo &ri, bool for_substeps,
                              float f[RP_MLP_IN]) {
    const double tau = for_substeps ? std::hypot(ri.t, ri.beta) : std::hypot(ri.tau, ri.taub);
    const double tol = ri.tol > 0 ? ri.tol : 1e-16;
    f[0] = (float)(std::log10(tol) / 8);
    f[1] = (float)(std::log1p(ri.h_l1 * tau) / 5);
    f[2] = (float)(std::log1p(ri.h_rms * tau) / 5);
    f[3] = (float)(ri.len / 32.0);
    f[4] = (float)(std::log((double)std::max(1L, ri.nterms)) / 10);
    f[5] = (float)(std::log(std::max(1.0, ri.apply_passes)) / 5);
    if (for_substeps) {
        f[6] = -0.25f;
        f[7] = 1.0f;
        f[8] = f[9] = f[10] = f[11] = 0.0f;
        return;
    }
    f[6] = (float)(ri.restart / 8.0);
    f[7] = ri.last_update > 0 ? (float)(std::log10(ri.last_update / tol) / 8) : 1.0f;
    f[8] = (ri.last_update > 0 && ri.prev_update > 0)
               ? (float)(std::log10(ri.last_update / ri.prev_update) / 8) : 0.0f;
    f[9] = (float)(ri.prev_krydim / 64.0);
    f[10] = (float)(ri.alpha0 / std::max(ri.h_rms, 1e-300));
    f[11] = (float)(std::log1p(std::abs(ri.beta0) * tau) / 5);
}

#ifndef REAPERS_NOGPU
class MLPSchedule : public KrylovSchedule {
    float *params = nullptr;   // device copy of the RP_MLP_NPARAMS parameters
    int min_dim, max_dim, max_sub;

    float eval(const RestartInfo &ri, bool sub, int o) const {
        float x[RP_MLP_IN], y[RP_MLP_OUT];
        proposer_features(ri, sub, x);
        RP_CALL(rp_mlp_infer(params, x, 1, y));
        return y[o];
    }

public:
    // params: RP_MLP_NPARAMS host floats
    MLPSchedule(const float *host_params, int min_dim = 3, int max_dim = 40, int max_sub = 8)
        : min_dim(min_dim), max_dim(max_dim), max_sub(max_sub) {
        if (min_dim < 2 || max_dim < min_dim || max_sub < 1)
            ThrowException(InvalidArgument, "bounds", "need 2 <= min_dim <= max_dim, max_sub >= 1");
        void *p = nullptr;
        RP_CALL(rp_malloc(&p, sizeof(float) * RP_MLP_NPARAMS));
        params = (float *)p;
        RP_CALL(rp_memcpy(params, host_params, sizeof(float) * RP_MLP_NPARAMS, RP_H2D));
    }
    MLPSchedule(const MLPSchedule &) = delete;
    MLPSchedule &operator=(const MLPSchedule &) = delete;
    ~MLPSchedule() override {
        if (params) rp_free(params);
    }

    int max_krydim() const { return max_dim; }

    int substeps(const RestartInfo &ri) override {
        double y = eval(ri, true, 1);
        long n = std::lround(std::exp2(std::max(0.0, std::min(y, 10.0))));
        return (int)std::max(1L, std::min<long>(n, max_sub));
    }
    int krydim(const RestartInfo &ri) override {
        double y = eval(ri, false, 0);
        long m = std::lround(std::exp2(std::max(1.0, std::min(y, 10.0))));
        return (int)std::max<long>(min_dim, std::min<long>(m, max_dim));
    }
};
#endif  // REAPERS_NOGPU
