// host_api.cc -- libreapers_host.so: the header-only C++ API instantiated
// for callers that cannot include C++ headers (Python's ctypes in the tests
// and the benchmark).  Every function catches the API's exceptions and
// returns a status; rph_last_error() gives the exception's message.
//
// Handles: a Hamiltonian handle owns a REAPERS::SumOps<Fp> (term list on the
// host plus its uploaded device layout); a state handle owns a
// REAPERS::State<Fp, GPUImpl>.  Fp is float for fp == 32, double for 64.
#include "reapers/reapers.h"

#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

using namespace REAPERS;
using namespace REAPERS::Model;

namespace {

thread_local std::string g_err;

template<typename F>
int guarded(F &&f)
{
    try {
        f();
        return 0;
    } catch (const DevOutOfMem &e) {
        g_err = e.what();
        return RP_ERR_OUT_OF_MEMORY;
    } catch (const InvalidArgument &e) {
        g_err = e.what();
        return RP_ERR_INVALID_ARGUMENT;
    } catch (const IndexTooLarge &e) {
        g_err = e.what();
        return RP_ERR_INVALID_ARGUMENT;
    } catch (const NoGpuDevice &e) {
        g_err = e.what();
        return RP_ERR_NO_DEVICE;
    } catch (const std::exception &e) {
        g_err = e.what();
        return RP_ERR_RUNTIME;
    } catch (...) {
        g_err = "unknown exception";
        return RP_ERR_RUNTIME;
    }
}

struct Ham {
    int fp;
    int len;
    SumOps<float> *f = nullptr;
    SumOps<double> *d = nullptr;
};

struct St {
    int fp;
    State<float, GPUImpl> *f = nullptr;
    State<double, GPUImpl> *d = nullptr;
};

template<RealScalar Fp>
int64_t export_ham(const HostSumOps<Fp> &h, uint64_t *bits, double *coeff, int64_t cap)
{
    int64_t n = (int64_t)h.size();
    if (n > cap) return -n;
    int64_t i = 0;
    for (auto &op : h) {
        bits[i] = op.get_bit_string();
        coeff[2 * i] = (double)op.get_coeff().real();
        coeff[2 * i + 1] = (double)op.get_coeff().imag();
        i++;
    }
    return n;
}

}  // namespace

extern "C" {

typedef struct {
    int64_t matvecs, restarts, gs_passes;
    int32_t breakdown, substeps;
    double last_nrm, host_ms, wall_ms;
    int32_t rejected, reserved;
} rph_stats;

const char *rph_last_error(void) { return g_err.c_str(); }

// SYK::gen_ham(std::mt19937(seed), regularize) in precision fp
int rph_syk_gen(int fp, int N, double sp, double J, int std_gamma, uint32_t seed, int regularize,
                uint64_t *bits, double *coeff, int64_t cap, int64_t *nterms)
{
    return guarded([&] {
        std::mt19937 rg(seed);
        if (fp == 32) {
            SYK<float, GPUImpl> syk(N, (float)sp, (float)J, std_gamma != 0);
            auto ham = syk.gen_ham(rg, regularize != 0);
            *nterms = export_ham<float>(ham, bits, coeff, cap);
        } else {
            SYK<double, GPUImpl> syk(N, sp, J, std_gamma != 0);
            auto ham = syk.gen_ham(rg, regularize != 0);
            *nterms = export_ham<double>(ham, bits, coeff, cap);
        }
    });
}

// XXZ(L, J, Delta, W).gen_ham(std::mt19937(seed))
int rph_xxz_gen(int fp, int L, double J, double Delta, double W, uint32_t seed, uint64_t *bits,
                double *coeff, int64_t cap, int64_t *nterms)
{
    return guarded([&] {
        std::mt19937 rg(seed);
        if (fp == 32) {
            XXZ<float, GPUImpl> m(L, (float)J, (float)Delta, (float)W);
            *nterms = export_ham<float>(m.gen_ham(rg), bits, coeff, cap);
        } else {
            XXZ<double, GPUImpl> m(L, J, Delta, W);
            *nterms = export_ham<double>(m.gen_ham(rg), bits, coeff, cap);
        }
    });
}

// FermionOpNonBlockForm(N, n, std_gamma) as one term
int rph_majorana(int fp, int N, int n, int std_gamma, uint64_t *bits, double *coeff)
{
    return guarded([&] {
        if (fp == 32) {
            FermionOpNonBlockForm<float, GPUImpl> op(N, n, std_gamma != 0);
            export_ham<float>(op, bits, coeff, 1);
        } else {
            FermionOpNonBlockForm<double, GPUImpl> op(N, n, std_gamma != 0);
            export_ham<double>(op, bits, coeff, 1);
        }
    });
}

// product of two term lists (HostSumOps::operator*), merged; for tests
int rph_ops_mul(int fp, const uint64_t *b0, const double *c0, int64_t n0, const uint64_t *b1,
                const double *c1, int64_t n1, uint64_t *bo, double *co, int64_t cap,
                int64_t *nout)
{
    return guarded([&] {
        if (fp == 32) {
            auto a = HostSumOps<float>::from_terms(b0, c0, n0);
            auto b = HostSumOps<float>::from_terms(b1, c1, n1);
            *nout = export_ham<float>(a * b, bo, co, cap);
        } else {
            auto a = HostSumOps<double>::from_terms(b0, c0, n0);
            auto b = HostSumOps<double>::from_terms(b1, c1, n1);
            *nout = export_ham<double>(a * b, bo, co, cap);
        }
    });
}

// raw <random> streams from std::mt19937(seed), as the SYK generators draw
// them (for comparing against other implementations of the same algorithms)
int rph_random_stream(int kind, int fp, uint32_t seed, double param_a, double param_b,
                      double *out, int64_t n)
{
    return guarded([&] {
        std::mt19937 g(seed);
        switch (kind) {
        case 0: {  // normal(0, 1)
            if (fp == 32) {
                std::normal_distribution<float> d(0.0, 1.0);
                for (int64_t i = 0; i < n; i++) out[i] = d(g);
            } else {
                std::normal_distribution<double> d(0.0, 1.0);
                for (int64_t i = 0; i < n; i++) out[i] = d(g);
            }
            break;
        }
        case 1: {  // bernoulli(param_a)
            std::bernoulli_distribution d(param_a);
            for (int64_t i = 0; i < n; i++) out[i] = d(g) ? 1.0 : 0.0;
            break;
        }
        case 2: {  // uniform_int(param_a, param_b)
            std::uniform_int_distribution<int> d((int)param_a, (int)param_b);
            for (int64_t i = 0; i < n; i++) out[i] = d(g);
            break;
        }
        case 3: {  // generate_canonical
            for (int64_t i = 0; i < n; i++)
                out[i] = fp == 32 ? (double)std::generate_canonical<float, 24>(g)
                                  : std::generate_canonical<double, 53>(g);
            break;
        }
        default:
            ThrowException(InvalidArgument, "kind", "must be 0..3");
        }
    });
}

int rph_init(int device)
{
    return guarded([&] {
        int st = rp_init(device);
        if (st) internal::throw_status(__FILE__, __LINE__, __func__, st);
    });
}

void *rph_ham_new(int fp, int len, const uint64_t *bits, const double *coeff, int64_t n)
{
    Ham *h = nullptr;
    int st = guarded([&] {
        h = new Ham{fp, len};
        if (fp == 32) h->f = new SumOps<float>(HostSumOps<float>::from_terms(bits, coeff, n));
        else h->d = new SumOps<double>(HostSumOps<double>::from_terms(bits, coeff, n));
    });
    if (st) { delete h; return nullptr; }
    return h;
}

int rph_ham_info(void *hp, rp_terms_info_t *info)
{
    return guarded([&] {
        Ham *h = (Ham *)hp;
        *info = h->fp == 32 ? h->f->device_info(h->len) : h->d->device_info(h->len);
    });
}

void rph_ham_free(void *hp)
{
    Ham *h = (Ham *)hp;
    if (!h) return;
    delete h->f;
    delete h->d;
    delete h;
}

void *rph_state_new(int fp, int len)
{
    St *s = nullptr;
    int st = guarded([&] {
        s = new St{fp};
        if (fp == 32) s->f = new State<float, GPUImpl>(len);
        else s->d = new State<double, GPUImpl>(len);
    });
    if (st) { delete s; return nullptr; }
    return s;
}

void rph_state_free(void *sp)
{
    St *s = (St *)sp;
    if (!s) return;
    delete s->f;
    delete s->d;
    delete s;
}

int rph_state_load(void *sp, const void *host)
{
    return guarded([&] {
        St *s = (St *)sp;
        if (s->fp == 32) s->f->load((const complex<float> *)host);
        else s->d->load((const complex<double> *)host);
    });
}

int rph_state_store(void *sp, void *host)
{
    return guarded([&] {
        St *s = (St *)sp;
        if (s->fp == 32) s->f->store((complex<float> *)host);
        else s->d->store((complex<double> *)host);
    });
}

int rph_state_random(void *sp, uint64_t seed)
{
    return guarded([&] {
        St *s = (St *)sp;
        if (s->fp == 32) s->f->random_state(seed);
        else s->d->random_state(seed);
    });
}

// dst = src with precision conversion (rounding applies to double -> float)
int rph_state_copy(void *dp, const void *sp, int rounding)
{
    return guarded([&] {
        St *d = (St *)dp;
        const St *s = (const St *)sp;
        if (d->fp == 32 && s->fp == 32) *d->f = *s->f;
        else if (d->fp == 64 && s->fp == 64) *d->d = *s->d;
        else if (d->fp == 64) *d->d = State<double, GPUImpl>(*s->f);
        else *d->f = State<float, GPUImpl>(*s->d, rounding);
    });
}

// s *= H (State::operator*=(SumOps))
int rph_state_apply(void *sp, void *hp)
{
    return guarded([&] {
        St *s = (St *)sp;
        Ham *h = (Ham *)hp;
        if (s->fp != h->fp) ThrowException(InvalidArgument, "precision of state and operator must match");
        if (s->fp == 32) *s->f *= *h->f;
        else *s->d *= *h->d;
    });
}

static void fill_stats(rph_stats *out, const KrylovStats &st, double wall)
{
    if (!out) return;
    out->matvecs = st.matvecs;
    out->restarts = st.restarts;
    out->gs_passes = st.gs_passes;
    out->breakdown = st.breakdown;
    out->substeps = st.substeps;
    out->last_nrm = st.last_nrm;
    out->host_ms = st.host_ms;
    out->wall_ms = wall;
    out->rejected = st.rejected;
    out->reserved = 0;
}

int rph_state_evolve(void *sp, void *hp, double t, double beta, int krydim, double tol,
                     int max_iters, int reorth, double substep, rph_stats *out);

}  // extern "C"

namespace {

// A schedule handle: the schedule plus a record of every decision it made
// (the proposer's 12 inputs, the decision's kind and value) when tracing.
struct Sched : KrylovSchedule {
    std::unique_ptr<KrylovSchedule> inner;
    int max_dim = 64;
    bool tracing = false;
    struct Rec {
        float f[RP_MLP_IN];
        int kind, value, restart, substep;
        double last_update;
    };
    std::vector<Rec> trace;

    void record(const RestartInfo &ri, bool sub, int value) {
        if (!tracing) return;
        Rec r;
        proposer_features(ri, sub, r.f);
        r.kind = sub ? 0 : 1;
        r.value = value;
        r.restart = ri.restart;
        r.substep = ri.substep;
        r.last_update = ri.last_update;
        trace.push_back(r);
    }
    int substeps(const RestartInfo &ri) override {
        int v = inner->substeps(ri);
        record(ri, true, v);
        return v;
    }
    int krydim(const RestartInfo &ri) override {
        int v = inner->krydim(ri);
        record(ri, false, std::max(2, std::min(v, max_dim)));
        return v;
    }
};

template<typename F>
int run_evolve(void *sp, void *hp, double t, double beta, KrylovOptions &opt, rph_stats *out,
               F &&extra)
{
    return guarded([&] {
        St *s = (St *)sp;
        Ham *h = (Ham *)hp;
        if (s->fp != h->fp) ThrowException(InvalidArgument, "precision of state and operator must match");
        extra(opt);
        auto t0 = std::chrono::steady_clock::now();
        KrylovStats st = s->fp == 32 ? s->f->evolve_ex(*h->f, t, beta, opt)
                                     : s->d->evolve_ex(*h->d, t, beta, opt);
        GPUImpl::synchronize();
        double wall = std::chrono::duration<double, std::milli>(
                          std::chrono::steady_clock::now() - t0).count();
        fill_stats(out, st, wall);
    });
}

}  // namespace

extern "C" {

int rph_state_evolve(void *sp, void *hp, double t, double beta, int krydim, double tol,
                     int max_iters, int reorth, double substep, rph_stats *out)
{
    KrylovOptions opt;
    opt.krydim = krydim;
    opt.tol = tol;
    opt.max_iters = max_iters;
    opt.reorth = reorth ? Reorth::Always : Reorth::Selective;
    opt.substep = substep;
    return run_evolve(sp, hp, t, beta, opt, out, [](KrylovOptions &) {});
}

// evolve with a list schedule: Krylov dimension dims[min(r, ndims-1)] at
// restart r, nsub equal sub-steps
int rph_state_evolve_list(void *sp, void *hp, double t, double beta, const int *dims,
                          int ndims, int nsub, double tol, int max_iters, int reorth,
                          rph_stats *out)
{
    if (!dims || ndims < 1) {
        g_err = "dims cannot be empty";
        return RP_ERR_INVALID_ARGUMENT;
    }
    ListSchedule sched(std::vector<int>(dims, dims + ndims), nsub);
    KrylovOptions opt;
    opt.krydim = dims[ndims - 1];
    opt.tol = tol;
    opt.max_iters = max_iters;
    opt.reorth = reorth ? Reorth::Always : Reorth::Selective;
    opt.schedule = &sched;
    opt.max_krydim = std::max(2, *std::max_element(dims, dims + ndims));
    return run_evolve(sp, hp, t, beta, opt, out, [](KrylovOptions &) {});
}

// OTOC F(t_k) = <i|A(t)BA(t)B|i> of the state handle sp (which becomes
// exp(-beta H/8)|psi> if beta != 0); out[2k], out[2k+1] = Re, Im F(t_k);
// sched (nullable) chooses each restart's Krylov dimension, krydim the
// fallback
int rph_otoc(void *hp, void *ap, void *bp, void *sp, double beta, const double *times,
             int ntimes, int krydim, double tol, double *out, rph_stats *stats, void *sched)
{
    return guarded([&] {
        Ham *h = (Ham *)hp, *a = (Ham *)ap, *b = (Ham *)bp;
        St *s = (St *)sp;
        if (h->fp != s->fp || a->fp != s->fp || b->fp != s->fp)
            ThrowException(InvalidArgument, "precisions of state and operators must match");
        if (ntimes < 0 || (ntimes && (!times || !out)))
            ThrowException(InvalidArgument, "times", "null or negative count");
        KrylovOptions opt;
        opt.krydim = krydim;
        opt.tol = tol;
        if (sched) {
            opt.schedule = (Sched *)sched;
            opt.max_krydim = std::max(krydim, ((Sched *)sched)->max_dim);
        }
        std::vector<double> ts(times, times + ntimes);
        KrylovStats st;
        auto t0 = std::chrono::steady_clock::now();
        auto v = s->fp == 32 ? otoc<float, GPUImpl>(*h->f, *a->f, *b->f, *s->f, ts, beta, opt, st)
                             : otoc<double, GPUImpl>(*h->d, *a->d, *b->d, *s->d, ts, beta, opt, st);
        GPUImpl::synchronize();
        double wall = std::chrono::duration<double, std::milli>(
                          std::chrono::steady_clock::now() - t0).count();
        for (int k = 0; k < ntimes; k++) {
            out[2 * k] = v[k].real();
            out[2 * k + 1] = v[k].imag();
        }
        fill_stats(stats, st, wall);
    });
}

// batches of right-hand sides ----------------------------------------------

struct Batch {
    int fp;
    BatchedStates<float> *f = nullptr;
    BatchedStates<double> *d = nullptr;
    ~Batch() { delete f; delete d; }
};

// nprob problems: chain length lens[p], terms bits/coeff[toff[p] .. toff[p+1])
void *rph_batch_new(int fp, int nprob, const int *lens, const uint64_t *bits, const double *coeff,
                    const int64_t *toff, int nrhs, int kmax, int flags)
{
    Batch *h = nullptr;
    int st = guarded([&] {
        if (fp != 32 && fp != 64) ThrowException(InvalidArgument, "fp", "must be 32 or 64");
        if (nprob < 1 || !lens || !toff) ThrowException(InvalidArgument, "problems", "none given");
        auto p = std::make_unique<Batch>();
        p->fp = fp;
        std::vector<int> L(lens, lens + nprob);
        if (fp == 32) {
            std::vector<HostSumOps<float>> hs;
            for (int q = 0; q < nprob; q++)
                hs.push_back(HostSumOps<float>::from_terms(bits + toff[q], coeff + 2 * toff[q],
                                                           (size_t)(toff[q + 1] - toff[q])));
            std::vector<const HostSumOps<float> *> ptr;
            for (auto &x : hs) ptr.push_back(&x);
            p->f = new BatchedStates<float>(ptr, L, nrhs, kmax, flags);
        } else {
            std::vector<HostSumOps<double>> hs;
            for (int q = 0; q < nprob; q++)
                hs.push_back(HostSumOps<double>::from_terms(bits + toff[q], coeff + 2 * toff[q],
                                                            (size_t)(toff[q + 1] - toff[q])));
            std::vector<const HostSumOps<double> *> ptr;
            for (auto &x : hs) ptr.push_back(&x);
            p->d = new BatchedStates<double>(ptr, L, nrhs, kmax, flags);
        }
        h = p.release();
    });
    return st ? nullptr : h;
}

void rph_batch_free(void *h) { delete (Batch *)h; }

int rph_batch_info(void *h, rp_batch_info_t *info, uint64_t *offsets)
{
    return guarded([&] {
        Batch *b = (Batch *)h;
        const rp_batch_info_t &i = b->fp == 32 ? b->f->layout() : b->d->layout();
        if (info) *info = i;
        if (offsets) {
            const auto &o = b->fp == 32 ? b->f->offsets() : b->d->offsets();
            std::copy(o.begin(), o.end(), offsets);
        }
    });
}

int rph_batch_load(void *h, const void *host)
{
    return guarded([&] {
        Batch *b = (Batch *)h;
        if (b->fp == 32) b->f->load((const complex<float> *)host);
        else b->d->load((const complex<double> *)host);
    });
}

int rph_batch_store(void *h, void *host)
{
    return guarded([&] {
        Batch *b = (Batch *)h;
        if (b->fp == 32) b->f->store((complex<float> *)host);
        else b->d->store((complex<double> *)host);
    });
}

int rph_batch_apply(void *h)
{
    return guarded([&] {
        Batch *b = (Batch *)h;
        if (b->fp == 32) b->f->apply();
        else b->d->apply();
        GPUImpl::synchronize();
    });
}

// sched: a schedule handle (rph_sched_*_new) or null for the fixed m
int rph_batch_evolve(void *h, double t, double beta, int m, double tol, int max_iters,
                     int use_graph, void *sched, rph_stats *out, int *restarts)
{
    return guarded([&] {
        Batch *b = (Batch *)h;
        std::vector<int> rs;
        KrylovSchedule *sc = (Sched *)sched;
        auto t0 = std::chrono::steady_clock::now();
        KrylovStats st = b->fp == 32
            ? b->f->evolve(t, beta, m, tol, max_iters, use_graph != 0, &rs, sc)
            : b->d->evolve(t, beta, m, tol, max_iters, use_graph != 0, &rs, sc);
        double wall = std::chrono::duration<double, std::milli>(
                          std::chrono::steady_clock::now() - t0).count();
        fill_stats(out, st, wall);
        if (restarts) std::copy(rs.begin(), rs.end(), restarts);
    });
}

// flags of later term uploads (RP_TERMS_AUTOTUNE, RP_TERMS_ACC_NATIVE)
int rph_set_upload_flags(int flags)
{
    GPUImpl::upload_flags() = flags;
    return 0;
}

// schedule handles --------------------------------------------------------

void *rph_sched_list_new(const int *dims, int ndims, int nsub)
{
    Sched *h = nullptr;
    int st = guarded([&] {
        if (!dims || ndims < 1) ThrowException(InvalidArgument, "dims", "cannot be empty");
        auto p = std::make_unique<Sched>();
        p->inner = std::make_unique<ListSchedule>(std::vector<int>(dims, dims + ndims), nsub);
        p->max_dim = std::max(2, *std::max_element(dims, dims + ndims));
        h = p.release();
    });
    return st ? nullptr : h;
}

void *rph_sched_mlp_new(const float *params, int min_dim, int max_dim, int max_sub)
{
    Sched *h = nullptr;
    int st = guarded([&] {
        if (!params) ThrowException(InvalidArgument, "params", "cannot be null");
        auto p = std::make_unique<Sched>();
        p->inner = std::make_unique<MLPSchedule>(params, min_dim, max_dim, max_sub);
        p->max_dim = max_dim;
        h = p.release();
    });
    return st ? nullptr : h;
}

void rph_sched_free(void *h) { delete (Sched *)h; }

int rph_sched_set_tracing(void *h, int on)
{
    if (!h) {
        g_err = "null schedule";
        return RP_ERR_INVALID_ARGUMENT;
    }
    ((Sched *)h)->tracing = on != 0;
    ((Sched *)h)->trace.clear();
    return 0;
}

// copies up to cap records: feats[cap][12]; meta[cap][4] = kind (0 sub-step
// count, 1 Krylov dimension), value, restart, sub-step; upd[cap] = the
// previous restart's update norm.  Returns the number of records in *count.
int rph_sched_trace(void *h, float *feats, int *meta, double *upd, int cap, int *count)
{
    if (!h || !count) {
        g_err = "null argument";
        return RP_ERR_INVALID_ARGUMENT;
    }
    auto &tr = ((Sched *)h)->trace;
    *count = (int)tr.size();
    for (int i = 0; i < cap && i < (int)tr.size(); i++) {
        if (feats) std::memcpy(feats + (size_t)i * RP_MLP_IN, tr[i].f, sizeof(tr[i].f));
        if (meta) {
            meta[4 * i] = tr[i].kind;
            meta[4 * i + 1] = tr[i].value;
            meta[4 * i + 2] = tr[i].restart;
            meta[4 * i + 3] = tr[i].substep;
        }
        if (upd) upd[i] = tr[i].last_update;
    }
    return 0;
}

// evolve with a schedule handle; krydim is the fixed dimension a rejected
// schedule falls back to
int rph_state_evolve_sched(void *sp, void *hp, void *sched, double t, double beta, int krydim,
                           double tol, int max_iters, int reorth, rph_stats *out)
{
    if (!sched) {
        g_err = "null schedule";
        return RP_ERR_INVALID_ARGUMENT;
    }
    KrylovOptions opt;
    opt.krydim = krydim;
    opt.tol = tol;
    opt.max_iters = max_iters;
    opt.reorth = reorth ? Reorth::Always : Reorth::Selective;
    opt.schedule = (Sched *)sched;
    opt.max_krydim = std::max(krydim, ((Sched *)sched)->max_dim);
    return run_evolve(sp, hp, t, beta, opt, out, [](KrylovOptions &) {});
}

int rph_state_inner(void *ap, void *bp, double *re, double *im)
{
    return guarded([&] {
        St *a = (St *)ap;
        St *b = (St *)bp;
        if (a->fp != b->fp) ThrowException(InvalidArgument, "precisions must match");
        auto c = a->fp == 32 ? a->f->inner64(*b->f) : a->d->inner64(*b->d);
        *re = c.real();
        *im = c.imag();
    });
}

int rph_state_norm(void *sp, double *nrm)
{
    return guarded([&] {
        St *s = (St *)sp;
        *nrm = s->fp == 32 ? s->f->norm64() : s->d->norm64();
    });
}

int rph_state_ground(void *sp, void *hp, int krydim, double eps, int maxiter, double *energy)
{
    return guarded([&] {
        St *s = (St *)sp;
        Ham *h = (Ham *)hp;
        if (s->fp == 32)
            *energy = s->f->ground_state(*h->f, krydim, (float)eps, maxiter, false);
        else
            *energy = s->d->ground_state(*h->d, krydim, eps, maxiter, false);
    });
}

}  // extern "C"

// parity-block form ---------------------------------------------------------

namespace {

struct BlkSt {
    int fp;
    BlockState<float, GPUImpl> *f = nullptr;
    BlockState<double, GPUImpl> *d = nullptr;
    ~BlkSt() { delete f; delete d; }
};

template<RealScalar Fp>
SumOps<Fp> &ham_of(Ham *h);
template<>
SumOps<float> &ham_of<float>(Ham *h) { return *h->f; }
template<>
SumOps<double> &ham_of<double>(Ham *h) { return *h->d; }

template<RealScalar Fp>
BlockState<Fp, GPUImpl> &blk_of(BlkSt *s);
template<>
BlockState<float, GPUImpl> &blk_of<float>(BlkSt *s) { return *s->f; }
template<>
BlockState<double, GPUImpl> &blk_of<double>(BlkSt *s) { return *s->d; }

void check_fp(int fp, std::initializer_list<Ham *> hs)
{
    for (Ham *h : hs)
        if (h && h->fp != fp) ThrowException(InvalidArgument, "precisions of state and operators must match");
}

// a block of a block operator: the handle's term list, or a null block
template<RealScalar Fp>
SumOps<Fp> block_of(Ham *h) { return h ? ham_of<Fp>(h) : SumOps<Fp>{}; }

template<RealScalar Fp>
void block_apply(BlkSt *s, Ham *LL, Ham *LR, Ham *RL, Ham *RR)
{
    auto &st = blk_of<Fp>(s);
    if (!LR && !RL) {
        st *= BlockDiag<SumOps<Fp>>(block_of<Fp>(LL), block_of<Fp>(RR));
    } else if (!LL && !RR) {
        st *= BlockAntiDiag<SumOps<Fp>>(block_of<Fp>(LR), block_of<Fp>(RL));
    } else {
        st *= BlockOp<SumOps<Fp>>(block_of<Fp>(LL), block_of<Fp>(LR), block_of<Fp>(RL),
                                  block_of<Fp>(RR));
    }
}

// A view of the term list of one precision converted to another
template<RealScalar To, RealScalar From>
SubspaceView<SumOps<To>> promote_view(const SubspaceView<SumOps<From>> &v)
{
    return SubspaceView<SumOps<To>>(v.spin_chain_length(), SumOps<To>(v));
}

template<RealScalar Fp>
void syk_block_npoint(int kind, int N, double sp, double J, uint32_t seed, int regularize,
                      int otoc_idx, const void *init, uint64_t state_seed, double beta,
                      double tmax, int nsteps, int krylov_dim, double tol, double *out,
                      double *e0, KrylovStats &st, void *sched, int promote)
{
    // the Hamiltonian and the Majorana block in Fp; with promote (Fp =
    // double) generated in FP32 and converted, the reference's FP32/FP64
    // cross-check (examples/syk.cc:440-474)
    using View = SubspaceView<SumOps<Fp>>;
    BlockDiag<View> ham;
    View op;
    std::mt19937 rg(seed);
    const int oi = (otoc_idx >= 0 && otoc_idx <= N - 2) ? otoc_idx : N - 2;
    if (promote && std::is_same_v<Fp, double>) {
        SYKBlockForm<float, GPUImpl> syk(N, (float)sp, (float)J);
        auto h = syk.gen_ham(rg, regularize != 0);
        ham = BlockDiag<View>(promote_view<Fp>(h.LL), promote_view<Fp>(h.RR));
        op = promote_view<Fp>(syk.fermion_ops()[oi].LR);
    } else {
        SYKBlockForm<Fp, GPUImpl> syk(N, (Fp)sp, (Fp)J);
        ham = syk.gen_ham(rg, regularize != 0);
        op = syk.fermion_ops()[oi].LR;
    }
    if (kind >= 2) {
        auto v = npoint_block_trace<Fp>(kind == 2 ? NPoint::Otoc : NPoint::Green, ham, op, tmax,
                                        nsteps, beta);
        for (size_t k = 0; k < v.size(); k++) {
            out[2 * k] = v[k].real();
            out[2 * k + 1] = v[k].imag();
        }
        if (e0) *e0 = 0.0;
        return;
    }
    State<Fp, GPUImpl> s0(N / 2 - 1);
    if (init) {
        s0.load((const complex<Fp> *)init);
    } else if (promote && std::is_same_v<Fp, double>) {
        State<float, GPUImpl> s32(N / 2 - 1);
        s32.random_state(state_seed);
        s0 = State<Fp, GPUImpl>(s32);
    } else {
        s0.random_state(state_seed);
    }
    KrylovOptions opt;
    opt.krydim = krylov_dim;
    opt.tol = tol;
    if (sched) {
        opt.schedule = (Sched *)sched;
        opt.max_krydim = std::max(krylov_dim, ((Sched *)sched)->max_dim);
    }
    auto v = npoint_block<Fp, GPUImpl>(kind ? NPoint::Green : NPoint::Otoc, ham, op, s0, tmax,
                                       nsteps, beta, krylov_dim, opt, st, e0);
    for (size_t k = 0; k < v.size(); k++) {
        out[2 * k] = v[k].real();
        out[2 * k + 1] = v[k].imag();
    }
}

}  // namespace

extern "C" {

// SYKBlockForm::gen_ham(std::mt19937(seed), regularize): which = 0 exports
// ham.LL, 1 ham.RR
int rph_syk_block_gen(int fp, int N, double sp, double J, int std_gamma, uint32_t seed,
                      int regularize, int which, uint64_t *bits, double *coeff, int64_t cap,
                      int64_t *nterms)
{
    return guarded([&] {
        std::mt19937 rg(seed);
        if (fp == 32) {
            SYKBlockForm<float, GPUImpl> syk(N, (float)sp, (float)J, std_gamma != 0);
            auto ham = syk.gen_ham(rg, regularize != 0);
            *nterms = export_ham<float>(which ? ham.RR : ham.LL, bits, coeff, cap);
        } else {
            SYKBlockForm<double, GPUImpl> syk(N, sp, J, std_gamma != 0);
            auto ham = syk.gen_ham(rg, regularize != 0);
            *nterms = export_ham<double>(which ? ham.RR : ham.LL, bits, coeff, cap);
        }
    });
}

// FermionOpBlockForm(N, n, std_gamma): which = 0 exports the LR block, 1 RL
int rph_majorana_block(int fp, int N, int n, int std_gamma, int which, uint64_t *bits,
                       double *coeff, int64_t cap, int64_t *nterms)
{
    return guarded([&] {
        if (fp == 32) {
            FermionOpBlockForm<float, GPUImpl> op(N, n, std_gamma != 0);
            *nterms = export_ham<float>(which ? op.RL : op.LR, bits, coeff, cap);
        } else {
            FermionOpBlockForm<double, GPUImpl> op(N, n, std_gamma != 0);
            *nterms = export_ham<double>(which ? op.RL : op.LR, bits, coeff, cap);
        }
    });
}

// chi_i chi_j + chi_j chi_i in block form (a BlockDiag): which = 0 exports
// LL, 1 RR; *null receives the block's null flag
int rph_block_anticomm(int fp, int N, int i, int j, int std_gamma, int which, uint64_t *bits,
                       double *coeff, int64_t cap, int64_t *nterms, int *null)
{
    return guarded([&] {
        auto run = [&](auto tag) {
            using Fp = decltype(tag);
            FermionOpBlockForm<Fp, GPUImpl> a(N, i, std_gamma != 0), b(N, j, std_gamma != 0);
            auto c = a * b + b * a;
            *null = which ? c.nullRR : c.nullLL;
            *nterms = export_ham<Fp>(which ? c.RR : c.LL, bits, coeff, cap);
        };
        if (fp == 32) run(float{});
        else run(double{});
    });
}

void *rph_block_state_new(int fp, int len)
{
    BlkSt *s = nullptr;
    int st = guarded([&] {
        s = new BlkSt{fp};
        if (fp == 32) s->f = new BlockState<float, GPUImpl>(len);
        else s->d = new BlockState<double, GPUImpl>(len);
    });
    if (st) { delete s; return nullptr; }
    return s;
}

void rph_block_state_free(void *sp) { delete (BlkSt *)sp; }

// load the two sectors from host vectors of 2^len amplitudes; a null
// pointer makes that sector null
int rph_block_state_load(void *sp, const void *L, const void *R)
{
    return guarded([&] {
        BlkSt *s = (BlkSt *)sp;
        auto run = [&](auto tag) {
            using Fp = decltype(tag);
            auto &st = blk_of<Fp>(s);
            int len = st.spin_chain_length();
            for (int k = 0; k < 2; k++) {
                const void *src = k ? R : L;
                auto &v = k ? st.R : st.L;
                bool &null = k ? st.nullR : st.nullL;
                if (src) {
                    if (!v.size()) v = State<Fp, GPUImpl>(len);
                    v.load((const complex<Fp> *)src);
                    null = false;
                } else {
                    v = State<Fp, GPUImpl>{};
                    null = true;
                }
            }
        };
        if (s->fp == 32) run(float{});
        else run(double{});
    });
}

// store the defined sectors; *nullL / *nullR receive the flags
int rph_block_state_store(void *sp, void *L, void *R, int *nullL, int *nullR)
{
    return guarded([&] {
        BlkSt *s = (BlkSt *)sp;
        auto run = [&](auto tag) {
            using Fp = decltype(tag);
            auto &st = blk_of<Fp>(s);
            *nullL = st.nullL;
            *nullR = st.nullR;
            if (!st.nullL && L) st.L.store((complex<Fp> *)L);
            if (!st.nullR && R) st.R.store((complex<Fp> *)R);
        };
        if (s->fp == 32) run(float{});
        else run(double{});
    });
}

int rph_block_state_random(void *sp, uint64_t seed)
{
    return guarded([&] {
        BlkSt *s = (BlkSt *)sp;
        if (s->fp == 32) s->f->random_state(seed);
        else s->d->random_state(seed);
    });
}

// amplitude i of the full vector (L sector first)
int rph_block_state_get(void *sp, uint64_t i, double *re, double *im)
{
    return guarded([&] {
        BlkSt *s = (BlkSt *)sp;
        auto c = s->fp == 32
                     ? std::complex<double>(static_cast<const BlockState<float, GPUImpl> &>(*s->f)[i])
                     : static_cast<const BlockState<double, GPUImpl> &>(*s->d)[i];
        *re = c.real();
        *im = c.imag();
    });
}

int rph_block_state_norm(void *sp, double *nrm)
{
    return guarded([&] {
        BlkSt *s = (BlkSt *)sp;
        *nrm = s->fp == 32 ? s->f->norm64() : s->d->norm64();
    });
}

// st *= op for the block operator with blocks LL, LR, RL, RR (null handles
// are null blocks): BlockDiag when LR and RL are null, BlockAntiDiag when LL
// and RR are, BlockOp otherwise
int rph_block_state_apply(void *sp, void *LL, void *LR, void *RL, void *RR)
{
    return guarded([&] {
        BlkSt *s = (BlkSt *)sp;
        check_fp(s->fp, {(Ham *)LL, (Ham *)LR, (Ham *)RL, (Ham *)RR});
        if (s->fp == 32) block_apply<float>(s, (Ham *)LL, (Ham *)LR, (Ham *)RL, (Ham *)RR);
        else block_apply<double>(s, (Ham *)LL, (Ham *)LR, (Ham *)RL, (Ham *)RR);
    });
}

// BlockState::evolve with BlockDiag(LL, RR)
int rph_block_state_evolve(void *sp, void *LL, void *RR, double t, double beta, int krydim,
                           double tol, rph_stats *out)
{
    return guarded([&] {
        BlkSt *s = (BlkSt *)sp;
        check_fp(s->fp, {(Ham *)LL, (Ham *)RR});
        KrylovOptions opt;
        opt.krydim = krydim;
        opt.tol = tol;
        auto t0 = std::chrono::steady_clock::now();
        KrylovStats st;
        auto run = [&](auto tag) {
            using Fp = decltype(tag);
            BlockDiag<SumOps<Fp>> h(block_of<Fp>((Ham *)LL), block_of<Fp>((Ham *)RR));
            st = blk_of<Fp>(s).evolve_ex(h, t, beta, opt);
        };
        if (s->fp == 32) run(float{});
        else run(double{});
        GPUImpl::synchronize();
        double wall = std::chrono::duration<double, std::milli>(
                          std::chrono::steady_clock::now() - t0).count();
        fill_stats(out, st, wall);
    });
}

// BlockState::ground_state with BlockDiag(LL, RR); *sector = 0 if the L
// sector was kept, 1 if R
int rph_block_state_ground(void *sp, void *LL, void *RR, int krydim, double eps, int maxiter,
                           double *energy, int *sector)
{
    return guarded([&] {
        BlkSt *s = (BlkSt *)sp;
        check_fp(s->fp, {(Ham *)LL, (Ham *)RR});
        auto run = [&](auto tag) {
            using Fp = decltype(tag);
            BlockDiag<SumOps<Fp>> h(block_of<Fp>((Ham *)LL), block_of<Fp>((Ham *)RR));
            auto &st = blk_of<Fp>(s);
            *energy = st.ground_state(h, krydim, (Fp)eps, maxiter);
            *sector = st.nullL ? 1 : 0;
        };
        if (s->fp == 32) run(float{});
        else run(double{});
    });
}

// One disorder realization of the reference's parity-block SYK driver:
// SYKBlockForm(N, sp, J).gen_ham(std::mt19937(seed), regularize), the
// initial L-sector state from init (2^(N/2-1) amplitudes) or random with
// state_seed, then npoint_block (kind 0 = OTOC, 1 = Green's function), or
// the trace definitions by exact diagonalization (kind 2 = OTOC, 3 = Green's
// function; no state, e0 = 0).
// out receives 2 (nsteps + 1) doubles (Re, Im of v_k / v_0), e0 the
// ground-state energy the Hamiltonian was shifted by; sched (nullable) is a
// schedule handle choosing each restart's Krylov dimension (krylov_dim is
// then the fallback; the Lanczos E0 always uses krylov_dim); promote (fp 64
// only) generates the Hamiltonian and the random initial state in FP32 and
// converts them, so that an FP32 run has an FP64 counterpart on the same
// couplings.
int rph_syk_block_npoint(int kind, int fp, int N, double sp, double J, uint32_t seed,
                         int regularize, int otoc_idx, const void *init, uint64_t state_seed,
                         double beta, double tmax, int nsteps, int krylov_dim, double tol,
                         double *out, double *e0, rph_stats *stats, void *sched, int promote)
{
    return guarded([&] {
        if (kind < 0 || kind > 3) ThrowException(InvalidArgument, "kind", "must be 0..3");
        if (!out || nsteps < 0) ThrowException(InvalidArgument, "out", "null or negative count");
        KrylovStats st;
        auto t0 = std::chrono::steady_clock::now();
        if (fp == 32)
            syk_block_npoint<float>(kind, N, sp, J, seed, regularize, otoc_idx, init, state_seed,
                                    beta, tmax, nsteps, krylov_dim, tol, out, e0, st, sched,
                                    0);
        else
            syk_block_npoint<double>(kind, N, sp, J, seed, regularize, otoc_idx, init, state_seed,
                                     beta, tmax, nsteps, krylov_dim, tol, out, e0, st, sched,
                                     promote);
        GPUImpl::synchronize();
        double wall = std::chrono::duration<double, std::milli>(
                          std::chrono::steady_clock::now() - t0).count();
        fill_stats(stats, st, wall);
    });
}

}  // extern "C"

// exact diagonalization ------------------------------------------------------

extern "C" {

// State::evolve<ExactDiagonalization>(H, t, beta)
int rph_state_evolve_ed(void *sp, void *hp, double t, double beta)
{
    return guarded([&] {
        St *s = (St *)sp;
        Ham *h = (Ham *)hp;
        if (s->fp != h->fp) ThrowException(InvalidArgument, "precision of state and operator must match");
        if (s->fp == 32) s->f->evolve<EvolutionAlgorithm::ExactDiagonalization>(*h->f, t, beta);
        else s->d->evolve<EvolutionAlgorithm::ExactDiagonalization>(*h->d, t, beta);
    });
}

// State::ground_state<ExactDiagonalization>(H)
int rph_state_ground_ed(void *sp, void *hp, double *energy)
{
    return guarded([&] {
        St *s = (St *)sp;
        Ham *h = (Ham *)hp;
        if (s->fp != h->fp) ThrowException(InvalidArgument, "precision of state and operator must match");
        if (s->fp == 32)
            *energy = s->f->ground_state<EvolutionAlgorithm::ExactDiagonalization>(*h->f);
        else
            *energy = s->d->ground_state<EvolutionAlgorithm::ExactDiagonalization>(*h->d);
    });
}

// host copies of the handle's dense forms on its chain length: what = 0
// the matrix, 1 exp(c H) (c = c_re + i c_im), 2 the eigenvectors; out holds
// 4^len complex values (column-major, interleaved doubles); eigvals (2^len
// doubles, may be null) receives the eigenvalues for what = 2
int rph_ham_dense(void *hp, int what, double c_re, double c_im, double *out, double *eigvals)
{
    return guarded([&] {
        Ham *h = (Ham *)hp;
        auto run = [&](auto tag) {
            using Fp = decltype(tag);
            auto &ops = ham_of<Fp>(h);
            const DevMatrix<Fp> *m = nullptr;
            if (what == 0) m = &ops.get_matrix(h->len);
            else if (what == 1) m = &ops.matexp(complex<Fp>{(Fp)c_re, (Fp)c_im}, h->len);
            else {
                auto es = ops.get_eigensystem(h->len);
                if (eigvals)
                    for (size_t i = 0; i < std::get<0>(es).size(); i++) eigvals[i] = std::get<0>(es)[i];
                m = &std::get<1>(es);
            }
            auto v = m->to_host();
            for (size_t i = 0; i < v.size(); i++) {
                out[2 * i] = v[i].real();
                out[2 * i + 1] = v[i].imag();
            }
        };
        if (what < 0 || what > 2) ThrowException(InvalidArgument, "what", "must be 0..2");
        if (h->fp == 32) run(float{});
        else run(double{});
    });
}

// entries in the handle's matexp cache
int rph_ham_matexp_entries(void *hp, int64_t *n)
{
    return guarded([&] {
        Ham *h = (Ham *)hp;
        *n = h->fp == 32 ? (int64_t)h->f->matexp_cache_entries()
                         : (int64_t)h->d->matexp_cache_entries();
    });
}

}  // extern "C"

// text forms ----------------------------------------------------------------

namespace {
int put_string(const std::string &s, char *buf, int64_t cap, int64_t *len)
{
    *len = (int64_t)s.size();
    if (buf && cap > 0) {
        int64_t n = std::min<int64_t>(cap - 1, (int64_t)s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = 0;
    }
    return 0;
}
}  // namespace

extern "C" {

// SumOps' operator<< of a term list; *len receives the full length (the
// text is cut to cap - 1 characters)
int rph_ham_to_string(int fp, const uint64_t *bits, const double *coeff, int64_t n, char *buf,
                      int64_t cap, int64_t *len)
{
    return guarded([&] {
        std::stringstream ss;
        if (fp == 32) ss << HostSumOps<float>::from_terms(bits, coeff, n);
        else ss << HostSumOps<double>::from_terms(bits, coeff, n);
        put_string(ss.str(), buf, cap, len);
    });
}

// the SYK driver's Hamiltonian dump of SYKBlockForm(N, sp, J).gen_ham(
// std::mt19937(seed), regularize)
int rph_syk_block_dump(int fp, int N, double sp, double J, uint32_t seed, int regularize,
                       char *buf, int64_t cap, int64_t *len)
{
    return guarded([&] {
        std::stringstream ss;
        std::mt19937 rg(seed);
        if (fp == 32) {
            SYKBlockForm<float, GPUImpl> syk(N, (float)sp, (float)J);
            dump_block_ham(ss, syk.gen_ham(rg, regularize != 0));
        } else {
            SYKBlockForm<double, GPUImpl> syk(N, sp, J);
            dump_block_ham(ss, syk.gen_ham(rg, regularize != 0));
        }
        put_string(ss.str(), buf, cap, len);
    });
}

// write_results of n values (interleaved re, im) at spacing dt
int rph_format_results(double dt, const double *v, int64_t n, char *buf, int64_t cap,
                       int64_t *len)
{
    return guarded([&] {
        std::vector<std::complex<double>> c(n);
        for (int64_t k = 0; k < n; k++) c[k] = {v[2 * k], v[2 * k + 1]};
        std::stringstream ss;
        write_results(ss, dt, c);
        put_string(ss.str(), buf, cap, len);
    });
}

}  // extern "C"

// element access and the matexp cache budget -----------------------------------

extern "C" {

// st[i] = re + i im through State's writable element (ElemRefType)
int rph_state_set(void *sp, uint64_t i, double re, double im)
{
    return guarded([&] {
        St *s = (St *)sp;
        if (s->fp == 32) (*s->f)[i] = complex<float>{(float)re, (float)im};
        else (*s->d)[i] = complex<double>{re, im};
    });
}

// GPUImpl::set_max_cache_size (bytes >= 0) and the budget's state
int rph_cache_budget(int64_t set_bytes, int64_t *max_bytes, int64_t *used_bytes)
{
    return guarded([&] {
        if (set_bytes >= 0) GPUImpl::set_max_cache_size((size_t)set_bytes);
        *max_bytes = (int64_t)GPUImpl::get_max_cache_size();
        *used_bytes = (int64_t)GPUImpl::get_current_cache_size();
    });
}

}  // extern "C"

extern "C" {

// GPUImpl::dense_apply(): apply through the dense matrix when the term count
// reaches 2^len (up to 15 sites); returns the previous setting
int rph_set_dense_apply(int on)
{
    int prev = GPUImpl::dense_apply() ? 1 : 0;
    GPUImpl::dense_apply() = on != 0;
    return prev;
}

// Krylov's Arnoldi steps through rp_apply_gs (apply + first Gram-Schmidt
// pass in one kernel); returns the previous setting
int rph_set_fused_first_pass(int on)
{
    int prev = GPUImpl::fused_first_pass() ? 1 : 0;
    GPUImpl::fused_first_pass() = on != 0;
    return prev;
}

}  // extern "C"

extern "C" {

// The same state on both backends (inc/state.h:114-118): the host vector psi
// becomes a State<Fp, CPUImpl>, is converted to a State<Fp, GPUImpl>, and
// both are evolved by exp(-(beta + i t) H) with their own backend (Krylov
// dimension krydim, tolerance tol).  out_cpu receives the CPU result,
// out_gpu the GPU result converted back into a CPU state, and out_copy the
// GPU state converted back before the evolution (a round trip of the
// conversions).  Vectors are complex<Fp> of 2^len elements.
int rph_cross_backend_evolve(int fp, int len, const uint64_t *bits, const double *coeff,
                             int64_t nterms, const void *psi, double t, double beta, int krydim,
                             double tol, void *out_cpu, void *out_gpu, void *out_copy)
{
    return guarded([&] {
        auto run = [&](auto tag) {
            using Fp = decltype(tag);
            HostSumOps<Fp> hops = HostSumOps<Fp>::from_terms(bits, coeff, (size_t)nterms);
            DevSumOps<Fp> dops(hops);
            State<Fp, CPUImpl> c(len);
            c.load((const complex<Fp> *)psi);
            State<Fp, GPUImpl> g(c);
            State<Fp, CPUImpl> back(g);
            back.store((complex<Fp> *)out_copy);
            KrylovOptions opt;
            opt.krydim = krydim;
            opt.tol = tol;
            c.evolve_ex(hops, t, beta, opt);
            g.evolve_ex(dops, t, beta, opt);
            c.store((complex<Fp> *)out_cpu);
            State<Fp, CPUImpl> gc(g);
            gc.store((complex<Fp> *)out_gpu);
        };
        if (fp == 32) run(float{});
        else if (fp == 64) run(double{});
        else ThrowException(InvalidArgument, "fp", "must be 32 or 64");
    });
}

}  // extern "C"
