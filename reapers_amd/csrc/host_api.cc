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

int rph_state_evolve(void *sp, void *hp, double t, double beta, int krydim, double tol,
                     int max_iters, int reorth, double substep, rph_stats *out)
{
    return guarded([&] {
        St *s = (St *)sp;
        Ham *h = (Ham *)hp;
        if (s->fp != h->fp) ThrowException(InvalidArgument, "precision of state and operator must match");
        KrylovOptions opt;
        opt.krydim = krydim;
        opt.tol = tol;
        opt.max_iters = max_iters;
        opt.reorth = reorth ? Reorth::Always : Reorth::Selective;
        opt.substep = substep;
        auto t0 = std::chrono::steady_clock::now();
        KrylovStats st = s->fp == 32 ? s->f->evolve_ex(*h->f, t, beta, opt)
                                     : s->d->evolve_ex(*h->d, t, beta, opt);
        GPUImpl::synchronize();
        double wall = std::chrono::duration<double, std::milli>(
                          std::chrono::steady_clock::now() - t0).count();
        if (out) {
            out->matvecs = st.matvecs;
            out->restarts = st.restarts;
            out->gs_passes = st.gs_passes;
            out->breakdown = st.breakdown;
            out->substeps = st.substeps;
            out->last_nrm = st.last_nrm;
            out->host_ms = st.host_ms;
            out->wall_ms = wall;
        }
    });
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
