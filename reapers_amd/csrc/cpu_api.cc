// cpu_api.cc -- libreapers_cpu.so: the C++ API built with REAPERS_NOGPU
// (DefImpl = CPUImpl), behind the C functions of include/reapers_cpu.h.
#define REAPERS_NOGPU
#include "reapers/reapers.h"

#include "reapers_cpu.h"

#include <string>

using namespace REAPERS;
using namespace REAPERS::EvolutionAlgorithm;

namespace {

thread_local std::string g_err;

template<typename F>
int guarded(F &&f)
{
    try {
        f();
        return RP_OK;
    } catch (const InvalidArgument &e) {
        g_err = e.what();
        return RP_ERR_INVALID_ARGUMENT;
    } catch (const IndexTooLarge &e) {
        g_err = e.what();
        return RP_ERR_INVALID_ARGUMENT;
    } catch (const std::bad_alloc &e) {
        g_err = e.what();
        return RP_ERR_OUT_OF_MEMORY;
    } catch (const std::exception &e) {
        g_err = e.what();
        return RP_ERR_RUNTIME;
    } catch (...) {
        g_err = "unknown exception";
        return RP_ERR_RUNTIME;
    }
}

template<RealScalar Fp>
SumOps<Fp> make_ops(const uint64_t *bits, const double *coeff, uint64_t n)
{
    return SumOps<Fp>::from_terms(bits, coeff, n);
}

template<RealScalar Fp>
State<Fp> load_state(int len, const void *psi)
{
    State<Fp> s(len);
    s.load((const complex<Fp> *)psi);
    return s;
}

void check_fp(int fp)
{
    if (fp != 32 && fp != 64) ThrowException(InvalidArgument, "fp", "must be 32 or 64");
}

}  // namespace

extern "C" {

const char *rpc_last_error(void) { return g_err.c_str(); }

int rpc_apply(int fp, int len, const uint64_t *bits, const double *coeff, uint64_t nterms,
              const void *in, void *out, int accumulate)
{
    return guarded([&] {
        check_fp(fp);
        auto run = [&](auto tag) {
            using Fp = decltype(tag);
            auto ops = make_ops<Fp>(bits, coeff, nterms);
            auto *o = (complex<Fp> *)out;
            const auto *i = (const complex<Fp> *)in;
            if (accumulate) CPUImpl::apply_ops<Fp>(len, o, ops, i);
            else CPUImpl::apply_ops_overwrite<Fp>(len, o, ops, i);
        };
        if (fp == 32) run(float{});
        else run(double{});
    });
}

int rpc_evolve(int fp, int len, const uint64_t *bits, const double *coeff, uint64_t nterms,
               void *psi, double t, double beta, int krydim, double tol, int64_t *matvecs)
{
    return guarded([&] {
        check_fp(fp);
        auto run = [&](auto tag) {
            using Fp = decltype(tag);
            auto ops = make_ops<Fp>(bits, coeff, nterms);
            State<Fp> s = load_state<Fp>(len, psi);
            KrylovOptions opt;
            opt.krydim = krydim;
            opt.tol = tol;
            KrylovStats st = s.evolve_ex(ops, t, beta, opt);
            s.store((complex<Fp> *)psi);
            if (matvecs) *matvecs = st.matvecs;
        };
        if (fp == 32) run(float{});
        else run(double{});
    });
}

int rpc_ed_evolve(int fp, int len, const uint64_t *bits, const double *coeff, uint64_t nterms,
                  void *psi, double t, double beta)
{
    return guarded([&] {
        check_fp(fp);
        auto run = [&](auto tag) {
            using Fp = decltype(tag);
            auto ops = make_ops<Fp>(bits, coeff, nterms);
            State<Fp> s = load_state<Fp>(len, psi);
            s.template evolve<ExactDiagonalization>(ops, (Fp)t, (Fp)beta);
            s.store((complex<Fp> *)psi);
        };
        if (fp == 32) run(float{});
        else run(double{});
    });
}

int rpc_ground_state(int fp, int len, const uint64_t *bits, const double *coeff,
                     uint64_t nterms, void *psi, int krydim, double eps, double *energy)
{
    return guarded([&] {
        check_fp(fp);
        auto run = [&](auto tag) {
            using Fp = decltype(tag);
            auto ops = make_ops<Fp>(bits, coeff, nterms);
            State<Fp> s = load_state<Fp>(len, psi);
            s.normalize();
            *energy = (double)Krylov<CPUImpl>::ground_state(s, ops, krydim, (Fp)eps, 0, false);
            s.store((complex<Fp> *)psi);
        };
        if (fp == 32) run(float{});
        else run(double{});
    });
}

int rpc_eigensystem(int len, const uint64_t *bits, const double *coeff, uint64_t nterms,
                    double *w, void *V)
{
    return guarded([&] {
        auto ops = make_ops<double>(bits, coeff, nterms);
        auto es = ops.get_eigensystem(len);
        const auto &vals = std::get<0>(es);
        const auto &vecs = std::get<1>(es);
        for (size_t i = 0; i < vals.size(); i++) w[i] = vals[i];
        std::memcpy(V, vecs.data(), vecs.bytes());
    });
}

int rpc_random_state(int fp, int len, uint64_t seed, void *out)
{
    return guarded([&] {
        check_fp(fp);
        auto run = [&](auto tag) {
            using Fp = decltype(tag);
            State<Fp> s(len);
            s.random_state(seed);
            s.store((complex<Fp> *)out);
        };
        if (fp == 32) run(float{});
        else run(double{});
    });
}

}  // extern "C"
