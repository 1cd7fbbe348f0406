// driver_api_check.cc -- the library calls of the reference's production
// driver (examples/syk.cc) and short example programs
// (examples/simple-syk-*.cc) that go beyond one evolution, exercised on this
// backend and checked:
//
//   1. FP32 -> FP64 promotion of a generated Hamiltonian and initial state
//      (ham64 = ham32; *init64 = *init32; examples/syk.cc:455-462), in the
//      parity-block form when built with REAPERS_USE_PARITY_BLOCKS;
//   2. a State moved to the CPU backend and back (the --swap pattern,
//      examples/syk.cc:203-220: State<F, CPUImpl>(s0), State<F>(host));
//   3. dense forms through the operator views: Matrix A = syk.fermion_ops(n),
//      ham.matexp(-t * 1_i), (A * A).trace() (examples/simple-syk-exdiag.cc);
//   4. the Krylov evolution against exact diagonalization
//      (State::evolve<ExactDiagonalization>).
//
// Built by the package for the default (GPU) backend; a test builds it again
// with REAPERS_NOGPU, where every State lives on the CPU backend.  Prints one
// line per check and exits with 1 if one fails.
//
// usage: driver_api_check [N]
#include <reapers.h>

#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <memory>

using namespace REAPERS;
using namespace REAPERS::Model;
using namespace REAPERS::EvolutionAlgorithm;

template<RealScalar FpType>
using HamOp = typename SYK<FpType>::HamOp;

static int failures = 0;

static void check(const char *what, double value, double bound)
{
    bool ok = value <= bound;
    std::printf("%-52s %10.3e  (bound %.1e)  %s\n", what, value, bound, ok ? "ok" : "FAIL");
    if (!ok) failures++;
}

template<RealScalar F>
static double max_diff(const State<F> &a, const State<double> &b)
{
    std::vector<complex<F>> x(a.dim());
    std::vector<complex<double>> y(b.dim());
    a.store(x.data());
    b.store(y.data());
    double d = 0;
    for (size_t i = 0; i < x.size(); i++)
        d = std::max(d, std::abs(complex<double>(x[i].real(), x[i].imag()) - y[i]));
    return d;
}

int main(int argc, char **argv)
{
    const int N = argc > 1 ? std::atoi(argv[1]) : 16;
    std::mt19937 rg(1234);
    SYK<float> syk32(N, 0.0f);
    SYK<double> syk64(N, 0.0);
    HamOp<float> ham32 = syk32.gen_ham(rg);
    HamOp<double> ham64;
    // 1. promotion of the Hamiltonian and of the initial state
    ham64 = ham32;
#ifdef REAPERS_USE_PARITY_BLOCKS
    const auto &h32 = ham32.LL;
    const auto &h64 = ham64.LL;
    const int len = N / 2 - 1;
#else
    const auto &h32 = ham32;
    const auto &h64 = ham64;
    const int len = N / 2;
#endif
    double dcoef = 0;
    size_t i = 0;
    auto it64 = h64.begin();
    for (auto it32 = h32.begin(); it32 != h32.end(); ++it32, ++it64, ++i) {
        if (it64 == h64.end() || it32->get_bit_string() != it64->get_bit_string()) {
            dcoef = 1;
            break;
        }
        dcoef = std::max(dcoef, std::abs(complex<double>(it32->get_coeff()) - it64->get_coeff()));
    }
    if (i != h64.size()) dcoef = 1;
    check("ham64 = ham32: couplings promoted exactly", dcoef, 0);

    auto init32 = std::make_unique<State<float>>(len);
    init32->random_state(99);
    auto init64 = std::make_unique<State<double>>(len);
    *init64 = *init32;
    check("*init64 = *init32: state promoted exactly", max_diff(*init32, *init64), 0);

    // FP32 and FP64 runs of the same promoted problem
    const double t = 1.0;
    State<float> s32(*init32);
    State<double> s64(*init64);
    s32.evolve(h32, t, 0.0, 20, 1e-7f);
    s64.evolve(h64, t, 0.0, 20, 1e-13);
    check("FP32 vs FP64 evolution to t=1 (max |diff|)", max_diff(s32, s64), 1e-5);

    // 2. a state offloaded to the host backend and restored
    auto s0 = std::make_unique<State<double>>(*init64);
    auto hostst = std::make_unique<State<double, CPUImpl>>(*s0);
    s0.reset();
    s0 = std::make_unique<State<double>>(*hostst);
    check("State -> State<F, CPUImpl> -> State round trip", max_diff(*s0, *init64), 0);

    // 3. dense forms through the operator views (12 sites at most)
    if (len <= 12) {
        using Matrix = typename SumOps<double>::MatrixType;
#ifdef REAPERS_USE_PARITY_BLOCKS
        Matrix A = syk64.fermion_ops(N - 2).LR;
#else
        Matrix A = syk64.fermion_ops(N - 1);
#endif
        const double dim = (double)((size_t)1 << len);
        // chi^2 = 1/2: tr(A A) = dim / 2
        check("tr(chi chi) - dim/2 for Matrix A = fermion_ops(n)",
              std::abs((A * A).trace() - complex<double>(dim / 2)), 1e-9 * dim);
        // exp(-iHt) exp(iHt) = 1 through matexp with the _i literals
        Matrix U = h64.matexp(-t * 1_id), V = h64.matexp(t * 1_id);
        check("tr(exp(-iHt) exp(iHt)) - dim", std::abs((U * V).trace() - complex<double>(dim)),
              1e-9 * dim);
        // 4. Krylov against exact diagonalization
        State<double> k(*init64), e(*init64);
        k.evolve(h64, t, 0.0, 20, 1e-13);
        e.evolve<ExactDiagonalization>(h64, t, 0.0);
        check("Krylov vs ExactDiagonalization at t=1", max_diff(k, e), 1e-11);
    }
    complex<float> li = 2_if;
    complex<double> ld = 2.5_id;
    complex<DefFpType> l = 3_i;
    check("literals 2_if, 2.5_id, 3_i",
          std::abs(li.imag() - 2.0f) + std::abs(ld.imag() - 2.5) + std::abs(l.imag() - 3.0), 0);
    try {
        ThrowException(FieldIndexTooSmall, 0, 1);
    } catch (const IndexTooSmall &ex) {
        check("FieldIndexTooSmall is an IndexTooSmall", 0, 0);
    }
    std::printf("%s\n", failures ? "FAILED" : "all checks passed");
    return failures ? 1 : 0;
}
