// syk_block_otoc.cc -- a user program of the REAPERS C++ API on this
// backend, written the way the reference's production driver uses the
// library (examples/syk.cc with REAPERS_USE_PARITY_BLOCKS): parity-block SYK
// from a seeded std::mt19937, the Lanczos ground-state energy of H_L, the
// shift, the beta/8 pre-evolution, then for t_k = k tmax / nsteps the OTOC
//     s = s0;  s *= op;  evolve(H_R, t, beta/4);  evolve(H_L, -t, beta/4);
//     s *= op; evolve(H_R, t, beta/4);  evolve(H_L, -t, 0);  v_k = <s0|s>
// and, with --trace, the trace definition by exact diagonalization.  Prints
// one "t Re Im" line per time point (v_k / v_0).
//
// usage: syk_block_otoc N sparsity seed beta tmax nsteps [--trace] [--fp64]
#define REAPERS_USE_PARITY_BLOCKS
#include <reapers.h>

#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>

using namespace REAPERS;
using namespace REAPERS::Model;
using namespace REAPERS::EvolutionAlgorithm;

template<RealScalar FpType>
static int run(int N, double sp, unsigned seed, double beta, double tmax, int nsteps, bool trace)
{
    using HamOp = typename SYK<FpType>::HamOp;
    SYK<FpType> syk(N, (FpType)sp);
    std::mt19937 rg(seed);
    HamOp ham = syk.gen_ham(rg);
    const auto &ops = syk.fermion_ops();
    auto op = ops[N - 2].LR;
    const int krylov_dim = 12;
    std::vector<complex<double>> v(nsteps + 1);
    double dt = tmax / nsteps;
    if (trace) {
        const auto &expmbH = ham.LL.matexp((FpType)-beta);
        for (int k = 0; k <= nsteps; k++) {
            FpType t = (FpType)(k * dt);
            const auto &m0 = ham.RR.matexp(complex<FpType>{0, -t});
            const auto &m1 = ham.LL.matexp(complex<FpType>{0, t});
            const auto &m2 = op.get_matrix();
            v[k] = complex<double>((m1 * m0 * m2 * m1 * m0 * m2 * expmbH).trace());
        }
    } else {
        State<FpType> s1(N / 2 - 1);
        FpType gnd = s1.ground_state(ham.LL, krylov_dim, (FpType)0.1);
        ham.LL += (-gnd) * SpinOp<FpType>::identity();
        ham.RR += (-gnd) * SpinOp<FpType>::identity();
        State<FpType> s0(N / 2 - 1);
        s0.random_state(seed + 1);
        s0.evolve(ham.LL, 0.0, beta / 8, krylov_dim);
        s0.gc();
        for (int k = 0; k <= nsteps; k++) {
            double t = k * dt;
            State<FpType> s(s0);
            s *= op;
            s.evolve(ham.RR, t, beta / 4, krylov_dim, (FpType)1e-7);
            s.evolve(ham.LL, -t, beta / 4, krylov_dim, (FpType)1e-7);
            s *= op;
            s.evolve(ham.RR, t, beta / 4, krylov_dim, (FpType)1e-7);
            s.evolve(ham.LL, -t, 0.0, krylov_dim, (FpType)1e-7);
            v[k] = s0.inner64(s);
        }
    }
    write_results(std::cout, dt, [&] {
        std::vector<complex<double>> w(v);
        for (auto &x : w) x /= v[0];
        return w;
    }());
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 7) {
        std::cerr << "usage: " << argv[0] << " N sparsity seed beta tmax nsteps [--trace] [--fp64]"
                  << std::endl;
        return 2;
    }
    int N = std::atoi(argv[1]);
    double sp = std::atof(argv[2]);
    unsigned seed = (unsigned)std::strtoul(argv[3], nullptr, 10);
    double beta = std::atof(argv[4]), tmax = std::atof(argv[5]);
    int nsteps = std::atoi(argv[6]);
    bool trace = false, fp64 = false;
    for (int i = 7; i < argc; i++) {
        if (!std::strcmp(argv[i], "--trace")) trace = true;
        else if (!std::strcmp(argv[i], "--fp64")) fp64 = true;
    }
    try {
        return fp64 ? run<double>(N, sp, seed, beta, tmax, nsteps, trace)
                    : run<float>(N, sp, seed, beta, tmax, nsteps, trace);
    } catch (const std::exception &e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
}
