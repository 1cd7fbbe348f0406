// xxz.h -- random-field XXZ chain, the second model family the Krylov
// schedule is trained on (the reference has SYK only):
//
//   H = J sum_{i<L-1} (sx_i sx_{i+1} + sy_i sy_{i+1} + Delta sz_i sz_{i+1})
//       + sum_i h_i sz_i,      h_i ~ U[-W, W]  (std::uniform_real_distribution)
//
// with Pauli matrices (not spin-1/2 operators), open boundary: 4L-3 terms.
// The strings are built with SpinOp products, so sigma_y carries its i as in
// the rest of the API.
#pragma once

#ifndef _REAPERS_H_
#error "Do not include this file directly. Include the master header file reapers.h"
#endif

template<RealScalar FpType = DefFpType, typename Impl = DefImpl>
class XXZ {
    int L;
    FpType J, Delta, W;

public:
    using HamOp = SubspaceView<typename Impl::template SumOps<FpType>>;

    XXZ(int L, FpType J, FpType Delta, FpType W) : L(L), J(J), Delta(Delta), W(W) {
        if (L < 2) ThrowException(InvalidArgument, "L", "must be at least two");
        if (L > NUM_SPIN_SITES) ThrowException(InvalidArgument, "L", "cannot exceed NUM_SPIN_SITES");
    }

    template<typename RandGen>
    HamOp gen_ham(RandGen &gen) const {
        using S = SpinOp<FpType>;
        typename Impl::template SumOps<FpType> ham;
        for (int i = 0; i + 1 < L; i++) {
            ham += complex<FpType>(J) * (S::sigma_x(i) * S::sigma_x(i + 1));
            ham += complex<FpType>(J) * (S::sigma_y(i) * S::sigma_y(i + 1));
            ham += complex<FpType>(J * Delta) * (S::sigma_z(i) * S::sigma_z(i + 1));
        }
        std::uniform_real_distribution<FpType> field(-W, W);
        for (int i = 0; i < L; i++) ham += complex<FpType>(field(gen)) * S::sigma_z(i);
        return HamOp(L, std::move(ham));
    }
};
