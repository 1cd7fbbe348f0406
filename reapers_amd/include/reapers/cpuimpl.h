// cpuimpl.h -- CPUImpl: the host backend of the Impl interface (the
// reference's CPUImpl, inc/cpuimpl.h:513-755), OpenMP over amplitudes.
//
// State<FpType, CPUImpl> and Krylov<CPUImpl> see the same static interface
// as with GPUImpl (VecType, BufType, zero_vec, copy_vec, add_vec, scale_vec,
// vec_prod, vec_norm, init_random, apply_ops, mat_mul), plus the fused
// helpers this build's Krylov driver calls (gs_pass, lincomb,
// apply_ops_overwrite, vec_prod64, vec_norm64, apply_passes).  It is the
// backend of programs built with REAPERS_NOGPU and can be used next to
// GPUImpl: a State converts between the two explicitly
// (State<F, GPUImpl> g(cpu_state); State<F, CPUImpl> c(g);), as in the
// reference (inc/state.h:114-118, inc/gpuimpl.h:438-462).
//
// Arithmetic: like the GPU backend, inner products, norms and the apply's
// sums accumulate in FP64 for both state precisions.  The apply computes
//     res[m] += sum_t c_t (-1)^popcount(m & z_t) vec[m ^ x_t]
// (x = flip mask, z = phase mask, c = coeff (-1)^popcount(x & z): the
// reference's SpinOp::apply, inc/ops.h:421-440) over blocks of 2^12 output
// amplitudes: one block's FP64 sums stay in cache while every term streams
// the input block at (m0 ^ x) with the low bits of x permuting inside it.
// As in the reference, an operator with at least 2^len terms is applied as
// a dense matrix (inc/cpuimpl.h:737-741).
// init_random draws the same Philox4x32-10 stream as the GPU backend's
// rp_init_random for a given seed (counter = amplitude index), so the two
// backends start from the same state up to the rounding of the Box-Muller
// transform.
#pragma once

#ifndef _REAPERS_H_
#error "Do not include this file directly. Include the master header file reapers.h"
#endif

namespace internal {
// Philox4x32-10 (Salmon, Moraes, Dror and Shaw, SC'11): ten rounds of the
// two 32x32->64 multiplies with the key bumped by the Weyl constants
inline void philox4x32_10(u32 (&c)[4], u32 k0, u32 k1)
{
    for (int r = 0; r < 10; r++) {
        const u64 p0 = (u64)0xD2511F53u * c[0], p1 = (u64)0xCD9E8D57u * c[2];
        const u32 hi0 = (u32)(p0 >> 32), lo0 = (u32)p0, hi1 = (u32)(p1 >> 32), lo1 = (u32)p1;
        c[0] = hi1 ^ c[1] ^ k0;
        c[1] = lo1;
        c[2] = hi0 ^ c[3] ^ k1;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}
}  // namespace internal

class CPUImpl {
public:
    // host vector owning dim complex elements (HostVec, inc/cpuimpl.h:520-541)
    template<RealScalar FpType>
    class HostVec {
        size_t dim = 0;
        std::unique_ptr<complex<FpType>[]> vec;
        HostVec(const HostVec &) = delete;
        HostVec &operator=(const HostVec &) = delete;
    public:
        explicit HostVec(size_t n) : dim(n), vec(new complex<FpType>[n]) {}
        HostVec(HostVec &&) = default;
        HostVec &operator=(HostVec &&) = default;
        complex<FpType> *get() { return vec.get(); }
        complex<FpType> *get() const { return vec.get(); }
        size_t dimension() const { return dim; }
        friend void swap(HostVec &a, HostVec &b) {
            std::swap(a.dim, b.dim);
            std::swap(a.vec, b.vec);
        }
    };

    template<RealScalar FpType> using VecType = HostVec<FpType>;
    template<RealScalar FpType> using VecSizeType = size_t;
    template<RealScalar FpType> using BufType = complex<FpType> *;
    template<RealScalar FpType> using ConstBufType = const complex<FpType> *;
    template<RealScalar FpType> using ElemRefType = complex<FpType> &;
    template<RealScalar FpType> using SumOps = HostSumOps<FpType>;

    // the matexp cache budget shared by all operators (inc/cpuimpl.h:570-600)
    static bool cache_reserve(size_t sz) {
        auto &b = internal::HostCacheBudget::get();
        if (b.used + sz > b.max) return false;
        b.used += sz;
        return true;
    }
    static void cache_release(size_t sz) {
        auto &b = internal::HostCacheBudget::get();
        b.used -= std::min(sz, b.used);
    }
    static void set_max_cache_size(size_t sz) { internal::HostCacheBudget::get().max = sz; }
    static size_t get_max_cache_size() { return internal::HostCacheBudget::get().max; }
    static size_t get_current_cache_size() { return internal::HostCacheBudget::get().used; }

    static void ensure_context() {}
    static void synchronize() {}

    template<RealScalar FpType>
    static void zero_vec(size_t size, BufType<FpType> v) {
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < size; i++) v[i] = 0;
    }

    // Re, Im ~ N(0, 1) from Philox4x32-10 keyed by the seed, then normalized
    // with an FP64 norm (inc/cpuimpl.h:600-611 draws from a global mt19937)
    template<RealScalar FpType>
    static void init_random(size_t size, BufType<FpType> v, u64 seed = 0) {
        static u64 counter = 0x5eed;
        if (!seed) seed = ++counter;
        const u32 k0 = (u32)seed, k1 = (u32)(seed >> 32);
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < size; i++) {
            u32 c[4] = {(u32)i, (u32)((u64)i >> 32), 0x52454150u, 0x45525321u};
            internal::philox4x32_10(c, k0, k1);
            const u64 a = ((u64)c[0] << 21) ^ (u64)(c[1] >> 11);
            const u64 b = ((u64)c[2] << 21) ^ (u64)(c[3] >> 11);
            const double u1 = ((double)(a & ((1ull << 53) - 1)) + 1.0) * (1.0 / 9007199254740992.0);
            const double u2 = (double)(b & ((1ull << 53) - 1)) * (1.0 / 9007199254740992.0);
            const double r = std::sqrt(-2.0 * std::log(u1));
            const double th = 2.0 * M_PI * u2;
            v[i] = complex<FpType>((FpType)(r * std::cos(th)), (FpType)(r * std::sin(th)));
        }
        scale_vec(size, v, (FpType)(1.0 / vec_norm64<FpType>(size, v)));
    }

    // v0 = v1, possibly converting precision; double -> float rounds to
    // nearest (as the reference's CPUImpl) or toward zero (its GPU path)
    template<RealScalar F0, RealScalar F1>
    static void copy_vec(size_t size, BufType<F0> v0, ConstBufType<F1> v1,
                         int rounding = RP_ROUND_NEAREST) {
        if constexpr (std::is_same_v<F0, float> && std::is_same_v<F1, double>) {
            if (rounding == RP_ROUND_TOWARD_ZERO) {
                auto rz = [](double d) {
                    float f = (float)d;
                    if (std::abs((double)f) > std::abs(d)) f = std::nextafter(f, 0.0f);
                    return f;
                };
#pragma omp parallel for schedule(static)
                for (size_t i = 0; i < size; i++)
                    v0[i] = complex<float>(rz(v1[i].real()), rz(v1[i].imag()));
                return;
            }
        }
        (void)rounding;
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < size; i++)
            v0[i] = complex<F0>((F0)v1[i].real(), (F0)v1[i].imag());
    }

    // res = column idx of a size x size matrix (inc/cpuimpl.h:635-642)
    template<RealScalar FpType>
    static void copy_vec(size_t size, BufType<FpType> res, const HostMatrix<FpType> &mat,
                         size_t idx) {
        if (mat.rows() != size || idx >= mat.cols())
            ThrowException(InvalidArgument, "column", "out of range");
        std::memcpy(res, mat.data() + idx * size, sizeof(complex<FpType>) * size);
    }

    template<RealScalar FpType>
    static void add_vec(size_t size, BufType<FpType> v0, ConstBufType<FpType> v1) {
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < size; i++) v0[i] += v1[i];
    }

    template<RealScalar FpType>
    static void add_vec(size_t size, BufType<FpType> v0, FpType s, ConstBufType<FpType> v1) {
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < size; i++) v0[i] += s * v1[i];
    }

    template<RealScalar FpType>
    static void add_vec(size_t size, BufType<FpType> v0, complex<FpType> s,
                        ConstBufType<FpType> v1) {
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < size; i++) v0[i] += s * v1[i];
    }

    template<RealScalar FpType>
    static void add_vec(size_t size, BufType<FpType> res, ConstBufType<FpType> v0,
                        ConstBufType<FpType> v1) {
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < size; i++) res[i] = v0[i] + v1[i];
    }

    template<RealScalar FpType>
    static void scale_vec(size_t size, BufType<FpType> v, FpType s) {
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < size; i++) v[i] *= s;
    }

    template<RealScalar FpType>
    static void scale_vec(size_t size, BufType<FpType> v, complex<FpType> s) {
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < size; i++) v[i] *= s;
    }

    // Sum of f(i) over [0, size) in FP64, deterministic: the range is cut
    // into kChunks pieces independent of the thread count, each summed in
    // order, and the pieces added in order
    static constexpr int kChunks = 256;
    template<typename T, typename Fn>
    static T det_sum(size_t size, Fn &&f) {
        T part[kChunks] = {};
        const size_t per = (size + kChunks - 1) / kChunks;
#pragma omp parallel for schedule(static)
        for (int c = 0; c < kChunks; c++) {
            T s{};
            const size_t lo = (size_t)c * per, hi = std::min(size, lo + per);
            for (size_t i = lo; i < hi; i++) s += f(i);
            part[c] = s;
        }
        T s{};
        for (int c = 0; c < kChunks; c++) s += part[c];
        return s;
    }

    // <v0|v1> and ||v|| accumulated in FP64
    template<RealScalar FpType>
    static std::complex<double> vec_prod64(size_t size, ConstBufType<FpType> v0,
                                           ConstBufType<FpType> v1) {
        return det_sum<std::complex<double>>(size, [&](size_t i) {
            return std::conj(std::complex<double>(v0[i])) * std::complex<double>(v1[i]);
        });
    }

    template<RealScalar FpType>
    static double vec_norm64(size_t size, ConstBufType<FpType> v) {
        return std::sqrt(det_sum<double>(size, [&](size_t i) {
            return std::norm(std::complex<double>(v[i]));
        }));
    }

    template<RealScalar FpType>
    static complex<FpType> vec_prod(size_t size, ConstBufType<FpType> v0,
                                    ConstBufType<FpType> v1) {
        auto c = vec_prod64<FpType>(size, v0, v1);
        return complex<FpType>{(FpType)c.real(), (FpType)c.imag()};
    }

    template<RealScalar FpType>
    static FpType vec_norm(size_t size, ConstBufType<FpType> v) {
        return (FpType)vec_norm64<FpType>(size, v);
    }

    // res = mat * vec (inc/cpuimpl.h:718-726); res and vec must differ
    template<RealScalar FpType>
    static void mat_mul(size_t dim, BufType<FpType> res, const HostMatrix<FpType> &mat,
                        ConstBufType<FpType> vec) {
        if (mat.rows() != dim || mat.cols() != dim)
            ThrowException(InvalidArgument, "matrix dimensions", "must match the state");
        matvec(dim, res, mat, vec, false);
    }

    // res += ops * vec (inc/cpuimpl.h:737-754)
    template<RealScalar FpType>
    static void apply_ops(typename SpinOp<FpType>::IndexType len, BufType<FpType> res,
                          const HostSumOps<FpType> &ops, ConstBufType<FpType> vec) {
        apply_impl(len, res, ops, vec, true);
    }

    // res = ops * vec
    template<RealScalar FpType>
    static void apply_ops_overwrite(typename SpinOp<FpType>::IndexType len, BufType<FpType> res,
                                    const HostSumOps<FpType> &ops, ConstBufType<FpType> vec) {
        apply_impl(len, res, ops, vec, false);
    }

    // memory traffic of one apply in state-vector passes (one read of the
    // input per term block sweep, one write): the proposer's cost input
    template<RealScalar FpType>
    static double apply_passes(typename SpinOp<FpType>::IndexType, const HostSumOps<FpType> &ops) {
        return 1.0 + (double)ops.size();
    }

    // one classical Gram-Schmidt pass of w against V (FP64 coefficients):
    // h = V^H w, w -= V h, returns ||w||; with fuse_min > 0 the pass also
    // normalizes w when its norm, known from ||w||^2 - |h|^2 without
    // cancellation (>= ||w||^2 / 2), is at least fuse_min
    template<RealScalar FpType>
    static double gs_pass(size_t size, const std::vector<const HostVec<FpType> *> &V,
                          BufType<FpType> w, std::vector<std::complex<double>> &h,
                          double fuse_min, bool &normalized) {
        const int k = (int)V.size();
        h.assign(k, {0.0, 0.0});
        for (int q = 0; q < k; q++) h[q] = vec_prod64<FpType>(size, V[q]->get(), w);
        const double wn2 = std::pow(vec_norm64<FpType>(size, w), 2);
        double hn2 = 0;
        for (auto &c : h) hn2 += std::norm(c);
        normalized = false;
        double scale = 1.0;
        if (fuse_min > 0) {
            const double b2 = wn2 - hn2;
            if (b2 >= wn2 / 2 && std::sqrt(b2) >= fuse_min) {
                normalized = true;
                scale = 1.0 / std::sqrt(b2);
            }
        }
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < size; i++) {
            std::complex<double> s(w[i]);
            for (int q = 0; q < k; q++) s -= h[q] * std::complex<double>(V[q]->get()[i]);
            s *= scale;
            w[i] = complex<FpType>((FpType)s.real(), (FpType)s.imag());
        }
        if (normalized) return std::sqrt(wn2 - hn2);
        return vec_norm64<FpType>(size, w);
    }

    template<RealScalar FpType>
    static double gs_pass(size_t size, const std::vector<const HostVec<FpType> *> &V,
                          BufType<FpType> w, std::vector<std::complex<double>> &h) {
        bool normalized;
        return gs_pass(size, V, w, h, 0.0, normalized);
    }

    // y += sum_i c[i] V[i] in one sweep (FP64 sums)
    template<RealScalar FpType>
    static void lincomb(size_t size, BufType<FpType> y,
                        const std::vector<const HostVec<FpType> *> &V,
                        const std::vector<std::complex<double>> &c) {
        const int k = (int)V.size();
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < size; i++) {
            std::complex<double> s(y[i]);
            for (int q = 0; q < k; q++) s += c[q] * std::complex<double>(V[q]->get()[i]);
            y[i] = complex<FpType>((FpType)s.real(), (FpType)s.imag());
        }
    }

private:
    template<RealScalar FpType>
    static void matvec(size_t dim, complex<FpType> *res, const HostMatrix<FpType> &mat,
                       const complex<FpType> *vec, bool accumulate) {
        const long n = (long)dim;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; i++) {
            std::complex<double> s = accumulate ? std::complex<double>(res[i]) : 0.0;
            for (size_t j = 0; j < dim; j++)
                s += std::complex<double>(mat(i, j)) * std::complex<double>(vec[j]);
            res[i] = complex<FpType>((FpType)s.real(), (FpType)s.imag());
        }
    }

    template<RealScalar FpType>
    static void apply_impl(int len, complex<FpType> *res, const HostSumOps<FpType> &ops,
                           const complex<FpType> *vec, bool accumulate) {
        if (len < 1 || len > NUM_SPIN_SITES) ThrowException(InvalidArgument, "len", "out of range");
        const size_t dim = (size_t)1 << len;
        if (res == vec) ThrowException(InvalidArgument, "res and vec", "must be different buffers");
        if (len <= 14 && ops.size() >= dim) {
            matvec(dim, res, ops.get_matrix(len), vec, accumulate);
            return;
        }
        struct Term { u32 x, z; std::complex<double> c; };
        std::vector<Term> terms;
        terms.reserve(ops.size());
        for (auto &op : ops) {
            const u64 b = op.get_bit_string();
            Term t{internal::flip_mask(b), internal::phase_mask(b),
                   std::complex<double>(op.get_coeff())};
            if (__builtin_popcount(t.x & t.z) & 1) t.c = -t.c;
            if (t.c != std::complex<double>(0)) terms.push_back(t);
        }
        const int bb = std::min(len, 12);
        const size_t B = (size_t)1 << bb;
        const u32 lowmask = (u32)(B - 1);
        const long nblk = (long)(dim >> bb);
#pragma omp parallel
        {
            std::vector<double> acc(2 * B);
#pragma omp for schedule(static)
            for (long blk = 0; blk < nblk; blk++) {
                const size_t m0 = (size_t)blk << bb;
                for (size_t j = 0; j < B; j++) {
                    acc[2 * j] = accumulate ? (double)res[m0 + j].real() : 0.0;
                    acc[2 * j + 1] = accumulate ? (double)res[m0 + j].imag() : 0.0;
                }
                for (const Term &t : terms) {
                    const u32 xl = t.x & lowmask, zl = t.z & lowmask;
                    const complex<FpType> *in = vec + ((m0 ^ t.x) & ~(size_t)lowmask);
                    const double sgn0 = (__builtin_popcountll(m0 & t.z) & 1) ? -1.0 : 1.0;
                    const double cr = sgn0 * t.c.real(), ci = sgn0 * t.c.imag();
                    for (size_t j = 0; j < B; j++) {
                        const complex<FpType> v = in[j ^ xl];
                        const double s = (__builtin_popcount((u32)j & zl) & 1) ? -1.0 : 1.0;
                        const double vr = v.real(), vi = v.imag();
                        acc[2 * j] += s * (cr * vr - ci * vi);
                        acc[2 * j + 1] += s * (cr * vi + ci * vr);
                    }
                }
                for (size_t j = 0; j < B; j++)
                    res[m0 + j] = complex<FpType>((FpType)acc[2 * j], (FpType)acc[2 * j + 1]);
            }
        }
    }
};
