// gpuimpl.h -- GPUImpl: the reference's GPU backend interface
// (inc/gpuimpl.h:396-1048) implemented on the C ABI of libreapers_hip.so.
//
// State<FpType, GPUImpl> and Krylov<GPUImpl> only see the static interface
// below (VecType, BufType, zero_vec, copy_vec, add_vec, scale_vec, vec_prod,
// vec_norm, init_random, apply_ops), as in the reference.  Additions that the
// reference does not have:
//   - vec_prod64 / vec_norm64 return the FP64-accumulated values;
//   - gs_pass / lincomb are one fused sweep each (Krylov uses them);
//   - copy_vec from double to float takes an explicit rounding mode
//     (default: round to nearest, as CPUImpl; the reference's GPU path
//     truncates, src/krnl.cu:33-39).
// Every call is checked and throws the reference's exception types.  There is
// no CPU fallback: if the HIP library cannot initialize a device, the first
// use throws NoGpuDevice.  CPUImpl (cpuimpl.h) is a separate backend that a
// program selects (REAPERS_NOGPU, or State<F, CPUImpl>); states convert
// between the two explicitly.
#pragma once

#ifndef _REAPERS_H_
#error "Do not include this file directly. Include the master header file reapers.h"
#endif

class GPUImpl;

namespace internal {
template<RealScalar FpType> constexpr int rp_prec() {
    return std::is_same_v<FpType, float> ? RP_C64 : RP_C128;
}
}  // namespace internal

template<RealScalar FpType> class DevMatrix;
namespace internal {
template<RealScalar FpType> struct DenseCache;
}

// A SumOps whose device layout is uploaded on first use and dropped on every
// host mutation (DevSumOps::upload/release, inc/gpuimpl.h:1064-1105).  Its
// dense forms (matrix, eigensystem, matrix exponentials; dense.h) are built
// on request and dropped the same way.
template<RealScalar FpType = DefFpType>
class DevSumOps : public HostSumOps<FpType> {
    friend class GPUImpl;
    mutable rp_terms_t handle = nullptr;
    mutable int handle_len = -1;
    mutable std::unique_ptr<internal::DenseCache<FpType>> dense;

    void free_handle() const {
        if (handle) {
            rp_terms_free(handle);
            handle = nullptr;
            handle_len = -1;
        }
    }

    void release() const override {
        HostSumOps<FpType>::release();
        free_handle();
        dense.reset();
    }

    internal::DenseCache<FpType> &dense_at(int len) const;
    DevMatrix<FpType> _matexp(complex<FpType> c, int len) const;

public:
    using HostSumOps<FpType>::HostSumOps;
    DevSumOps(const HostSumOps<FpType> &ops) : HostSumOps<FpType>(ops) {}
    DevSumOps(const DevSumOps &ops) : HostSumOps<FpType>(ops) {}
    DevSumOps(DevSumOps &&ops) : HostSumOps<FpType>(std::move(ops)) {
        std::swap(handle, ops.handle);
        std::swap(handle_len, ops.handle_len);
        std::swap(dense, ops.dense);
    }
    template<RealScalar F1>
    explicit DevSumOps(const DevSumOps<F1> &ops) : HostSumOps<FpType>(ops) {}
    DevSumOps &operator=(const DevSumOps &rhs) {
        HostSumOps<FpType>::operator=(rhs);
        return *this;
    }
    DevSumOps &operator=(DevSumOps &&rhs) {
        if (this == &rhs) return *this;
        HostSumOps<FpType>::operator=(static_cast<const HostSumOps<FpType> &>(rhs));
        return *this;
    }
    using HostSumOps<FpType>::operator=;
    ~DevSumOps() override { release(); }

    // the uploaded layout for chain length len (uploads on first use)
    rp_terms_t device_terms(int len) const;

    // layout statistics of the uploaded form (groups, bytes per apply)
    rp_terms_info_t device_info(int len) const {
        rp_terms_info_t info;
        RP_CALL(rp_terms_info(device_terms(len), &info));
        return info;
    }

    // dense forms on len sites (get_matrix, get_eigensystem, matexp;
    // inc/gpuimpl.h:370-388): cached until the term list changes
    using MatrixType = DevMatrix<FpType>;
    using EigenVals = std::vector<FpType>;
    using EigenSystem = std::tuple<const EigenVals &, const MatrixType &>;
    const MatrixType &get_matrix(int len) const;
    EigenSystem get_eigensystem(int len) const;
    // exp(c H), cached by c (MatexpCache, inc/cpuimpl.h:29-88)
    const MatrixType &matexp(complex<FpType> c, int len) const;
    size_t matexp_cache_entries() const;
};

class GPUImpl {
    template<RealScalar F> friend class DevSumOps;

    struct Context {
        Context() {
            int dev = 0;
            if (const char *e = std::getenv("REAPERS_DEVICE")) dev = std::atoi(e);
            int st = rp_init(dev);
            if (st != RP_OK) internal::throw_status(__FILE__, __LINE__, __func__, st);
        }
    };

public:
    static void ensure_context() { static Context c; (void)c; }

    // device vector owning 2^len complex elements (DeviceVec,
    // inc/gpuimpl.h:404-494)
    template<RealScalar FpType>
    class DeviceVec {
        void *ptr = nullptr;
        size_t dim = 0;
        friend class GPUImpl;
        DeviceVec(const DeviceVec &) = delete;
        DeviceVec &operator=(const DeviceVec &) = delete;
    public:
        explicit DeviceVec(size_t n) : dim(n) {
            ensure_context();
            int st = rp_malloc(&ptr, n * sizeof(complex<FpType>));
            if (st != RP_OK)
                internal::throw_status(__FILE__, __LINE__, __func__, st,
                                       n * sizeof(complex<FpType>));
        }
        DeviceVec(DeviceVec &&v) : ptr(v.ptr), dim(v.dim) { v.ptr = nullptr; v.dim = 0; }
        DeviceVec &operator=(DeviceVec &&v) {
            if (this != &v) {
                if (ptr) rp_free(ptr);
                ptr = v.ptr; dim = v.dim;
                v.ptr = nullptr; v.dim = 0;
            }
            return *this;
        }
        ~DeviceVec() { if (ptr) rp_free(ptr); }
        DeviceVec &get() { return *this; }
        const DeviceVec &get() const { return *this; }
        void *data() { return ptr; }
        const void *data() const { return ptr; }
        size_t dimension() const { return dim; }
        friend void swap(DeviceVec &a, DeviceVec &b) {
            std::swap(a.ptr, b.ptr);
            std::swap(a.dim, b.dim);
        }
        complex<FpType> operator[](size_t idx) const {
            complex<FpType> c;
            RP_CALL(rp_memcpy(&c, (const complex<FpType> *)ptr + idx, sizeof(c), RP_D2H));
            return c;
        }
        // a writable element (DeviceVec::ElemRefType, inc/gpuimpl.h:413-424):
        // assignment copies one value to the device, reading copies it back
        class ElemRefType {
            DeviceVec &vec;
            size_t idx;
        public:
            ElemRefType(DeviceVec &v, size_t i) : vec(v), idx(i) {}
            ElemRefType &operator=(complex<FpType> s) {
                RP_CALL(rp_memcpy((complex<FpType> *)vec.ptr + idx, &s, sizeof(s), RP_H2D));
                return *this;
            }
            operator complex<FpType>() const {
                return static_cast<const DeviceVec &>(vec)[idx];
            }
        };
        ElemRefType operator[](size_t idx) { return ElemRefType(*this, idx); }
        // conversions to and from the CPU backend's vectors (DeviceVec's
        // CPUImpl::VecType constructor and conversion operator,
        // inc/gpuimpl.h:438-462): one host<->device copy each
        explicit DeviceVec(const CPUImpl::HostVec<FpType> &v) : DeviceVec(v.dimension()) {
            upload(v.get());
        }
        explicit operator CPUImpl::HostVec<FpType>() const {
            CPUImpl::HostVec<FpType> v(dim);
            download(v.get());
            return v;
        }
        // host <-> device transfers of the whole vector
        void upload(const complex<FpType> *host) {
            RP_CALL(rp_memcpy(ptr, host, dim * sizeof(complex<FpType>), RP_H2D));
        }
        void download(complex<FpType> *host) const {
            RP_CALL(rp_memcpy(host, ptr, dim * sizeof(complex<FpType>), RP_D2H));
        }
    };

    template<RealScalar FpType> using VecType = DeviceVec<FpType>;
    template<RealScalar FpType> using VecSizeType = size_t;
    template<RealScalar FpType> using BufType = DeviceVec<FpType> &;
    template<RealScalar FpType> using ConstBufType = const DeviceVec<FpType> &;
    template<RealScalar FpType> using SumOps = DevSumOps<FpType>;
    template<RealScalar FpType> using ElemRefType = typename DeviceVec<FpType>::ElemRefType;

private:
    struct CacheBudget {
        size_t max = 0, used = 0;
        bool set = false;
    };
    static CacheBudget &cache_budget() {
        static CacheBudget b;
        if (!b.set) {
            ensure_context();
            size_t fr = 0, tot = 0;
            b.max = rp_mem_get_info(&fr, &tot) == RP_OK ? fr / 2 : ((size_t)1 << 34);
            b.set = true;
        }
        return b;
    }

public:
    // the matexp cache budget shared by all operators (inc/gpuimpl.h:
    // 560-597): half of the free device memory at first use by default
    static bool cache_reserve(size_t sz) {
        auto &b = cache_budget();
        if (b.used + sz > b.max) return false;
        b.used += sz;
        return true;
    }
    static void cache_release(size_t sz) {
        auto &b = cache_budget();
        b.used -= std::min(sz, b.used);
    }
    static void set_max_cache_size(size_t sz) { cache_budget().max = sz; }
    static void set_max_cache_size(float fraction) {
        size_t fr = 0, tot = 0;
        ensure_context();
        RP_CALL(rp_mem_get_info(&fr, &tot));
        cache_budget().max = (size_t)(fr * (double)fraction);
    }
    static size_t get_max_cache_size() { return cache_budget().max; }
    static size_t get_current_cache_size() { return cache_budget().used; }

    static void synchronize() { RP_CALL(rp_synchronize()); }

    // flags of every later term upload: RP_TERMS_AUTOTUNE times the apply
    // tilings once per Hamiltonian, RP_TERMS_ACC_NATIVE accumulates FP32
    // applies in FP32 (no counterpart in the reference, whose kernel has one
    // launch shape, src/krnl.cu:236-246)
    static int &upload_flags() {
        static int flags = 0;
        return flags;
    }

    template<RealScalar FpType>
    static void zero_vec(size_t size, BufType<FpType> v) {
        RP_CALL(rp_zero_vec(internal::rp_prec<FpType>(), v.ptr, size));
    }

    // Haar random state from a seeded Philox stream (the reference seeds
    // cuRAND from std::random_device, inc/gpuimpl.h:540-551)
    template<RealScalar FpType>
    static void init_random(size_t size, BufType<FpType> v, u64 seed = 0) {
        static u64 counter = 0x5eed;
        if (!seed) seed = ++counter;
        RP_CALL(rp_init_random(internal::rp_prec<FpType>(), v.ptr, size, seed));
    }

    template<RealScalar F0, RealScalar F1>
    static void copy_vec(size_t size, BufType<F0> v0, ConstBufType<F1> v1,
                         int rounding = RP_ROUND_NEAREST) {
        RP_CALL(rp_copy_vec(internal::rp_prec<F0>(), v0.ptr, internal::rp_prec<F1>(), v1.ptr,
                            size, rounding));
    }

    template<RealScalar FpType>
    static void add_vec(size_t size, BufType<FpType> v0, ConstBufType<FpType> v1) {
        RP_CALL(rp_add_vec(internal::rp_prec<FpType>(), v0.ptr, v1.ptr, size));
    }

    template<RealScalar FpType>
    static void add_vec(size_t size, BufType<FpType> res, ConstBufType<FpType> v0,
                        ConstBufType<FpType> v1) {
        RP_CALL(rp_add_vec3(internal::rp_prec<FpType>(), res.ptr, v0.ptr, v1.ptr, size));
    }

    template<RealScalar FpType>
    static void add_vec(size_t size, BufType<FpType> v0, complex<FpType> s,
                        ConstBufType<FpType> v1) {
        RP_CALL(rp_axpy(internal::rp_prec<FpType>(), v0.ptr, (double)s.real(), (double)s.imag(),
                        v1.ptr, size));
    }

    template<RealScalar FpType>
    static void add_vec(size_t size, BufType<FpType> v0, FpType s, ConstBufType<FpType> v1) {
        RP_CALL(rp_axpy(internal::rp_prec<FpType>(), v0.ptr, (double)s, 0.0, v1.ptr, size));
    }

    template<RealScalar FpType>
    static void scale_vec(size_t size, BufType<FpType> v, FpType s) {
        RP_CALL(rp_scale_vec(internal::rp_prec<FpType>(), v.ptr, (double)s, 0.0, size));
    }

    template<RealScalar FpType>
    static void scale_vec(size_t size, BufType<FpType> v, complex<FpType> s) {
        RP_CALL(rp_scale_vec(internal::rp_prec<FpType>(), v.ptr, (double)s.real(),
                             (double)s.imag(), size));
    }

    // FP64-accumulated <v0|v1> and ||v|| for either element precision
    template<RealScalar FpType>
    static std::complex<double> vec_prod64(size_t size, ConstBufType<FpType> v0,
                                           ConstBufType<FpType> v1) {
        double re, im;
        RP_CALL(rp_vec_prod(internal::rp_prec<FpType>(), v0.ptr, v1.ptr, size, &re, &im));
        return {re, im};
    }

    template<RealScalar FpType>
    static double vec_norm64(size_t size, ConstBufType<FpType> v) {
        double n;
        RP_CALL(rp_vec_norm(internal::rp_prec<FpType>(), v.ptr, size, &n));
        return n;
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

    // res += ops * vec (GPUImpl::apply_ops, src/krnl.cu:220-257)
    template<RealScalar FpType>
    static void apply_ops(typename SpinOp<FpType>::IndexType len, BufType<FpType> res,
                          const DevSumOps<FpType> &ops, ConstBufType<FpType> vec) {
        if (use_dense(len, ops)) return dense_apply_impl(len, res, ops, vec, true);
        RP_CALL(rp_apply_ops(ops.device_terms(len), res.ptr, vec.ptr, 1));
    }

    // res = ops * vec, with the zeroing fused into the apply
    template<RealScalar FpType>
    static void apply_ops_overwrite(typename SpinOp<FpType>::IndexType len, BufType<FpType> res,
                                    const DevSumOps<FpType> &ops, ConstBufType<FpType> vec) {
        if (use_dense(len, ops)) return dense_apply_impl(len, res, ops, vec, false);
        RP_CALL(rp_apply_ops(ops.device_terms(len), res.ptr, vec.ptr, 0));
    }

    // Apply through the dense matrix when the term count reaches 2^len (the
    // reference's switch in DevSumOps::upload, inc/gpuimpl.h:1068-1084), up
    // to 15 sites.  Off by default: at those sizes the term-split kernel is
    // faster (a dense SYK N=24 evolution: 3.8 ms against 11.2 ms in FP64,
    // 7.3 ms against 37.8 ms in FP32) and sums FP32 states in FP64, where
    // the matrix-vector product sums in the state's precision.
    static bool &dense_apply() {
        static bool on = false;
        return on;
    }

private:
    template<RealScalar FpType>
    static bool use_dense(int len, const DevSumOps<FpType> &ops) {
        return dense_apply() && len <= 15 && ops.size() >= ((size_t)1 << len);
    }

    template<RealScalar FpType>
    static void dense_apply_impl(int len, BufType<FpType> res, const DevSumOps<FpType> &ops,
                                 ConstBufType<FpType> vec, bool accumulate) {
        const auto &m = ops.get_matrix(len);
        size_t dim = (size_t)1 << len;
        RP_CALL(rp_dense_gemm(internal::rp_prec<FpType>(), dim, 1, dim, 0, 0, 1.0, 0.0, m.data(),
                              dim, vec.ptr, dim, accumulate ? 1.0 : 0.0, 0.0, res.ptr, dim));
    }

public:

    // HBM traffic of one apply in units of one state vector (the uploaded
    // layout's tile groups plus the output write)
    template<RealScalar FpType>
    // the apply's cost in state-vector passes: as the autotuner measured it
    // at upload (kept tiling's time over one streaming pass), else the tile
    // stream's count G_b + 1
    static double apply_passes(typename SpinOp<FpType>::IndexType len,
                               const DevSumOps<FpType> &ops) {
        rp_terms_info_t info = ops.device_info(len);
        if (info.apply_passes > 0) return info.apply_passes;
        double vec = (double)((size_t)1 << len) * 2 * sizeof(FpType);
        return std::max(1.0, (double)info.bytes_per_apply / vec);
    }

    // one fused Gram-Schmidt pass of w against V[0..k): h = V^H w (FP64),
    // w -= V h, returns ||w|| (FP64)
    template<RealScalar FpType>
    static double gs_pass(size_t size, const std::vector<const DeviceVec<FpType> *> &V,
                          BufType<FpType> w, std::vector<std::complex<double>> &h) {
        bool normalized;
        return gs_pass(size, V, w, h, 0.0, normalized);
    }

    // the same pass, writing w / ||w|| when the norm is known to be at least
    // fuse_min and free of cancellation (normalized is then true)
    template<RealScalar FpType>
    static double gs_pass(size_t size, const std::vector<const DeviceVec<FpType> *> &V,
                          BufType<FpType> w, std::vector<std::complex<double>> &h,
                          double fuse_min, bool &normalized) {
        std::vector<const void *> p(V.size());
        for (size_t i = 0; i < V.size(); i++) p[i] = V[i]->ptr;
        h.resize(V.size());
        double nrm;
        int nz = 0;
        RP_CALL(rp_gs_pass_fused(internal::rp_prec<FpType>(), p.data(), (int)V.size(), w.ptr,
                                 size, (double *)h.data(), &nrm, fuse_min, &nz));
        normalized = nz != 0;
        return nrm;
    }

    // Whether Krylov's Arnoldi steps use apply_gs for the apply and their
    // first Gram-Schmidt pass.  Off by default: at 24 qubits the epilogue's
    // reads of the basis cost the apply more (65.9 -> 77.6 ms per N=48
    // evolution) than the separate dot sweep it saves (21.0 -> 11.8 ms).
    static bool &fused_first_pass() {
        static bool on = false;
        return on;
    }

    // w = ops * v fused with the first Gram-Schmidt pass of w against V
    // (rp_apply_gs: the pass's coefficients formed by the apply kernel as it
    // writes w); returns ||w|| after the pass, or beta when normalized
    template<RealScalar FpType>
    static double apply_gs(typename SpinOp<FpType>::IndexType len, BufType<FpType> w,
                           const DevSumOps<FpType> &ops, ConstBufType<FpType> v,
                           const std::vector<const DeviceVec<FpType> *> &V,
                           std::vector<std::complex<double>> &h, double fuse_min,
                           bool &normalized) {
        if (use_dense(len, ops)) {
            dense_apply_impl(len, w, ops, v, false);
            return gs_pass((size_t)1 << len, V, w, h, fuse_min, normalized);
        }
        std::vector<const void *> p(V.size());
        for (size_t i = 0; i < V.size(); i++) p[i] = V[i]->ptr;
        h.resize(V.size());
        double nrm;
        int nz = 0;
        RP_CALL(rp_apply_gs(ops.device_terms(len), w.ptr, v.ptr, p.data(), (int)V.size(),
                            (double *)h.data(), &nrm, fuse_min, &nz));
        normalized = nz != 0;
        return nrm;
    }

    // res = mat * vec for a dim x dim device matrix (mat_mul,
    // inc/gpuimpl.h:1009-1022); res and vec must differ
    template<RealScalar FpType, typename Mat>
    static void mat_mul(size_t dim, BufType<FpType> res, const Mat &mat, ConstBufType<FpType> vec) {
        if (mat.rows() != dim || mat.cols() != dim)
            ThrowException(InvalidArgument, "matrix dimensions", "must match the state");
        RP_CALL(rp_dense_gemm(internal::rp_prec<FpType>(), dim, 1, dim, 0, 0, 1.0, 0.0,
                              mat.data(), dim, vec.ptr, dim, 0.0, 0.0, res.ptr, dim));
    }

    // v = column idx of a dim x dim device matrix (copy_vec,
    // inc/gpuimpl.h:891-906)
    template<RealScalar FpType, typename Mat>
        requires requires(const Mat &m) { m.rows(); m.cols(); m.data(); }
    static void copy_vec(size_t dim, BufType<FpType> v, const Mat &mat, size_t idx) {
        if (mat.rows() != dim || idx >= mat.cols())
            ThrowException(InvalidArgument, "column", "out of range");
        RP_CALL(rp_memcpy(v.ptr, (const complex<FpType> *)mat.data() + idx * dim,
                          dim * sizeof(complex<FpType>), RP_D2D));
    }

    // y += sum_i c[i] V[i] in one sweep
    template<RealScalar FpType>
    static void lincomb(size_t size, BufType<FpType> y,
                        const std::vector<const DeviceVec<FpType> *> &V,
                        const std::vector<std::complex<double>> &c) {
        std::vector<const void *> p(V.size());
        for (size_t i = 0; i < V.size(); i++) p[i] = V[i]->ptr;
        RP_CALL(rp_lincomb(internal::rp_prec<FpType>(), y.ptr, p.data(), (int)V.size(),
                           (const double *)c.data(), size));
    }
};

template<RealScalar FpType>
inline rp_terms_t DevSumOps<FpType>::device_terms(int len) const {
    if (handle && handle_len == len) return handle;
    free_handle();
    GPUImpl::ensure_context();
    std::vector<u64> bits;
    std::vector<double> coeff;
    this->export_terms(bits, coeff);
    RP_CALL(rp_terms_upload_ex(&handle, internal::rp_prec<FpType>(), len, bits.data(),
                               coeff.data(), bits.size(), GPUImpl::upload_flags()));
    handle_len = len;
    return handle;
}
