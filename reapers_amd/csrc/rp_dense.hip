// rp_dense.hip -- dense matrices and exact diagonalization: the dense form of
// a term list, the Hermitian eigensystem, the matrix exponential from it,
// products, traces (the reference's DevSumOps::MatrixType, get_matrix,
// get_eigensystem and _matexp, inc/gpuimpl.h:139-300 and 1179-1263, with the
// kernels krnl_vec_exp / krnl_scale_vecs / krnl_eye, src/krnl.cu:113-154).
//
// This is the small-N path (up to 2^15 amplitudes): a matrix of n x n
// complex elements, column-major with leading dimension n, in caller-owned
// device memory.  The O(n^3) parts -- the eigensolver and the products --
// are rocSOLVER's heevd and rocBLAS's gemm on the library stream; the
// library's own kernels build the matrix from the terms, scale the
// eigenvector columns by exp(c w_j) (one pass, the exponential computed in
// FP64 for both precisions) and reduce traces in FP64.
#include "rp_internal.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

namespace rp {

template <typename R> struct Cx2;
template <> struct Cx2<float> { using T = float2; };
template <> struct Cx2<double> { using T = double2; };

constexpr int kDenseMaxLen = 15;
constexpr int kDenseThreads = 256;

// one thread per column c: H[c ^ x][c] = sum over the terms with flip mask x
// of coeff (-1)^popcount(c & z) -- the Pauli string maps |c> to
// (-1)^popcount(c & z) |c ^ x> up to the coefficient (inc/ops.h:421-440).
// Terms arrive sorted by x, so each element is written once, summed in FP64.
template <typename R>
__global__ void __launch_bounds__(kDenseThreads)
k_dense_build(typename Cx2<R>::T *__restrict__ mat, uint64_t n, const uint64_t *__restrict__ xs,
              const uint64_t *__restrict__ zs, const double *__restrict__ co, uint64_t nterms)
{
    uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (c >= n) return;
    typename Cx2<R>::T *col = mat + c * n;
    uint64_t j = 0;
    while (j < nterms) {
        uint64_t x = xs[j];
        double sr = 0.0, si = 0.0;
        for (; j < nterms && xs[j] == x; j++) {
            double s = (__popcll(c & zs[j]) & 1) ? -1.0 : 1.0;
            sr += s * co[2 * j];
            si += s * co[2 * j + 1];
        }
        typename Cx2<R>::T v;
        v.x = (R)sr;
        v.y = (R)si;
        col[c ^ x] = v;
    }
}

// tmp[:, j] = V[:, j] * exp(c w_j); grid.y = column
template <typename R>
__global__ void __launch_bounds__(kDenseThreads)
k_scale_cols_exp(typename Cx2<R>::T *__restrict__ out, const typename Cx2<R>::T *__restrict__ V,
                 const R *__restrict__ w, uint64_t n, double cr, double ci)
{
    uint64_t j = blockIdx.y;
    double wj = (double)w[j];
    double mag = exp(cr * wj);
    double er = mag * cos(ci * wj), ei = mag * sin(ci * wj);
    R fr = (R)er, fi = (R)ei;
    const typename Cx2<R>::T *src = V + j * n;
    typename Cx2<R>::T *dst = out + j * n;
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < n;
         r += (uint64_t)gridDim.x * blockDim.x) {
        typename Cx2<R>::T v = src[r], o;
        o.x = v.x * fr - v.y * fi;
        o.y = v.x * fi + v.y * fr;
        dst[r] = o;
    }
}

template <typename R>
__global__ void __launch_bounds__(kDenseThreads)
k_eye(typename Cx2<R>::T *__restrict__ mat, uint64_t n)
{
    uint64_t total = n * n;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        typename Cx2<R>::T v;
        v.x = (i % (n + 1) == 0) ? (R)1 : (R)0;
        v.y = (R)0;
        mat[i] = v;
    }
}

// U[r + c n] = (-1)^popcount(r & c) / sqrt(n) * exp(i theta_c): a
// Walsh-Hadamard matrix times random phases (theta_c from a splitmix64 hash
// of c and the seed), unitary for n a power of two
template <typename R>
__global__ void __launch_bounds__(kDenseThreads)
k_mixing_unitary(typename Cx2<R>::T *__restrict__ U, uint64_t n, uint64_t seed)
{
    uint64_t total = n * n;
    double inv = 1.0 / sqrt((double)n);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t r = i % n, c = i / n;
        uint64_t z = (c + 1) * 0x9e3779b97f4a7c15ull + seed;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        double theta = 6.283185307179586 * (double)(z >> 11) * (1.0 / 9007199254740992.0);
        double sg = (__popcll(r & c) & 1) ? -inv : inv;
        typename Cx2<R>::T v;
        v.x = (R)(sg * cos(theta));
        v.y = (R)(sg * sin(theta));
        U[i] = v;
    }
}

// one workgroup: the diagonal summed in FP64 (fixed order per n)
template <typename R>
__global__ void __launch_bounds__(kDenseThreads)
k_trace(const typename Cx2<R>::T *__restrict__ mat, uint64_t n, double *__restrict__ out)
{
    __shared__ double sr[kDenseThreads], si[kDenseThreads];
    double ar = 0.0, ai = 0.0;
    for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) {
        typename Cx2<R>::T v = mat[i * (n + 1)];
        ar += (double)v.x;
        ai += (double)v.y;
    }
    sr[threadIdx.x] = ar;
    si[threadIdx.x] = ai;
    __syncthreads();
    for (int s = kDenseThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sr[threadIdx.x] += sr[threadIdx.x + s];
            si[threadIdx.x] += si[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sr[0];
        out[1] = si[0];
    }
}

// the rocBLAS handle of the library stream (created on first use; the
// stream is re-bound on every call since rp_shutdown/rp_init recreate it).
// rocBLAS and rocSOLVER allocate their workspaces themselves, so kept device
// blocks are given back first when fewer than 4 GiB are free.
static rocblas_handle blas()
{
    ensure_device_headroom(size_t(4) << 30);
    static rocblas_handle h = nullptr;
    if (!h && rocblas_create_handle(&h) != rocblas_status_success) h = nullptr;
    if (h) rocblas_set_stream(h, ctx().stream);
    return h;
}

static int blas_error(const char *what, rocblas_status s)
{
    return set_error(RP_ERR_RUNTIME, "%s failed: %s", what, rocblas_status_to_string(s));
}

static inline int grid_of(uint64_t units)
{
    uint64_t g = (units + kDenseThreads - 1) / kDenseThreads;
    return (int)std::min<uint64_t>(std::max<uint64_t>(g, 1), 65536);
}

#define RP_DENSE_LAUNCH(kernel, grid, block, ...)                                 \
    do {                                                                          \
        hipLaunchKernelGGL(kernel, (grid), (block), 0, ::rp::ctx().stream, __VA_ARGS__); \
        RP_CHECK_LAUNCH();                                                        \
    } while (0)

static bool dense_size_ok(uint64_t n)
{
    return n >= 1 && n <= (1ull << kDenseMaxLen);
}

}  // namespace rp

using namespace rp;

extern "C" {

int rp_dense_build(int prec, int len, const uint64_t *bits, const double *coeff, uint64_t nterms,
                   void *mat)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(valid_prec(prec), "invalid precision");
    RP_REQUIRE(len >= 1 && len <= kDenseMaxLen, "len must be 1..15");
    RP_REQUIRE(mat != nullptr, "null matrix");
    RP_REQUIRE(nterms == 0 || (bits && coeff), "null term arrays");
    uint64_t n = 1ull << len;
    uint64_t mask = n - 1;
    // flip and phase masks; terms sorted by x (stable: equal x keep their order)
    std::vector<uint64_t> x(nterms), z(nterms);
    for (uint64_t j = 0; j < nterms; j++) {
        uint64_t b = bits[j], xm = 0, zm = 0;
        for (int s = 0; s < 32 && b; s++, b >>= 2) {
            xm |= (b & 1ull) << s;
            zm |= ((b >> 1) & 1ull) << s;
        }
        RP_REQUIRE((xm & ~mask) == 0 && (zm & ~mask) == 0, "a term acts beyond len sites");
        x[j] = xm;
        z[j] = zm;
    }
    std::vector<uint64_t> order(nterms);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return x[a] < x[b]; });
    std::vector<uint64_t> hx(nterms), hz(nterms);
    std::vector<double> hc(2 * nterms);
    for (uint64_t k = 0; k < nterms; k++) {
        hx[k] = x[order[k]];
        hz[k] = z[order[k]];
        hc[2 * k] = coeff[2 * order[k]];
        hc[2 * k + 1] = coeff[2 * order[k] + 1];
    }
    size_t ebytes = elem_bytes(prec);
    Context &c = ctx();
    RP_CHECK_HIP(hipMemsetAsync(mat, 0, n * n * ebytes, c.stream));
    if (nterms == 0) return RP_OK;
    uint64_t *dx = nullptr;
    void *dbuf = nullptr;
    size_t tbytes = nterms * (2 * sizeof(uint64_t) + 2 * sizeof(double));
    RP_CHECK_HIP(dev_malloc((void **)&dbuf, tbytes));
    dx = (uint64_t *)dbuf;
    uint64_t *dz = dx + nterms;
    double *dc = (double *)(dz + nterms);
    int rc = RP_OK;
    do {
        hipError_t e = hipMemcpyAsync(dx, hx.data(), nterms * 8, hipMemcpyHostToDevice, c.stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(dz, hz.data(), nterms * 8, hipMemcpyHostToDevice, c.stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(dc, hc.data(), nterms * 16, hipMemcpyHostToDevice, c.stream);
        if (e != hipSuccess) {
            rc = set_error(RP_ERR_RUNTIME, "rp_dense_build: copy failed: %s", hipGetErrorString(e));
            break;
        }
        if (prec == RP_C64)
            hipLaunchKernelGGL(k_dense_build<float>, dim3(grid_of(n)), dim3(kDenseThreads), 0,
                               c.stream, (float2 *)mat, n, dx, dz, dc, nterms);
        else
            hipLaunchKernelGGL(k_dense_build<double>, dim3(grid_of(n)), dim3(kDenseThreads), 0,
                               c.stream, (double2 *)mat, n, dx, dz, dc, nterms);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
        if (e != hipSuccess)
            rc = set_error(RP_ERR_RUNTIME, "rp_dense_build: %s", hipGetErrorString(e));
    } while (0);
    (void)hipFree(dbuf);
    return rc;
}

// heevd (divide and conquer) first.  rocSOLVER's heevd returns non-finite
// eigenvectors for matrices that split into decoupled blocks, such as a
// parity-conserving Hamiltonian written in the computational basis (its
// tridiagonal form splits).  Then the saved matrix is rotated to
// A' = U^H A U with a Walsh-Hadamard matrix times random phases, whose
// tridiagonal form does not split, heevd solves A' and the eigenvectors are
// U V'.  QR iteration (heev) is the last resort.
static int eigh_call(int prec, rocblas_handle h, bool dc, uint64_t n, void *mat, void *w,
                     void *e, rocblas_int *dinfo, rocblas_status *s)
{
    if (prec == RP_C64)
        *s = dc ? rocsolver_cheevd(h, rocblas_evect_original, rocblas_fill_lower, (rocblas_int)n,
                                   (rocblas_float_complex *)mat, (rocblas_int)n, (float *)w,
                                   (float *)e, dinfo)
                : rocsolver_cheev(h, rocblas_evect_original, rocblas_fill_lower, (rocblas_int)n,
                                  (rocblas_float_complex *)mat, (rocblas_int)n, (float *)w,
                                  (float *)e, dinfo);
    else
        *s = dc ? rocsolver_zheevd(h, rocblas_evect_original, rocblas_fill_lower, (rocblas_int)n,
                                   (rocblas_double_complex *)mat, (rocblas_int)n, (double *)w,
                                   (double *)e, dinfo)
                : rocsolver_zheev(h, rocblas_evect_original, rocblas_fill_lower, (rocblas_int)n,
                                  (rocblas_double_complex *)mat, (rocblas_int)n, (double *)w,
                                  (double *)e, dinfo);
    if (*s != rocblas_status_success) return blas_error(dc ? "rocsolver heevd" : "rocsolver heev", *s);
    rocblas_int info = 0;
    RP_CHECK_HIP(hipMemcpyAsync(ctx().pinned, dinfo, sizeof(info), hipMemcpyDeviceToHost,
                                ctx().stream));
    RP_CHECK_HIP(hipStreamSynchronize(ctx().stream));
    std::memcpy(&info, ctx().pinned, sizeof(info));
    if (info != 0)
        return set_error(RP_ERR_RUNTIME, "rocsolver %s: info = %d", dc ? "heevd" : "heev", (int)info);
    return RP_OK;
}

int rp_dense_eigh(int prec, uint64_t n, void *mat, void *w)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(valid_prec(prec), "invalid precision");
    RP_REQUIRE(dense_size_ok(n), "n must be 1..32768");
    RP_REQUIRE(mat && w, "null pointer");
    rocblas_handle h = blas();
    if (!h) return set_error(RP_ERR_RUNTIME, "rocblas_create_handle failed");
    Context &c = ctx();
    size_t rbytes = prec == RP_C64 ? sizeof(float) : sizeof(double);
    size_t mbytes = n * n * elem_bytes(prec);
    void *work = nullptr, *saved = nullptr, *U = nullptr, *tmp = nullptr;
    RP_CHECK_HIP(dev_malloc((void **)&work, n * rbytes + sizeof(rocblas_int)));
    if (dev_malloc((void **)&saved, mbytes) != hipSuccess) {
        (void)hipFree(work);
        return set_error(RP_ERR_OUT_OF_MEMORY, "rp_dense_eigh: no memory for the matrix copy");
    }
    rocblas_int *dinfo = (rocblas_int *)((char *)work + n * rbytes);
    int rc = RP_OK;
    rocblas_status s;
    auto finite = [&](bool &ok) {
        double nrm = 0;
        int r = rp_vec_norm(prec, mat, n * n, &nrm);
        ok = std::isfinite(nrm);
        return r;
    };
    auto hip_ok = [&](hipError_t e, const char *what) -> int {
        if (e == hipSuccess) return RP_OK;
        return set_error(RP_ERR_RUNTIME, "rp_dense_eigh: %s failed: %s", what, hipGetErrorString(e));
    };
    auto restore = [&]() -> int {
        return hip_ok(hipMemcpyAsync(mat, saved, mbytes, hipMemcpyDeviceToDevice, c.stream), "copy");
    };
    do {
        if ((rc = hip_ok(hipMemcpyAsync(saved, mat, mbytes, hipMemcpyDeviceToDevice, c.stream),
                         "copy")))
            break;
        rc = eigh_call(prec, h, true, n, mat, w, work, dinfo, &s);
        if (rc) break;
        bool ok = false;
        if ((rc = finite(ok)) || ok) break;
        bool pow2 = (n & (n - 1)) == 0;
        // rotated (n a power of two): tmp = A U, mat = U^H tmp, heevd, then
        // V = U V' into mat
        if (pow2 && (dev_malloc((void **)&U, mbytes) != hipSuccess ||
                     dev_malloc((void **)&tmp, mbytes) != hipSuccess)) {
            rc = set_error(RP_ERR_OUT_OF_MEMORY, "rp_dense_eigh: no memory for the rotation");
            break;
        }
        if (!pow2) {
            if ((rc = restore())) break;
            rc = eigh_call(prec, h, false, n, mat, w, work, dinfo, &s);
            if (!rc && !(rc = finite(ok)) && !ok)
                rc = set_error(RP_ERR_RUNTIME, "rp_dense_eigh: eigenvectors are not finite");
            break;
        }
        if (prec == RP_C64)
            hipLaunchKernelGGL(k_mixing_unitary<float>, dim3(grid_of(n * n)), dim3(kDenseThreads), 0,
                               c.stream, (float2 *)U, n, (uint64_t)0x5eedf00dull);
        else
            hipLaunchKernelGGL(k_mixing_unitary<double>, dim3(grid_of(n * n)), dim3(kDenseThreads), 0,
                               c.stream, (double2 *)U, n, (uint64_t)0x5eedf00dull);
        if ((rc = hip_ok(hipGetLastError(), "launch"))) break;
        if ((rc = rp_dense_gemm(prec, n, n, n, 0, 0, 1, 0, saved, n, U, n, 0, 0, tmp, n))) break;
        if ((rc = rp_dense_gemm(prec, n, n, n, 1, 0, 1, 0, U, n, tmp, n, 0, 0, mat, n))) break;
        rc = eigh_call(prec, h, true, n, mat, w, work, dinfo, &s);
        if (rc) break;
        if ((rc = finite(ok))) break;
        if (ok) {
            if ((rc = hip_ok(hipMemcpyAsync(tmp, mat, mbytes, hipMemcpyDeviceToDevice, c.stream),
                             "copy")))
                break;
            rc = rp_dense_gemm(prec, n, n, n, 0, 0, 1, 0, U, n, tmp, n, 0, 0, mat, n);
            break;
        }
        // last resort: QR iteration on the original matrix
        if ((rc = restore())) break;
        rc = eigh_call(prec, h, false, n, mat, w, work, dinfo, &s);
        if (rc) break;
        if ((rc = finite(ok))) break;
        if (!ok) rc = set_error(RP_ERR_RUNTIME, "rp_dense_eigh: eigenvectors are not finite");
    } while (0);
    (void)hipStreamSynchronize(c.stream);
    if (tmp) (void)hipFree(tmp);
    if (U) (void)hipFree(U);
    (void)hipFree(saved);
    (void)hipFree(work);
    return rc;
}

int rp_dense_expm_eig(int prec, uint64_t n, const void *V, const void *w, double c_re,
                      double c_im, void *res, void *work)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(valid_prec(prec), "invalid precision");
    RP_REQUIRE(dense_size_ok(n), "n must be 1..32768");
    RP_REQUIRE(V && w && res && work, "null pointer");
    RP_REQUIRE(res != V && res != work && work != V, "res, work and V must be distinct");
    Context &c = ctx();
    dim3 grid((unsigned)std::min<uint64_t>((n + kDenseThreads - 1) / kDenseThreads, 64),
              (unsigned)n);
    if (prec == RP_C64)
        RP_DENSE_LAUNCH(k_scale_cols_exp<float>, grid, dim3(kDenseThreads),
                  (float2 *)work, (const float2 *)V, (const float *)w, n, c_re, c_im);
    else
        RP_DENSE_LAUNCH(k_scale_cols_exp<double>, grid, dim3(kDenseThreads),
                  (double2 *)work, (const double2 *)V, (const double *)w, n, c_re, c_im);
    (void)c;
    // res = work * V^H
    return rp_dense_gemm(prec, n, n, n, 0, 1, 1.0, 0.0, work, n, V, n, 0.0, 0.0, res, n);
}

int rp_dense_gemm(int prec, uint64_t m, uint64_t n, uint64_t k, int op_a, int op_b, double a_re,
                  double a_im, const void *A, uint64_t lda, const void *B, uint64_t ldb,
                  double b_re, double b_im, void *C, uint64_t ldc)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(valid_prec(prec), "invalid precision");
    RP_REQUIRE(m >= 1 && n >= 1 && k >= 1 && m < (1ull << 31) && n < (1ull << 31) &&
               k < (1ull << 31), "dimensions must be 1..2^31-1");
    RP_REQUIRE((op_a == 0 || op_a == 1) && (op_b == 0 || op_b == 1), "op must be 0 or 1");
    RP_REQUIRE(A && B && C, "null pointer");
    RP_REQUIRE(lda >= (op_a ? k : m) && ldb >= (op_b ? n : k) && ldc >= m, "leading dimension too small");
    rocblas_handle h = blas();
    if (!h) return set_error(RP_ERR_RUNTIME, "rocblas_create_handle failed");
    rocblas_operation ta = op_a ? rocblas_operation_conjugate_transpose : rocblas_operation_none;
    rocblas_operation tb = op_b ? rocblas_operation_conjugate_transpose : rocblas_operation_none;
    rocblas_status s;
    if (prec == RP_C64) {
        rocblas_float_complex al{(float)a_re, (float)a_im}, be{(float)b_re, (float)b_im};
        s = rocblas_cgemm(h, ta, tb, (rocblas_int)m, (rocblas_int)n, (rocblas_int)k, &al,
                          (const rocblas_float_complex *)A, (rocblas_int)lda,
                          (const rocblas_float_complex *)B, (rocblas_int)ldb, &be,
                          (rocblas_float_complex *)C, (rocblas_int)ldc);
    } else {
        rocblas_double_complex al{a_re, a_im}, be{b_re, b_im};
        s = rocblas_zgemm(h, ta, tb, (rocblas_int)m, (rocblas_int)n, (rocblas_int)k, &al,
                          (const rocblas_double_complex *)A, (rocblas_int)lda,
                          (const rocblas_double_complex *)B, (rocblas_int)ldb, &be,
                          (rocblas_double_complex *)C, (rocblas_int)ldc);
    }
    if (s != rocblas_status_success) return blas_error("rocblas gemm", s);
    return RP_OK;
}

int rp_dense_trace(int prec, uint64_t n, const void *mat, double *re, double *im)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(valid_prec(prec), "invalid precision");
    RP_REQUIRE(dense_size_ok(n), "n must be 1..32768");
    RP_REQUIRE(mat && re && im, "null pointer");
    Context &c = ctx();
    double *out = (double *)c.work;
    if (prec == RP_C64)
        RP_DENSE_LAUNCH(k_trace<float>, dim3(1), dim3(kDenseThreads),
                  (const float2 *)mat, n, out);
    else
        RP_DENSE_LAUNCH(k_trace<double>, dim3(1), dim3(kDenseThreads),
                  (const double2 *)mat, n, out);
    RP_CHECK_HIP(hipMemcpyAsync(c.pinned, out, 2 * sizeof(double), hipMemcpyDeviceToHost,
                                c.stream));
    RP_CHECK_HIP(hipStreamSynchronize(c.stream));
    *re = c.pinned[0];
    *im = c.pinned[1];
    return RP_OK;
}

int rp_dense_eye(int prec, uint64_t n, void *mat)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(valid_prec(prec), "invalid precision");
    RP_REQUIRE(dense_size_ok(n), "n must be 1..32768");
    RP_REQUIRE(mat != nullptr, "null matrix");
    if (prec == RP_C64)
        RP_DENSE_LAUNCH(k_eye<float>, dim3(grid_of(n * n)), dim3(kDenseThreads),
                  (float2 *)mat, n);
    else
        RP_DENSE_LAUNCH(k_eye<double>, dim3(grid_of(n * n)), dim3(kDenseThreads),
                  (double2 *)mat, n);
    return RP_OK;
}

}  // extern "C"
