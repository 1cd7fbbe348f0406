// rp_vec.hip -- vector kernels of libreapers_hip.so: the BLAS-1 operations
// of the reference's GPUImpl (inc/gpuimpl.h:806-1062, src/krnl.cu:25-58),
// reductions accumulated in FP64 for both element precisions, and the fused
// Gram-Schmidt / linear-combination sweeps the Krylov driver uses in place of
// one cuBLAS call per vector (inc/algo.h:86-101, inc/algo.h:342-345).
//
// All kernels are memory-bound streams with grid-stride loops.  Reductions
// keep FP64 partial sums per thread, reduce them within a wavefront with
// cross-lane shuffles, across the workgroup in LDS, and across workgroups in
// a second small kernel, so every reduction is deterministic for a given n.
#include "rp_internal.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace rp {

template <typename R> struct Cx;
template <> struct Cx<float> { using T = float2; };
template <> struct Cx<double> { using T = double2; };

static inline int grid_for(uint64_t units, int threads, int cap)
{
    uint64_t g = (units + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > (uint64_t)cap) g = cap;
    return (int)g;
}

constexpr int kEwThreads = 256;
constexpr int kEwMaxBlocks = 65536;
constexpr int kMaxVecs = RP_MAX_SWEEP_VECS;   // vectors per fused sweep (API limit)
constexpr int kDotChunk = 8;      // vectors whose dots one launch accumulates
constexpr int kUpdChunk = 16;     // vectors one update / lincomb launch reads
constexpr int kWaves = kRedThreads / 64;

struct VecPtrs16 { const void *p[kUpdChunk]; };
struct Coefs16 { double c[2 * kUpdChunk]; };

// ---------------------------------------------------------------------------
// elementwise kernels
// ---------------------------------------------------------------------------
#define GRID_STRIDE(i, m)                                                        \
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < (m); \
         i += (uint64_t)gridDim.x * blockDim.x)

template <typename R>
__global__ void k_add(R *__restrict__ a, const R *__restrict__ b, uint64_t m)
{
    GRID_STRIDE(i, m) a[i] += b[i];
}

template <typename R>
__global__ void k_add3(R *__restrict__ r, const R *__restrict__ a, const R *__restrict__ b,
                       uint64_t m)
{
    GRID_STRIDE(i, m) r[i] = a[i] + b[i];
}

template <typename R>
__global__ void k_axpy(typename Cx<R>::T *__restrict__ y, R sr, R si,
                       const typename Cx<R>::T *__restrict__ x, uint64_t n)
{
    GRID_STRIDE(i, n) {
        auto v = x[i];
        auto o = y[i];
        o.x += sr * v.x - si * v.y;
        o.y += sr * v.y + si * v.x;
        y[i] = o;
    }
}

template <typename R>
__global__ void k_scale(typename Cx<R>::T *__restrict__ y, R sr, R si, uint64_t n)
{
    GRID_STRIDE(i, n) {
        auto v = y[i];
        typename Cx<R>::T o;
        o.x = sr * v.x - si * v.y;
        o.y = sr * v.y + si * v.x;
        y[i] = o;
    }
}

__global__ void k_f2d(double *__restrict__ d, const float *__restrict__ s, uint64_t m)
{
    GRID_STRIDE(i, m) d[i] = (double)s[i];
}

// double -> float, round to nearest or toward zero.  Toward zero is derived
// from the round-to-nearest result, which either is the truncation already or
// lies one ulp further from zero.
__global__ void k_d2f(float *__restrict__ d, const double *__restrict__ s, uint64_t m, int rz)
{
    GRID_STRIDE(i, m) {
        double v = s[i];
        float f = (float)v;
        if (rz && fabs((double)f) > fabs(v)) f = nextafterf(f, 0.0f);
        d[i] = f;
    }
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) + Box-Muller
// ---------------------------------------------------------------------------
__device__ inline void philox_round(uint32_t (&c)[4], uint32_t k0, uint32_t k1)
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    uint64_t p0 = (uint64_t)M0 * c[0];
    uint64_t p1 = (uint64_t)M1 * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c[0] = hi1 ^ c[1] ^ k0;
    c[1] = lo1;
    c[2] = hi0 ^ c[3] ^ k1;
    c[3] = lo0;
}

__device__ inline void philox4x32_10(uint32_t (&c)[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        philox_round(c, k0, k1);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

template <typename R>
__global__ void k_fill_normal(typename Cx<R>::T *__restrict__ v, uint64_t n, uint32_t k0,
                              uint32_t k1)
{
    GRID_STRIDE(i, n) {
        uint32_t c[4] = {(uint32_t)i, (uint32_t)(i >> 32), 0x52454150u, 0x45525321u};
        philox4x32_10(c, k0, k1);
        // two uniforms with 53 random bits each, u1 in (0, 1]
        uint64_t a = ((uint64_t)c[0] << 21) ^ (uint64_t)(c[1] >> 11);
        uint64_t b = ((uint64_t)c[2] << 21) ^ (uint64_t)(c[3] >> 11);
        double u1 = ((double)(a & ((1ull << 53) - 1)) + 1.0) * (1.0 / 9007199254740992.0);
        double u2 = (double)(b & ((1ull << 53) - 1)) * (1.0 / 9007199254740992.0);
        double r = sqrt(-2.0 * log(u1));
        double s, co;
        sincospi(2.0 * u2, &s, &co);
        typename Cx<R>::T o;
        o.x = (R)(r * co);
        o.y = (R)(r * s);
        v[i] = o;
    }
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Reduce NV per-thread doubles over the workgroup; the NV sums end up in
// sh[kWaves*NV + q].  sh must hold (kWaves + 1) * NV doubles.
template <int NV>
__device__ inline void block_reduce(double (&v)[NV], double *sh)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; q++) v[q] = wave_sum(v[q]);
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) sh[wid * NV + q] = v[q];
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0;
#pragma unroll
        for (int w = 0; w < kWaves; w++) s += sh[w * NV + threadIdx.x];
        sh[kWaves * NV + threadIdx.x] = s;
    }
    __syncthreads();
}

// sum of `nparts` rows: out[j] = sum_p partial[p*width + j], one block per j
__global__ void __launch_bounds__(kRedThreads)
k_reduce_cols(const double *__restrict__ partial, int nparts, int width, double *__restrict__ out)
{
    __shared__ double sh[(kWaves + 1) * 1];
    const int j = blockIdx.x;
    double v[1] = {0.0};
    for (int p = threadIdx.x; p < nparts; p += blockDim.x) v[0] += partial[(size_t)p * width + j];
    block_reduce<1>(v, sh);
    if (threadIdx.x == 0) out[j] = sh[kWaves];
}

// <V[q]|w> for the KC vectors of one chunk, FP64 partials per block written
// to partial[block*width + 2*(q0+q) + {0,1}]; with wnorm, also ||w||^2 to
// partial[block*width + 2*kMaxVecs].  All KC loads of an element are issued
// before any use.
template <typename R, int KC>
__global__ void __launch_bounds__(kRedThreads)
k_dots(VecPtrs16 V, const typename Cx<R>::T *__restrict__ w, uint64_t n,
       double *__restrict__ partial, int width, int q0, int wnorm)
{
    using T = typename Cx<R>::T;
    __shared__ double sh[(kWaves + 1) * (2 * KC + 1)];
    double acc[2 * KC + 1];
#pragma unroll
    for (int q = 0; q < 2 * KC + 1; q++) acc[q] = 0.0;
    GRID_STRIDE(i, n) {
        const T y = w[i];
        T x[KC];
#pragma unroll
        for (int q = 0; q < KC; q++) x[q] = ((const T *)V.p[q])[i];
        const double yr = y.x, yi = y.y;
#pragma unroll
        for (int q = 0; q < KC; q++) {
            const double xr = x[q].x, xi = x[q].y;
            acc[2 * q] += xr * yr + xi * yi;
            acc[2 * q + 1] += xr * yi - xi * yr;
        }
        acc[2 * KC] += yr * yr + yi * yi;
    }
    block_reduce<2 * KC + 1>(acc, sh);
    if (threadIdx.x < 2 * KC)
        partial[(size_t)blockIdx.x * width + 2 * q0 + threadIdx.x] =
            sh[kWaves * (2 * KC + 1) + threadIdx.x];
    if (wnorm && threadIdx.x == 2 * KC)
        partial[(size_t)blockIdx.x * width + 2 * kMaxVecs] = sh[kWaves * (2 * KC + 1) + 2 * KC];
}

// w -= sum_q h[q0+q] V[q] for the KC vectors of one chunk (FP64 arithmetic,
// one rounding); the last chunk also writes ||w||^2 partials.
// Fused normalization (last chunk, fuse_min > 0): with all k coefficients h
// and ||w||^2 before the pass in res, beta^2 = ||w||^2 - sum |h_q|^2 is the
// norm of the result when the basis is orthonormal; if beta^2 >= ||w||^2 / 2
// (no cancellation to speak of) and beta >= fuse_min, the result is written
// scaled by 1/beta and beta is stored to res[2 kMaxVecs + 2] (else 0).
template <typename R, int KC>
__global__ void __launch_bounds__(kRedThreads)
k_gs_update(VecPtrs16 V, const double *__restrict__ h, typename Cx<R>::T *__restrict__ w,
            uint64_t n, double *__restrict__ partial, const double *__restrict__ res, int k,
            double fuse_min)
{
    using T = typename Cx<R>::T;
    __shared__ double hs[2 * KC];
    __shared__ double sh[(kWaves + 1) * 1];
    __shared__ double fscale;
    if (threadIdx.x < 2 * KC) hs[threadIdx.x] = h[threadIdx.x];
    if (threadIdx.x == 0) {
        double scale = 1.0, beta = 0.0;
        if (fuse_min > 0) {
            double s2 = 0;
            for (int q = 0; q < k; q++) s2 += res[2 * q] * res[2 * q] + res[2 * q + 1] * res[2 * q + 1];
            const double wn2 = res[2 * kMaxVecs + 1];
            const double b2 = wn2 - s2;
            if (b2 > 0 && b2 >= 0.5 * wn2 && sqrt(b2) >= fuse_min) {
                beta = sqrt(b2);
                scale = 1.0 / beta;
            }
            if (blockIdx.x == 0) ((double *)res)[2 * kMaxVecs + 2] = beta;
        }
        fscale = scale;
    }
    __syncthreads();
    const double f = fscale;
    double nacc[1] = {0.0};
    GRID_STRIDE(i, n) {
        const T y = w[i];
        T x[KC];
#pragma unroll
        for (int q = 0; q < KC; q++) x[q] = ((const T *)V.p[q])[i];
        double yr = y.x, yi = y.y;
#pragma unroll
        for (int q = 0; q < KC; q++) {
            const double xr = x[q].x, xi = x[q].y, hr = hs[2 * q], hi = hs[2 * q + 1];
            yr -= hr * xr - hi * xi;
            yi -= hr * xi + hi * xr;
        }
        T o;
        o.x = (R)(yr * f);
        o.y = (R)(yi * f);
        w[i] = o;
        const double orr = o.x, oi = o.y;
        nacc[0] += orr * orr + oi * oi;
    }
    if (partial) {
        block_reduce<1>(nacc, sh);
        if (threadIdx.x == 0) partial[blockIdx.x] = sh[kWaves];
    }
}

// y += sum_q c[q] V[q] for the KC vectors of one chunk (FP64 arithmetic)
template <typename R, int KC>
__global__ void __launch_bounds__(kRedThreads)
k_lincomb(VecPtrs16 V, Coefs16 c, typename Cx<R>::T *__restrict__ y, uint64_t n)
{
    using T = typename Cx<R>::T;
    GRID_STRIDE(i, n) {
        T o = y[i];
        T x[KC];
#pragma unroll
        for (int q = 0; q < KC; q++) x[q] = ((const T *)V.p[q])[i];
        double yr = o.x, yi = o.y;
#pragma unroll
        for (int q = 0; q < KC; q++) {
            const double xr = x[q].x, xi = x[q].y, cr = c.c[2 * q], ci = c.c[2 * q + 1];
            yr += cr * xr - ci * xi;
            yi += cr * xi + ci * xr;
        }
        o.x = (R)yr;
        o.y = (R)yi;
        y[i] = o;
    }
}

// norm^2 / sum |x|^2 partials
template <typename R>
__global__ void __launch_bounds__(kRedThreads)
k_nrm2_partial(const typename Cx<R>::T *__restrict__ a, uint64_t n, double *__restrict__ partial)
{
    __shared__ double sh[(kWaves + 1) * 1];
    double acc[1] = {0.0};
    GRID_STRIDE(i, n) {
        auto x = a[i];
        double xr = x.x, xi = x.y;
        acc[0] += xr * xr + xi * xi;
    }
    block_reduce<1>(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[kWaves];
}

template <typename R>
__global__ void __launch_bounds__(kRedThreads)
k_dot_partial(const typename Cx<R>::T *__restrict__ a, const typename Cx<R>::T *__restrict__ b,
              uint64_t n, double *__restrict__ partial)
{
    __shared__ double sh[(kWaves + 1) * 2];
    double acc[2] = {0.0, 0.0};
    GRID_STRIDE(i, n) {
        auto x = a[i];
        auto y = b[i];
        double xr = x.x, xi = x.y, yr = y.x, yi = y.y;
        acc[0] += xr * yr + xi * yi;
        acc[1] += xr * yi - xi * yr;
    }
    block_reduce<2>(acc, sh);
    if (threadIdx.x < 2) partial[2 * blockIdx.x + threadIdx.x] = sh[kWaves * 2 + threadIdx.x];
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
static inline double *ws_partial() { return (double *)ctx().work; }
// [0, B*(2*kMaxVecs+1)): dot partials (2 kMaxVecs columns of coefficients
// and one of ||w||^2 per block); [B*2*kMaxVecs, B*(2*kMaxVecs+1)): norm
// partials of the update, written after the dot partials are reduced; then
// the reduced results (B = kRedBlocks)
static inline double *ws_result()
{
    return (double *)ctx().work + (size_t)kRedBlocks * (2 * kMaxVecs + 1);
}

// The Gram-Schmidt pass's workspace in a file-scope device array, for
// rp_gs_pass_fused_noalloc: the partials and results of the library
// workspace layout above, 1.06 MB.  Each device holds its own copy; the
// host reaches it only through the address hipGetSymbolAddress returns for
// the current device (taken once per device at rp_init).
constexpr size_t kGsScratchDoubles = (size_t)kRedBlocks * (2 * kMaxVecs + 1) + 2 * kMaxVecs + 8;
static __device__ __align__(256) double scratch[kGsScratchDoubles];
static std::vector<double *> g_scratch_ptr;     // per device, null until set up

int vec_scratch_setup(int device)
{
    if (device < 0) return set_error(RP_ERR_INVALID_ARGUMENT, "bad device %d", device);
    if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize(device + 1, nullptr);
    if (g_scratch_ptr[device]) return RP_OK;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_error(RP_ERR_RUNTIME, "hipGetSymbolAddress(scratch) failed: %s",
                         hipGetErrorString(e));
    g_scratch_ptr[device] = (double *)p;
    return RP_OK;
}

static double *vec_scratch(int device)
{
    if (device < 0 || (size_t)device >= g_scratch_ptr.size()) return nullptr;
    return g_scratch_ptr[device];
}

static inline int red_blocks(uint64_t n)
{
    return grid_for(n, kRedThreads, kRedBlocks);
}

// copy `cnt` doubles from a device result area (default: the workspace's)
// to the host (synchronizes)
static int fetch_results(double *dst, int cnt, const double *src = nullptr)
{
    Context &c = ctx();
    RP_CHECK_HIP(hipMemcpyAsync(c.pinned, src ? src : ws_result(), sizeof(double) * cnt,
                                hipMemcpyDeviceToHost, c.stream));
    RP_CHECK_HIP(hipStreamSynchronize(c.stream));
    prof_collect();
    memcpy(dst, c.pinned, sizeof(double) * cnt);
    return RP_OK;
}

template <typename R>
static int dot_impl(const void *a, const void *b, uint64_t n, double *out2)
{
    using T = typename Cx<R>::T;
    int nb = red_blocks(n);
    RP_LAUNCH(RP_KERNEL_OTHER, k_dot_partial<R>, dim3(nb), dim3(kRedThreads), 0,
              (const T *)a, (const T *)b, n, ws_partial());
    RP_LAUNCH(RP_KERNEL_OTHER, k_reduce_cols, dim3(2), dim3(kRedThreads), 0,
              (const double *)ws_partial(), nb, 2, ws_result());
    return fetch_results(out2, 2);
}

template <typename R>
static int nrm2_launch(const void *a, uint64_t n, double *dst_dev, double *part = nullptr)
{
    using T = typename Cx<R>::T;
    int nb = red_blocks(n);
    if (!part) part = ws_partial();
    RP_LAUNCH(RP_KERNEL_OTHER, k_nrm2_partial<R>, dim3(nb), dim3(kRedThreads), 0,
              (const T *)a, n, part);
    RP_LAUNCH(RP_KERNEL_OTHER, k_reduce_cols, dim3(1), dim3(kRedThreads), 0,
              (const double *)part, nb, 1, dst_dev);
    return RP_OK;
}

// launch kernel K<R, KC> for a runtime chunk width kc in [1, 16]
#define RP_KC_SWITCH(kc, LAUNCH)                                                  \
    switch (kc) {                                                                 \
    case 1: LAUNCH(1); break;   case 2: LAUNCH(2); break;                         \
    case 3: LAUNCH(3); break;   case 4: LAUNCH(4); break;                         \
    case 5: LAUNCH(5); break;   case 6: LAUNCH(6); break;                         \
    case 7: LAUNCH(7); break;   case 8: LAUNCH(8); break;                         \
    case 9: LAUNCH(9); break;   case 10: LAUNCH(10); break;                       \
    case 11: LAUNCH(11); break; case 12: LAUNCH(12); break;                       \
    case 13: LAUNCH(13); break; case 14: LAUNCH(14); break;                       \
    case 15: LAUNCH(15); break; default: LAUNCH(16); break;                       \
    }

// One Gram-Schmidt pass; res layout: [0, 2k) h, [2 kMaxVecs] ||w_out||^2,
// [2 kMaxVecs + 1] ||w_in||^2, [2 kMaxVecs + 2] fused beta (0: not fused)
// pre / pre_rows: the dot partials of w already formed by the apply's
// epilogue (rows of width doubles, ||w||^2 in column 2 kMaxVecs); the pass
// then starts from their reduction
template <typename R>
static int gs_pass_impl(const void *const *V, int k, void *w, uint64_t n, double *h,
                        double *norm, double fuse_min, int *normalized, double *work = nullptr,
                        const double *pre = nullptr, uint64_t pre_rows = 0)
{
    using T = typename Cx<R>::T;
    const int nb = red_blocks(n);
    const int width = 2 * kMaxVecs + 1;
    // the workspace: the library's, or the caller-chosen one of the same layout
    double *const part = work ? work : ws_partial();
    double *res = part + (size_t)kRedBlocks * (2 * kMaxVecs + 1);
    double *npart = part + (size_t)kRedBlocks * (2 * kMaxVecs);   // norm partials
    const bool fuse = fuse_min > 0 && k > 0;
    if (pre) {
        if (k > 0)
            RP_LAUNCH(RP_KERNEL_GS, k_reduce_cols, dim3(2 * k), dim3(kRedThreads), 0, pre,
                      (int)pre_rows, width, res);
        RP_LAUNCH(RP_KERNEL_GS, k_reduce_cols, dim3(1), dim3(kRedThreads), 0,
                  pre + 2 * kMaxVecs, (int)pre_rows, width,
                  res + 2 * kMaxVecs + (k > 0 ? 1 : 0));
    } else if (k > 0) {
        for (int q0 = 0; q0 < k; q0 += kDotChunk) {
            const int kc = k - q0 < kDotChunk ? k - q0 : kDotChunk;
            VecPtrs16 P{};
            for (int q = 0; q < kc; q++) P.p[q] = V[q0 + q];
            const int wn = fuse && q0 == 0;
#define RP_DOTS(KC)                                                               \
    RP_LAUNCH(RP_KERNEL_GS, (k_dots<R, KC>), dim3(nb), dim3(kRedThreads), 0, P, (const T *)w, n, \
              part, width, q0, wn)
            RP_KC_SWITCH(kc, RP_DOTS)
#undef RP_DOTS
        }
        RP_LAUNCH(RP_KERNEL_GS, k_reduce_cols, dim3(2 * k), dim3(kRedThreads), 0,
                  (const double *)part, nb, width, res);
        if (fuse)
            RP_LAUNCH(RP_KERNEL_GS, k_reduce_cols, dim3(1), dim3(kRedThreads), 0,
                      (const double *)part + 2 * kMaxVecs, nb, width,
                      res + 2 * kMaxVecs + 1);
    }
    if (k == 0) {
        if (!pre) {
            int st = nrm2_launch<R>(w, n, res + 2 * kMaxVecs, part);
            if (st) return st;
        }
    } else {
        for (int q0 = 0; q0 < k; q0 += kUpdChunk) {
            const int kc = k - q0 < kUpdChunk ? k - q0 : kUpdChunk;
            const bool last = q0 + kc == k;
            VecPtrs16 P{};
            for (int q = 0; q < kc; q++) P.p[q] = V[q0 + q];
#define RP_UPD(KC)                                                                \
    RP_LAUNCH(RP_KERNEL_GS, (k_gs_update<R, KC>), dim3(nb), dim3(kRedThreads), 0, P,          \
              (const double *)(res + 2 * q0), (T *)w, n, last ? npart : (double *)nullptr,   \
              (const double *)res, k, (last && fuse) ? fuse_min : 0.0)
            RP_KC_SWITCH(kc, RP_UPD)
#undef RP_UPD
        }
        RP_LAUNCH(RP_KERNEL_GS, k_reduce_cols, dim3(1), dim3(kRedThreads), 0,
                  (const double *)npart, nb, 1, res + 2 * kMaxVecs);
    }
    double tmp[2 * kMaxVecs + 3];
    int st = fetch_results(tmp, 2 * kMaxVecs + 3, res);
    if (st) return st;
    for (int q = 0; q < 2 * k; q++) h[q] = tmp[q];
    const double beta = fuse ? tmp[2 * kMaxVecs + 2] : 0.0;
    if (normalized) *normalized = beta > 0 ? 1 : 0;
    *norm = beta > 0 ? beta : sqrt(tmp[2 * kMaxVecs]);
    return RP_OK;
}

template <typename R>
static int lincomb_impl(void *y, const void *const *V, int k, const double *c, uint64_t n)
{
    using T = typename Cx<R>::T;
    const int nb = red_blocks(n);
    for (int q0 = 0; q0 < k; q0 += kUpdChunk) {
        const int kc = k - q0 < kUpdChunk ? k - q0 : kUpdChunk;
        VecPtrs16 P{};
        Coefs16 C{};
        for (int q = 0; q < kc; q++) {
            P.p[q] = V[q0 + q];
            C.c[2 * q] = c[2 * (q0 + q)];
            C.c[2 * q + 1] = c[2 * (q0 + q) + 1];
        }
#define RP_LC(KC)                                                                 \
    RP_LAUNCH(RP_KERNEL_OTHER, (k_lincomb<R, KC>), dim3(nb), dim3(kRedThreads), 0, P, C, (T *)y, n)
        RP_KC_SWITCH(kc, RP_LC)
#undef RP_LC
    }
    return RP_OK;
}

}  // namespace rp

using namespace rp;

#define RP_PREC_DISPATCH(prec, FN, ...)                                           \
    ((prec) == RP_C64 ? FN<float>(__VA_ARGS__) : FN<double>(__VA_ARGS__))

extern "C" {

int rp_zero_vec(int prec, void *v, uint64_t n)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(valid_prec(prec), "bad precision");
    if (n == 0) return RP_OK;
    RP_REQUIRE(v, "null vector");
    RP_CHECK_HIP(hipMemsetAsync(v, 0, n * elem_bytes(prec), ctx().stream));
    return RP_OK;
}

int rp_copy_vec(int dst_prec, void *dst, int src_prec, const void *src, uint64_t n, int rounding)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(valid_prec(dst_prec) && valid_prec(src_prec), "bad precision");
    if (n == 0) return RP_OK;
    RP_REQUIRE(dst && src, "null vector");
    uint64_t m = 2 * n;
    int g = grid_for(m, kEwThreads, kEwMaxBlocks);
    if (dst_prec == src_prec) {
        RP_CHECK_HIP(hipMemcpyAsync(dst, src, n * elem_bytes(dst_prec), hipMemcpyDeviceToDevice,
                                    ctx().stream));
    } else if (dst_prec == RP_C128) {
        RP_LAUNCH(RP_KERNEL_OTHER, k_f2d, dim3(g), dim3(kEwThreads), 0, (double *)dst,
                  (const float *)src, m);
    } else {
        RP_REQUIRE(rounding == RP_ROUND_NEAREST || rounding == RP_ROUND_TOWARD_ZERO,
                   "bad rounding mode");
        RP_LAUNCH(RP_KERNEL_OTHER, k_d2f, dim3(g), dim3(kEwThreads), 0, (float *)dst,
                  (const double *)src, m, rounding == RP_ROUND_TOWARD_ZERO ? 1 : 0);
    }
    return RP_OK;
}

int rp_add_vec(int prec, void *v0, const void *v1, uint64_t n)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(valid_prec(prec), "bad precision");
    if (n == 0) return RP_OK;
    RP_REQUIRE(v0 && v1, "null vector");
    uint64_t m = 2 * n;
    int g = grid_for(m, kEwThreads, kEwMaxBlocks);
    if (prec == RP_C64)
        RP_LAUNCH(RP_KERNEL_OTHER, k_add<float>, dim3(g), dim3(kEwThreads), 0, (float *)v0,
                  (const float *)v1, m);
    else
        RP_LAUNCH(RP_KERNEL_OTHER, k_add<double>, dim3(g), dim3(kEwThreads), 0, (double *)v0,
                  (const double *)v1, m);
    return RP_OK;
}

int rp_add_vec3(int prec, void *res, const void *v0, const void *v1, uint64_t n)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(valid_prec(prec), "bad precision");
    if (n == 0) return RP_OK;
    RP_REQUIRE(res && v0 && v1, "null vector");
    uint64_t m = 2 * n;
    int g = grid_for(m, kEwThreads, kEwMaxBlocks);
    if (prec == RP_C64)
        RP_LAUNCH(RP_KERNEL_OTHER, k_add3<float>, dim3(g), dim3(kEwThreads), 0, (float *)res,
                  (const float *)v0, (const float *)v1, m);
    else
        RP_LAUNCH(RP_KERNEL_OTHER, k_add3<double>, dim3(g), dim3(kEwThreads), 0, (double *)res,
                  (const double *)v0, (const double *)v1, m);
    return RP_OK;
}

int rp_axpy(int prec, void *v0, double s_re, double s_im, const void *v1, uint64_t n)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(valid_prec(prec), "bad precision");
    if (n == 0) return RP_OK;
    RP_REQUIRE(v0 && v1, "null vector");
    int g = grid_for(n, kEwThreads, kEwMaxBlocks);
    if (prec == RP_C64)
        RP_LAUNCH(RP_KERNEL_OTHER, k_axpy<float>, dim3(g), dim3(kEwThreads), 0, (float2 *)v0,
                  (float)s_re, (float)s_im, (const float2 *)v1, n);
    else
        RP_LAUNCH(RP_KERNEL_OTHER, k_axpy<double>, dim3(g), dim3(kEwThreads), 0, (double2 *)v0,
                  s_re, s_im, (const double2 *)v1, n);
    return RP_OK;
}

int rp_scale_vec(int prec, void *v, double s_re, double s_im, uint64_t n)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(valid_prec(prec), "bad precision");
    if (n == 0) return RP_OK;
    RP_REQUIRE(v, "null vector");
    int g = grid_for(n, kEwThreads, kEwMaxBlocks);
    if (prec == RP_C64)
        RP_LAUNCH(RP_KERNEL_OTHER, k_scale<float>, dim3(g), dim3(kEwThreads), 0, (float2 *)v,
                  (float)s_re, (float)s_im, n);
    else
        RP_LAUNCH(RP_KERNEL_OTHER, k_scale<double>, dim3(g), dim3(kEwThreads), 0, (double2 *)v,
                  s_re, s_im, n);
    return RP_OK;
}

int rp_vec_prod(int prec, const void *v0, const void *v1, uint64_t n, double *re, double *im)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(valid_prec(prec), "bad precision");
    RP_REQUIRE(re && im, "null output");
    if (n == 0) { *re = 0; *im = 0; return RP_OK; }
    RP_REQUIRE(v0 && v1, "null vector");
    double out[2];
    st = prec == RP_C64 ? dot_impl<float>(v0, v1, n, out) : dot_impl<double>(v0, v1, n, out);
    if (st) return st;
    *re = out[0];
    *im = out[1];
    return RP_OK;
}

int rp_vec_norm(int prec, const void *v, uint64_t n, double *nrm)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(valid_prec(prec), "bad precision");
    RP_REQUIRE(nrm, "null output");
    if (n == 0) { *nrm = 0; return RP_OK; }
    RP_REQUIRE(v, "null vector");
    st = prec == RP_C64 ? nrm2_launch<float>(v, n, ws_result())
                        : nrm2_launch<double>(v, n, ws_result());
    if (st) return st;
    double s;
    st = fetch_results(&s, 1);
    if (st) return st;
    *nrm = sqrt(s);
    return RP_OK;
}

int rp_init_random(int prec, void *v, uint64_t n, uint64_t seed)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(valid_prec(prec), "bad precision");
    if (n == 0) return RP_OK;
    RP_REQUIRE(v, "null vector");
    int g = grid_for(n, kEwThreads, kEwMaxBlocks);
    if (prec == RP_C64)
        RP_LAUNCH(RP_KERNEL_OTHER, k_fill_normal<float>, dim3(g), dim3(kEwThreads), 0,
                  (float2 *)v, n, (uint32_t)seed, (uint32_t)(seed >> 32));
    else
        RP_LAUNCH(RP_KERNEL_OTHER, k_fill_normal<double>, dim3(g), dim3(kEwThreads), 0,
                  (double2 *)v, n, (uint32_t)seed, (uint32_t)(seed >> 32));
    double nrm;
    st = rp_vec_norm(prec, v, n, &nrm);
    if (st) return st;
    return rp_scale_vec(prec, v, 1.0 / nrm, 0.0, n);
}

int rp_gs_pass(int prec, const void *const *V, int k, void *w, uint64_t n, double *h,
               double *norm)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(valid_prec(prec), "bad precision");
    RP_REQUIRE(k >= 0 && k <= kMaxVecs, "k must be in [0, 64]");
    RP_REQUIRE(norm && (k == 0 || (h && V)), "null argument");
    RP_REQUIRE(w || n == 0, "null vector");
    for (int q = 0; q < k; q++) RP_REQUIRE(V[q], "null basis vector");
    if (n == 0) {
        for (int q = 0; q < 2 * k; q++) h[q] = 0;
        *norm = 0;
        return RP_OK;
    }
    return prec == RP_C64 ? gs_pass_impl<float>(V, k, w, n, h, norm, 0.0, nullptr)
                          : gs_pass_impl<double>(V, k, w, n, h, norm, 0.0, nullptr);
}

int rp_gs_pass_fused(int prec, const void *const *V, int k, void *w, uint64_t n, double *h,
                     double *norm, double fuse_min, int *normalized)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(valid_prec(prec), "bad precision");
    RP_REQUIRE(k >= 0 && k <= kMaxVecs, "k must be in [0, 64]");
    RP_REQUIRE(norm && normalized && (k == 0 || (h && V)), "null argument");
    RP_REQUIRE(w || n == 0, "null vector");
    for (int q = 0; q < k; q++) RP_REQUIRE(V[q], "null basis vector");
    *normalized = 0;
    if (n == 0) {
        for (int q = 0; q < 2 * k; q++) h[q] = 0;
        *norm = 0;
        return RP_OK;
    }
    return prec == RP_C64 ? gs_pass_impl<float>(V, k, w, n, h, norm, fuse_min, normalized)
                          : gs_pass_impl<double>(V, k, w, n, h, norm, fuse_min, normalized);
}

int rp_gs_pass_fused_noalloc(int prec, const void *const *V, int k, void *w, uint64_t n,
                             double *h, double *norm, double fuse_min, int *normalized)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(valid_prec(prec), "bad precision");
    RP_REQUIRE(k >= 0 && k <= kMaxVecs,
               "k must be in [0, 64]: the fixed scratch holds partials of 64 basis vectors");
    RP_REQUIRE(norm && normalized && (k == 0 || (h && V)), "null argument");
    RP_REQUIRE(w || n == 0, "null vector");
    for (int q = 0; q < k; q++) RP_REQUIRE(V[q], "null basis vector");
    double *work = vec_scratch(ctx().device);
    if (!work)
        return set_error(RP_ERR_RUNTIME, "rp_gs_pass_fused_noalloc: no scratch on device %d",
                         ctx().device);
    *normalized = 0;
    if (n == 0) {
        for (int q = 0; q < 2 * k; q++) h[q] = 0;
        *norm = 0;
        return RP_OK;
    }
    return prec == RP_C64 ? gs_pass_impl<float>(V, k, w, n, h, norm, fuse_min, normalized, work)
                          : gs_pass_impl<double>(V, k, w, n, h, norm, fuse_min, normalized, work);
}

int rp_apply_gs(rp_terms_t hd, void *res, const void *vec, const void *const *V, int k,
                double *h, double *norm, double fuse_min, int *normalized)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(hd, "null term handle");
    RP_REQUIRE(res && vec && res != vec, "res and vec must be different non-null buffers");
    RP_REQUIRE(k >= 0 && k <= kMaxVecs, "k must be in [0, 64]");
    RP_REQUIRE(norm && normalized && (k == 0 || (h && V)), "null argument");
    for (int q = 0; q < k; q++) {
        RP_REQUIRE(V[q], "null basis vector");
        RP_REQUIRE(V[q] != res, "res must not be one of the basis vectors");
    }
    rp_terms_info_t info;
    if ((st = rp_terms_info(hd, &info))) return st;
    const uint64_t n = 1ull << info.len;
    *normalized = 0;
    Context &c = ctx();
    const int width = 2 * kMaxVecs + 1;
    // partial rows for up to 2^16 workgroups (2^28 amplitudes at 2^12 per tile)
    const uint64_t max_rows = 1ull << 16;
    const uint64_t want = (n >> 12) < max_rows ? (n >> 12) : max_rows;
    const size_t need = (size_t)want * width * sizeof(double);
    if (want && c.epi_bytes < need) {
        if (c.epi_work) {
            RP_CHECK_HIP(hipStreamSynchronize(c.stream));
            (void)hipFree(c.epi_work);
        }
        c.epi_work = nullptr;
        c.epi_bytes = 0;
        if (dev_malloc(&c.epi_work, need) != hipSuccess) {
            (void)hipGetLastError();
            c.epi_work = nullptr;
        } else {
            c.epi_bytes = need;
        }
    }
    uint64_t rows = 0;
    if (c.epi_work) {
        st = apply_dots_launch(hd, res, vec, V, k, (double *)c.epi_work, width,
                               c.epi_bytes / (width * sizeof(double)), &rows);
        if (st) return st;
    }
    if (rows == 0) {
        if ((st = rp_apply_ops(hd, res, vec, 0))) return st;
        return rp_gs_pass_fused(info.prec, V, k, res, n, h, norm, fuse_min, normalized);
    }
    return info.prec == RP_C64
               ? gs_pass_impl<float>(V, k, res, n, h, norm, fuse_min, normalized, nullptr,
                                     (const double *)c.epi_work, rows)
               : gs_pass_impl<double>(V, k, res, n, h, norm, fuse_min, normalized, nullptr,
                                      (const double *)c.epi_work, rows);
}

int rp_lincomb(int prec, void *y, const void *const *V, int k, const double *c, uint64_t n)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(valid_prec(prec), "bad precision");
    RP_REQUIRE(k >= 0, "k must not be negative");
    if (n == 0 || k == 0) return RP_OK;
    RP_REQUIRE(y && V && c, "null argument");
    for (int q = 0; q < k; q++) RP_REQUIRE(V[q], "null basis vector");
    return prec == RP_C64 ? lincomb_impl<float>(y, V, k, c, n)
                          : lincomb_impl<double>(y, V, k, c, n);
}

}  // extern "C"
