// rp_batch.hip -- many right-hand sides of uneven size in one launch per
// operation: the batched Pauli-sum apply and the batched Arnoldi step that a
// restarted Krylov evolution of many disorder realizations (each with a few
// initial states) is built from.
//
// Block-state layout.  Problem p (one Hamiltonian on len[p] qubits) carries
// NS states (1, 2 or 4); a batch vector is one allocation in which problem
// p's block starts at basis offset off[p] and holds amplitude m of state s at
// element (off[p] + m) * NS + s.  One gather of the input at m ^ x then
// serves all NS states of the problem: the term list is walked once for all
// of them.
//
// Apply.  The work items are (problem, output tile of 2^b amplitudes) pairs
// of all problems; a persistent grid walks them, each XCD a contiguous range
// so that the input tiles a workgroup reads overlap in that XCD's L2.  An
// item applies its problem's groups exactly as the single-state tiled kernel
// does (input tile staged in LDS by asynchronous global->LDS loads, double
// buffered, every sub of the group applied from LDS with the sign split into
// lane, row and workgroup parts), with NS accumulators per amplitude.
//
// Arnoldi step.  V_{j+1} = A V_j, one classical Gram-Schmidt pass against
// V_0..V_j with FP64 dot products and a second one for the problems with a
// state whose first-pass norm asks for it (the reference's selective
// criterion, decided on the device: the step stays a fixed sequence of
// kernels that a HIP graph can capture, the second pass's kernels return at
// once for the other problems), H(0..j, j) and
// H(j+1, j) = ||V_{j+1}|| written to a device matrix per right-hand side,
// and V_{j+1} scaled by 1/||V_{j+1}|| on the device.  Reductions are two
// stage (per chunk of 2^cb amplitudes, then per problem in chunk order):
// deterministic, no atomics.  Every kernel skips problems whose active flag
// is zero.
#include "rp_internal.h"
#include "rp_tables.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

namespace rp {

struct ProbDesc {
    uint64_t off;                 // basis offset of the problem's block
    uint32_t len, tile0, ntiles;  // first work item, items (tiles) of the problem
    uint32_t g0, g1;              // groups [g0, g1) in the batch's group table
    uint32_t chunk0, nchunks;     // reduction chunks of the problem
    uint32_t pad;
};

}  // namespace rp

struct rp_batch {
    int prec = RP_C64;
    int nprob = 0, ns = 1, kmax = 0;
    int tile_bits = 0, threads = 0, chunk_bits = 0;
    int group_acc = 0;                 // FP32 sums per pair of tile groups (complex<float>)
    uint64_t total_tiles = 0, total_chunks = 0, nbasis = 0;
    uint64_t nterms = 0, ngroups = 0;
    std::vector<int> len;
    std::vector<uint64_t> off;        // nprob + 1 basis offsets
    std::vector<rp::ProbDesc> probs;
    void *d_probs = nullptr, *d_tile_prob = nullptr, *d_chunk_prob = nullptr;
    void *d_groups = nullptr, *d_subs = nullptr, *d_terms = nullptr;
    void *d_active = nullptr;          // int per problem
    void *d_V = nullptr;               // kmax + 2 basis pointers
    void *d_part = nullptr;            // chunk partials: (kmax+1) * ns complex double
    void *d_npart = nullptr;           // chunk norm partials: ns double
    void *d_h = nullptr;               // per rhs: kmax+1 complex double (current pass)
    void *d_inv = nullptr;             // per rhs: 1 / norm
    void *d_need2 = nullptr;           // per rhs: the Arnoldi step's second pass is needed
    std::vector<void *> V;             // host copy of the basis table
    int grid = 0;                      // persistent apply grid
    // state planes (one problem of >= kPlanesMinLen qubits): the apply
    // gathers each state into a plane of its own, applies the single-state
    // kernel to every plane and scatters the results back
    int planes = 0;
    rp_terms_t plane_terms = nullptr;
    void *d_plane_in = nullptr, *d_plane_out = nullptr;
};

namespace rp {
namespace {

constexpr int kRedT = 256;            // threads of the reduction kernels
constexpr int kKC = 8;                // basis vectors per sweep of the dot kernel

template <typename R, int NS> struct Elt { typename CxA<R>::T v[NS]; };

constexpr int kPlanesMinLen = 20;     // smallest problem the plane apply takes
constexpr int kPlaneT = 256;          // threads of the gather / scatter kernels

// plane s of the NS-state block: planes[s * n + m] = in[m].v[s]; nothing
// when the (only) problem is inactive
template <typename R, int NS>
__global__ void __launch_bounds__(kPlaneT)
k_block_to_planes(const Elt<R, NS> *__restrict__ in, typename CxA<R>::T *__restrict__ planes,
                  uint64_t n, const int *__restrict__ active)
{
    if (!active[0]) return;
    const uint64_t m = (uint64_t)blockIdx.x * kPlaneT + threadIdx.x;
    if (m >= n) return;
    const Elt<R, NS> e = in[m];
#pragma unroll
    for (int q = 0; q < NS; q++) planes[q * n + m] = e.v[q];
}

// out[m].v[s] (+)= planes[s * n + m]
template <typename R, int NS>
__global__ void __launch_bounds__(kPlaneT)
k_planes_to_block(const typename CxA<R>::T *__restrict__ planes, Elt<R, NS> *__restrict__ out,
                  uint64_t n, int accumulate, const int *__restrict__ active)
{
    if (!active[0]) return;
    const uint64_t m = (uint64_t)blockIdx.x * kPlaneT + threadIdx.x;
    if (m >= n) return;
    Elt<R, NS> e;
#pragma unroll
    for (int q = 0; q < NS; q++) e.v[q] = planes[q * n + m];
    if (accumulate) {
        const Elt<R, NS> o = out[m];
#pragma unroll
        for (int q = 0; q < NS; q++) {
            e.v[q].x += o.v[q].x;
            e.v[q].y += o.v[q].y;
        }
    }
    out[m] = e;
}

template <typename R, typename A, int TB, int NT, int NS>
constexpr int batch_waves_per_eu()
{
    constexpr int lds = 2 * (1 << TB) * NS * 2 * (int)sizeof(R);
    constexpr int wg_by_lds = (160 * 1024) / lds;
    constexpr int wg_by_waves = 1024 / NT;
    constexpr int wg = wg_by_lds < wg_by_waves ? wg_by_lds : wg_by_waves;
    constexpr int w = (wg < 1 ? 1 : wg) * (NT / 64) / 4;
    return w < 1 ? 1 : w;
}

// asynchronous global->LDS copy of BYTES bytes (a multiple of 16 * NT)
template <int BYTES, int NT>
__device__ inline void dma_bytes(const char *__restrict__ g, char *l)
{
    constexpr int PER_THREAD = BYTES / (16 * NT);
    static_assert(PER_THREAD >= 1 && BYTES % (16 * NT) == 0, "tile too small for DMA");
    const uint32_t tid = threadIdx.x, wave_base = tid & ~63u;
#pragma unroll
    for (int c = 0; c < PER_THREAD; c++) {
        const uint32_t chunk = c * NT;
        __builtin_amdgcn_global_load_lds(
            (const void *)(g + (size_t)(chunk + tid) * 16),
            (__attribute__((address_space(3))) void *)(l + (size_t)(chunk + wave_base) * 16),
            16, 0, 0);
    }
}

// a sub's coefficient in the partial-sum type P (the float copies when FP32
// partial sums are formed from an FP64-coefficient layout)
template <typename P, typename A>
__device__ __forceinline__ P bsub_re(const SubRec<A> &S)
{
    if constexpr (std::is_same_v<P, float> && !std::is_same_v<A, float>) return S.fr;
    else return (P)S.cr;
}
template <typename P, typename A>
__device__ __forceinline__ P bsub_im(const SubRec<A> &S)
{
    if constexpr (std::is_same_v<P, float> && !std::is_same_v<A, float>) return S.fi;
    else return (P)S.ci;
}

// Apply every sub of group G from the staged tile `buf` to the RP rows x NS
// states of this thread, summing in P.
template <typename R, typename A, int TB, int NT, int NS, typename P = A>
__device__ __forceinline__ void batch_group(const GroupRuns &G, const Elt<R, NS> *buf,
                                            const SubRec<A> *__restrict__ subs,
                                            const TermRec<A> *__restrict__ terms, uint32_t t0,
                                            uint32_t tid, P (&ar)[(1 << TB) / NT][NS],
                                            P (&ai)[(1 << TB) / NT][NS])
{
    using E = Elt<R, NS>;
    constexpr int RP = (1 << TB) / NT;
    constexpr int LB = NT == 1024 ? 10 : NT == 512 ? 9 : NT == 256 ? 8 : 7;
    // real c (both row-parity classes): c * v
    for (uint32_t s = G.b[kRunReal0]; s < G.b[kRunReal1 + 1]; s++) {
        const SubRec<A> S = subs[s];
        const E *row = buf + (tid ^ (S.xl & (NT - 1)));
        const uint32_t xr = S.xl >> LB, zr = (S.z >> LB) & (RP - 1);
        const P c0 = bsub_re<P>(S), c = parity32((t0 | tid) & S.z) ? -c0 : c0;
#pragma unroll
        for (int r = 0; r < RP; r++) {
            const E e = row[(r ^ xr) * NT];
            const P cc = parity32(r & zr) ? -c : c;
#pragma unroll
            for (int q = 0; q < NS; q++) {
                ar[r][q] += cc * (P)e.v[q].x;
                ai[r][q] += cc * (P)e.v[q].y;
            }
        }
    }
    // imaginary c = i ci: -ci vi + i ci vr
    for (uint32_t s = G.b[kRunImag0]; s < G.b[kRunImag1 + 1]; s++) {
        const SubRec<A> S = subs[s];
        const E *row = buf + (tid ^ (S.xl & (NT - 1)));
        const uint32_t xr = S.xl >> LB, zr = (S.z >> LB) & (RP - 1);
        const P c0 = bsub_im<P>(S), c = parity32((t0 | tid) & S.z) ? -c0 : c0;
#pragma unroll
        for (int r = 0; r < RP; r++) {
            const E e = row[(r ^ xr) * NT];
            const P cc = parity32(r & zr) ? -c : c;
#pragma unroll
            for (int q = 0; q < NS; q++) {
                ar[r][q] -= cc * (P)e.v[q].y;
                ai[r][q] += cc * (P)e.v[q].x;
            }
        }
    }
    for (uint32_t s = G.b[kRunCplx]; s < G.b[kRunCplx + 1]; s++) {
        const SubRec<A> S = subs[s];
        const E *row = buf + (tid ^ (S.xl & (NT - 1)));
        const uint32_t xr = S.xl >> LB, zr = (S.z >> LB) & (RP - 1);
        const bool ln = parity32((t0 | tid) & S.z);
        const P r0 = bsub_re<P>(S), i0 = bsub_im<P>(S);
        const P cr = ln ? -r0 : r0, ci = ln ? -i0 : i0;
#pragma unroll
        for (int r = 0; r < RP; r++) {
            const E e = row[(r ^ xr) * NT];
            const bool n = parity32(r & zr);
            const P a = n ? -cr : cr, b = n ? -ci : ci;
#pragma unroll
            for (int q = 0; q < NS; q++) {
                ar[r][q] += a * (P)e.v[q].x - b * (P)e.v[q].y;
                ai[r][q] += a * (P)e.v[q].y + b * (P)e.v[q].x;
            }
        }
    }
    // several terms share x: d(m) = sum_t c_t (-1)^popcount(m & z_t), formed
    // in A, applied in P
    for (uint32_t s = G.b[kRunMulti]; s < G.b[kRunMulti + 1]; s++) {
        const SubRec<A> S = subs[s];
        const E *row = buf + (tid ^ (S.xl & (NT - 1)));
        const uint32_t xr = S.xl >> LB;
#pragma unroll
        for (int r = 0; r < RP; r++) {
            A dr = 0, di = 0;
            for (uint32_t t = S.term_begin; t < S.term_end; t++) {
                const TermRec<A> C = terms[t];
                const bool n = parity32((t0 | tid) & C.z) ^ parity32(r & ((C.z >> LB) & (RP - 1)));
                dr += n ? -C.cr : C.cr;
                di += n ? -C.ci : C.ci;
            }
            const P pr = (P)dr, pi = (P)di;
            const E e = row[(r ^ xr) * NT];
#pragma unroll
            for (int q = 0; q < NS; q++) {
                ar[r][q] += pr * (P)e.v[q].x - pi * (P)e.v[q].y;
                ai[r][q] += pr * (P)e.v[q].y + pi * (P)e.v[q].x;
            }
        }
    }
}

// GA: the terms of each pair of tile groups are summed in FP32 partials and
// those added to the FP64 sums (complex<float> states with FP64 layouts)
template <typename R, typename A, int TB, int NT, int NS, bool GA = false>
__global__ void __launch_bounds__(NT, (batch_waves_per_eu<R, A, TB, NT, NS>()))
k_batch_apply(const Elt<R, NS> *__restrict__ in, Elt<R, NS> *__restrict__ out,
              const ProbDesc *__restrict__ probs, const uint32_t *__restrict__ tile_prob,
              uint32_t total, const GroupRuns *__restrict__ groups,
              const SubRec<A> *__restrict__ subs, const TermRec<A> *__restrict__ terms,
              const int *__restrict__ active, int accumulate)
{
    using E = Elt<R, NS>;
    constexpr int TILE = 1 << TB;
    constexpr int RP = TILE / NT;
    constexpr int BYTES = TILE * (int)sizeof(E);
    __shared__ E tile0[TILE];
    __shared__ E tile1[TILE];
    const uint32_t tid = threadIdx.x;
    // XCD-aware persistent walk: the workgroups of XCD x (ids = x mod 8)
    // take the items of the x-th eighth of the item range in order
    const uint32_t G = gridDim.x, b = blockIdx.x;
    uint32_t first, stride, end;
    if ((G & 7u) == 0) {
        const uint32_t x = b & 7u, per = (total + 7) / 8;
        first = x * per + (b >> 3);
        stride = G >> 3;
        end = min(total, (x + 1) * per);
    } else {
        first = b;
        stride = G;
        end = total;
    }
    for (uint32_t item = first; item < end; item += stride) {
        const uint32_t p = tile_prob[item];
        if (!active[p]) continue;
        const ProbDesc P = probs[p];
        const uint32_t t0 = (item - P.tile0) << TB;
        const E *pin = in + P.off;
        E *pout = out + P.off;
        A ar[RP][NS], ai[RP][NS];
        float pr[RP][NS], pi[RP][NS];
#pragma unroll
        for (int r = 0; r < RP; r++)
#pragma unroll
            for (int q = 0; q < NS; q++) {
                ar[r][q] = 0;
                ai[r][q] = 0;
                pr[r][q] = 0;
                pi[r][q] = 0;
            }
        auto flush = [&]() {
#pragma unroll
            for (int r = 0; r < RP; r++)
#pragma unroll
                for (int q = 0; q < NS; q++) {
                    ar[r][q] += (A)pr[r][q];
                    ai[r][q] += (A)pi[r][q];
                    pr[r][q] = 0;
                    pi[r][q] = 0;
                }
        };
        auto group = [&](const GroupRuns &Gr, const E *tile) {
            if constexpr (GA) batch_group<R, A, TB, NT, NS, float>(Gr, tile, subs, terms, t0, tid, pr, pi);
            else batch_group<R, A, TB, NT, NS>(Gr, tile, subs, terms, t0, tid, ar, ai);
        };
        const int ng = (int)(P.g1 - P.g0);
        const GroupRuns *gr = groups + P.g0;
        // the previous item's reads of both tiles are complete
        __syncthreads();
        if (ng > 0) dma_bytes<BYTES, NT>((const char *)(pin + (t0 ^ (gr[0].xh << TB))), (char *)tile0);
        for (int g = 0; g < ng; g += 2) {
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            if (g + 1 < ng)
                dma_bytes<BYTES, NT>((const char *)(pin + (t0 ^ (gr[g + 1].xh << TB))), (char *)tile1);
            group(gr[g], tile0);
            if (g + 1 >= ng) break;
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            if (g + 2 < ng)
                dma_bytes<BYTES, NT>((const char *)(pin + (t0 ^ (gr[g + 2].xh << TB))), (char *)tile0);
            group(gr[g + 1], tile1);
            if constexpr (GA) flush();
        }
        if constexpr (GA) flush();
#pragma unroll
        for (int r = 0; r < RP; r++) {
            const uint32_t m = t0 + r * NT + tid;
            E o;
            if (accumulate) o = pout[m];
#pragma unroll
            for (int q = 0; q < NS; q++) {
                A xr = ar[r][q], xi = ai[r][q];
                if (accumulate) { xr += (A)o.v[q].x; xi += (A)o.v[q].y; }
                o.v[q].x = (R)xr;
                o.v[q].y = (R)xi;
            }
            pout[m] = o;
        }
    }
}

// ---- reductions over chunks ---------------------------------------------
// A chunk is 2^cb amplitudes of one problem = 2^cb * NS complex elements;
// thread t of a 256-thread block handles elements t, t + 256, ... of the
// chunk, all of state s = t mod NS.

__device__ inline double block_sum_by_state(double v, int ns, double *red)
{
    // returns, in thread s < ns, the sum over threads t = s mod ns
    red[threadIdx.x] = v;
    __syncthreads();
    double a = 0;
    if ((int)threadIdx.x < ns)
        for (int t = threadIdx.x; t < kRedT; t += ns) a += red[t];
    __syncthreads();
    return a;
}

// a Gram-Schmidt pass after the first runs for problem p when one of its
// states asked for it (need: per rhs flags, or null: always)
__device__ inline bool pass_needed(const int *__restrict__ need, uint32_t p, int ns)
{
    if (!need) return true;
    for (int s = 0; s < ns; s++)
        if (need[(size_t)p * ns + s]) return true;
    return false;
}

// part[c][i][s] = sum over the chunk of conj(V_i) w   (i < k, complex double)
template <typename R>
__global__ void __launch_bounds__(kRedT)
k_batch_dots(const void *const *__restrict__ Vt, int k, const typename CxA<R>::T *__restrict__ w,
             const ProbDesc *__restrict__ probs, const uint32_t *__restrict__ chunk_prob,
             const int *__restrict__ active, int ns, int cb, double *__restrict__ part,
             const int *__restrict__ need)
{
    using T = typename CxA<R>::T;
    __shared__ double red[kRedT];
    const uint32_t c = blockIdx.x;
    const uint32_t p = chunk_prob[c];
    if (!active[p] || !pass_needed(need, p, ns)) return;
    const ProbDesc P = probs[p];
    const uint64_t base = (P.off + ((uint64_t)(c - P.chunk0) << cb)) * ns;
    const uint32_t nel = (1u << cb) * ns;
    double *out = part + (size_t)c * (2 * (size_t)ns * 65);
    for (int i0 = 0; i0 < k; i0 += kKC) {
        const int kc = min(kKC, k - i0);
        double sr[kKC], si[kKC];
#pragma unroll
        for (int q = 0; q < kKC; q++) { sr[q] = 0; si[q] = 0; }
        for (uint32_t e = threadIdx.x; e < nel; e += kRedT) {
            const T x = w[base + e];
#pragma unroll
            for (int q = 0; q < kKC; q++) {
                if (q < kc) {
                    const T v = ((const T *)Vt[i0 + q])[base + e];
                    sr[q] += (double)v.x * x.x + (double)v.y * x.y;
                    si[q] += (double)v.x * x.y - (double)v.y * x.x;
                }
            }
        }
        for (int q = 0; q < kc; q++) {
            double a = block_sum_by_state(sr[q], ns, red);
            double b = block_sum_by_state(si[q], ns, red);
            if ((int)threadIdx.x < ns) {
                out[2 * ((size_t)(i0 + q) * ns + threadIdx.x)] = a;
                out[2 * ((size_t)(i0 + q) * ns + threadIdx.x) + 1] = b;
            }
        }
    }
}

// Sum of the values src[c * stride] over chunks c in [c0, c0 + nc) by the
// block's kRedT threads: thread t takes chunks c0 + t, c0 + t + kRedT, ...
// in order, then a fixed pairwise tree over the threads (deterministic for a
// given chunk count).  The result is valid in thread 0.
__device__ inline double block_chunk_sum(const double *__restrict__ src, size_t stride,
                                         uint32_t c0, uint32_t nc, double *red)
{
    double a = 0;
    for (uint32_t c = threadIdx.x; c < nc; c += kRedT) a += src[(size_t)(c0 + c) * stride];
    red[threadIdx.x] = a;
    __syncthreads();
#pragma unroll
    for (int w = kRedT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    return red[0];
}

// h[rhs][i] = sum over the problem's chunks of part; the column of the
// Hessenberg matrix accumulates h over the passes.  Block (p, e) reduces
// element e = i * ns + s of problem p.
__global__ void __launch_bounds__(kRedT)
k_batch_reduce_dots(const double *__restrict__ part, int k, const ProbDesc *__restrict__ probs,
                    const int *__restrict__ active, int ns, double *__restrict__ h, int hstride,
                    double *__restrict__ H, int j, int kmax, int first_pass,
                    const int *__restrict__ need)
{
    __shared__ double red[kRedT];
    const int p = blockIdx.x, e = blockIdx.y;
    if (!active[p] || e >= k * ns || !pass_needed(need, p, ns)) return;
    const ProbDesc P = probs[p];
    const int i = e / ns, s = e % ns;
    const size_t stride = 2 * (size_t)ns * 65;
    const double *q = part + 2 * ((size_t)i * ns + s);
    const double a = block_chunk_sum(q, stride, P.chunk0, P.nchunks, red);
    __syncthreads();
    const double b = block_chunk_sum(q + 1, stride, P.chunk0, P.nchunks, red);
    if (threadIdx.x == 0) {
        const size_t rhs = (size_t)p * ns + s;
        h[rhs * hstride + 2 * i] = a;
        h[rhs * hstride + 2 * i + 1] = b;
        // H(i, j) column-major per rhs, (kmax + 1) rows
        double *Hij = H + rhs * (size_t)(2 * (kmax + 1) * kmax) + 2 * ((size_t)j * (kmax + 1) + i);
        if (first_pass) { Hij[0] = a; Hij[1] = b; }
        else { Hij[0] += a; Hij[1] += b; }
    }
}

// w -= sum_i h_i V_i; npart[c][s] = sum over the chunk of |w|^2
template <typename R>
__global__ void __launch_bounds__(kRedT)
k_batch_update(const void *const *__restrict__ Vt, int k, typename CxA<R>::T *__restrict__ w,
               const ProbDesc *__restrict__ probs, const uint32_t *__restrict__ chunk_prob,
               const int *__restrict__ active, int ns, int cb, const double *__restrict__ h,
               int hstride, double *__restrict__ npart, const int *__restrict__ need)
{
    using T = typename CxA<R>::T;
    __shared__ double red[kRedT];
    __shared__ double hs[2 * 65 * 4];
    const uint32_t c = blockIdx.x;
    const uint32_t p = chunk_prob[c];
    if (!active[p] || !pass_needed(need, p, ns)) return;
    const ProbDesc P = probs[p];
    for (int e = threadIdx.x; e < k * ns; e += kRedT) {
        const int i = e / ns, s = e % ns;
        const size_t rhs = (size_t)p * ns + s;
        hs[2 * (i * ns + s)] = h[rhs * hstride + 2 * i];
        hs[2 * (i * ns + s) + 1] = h[rhs * hstride + 2 * i + 1];
    }
    __syncthreads();
    const uint64_t base = (P.off + ((uint64_t)(c - P.chunk0) << cb)) * ns;
    const uint32_t nel = (1u << cb) * ns;
    const int s = threadIdx.x % ns;
    double nrm = 0;
    for (uint32_t e = threadIdx.x; e < nel; e += kRedT) {
        T x = w[base + e];
        double xr = x.x, xi = x.y;
        for (int i = 0; i < k; i++) {
            const T v = ((const T *)Vt[i])[base + e];
            const double hr = hs[2 * (i * ns + s)], hi = hs[2 * (i * ns + s) + 1];
            xr -= hr * v.x - hi * v.y;
            xi -= hr * v.y + hi * v.x;
        }
        x.x = (R)xr;
        x.y = (R)xi;
        w[base + e] = x;
        nrm += (double)x.x * x.x + (double)x.y * x.y;
    }
    const double a = block_sum_by_state(nrm, ns, red);
    if ((int)threadIdx.x < ns) npart[(size_t)c * ns + threadIdx.x] = a;
}

// norm of the chunk: npart[c][s] = sum |v|^2
template <typename R>
__global__ void __launch_bounds__(kRedT)
k_batch_norm_part(const typename CxA<R>::T *__restrict__ v, const ProbDesc *__restrict__ probs,
                  const uint32_t *__restrict__ chunk_prob, const int *__restrict__ active, int ns,
                  int cb, double *__restrict__ npart)
{
    using T = typename CxA<R>::T;
    __shared__ double red[kRedT];
    const uint32_t c = blockIdx.x;
    const uint32_t p = chunk_prob[c];
    if (!active[p]) return;
    const ProbDesc P = probs[p];
    const uint64_t base = (P.off + ((uint64_t)(c - P.chunk0) << cb)) * ns;
    const uint32_t nel = (1u << cb) * ns;
    double nrm = 0;
    for (uint32_t e = threadIdx.x; e < nel; e += kRedT) {
        const T x = v[base + e];
        nrm += (double)x.x * x.x + (double)x.y * x.y;
    }
    const double a = block_sum_by_state(nrm, ns, red);
    if ((int)threadIdx.x < ns) npart[(size_t)c * ns + threadIdx.x] = a;
}

// per rhs: nrm = sqrt(sum of the problem's chunk partials); inv = 1/nrm
// (0 if nrm == 0); optionally H(j+1, j) = nrm and nrm_out[rhs] = nrm
// im_only: write the norm to Im H(j+1, j) only (the norm after the first
// Gram-Schmidt pass, kept for the breakdown test) and leave inv alone; with
// need, also whether this state needs a second pass: the reference's
// selective criterion (inc/algo.h:79-101) starts from old norm 1 and norm
// eta/2, so after the first pass a second one runs iff 0 < norm < eta^2/2
// ||v|| of every rhs from the chunk partials npart[c][s]: block rhs
__global__ void __launch_bounds__(kRedT)
k_batch_norm_final(const double *__restrict__ npart, const ProbDesc *__restrict__ probs,
                   const int *__restrict__ active, int nprob, int ns,
                   double *__restrict__ inv, double *__restrict__ nrm_out,
                   double *__restrict__ H, int j, int kmax, int im_only,
                   int *__restrict__ need)
{
    __shared__ double red[kRedT];
    const int rhs = blockIdx.x;
    if (rhs >= nprob * ns) return;
    const int p = rhs / ns, s = rhs % ns;
    if (!active[p]) return;
    const ProbDesc P = probs[p];
    const double a = block_chunk_sum(npart + s, (size_t)ns, P.chunk0, P.nchunks, red);
    if (threadIdx.x != 0) return;
    const double n = sqrt(a);
    double *Hij = H ? H + (size_t)rhs * (2 * (kmax + 1) * kmax) + 2 * ((size_t)j * (kmax + 1) + j + 1)
                    : nullptr;
    if (im_only) {
        if (Hij) Hij[1] = n;
        if (need) need[rhs] = n > 0 && n < 0.25;
        return;
    }
    inv[rhs] = n > 0 ? 1.0 / n : 0.0;
    if (nrm_out) nrm_out[rhs] = n;
    if (Hij) Hij[0] = n;
}

// v *= inv[rhs]
template <typename R>
__global__ void __launch_bounds__(kRedT)
k_batch_scale(typename CxA<R>::T *__restrict__ v, const ProbDesc *__restrict__ probs,
              const uint32_t *__restrict__ chunk_prob, const int *__restrict__ active, int ns,
              int cb, const double *__restrict__ inv)
{
    using T = typename CxA<R>::T;
    const uint32_t c = blockIdx.x;
    const uint32_t p = chunk_prob[c];
    if (!active[p]) return;
    const ProbDesc P = probs[p];
    const uint64_t base = (P.off + ((uint64_t)(c - P.chunk0) << cb)) * ns;
    const uint32_t nel = (1u << cb) * ns;
    for (uint32_t e = threadIdx.x; e < nel; e += kRedT) {
        T x = v[base + e];
        const double f = inv[(size_t)p * ns + e % ns];
        x.x = (R)(x.x * f);
        x.y = (R)(x.y * f);
        v[base + e] = x;
    }
}

// y += sum_{i<k} c[rhs][i] V_i   (c: complex double, stride cstride doubles)
template <typename R>
__global__ void __launch_bounds__(kRedT)
k_batch_lincomb(const void *const *__restrict__ Vt, int k, typename CxA<R>::T *__restrict__ y,
                const ProbDesc *__restrict__ probs, const uint32_t *__restrict__ chunk_prob,
                const int *__restrict__ active, int ns, int cb, const double *__restrict__ cf,
                int cstride)
{
    using T = typename CxA<R>::T;
    __shared__ double cs[2 * 65 * 4];
    const uint32_t c = blockIdx.x;
    const uint32_t p = chunk_prob[c];
    if (!active[p]) return;
    const ProbDesc P = probs[p];
    for (int e = threadIdx.x; e < k * ns; e += kRedT) {
        const int i = e / ns, s = e % ns;
        const size_t rhs = (size_t)p * ns + s;
        cs[2 * (i * ns + s)] = cf[rhs * cstride + 2 * i];
        cs[2 * (i * ns + s) + 1] = cf[rhs * cstride + 2 * i + 1];
    }
    __syncthreads();
    const uint64_t base = (P.off + ((uint64_t)(c - P.chunk0) << cb)) * ns;
    const uint32_t nel = (1u << cb) * ns;
    const int s = threadIdx.x % ns;
    for (uint32_t e = threadIdx.x; e < nel; e += kRedT) {
        T x = y[base + e];
        double xr = x.x, xi = x.y;
        for (int i = 0; i < k; i++) {
            const T v = ((const T *)Vt[i])[base + e];
            const double a = cs[2 * (i * ns + s)], b = cs[2 * (i * ns + s) + 1];
            xr += a * v.x - b * v.y;
            xi += a * v.y + b * v.x;
        }
        x.x = (R)xr;
        x.y = (R)xi;
        y[base + e] = x;
    }
}

// ---- host side -------------------------------------------------------------

int batch_alloc(void **p, size_t bytes)
{
    hipError_t e = dev_malloc((void **)p, bytes ? bytes : 16);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return set_error(RP_ERR_OUT_OF_MEMORY, "Failed to allocate %zu bytes on GPU device.",
                         bytes);
    }
    return RP_OK;
}

void release_planes(rp_batch *b)
{
    if (b->plane_terms) (void)rp_terms_free(b->plane_terms);
    for (void *p : {b->d_plane_in, b->d_plane_out})
        if (p) (void)hipFree(p);
    b->plane_terms = nullptr;
    b->d_plane_in = b->d_plane_out = nullptr;
    b->planes = 0;
}

// the plane apply's resources: the single-state term handle (its own
// tiling autotuner when autotune, group accumulation as the batch) and two
// NS-plane scratch vectors
int setup_planes(rp_batch *b, const uint64_t *bits, const double *coeff, uint64_t nterms,
                 int autotune)
{
    const int fl = (autotune ? RP_TERMS_AUTOTUNE : 0) | (b->group_acc ? RP_TERMS_ACC_GROUP : 0);
    int st = rp_terms_upload_ex(&b->plane_terms, b->prec, b->len[0], bits, coeff, nterms, fl);
    // no term left (H = 0): the single-state apply would clear its output
    // with a memset, which a captured restart must not hold; the batched
    // kernel takes it
    rp_terms_info_t ti;
    if (!st && !(st = rp_terms_info(b->plane_terms, &ti)) && ti.nterms == 0) {
        release_planes(b);
        return RP_OK;
    }
    const size_t bytes = ((size_t)b->ns * elem_bytes(b->prec)) << b->len[0];
    if (!st) st = batch_alloc(&b->d_plane_in, bytes);
    if (!st) st = batch_alloc(&b->d_plane_out, bytes);
    if (st) release_planes(b);
    return st;
}

int launch_apply(rp_batch *b, void *res, const void *vec, int accumulate);

// average time of 3 batched applies (after a warm-up) on zeroed scratch
// vectors with the plane apply on or off
int time_batch_apply(rp_batch *b, int planes, float *ms)
{
    Context &c = ctx();
    void *x = nullptr, *y = nullptr;
    const size_t bytes = b->nbasis * b->ns * elem_bytes(b->prec);
    int st = batch_alloc(&x, bytes);
    if (!st) st = batch_alloc(&y, bytes);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!st && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        st = set_error(RP_ERR_RUNTIME, "hipEventCreate failed");
    const int keep = b->planes;
    b->planes = planes;
    if (!st && hipMemsetAsync(x, 0, bytes, c.stream) != hipSuccess)
        st = set_error(RP_ERR_RUNTIME, "hipMemsetAsync failed");
    if (!st) st = launch_apply(b, y, x, 0);
    if (!st && hipEventRecord(e0, c.stream) != hipSuccess)
        st = set_error(RP_ERR_RUNTIME, "hipEventRecord failed");
    for (int r = 0; r < 3 && !st; r++) st = launch_apply(b, y, x, 0);
    if (!st && (hipEventRecord(e1, c.stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(ms, e0, e1) != hipSuccess))
        st = set_error(RP_ERR_RUNTIME, "timing the batched apply failed");
    b->planes = keep;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (x) (void)hipFree(x);
    if (y) (void)hipFree(y);
    return st;
}

void batch_release(rp_batch *b)
{
    if (!b) return;
    for (void *p : {b->d_probs, b->d_tile_prob, b->d_chunk_prob, b->d_groups, b->d_subs,
                    b->d_terms, b->d_active, b->d_V, b->d_part, b->d_npart, b->d_h, b->d_inv,
                    b->d_need2})
        if (p) (void)hipFree(p);
    release_planes(b);
    delete b;
}

// The apply is instantiated for two tile sizes per element type: 32 KiB
// (256 threads) and 64 KiB (512 threads) of LDS per tile, i.e. tile bits
// b = log2(32 KiB / element bytes) and b + 1, element = nrhs complex values.
constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v / 2); }
template <typename R, int NS> constexpr int small_tb() { return ilog2c(32768 / (2 * (int)sizeof(R) * NS)); }

int tile_bits_for(int prec, int ns, int large)
{
    const int esz = (prec == RP_C64 ? 8 : 16) * ns;
    return ilog2c(32768 / esz) + (large ? 1 : 0);
}

template <typename R, typename A, int TB, int NT, int NS>
int launch_apply_t(rp_batch *b, void *res, const void *vec, int accumulate)
{
    using E = Elt<R, NS>;
    if constexpr (sizeof(R) == 4 && sizeof(A) == 8) {
        if (b->group_acc) {
            RP_LAUNCH(RP_KERNEL_APPLY, (k_batch_apply<R, A, TB, NT, NS, true>), dim3(b->grid),
                      dim3(NT), 0, (const E *)vec, (E *)res, (const ProbDesc *)b->d_probs,
                      (const uint32_t *)b->d_tile_prob, (uint32_t)b->total_tiles,
                      (const GroupRuns *)b->d_groups, (const SubRec<A> *)b->d_subs,
                      (const TermRec<A> *)b->d_terms, (const int *)b->d_active, accumulate);
            return RP_OK;
        }
    }
    RP_LAUNCH(RP_KERNEL_APPLY, (k_batch_apply<R, A, TB, NT, NS>), dim3(b->grid), dim3(NT), 0,
              (const E *)vec, (E *)res, (const ProbDesc *)b->d_probs,
              (const uint32_t *)b->d_tile_prob, (uint32_t)b->total_tiles,
              (const GroupRuns *)b->d_groups, (const SubRec<A> *)b->d_subs,
              (const TermRec<A> *)b->d_terms, (const int *)b->d_active, accumulate);
    return RP_OK;
}

template <typename R, typename A, int NS>
int launch_apply_ns(rp_batch *b, void *res, const void *vec, int accumulate)
{
    constexpr int S = small_tb<R, NS>();
    if (b->tile_bits == S && b->threads == 256)
        return launch_apply_t<R, A, S, 256, NS>(b, res, vec, accumulate);
    if (b->tile_bits == S && b->threads == 512)
        return launch_apply_t<R, A, S, 512, NS>(b, res, vec, accumulate);
    if (b->tile_bits == S + 1) return launch_apply_t<R, A, S + 1, 512, NS>(b, res, vec, accumulate);
    return set_error(RP_ERR_RUNTIME, "batched apply: tiling b=%d not instantiated for %d states",
                     b->tile_bits, NS);
}

template <typename R, int NS>
int launch_planes_ns(rp_batch *b, void *res, const void *vec, int accumulate)
{
    using T = typename CxA<R>::T;
    const uint64_t n = 1ull << b->len[0];
    const dim3 grid((unsigned)((n + kPlaneT - 1) / kPlaneT));
    const int *act = (const int *)b->d_active;
    RP_LAUNCH(RP_KERNEL_APPLY, (k_block_to_planes<R, NS>), grid, dim3(kPlaneT), 0,
              (const Elt<R, NS> *)vec, (T *)b->d_plane_in, n, act);
    for (int q = 0; q < NS; q++) {
        int st = rp_apply_ops(b->plane_terms, (T *)b->d_plane_out + q * n,
                              (const T *)b->d_plane_in + q * n, 0);
        if (st) return st;
    }
    RP_LAUNCH(RP_KERNEL_APPLY, (k_planes_to_block<R, NS>), grid, dim3(kPlaneT), 0,
              (const T *)b->d_plane_out, (Elt<R, NS> *)res, n, accumulate, act);
    return RP_OK;
}

template <typename R>
int launch_planes(rp_batch *b, void *res, const void *vec, int accumulate)
{
    switch (b->ns) {
    case 1: return launch_planes_ns<R, 1>(b, res, vec, accumulate);
    case 2: return launch_planes_ns<R, 2>(b, res, vec, accumulate);
    case 4: return launch_planes_ns<R, 4>(b, res, vec, accumulate);
    }
    return set_error(RP_ERR_INVALID_ARGUMENT, "batched apply: nrhs must be 1, 2 or 4");
}

int launch_apply(rp_batch *b, void *res, const void *vec, int accumulate)
{
    if (b->planes)
        return b->prec == RP_C64 ? launch_planes<float>(b, res, vec, accumulate)
                                 : launch_planes<double>(b, res, vec, accumulate);
    if (b->prec == RP_C64) {
        switch (b->ns) {
        case 1: return launch_apply_ns<float, double, 1>(b, res, vec, accumulate);
        case 2: return launch_apply_ns<float, double, 2>(b, res, vec, accumulate);
        case 4: return launch_apply_ns<float, double, 4>(b, res, vec, accumulate);
        }
    } else {
        switch (b->ns) {
        case 1: return launch_apply_ns<double, double, 1>(b, res, vec, accumulate);
        case 2: return launch_apply_ns<double, double, 2>(b, res, vec, accumulate);
        case 4: return launch_apply_ns<double, double, 4>(b, res, vec, accumulate);
        }
    }
    return set_error(RP_ERR_INVALID_ARGUMENT, "batched apply: nrhs must be 1, 2 or 4");
}

// (re)build the group tables of every problem for tile bits tb
int build_batch_tables(rp_batch *b, int tb, int nt, const uint64_t *bits, const double *coeff,
                       const int64_t *toff)
{
    auto build = [&](auto tag) -> int {
        using A = double;
        using R = decltype(tag);
        std::vector<GroupRuns> groups;
        std::vector<SubRec<A>> subs;
        std::vector<TermRec<A>> terms;
        std::vector<uint32_t> tile_prob;
        uint64_t nterms = 0;
        for (int p = 0; p < b->nprob; p++) {
            Tables<A> t;
            int st = build_tables<R, A>(b->len[p], tb, nt, bits + toff[p], coeff + 2 * toff[p],
                                        (uint64_t)(toff[p + 1] - toff[p]), t);
            if (st) return st;
            ProbDesc &P = b->probs[p];
            P.g0 = (uint32_t)groups.size();
            const uint32_t s0 = (uint32_t)subs.size(), t0 = (uint32_t)terms.size();
            for (auto G : t.groups) {
                for (int k = 0; k <= kRuns; k++) G.b[k] += s0;
                groups.push_back(G);
            }
            for (auto S : t.subs) {
                if (S.kind == kSubMulti) { S.term_begin += t0; S.term_end += t0; }
                subs.push_back(S);
            }
            terms.insert(terms.end(), t.terms.begin(), t.terms.end());
            P.g1 = (uint32_t)groups.size();
            P.ntiles = 1u << (b->len[p] - tb);
            P.tile0 = (uint32_t)tile_prob.size();
            for (uint32_t q = 0; q < P.ntiles; q++) tile_prob.push_back((uint32_t)p);
            nterms += t.nterms;
        }
        b->nterms = nterms;
        b->ngroups = groups.size();
        b->total_tiles = tile_prob.size();
        for (void **p : {&b->d_groups, &b->d_subs, &b->d_terms, &b->d_tile_prob, &b->d_probs})
            if (*p) { (void)hipFree(*p); *p = nullptr; }
        int st;
        Context &c = ctx();
        auto up = [&](void **dst, const void *src, size_t bytes) -> int {
            int r = batch_alloc(dst, bytes);
            if (r) return r;
            if (bytes)
                RP_CHECK_HIP(hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, c.stream));
            return RP_OK;
        };
        if ((st = up(&b->d_groups, groups.data(), groups.size() * sizeof(GroupRuns)))) return st;
        if ((st = up(&b->d_subs, subs.data(), subs.size() * sizeof(SubRec<A>)))) return st;
        if ((st = up(&b->d_terms, terms.data(), terms.size() * sizeof(TermRec<A>)))) return st;
        if ((st = up(&b->d_tile_prob, tile_prob.data(), tile_prob.size() * 4))) return st;
        if ((st = up(&b->d_probs, b->probs.data(), b->probs.size() * sizeof(ProbDesc)))) return st;
        RP_CHECK_HIP(hipStreamSynchronize(c.stream));
        return RP_OK;
    };
    b->tile_bits = tb;
    b->threads = nt;
    int st = b->prec == RP_C64 ? build(float{}) : build(double{});
    if (st) return st;
    // persistent grid: a multiple of 8 (one share per XCD), at most 8 waves
    // of workgroups per CU
    uint64_t g = std::min<uint64_t>(b->total_tiles, (uint64_t)std::max(1, ctx().num_cus) * 4);
    if (g >= 8) g &= ~7ull;
    b->grid = (int)std::max<uint64_t>(1, g);
    return RP_OK;
}

}  // namespace
}  // namespace rp

using namespace rp;

extern "C" {

int rp_batch_create(rp_batch_t *out, int prec, int nprob, const int *len, const uint64_t *bits,
                    const double *coeff, const int64_t *term_off, int nrhs, int kmax, int flags)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(out, "null output handle");
    *out = nullptr;
    RP_REQUIRE(valid_prec(prec), "bad precision");
    RP_REQUIRE(nprob >= 1 && nprob <= 65536, "nprob must be in [1, 65536]");
    RP_REQUIRE(len && term_off && (term_off[nprob] == 0 || (bits && coeff)), "null argument");
    RP_REQUIRE(nrhs == 1 || nrhs == 2 || nrhs == 4, "nrhs must be 1, 2 or 4");
    RP_REQUIRE(kmax >= 2 && kmax <= 64, "kmax must be in [2, 64]");
    RP_REQUIRE((flags & ~(RP_BATCH_AUTOTUNE | RP_BATCH_ACC_GROUP | RP_BATCH_WIDE |
                          RP_BATCH_PLANES | 0xff00)) == 0,
               "unknown flags");
    RP_REQUIRE(!(flags & RP_BATCH_PLANES) || (nprob == 1 && len[0] >= kPlanesMinLen),
               "RP_BATCH_PLANES needs one problem of at least 20 qubits");
    const int forced_tb = (flags >> 8) & 0xff;
    RP_REQUIRE(forced_tb == 0 || forced_tb == tile_bits_for(prec, nrhs, 0) ||
                   forced_tb == tile_bits_for(prec, nrhs, 1),
               "forced tile bits must be log2(32 KiB or 64 KiB / (nrhs * element bytes))");
    int minlen = 32;
    for (int p = 0; p < nprob; p++) {
        RP_REQUIRE(len[p] >= 12 && len[p] <= 30, "every len must be in [12, 30]");
        RP_REQUIRE(term_off[p + 1] >= term_off[p], "term offsets must not decrease");
        minlen = std::min(minlen, len[p]);
    }
    rp_batch *b = new rp_batch;
    b->prec = prec;
    b->group_acc = (prec == RP_C64 && (flags & RP_BATCH_ACC_GROUP)) ? 1 : 0;
    b->nprob = nprob;
    b->ns = nrhs;
    b->kmax = kmax;
    b->len.assign(len, len + nprob);
    b->chunk_bits = std::min(12, minlen);
    b->probs.resize(nprob);
    b->off.resize(nprob + 1);
    uint64_t off = 0, chunks = 0;
    std::vector<uint32_t> chunk_prob;
    for (int p = 0; p < nprob; p++) {
        ProbDesc &P = b->probs[p];
        memset(&P, 0, sizeof(P));
        P.off = off;
        P.len = (uint32_t)len[p];
        P.chunk0 = (uint32_t)chunks;
        P.nchunks = 1u << (len[p] - b->chunk_bits);
        for (uint32_t q = 0; q < P.nchunks; q++) chunk_prob.push_back((uint32_t)p);
        chunks += P.nchunks;
        b->off[p] = off;
        // blocks start at multiples of 2^chunk_bits amplitudes (256 B aligned)
        off += 1ull << len[p];
    }
    b->off[nprob] = off;
    b->nbasis = off;
    b->total_chunks = chunks;
    const size_t nr = (size_t)nprob * nrhs;
    if ((st = batch_alloc(&b->d_chunk_prob, chunk_prob.size() * 4)) ||
        (st = batch_alloc(&b->d_active, (size_t)nprob * sizeof(int))) ||
        (st = batch_alloc(&b->d_V, (size_t)(kmax + 2) * sizeof(void *))) ||
        (st = batch_alloc(&b->d_part, (size_t)chunks * nrhs * 65 * 2 * sizeof(double))) ||
        (st = batch_alloc(&b->d_npart, (size_t)chunks * nrhs * sizeof(double))) ||
        (st = batch_alloc(&b->d_h, nr * 65 * 2 * sizeof(double))) ||
        (st = batch_alloc(&b->d_inv, nr * sizeof(double))) ||
        (st = batch_alloc(&b->d_need2, nr * sizeof(int)))) {
        batch_release(b);
        return st;
    }
    Context &c = ctx();
    RP_CHECK_HIP(hipMemcpyAsync(b->d_chunk_prob, chunk_prob.data(), chunk_prob.size() * 4,
                                hipMemcpyHostToDevice, c.stream));
    std::vector<int> act(nprob, 1);
    RP_CHECK_HIP(hipMemcpyAsync(b->d_active, act.data(), act.size() * sizeof(int),
                                hipMemcpyHostToDevice, c.stream));
    RP_CHECK_HIP(hipStreamSynchronize(c.stream));
    b->V.assign(kmax + 2, nullptr);
    // default tiling: 64 KiB tiles if every problem has that many amplitudes;
    // RP_BATCH_TILE_BITS of the 32 KiB tile forces that (256 threads, or 512
    // with RP_BATCH_WIDE)
    int tb = tile_bits_for(prec, nrhs, 1), nt = 512;
    if (tb > minlen || forced_tb == tile_bits_for(prec, nrhs, 0)) {
        tb = tile_bits_for(prec, nrhs, 0);
        nt = (flags & RP_BATCH_WIDE) ? 512 : 256;
    }
    if (tb > minlen) {
        batch_release(b);
        return set_error(RP_ERR_INVALID_ARGUMENT, "rp_batch_create: problems of %d qubits are "
                         "smaller than one %d-amplitude tile", minlen, 1 << tb);
    }
    if ((st = build_batch_tables(b, tb, nt, bits, coeff, term_off))) {
        batch_release(b);
        return st;
    }
    if ((flags & RP_BATCH_AUTOTUNE) && !forced_tb) {
        // time every instantiated tiling on scratch vectors; keep the fastest
        void *x = nullptr, *y = nullptr;
        const size_t bytes = b->nbasis * nrhs * elem_bytes(prec);
        if ((st = batch_alloc(&x, bytes)) || (st = batch_alloc(&y, bytes))) {
            if (x) (void)hipFree(x);
            batch_release(b);
            return st;
        }
        RP_CHECK_HIP(hipMemsetAsync(x, 0, bytes, c.stream));
        hipEvent_t e0, e1;
        RP_CHECK_HIP(hipEventCreate(&e0));
        RP_CHECK_HIP(hipEventCreate(&e1));
        float best = 1e30f;
        struct Tiling { int tb, nt; };
        Tiling pick{tb, nt};
        // 32 KiB tiles with 256 or 512 threads (4 or 2 rows per thread: the
        // latter doubles the waves per CU), 64 KiB tiles with 512 threads
        const Tiling cands[3] = {{tile_bits_for(prec, nrhs, 0), 256},
                                 {tile_bits_for(prec, nrhs, 0), 512},
                                 {tile_bits_for(prec, nrhs, 1), 512}};
        for (const Tiling &T : cands) {
            if (T.tb > minlen) continue;
            if ((st = build_batch_tables(b, T.tb, T.nt, bits, coeff, term_off))) break;
            if ((st = launch_apply(b, y, x, 0))) break;   // warm-up
            RP_CHECK_HIP(hipEventRecord(e0, c.stream));
            for (int r = 0; r < 3 && !st; r++) st = launch_apply(b, y, x, 0);
            RP_CHECK_HIP(hipEventRecord(e1, c.stream));
            RP_CHECK_HIP(hipEventSynchronize(e1));
            float ms = 0;
            RP_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) { best = ms; pick = T; }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipFree(x);
        (void)hipFree(y);
        if (!st && (pick.tb != b->tile_bits || pick.nt != b->threads))
            st = build_batch_tables(b, pick.tb, pick.nt, bits, coeff, term_off);
        if (st) {
            batch_release(b);
            return st;
        }
    }
    // one problem of kPlanesMinLen qubits or more: the plane apply, forced,
    // or kept when the autotuner times it faster than the batched kernel,
    // or from 22 qubits without the autotuner
    const bool autot = (flags & RP_BATCH_AUTOTUNE) && !forced_tb;
    if (nprob == 1 && len[0] >= kPlanesMinLen && !forced_tb &&
        ((flags & RP_BATCH_PLANES) || autot || len[0] >= 22)) {
        const int64_t t0 = term_off[0];
        if ((st = setup_planes(b, bits + t0, coeff + 2 * t0, (uint64_t)(term_off[1] - t0),
                               autot))) {
            batch_release(b);
            return st;
        }
        b->planes = b->plane_terms != nullptr;
        if (b->planes && autot && !(flags & RP_BATCH_PLANES)) {
            float ms[2] = {0, 0};
            if ((st = time_batch_apply(b, 0, &ms[0])) || (st = time_batch_apply(b, 1, &ms[1]))) {
                batch_release(b);
                return st;
            }
            if (ms[0] <= ms[1]) release_planes(b);
            else b->planes = 1;
        }
    }
    *out = b;
    return RP_OK;
}

int rp_batch_free(rp_batch_t b)
{
    int st = check_init();
    if (st) return st;
    if (!b) return RP_OK;
    RP_CHECK_HIP(hipStreamSynchronize(ctx().stream));
    batch_release(b);
    return RP_OK;
}

int rp_batch_info(rp_batch_t b, rp_batch_info_t *info)
{
    RP_REQUIRE(b && info, "null argument");
    info->prec = b->prec;
    info->nprob = b->nprob;
    info->nrhs = b->ns;
    info->kmax = b->kmax;
    info->tile_bits = b->tile_bits;
    info->threads = b->threads;
    info->grid = b->grid;
    info->planes = b->planes;
    info->nbasis = b->nbasis;
    info->vec_bytes = b->nbasis * b->ns * elem_bytes(b->prec);
    info->nterms = b->nterms;
    info->ngroups = b->ngroups;
    info->bytes_per_apply = 0;
    for (int p = 0; p < b->nprob; p++) {
        const ProbDesc &P = b->probs[p];
        info->bytes_per_apply += (uint64_t)(P.g1 - P.g0 + 1) << P.len;
    }
    info->bytes_per_apply *= b->ns * elem_bytes(b->prec);
    if (b->planes) {   // the single-state kernel's layout
        rp_terms_info_t ti;
        int st = rp_terms_info(b->plane_terms, &ti);
        if (st) return st;
        info->tile_bits = ti.tile_bits;
        info->threads = ti.threads;
        info->grid = (int)((1ull << b->len[0]) >> (ti.tile_bits ? ti.tile_bits : 8));
        info->ngroups = ti.ngroups;
        info->bytes_per_apply = ti.bytes_per_apply * b->ns;
    }
    return RP_OK;
}

int rp_batch_offsets(rp_batch_t b, uint64_t *off)
{
    RP_REQUIRE(b && off, "null argument");
    for (int p = 0; p <= b->nprob; p++) off[p] = b->off[p];
    return RP_OK;
}

int rp_batch_groups(rp_batch_t b, uint64_t *groups)
{
    RP_REQUIRE(b && groups, "null argument");
    for (int p = 0; p < b->nprob; p++) groups[p] = b->probs[p].g1 - b->probs[p].g0;
    return RP_OK;
}

int rp_batch_set_active(rp_batch_t b, const int *active)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(b && active, "null argument");
    Context &c = ctx();
    RP_CHECK_HIP(hipMemcpyAsync(b->d_active, active, (size_t)b->nprob * sizeof(int),
                                hipMemcpyHostToDevice, c.stream));
    RP_CHECK_HIP(hipStreamSynchronize(c.stream));
    return RP_OK;
}

int rp_batch_set_basis(rp_batch_t b, void *const *V, int n)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(b && V, "null argument");
    RP_REQUIRE(n >= 1 && n <= b->kmax + 2, "n must be in [1, kmax + 2]");
    for (int i = 0; i < n; i++) {
        RP_REQUIRE(V[i], "null basis vector");
        b->V[i] = V[i];
    }
    Context &c = ctx();
    RP_CHECK_HIP(hipMemcpyAsync(b->d_V, b->V.data(), (size_t)(b->kmax + 2) * sizeof(void *),
                                hipMemcpyHostToDevice, c.stream));
    RP_CHECK_HIP(hipStreamSynchronize(c.stream));
    return RP_OK;
}

int rp_batch_apply(rp_batch_t b, void *res, const void *vec, int accumulate)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(b, "null batch handle");
    RP_REQUIRE(res && vec && res != vec, "null or aliased vectors");
    if (b->total_tiles == 0) return RP_OK;
    return launch_apply(b, res, vec, accumulate);
}

// V_i *= 1 / ||V_i|| per right-hand side; nrm_out[rhs] = ||V_i|| (device)
int rp_batch_normalize(rp_batch_t b, int i, double *nrm_out)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(b, "null batch handle");
    RP_REQUIRE(i >= 0 && i <= b->kmax + 1 && b->V[i], "basis vector not set");
    const int nr = b->nprob * b->ns;
    const ProbDesc *P = (const ProbDesc *)b->d_probs;
    const uint32_t *cp = (const uint32_t *)b->d_chunk_prob;
    const int *act = (const int *)b->d_active;
    const dim3 gc((unsigned)b->total_chunks);
    if (b->prec == RP_C64) {
        RP_LAUNCH(RP_KERNEL_GS, k_batch_norm_part<float>, gc, dim3(kRedT), 0,
                  (const float2 *)b->V[i], P, cp, act, b->ns, b->chunk_bits, (double *)b->d_npart);
    } else {
        RP_LAUNCH(RP_KERNEL_GS, k_batch_norm_part<double>, gc, dim3(kRedT), 0,
                  (const double2 *)b->V[i], P, cp, act, b->ns, b->chunk_bits, (double *)b->d_npart);
    }
    RP_LAUNCH(RP_KERNEL_GS, k_batch_norm_final, dim3(nr), dim3(kRedT), 0,
              (const double *)b->d_npart, P, act, b->nprob, b->ns, (double *)b->d_inv, nrm_out,
              (double *)nullptr, 0, b->kmax, 0, (int *)nullptr);
    if (b->prec == RP_C64) {
        RP_LAUNCH(RP_KERNEL_GS, k_batch_scale<float>, gc, dim3(kRedT), 0, (float2 *)b->V[i], P,
                  cp, act, b->ns, b->chunk_bits, (const double *)b->d_inv);
    } else {
        RP_LAUNCH(RP_KERNEL_GS, k_batch_scale<double>, gc, dim3(kRedT), 0, (double2 *)b->V[i],
                  P, cp, act, b->ns, b->chunk_bits, (const double *)b->d_inv);
    }
    return RP_OK;
}

// One Arnoldi step j on the basis table: V_{j+1} = A V_j, two classical
// Gram-Schmidt passes against V_0..V_j (FP64 coefficients), H(0..j, j) and
// H(j+1, j) written to H (per rhs a column-major (kmax+1) x kmax complex
// double matrix), V_{j+1} normalized.  Only kernel launches: capturable.
int rp_batch_arnoldi_step(rp_batch_t b, int j, double *H)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(b && H, "null argument");
    RP_REQUIRE(j >= 0 && j < b->kmax, "j must be in [0, kmax)");
    RP_REQUIRE(b->V[j] && b->V[j + 1], "basis vectors j and j+1 not set");
    if ((st = launch_apply(b, b->V[j + 1], b->V[j], 0))) return st;
    const ProbDesc *P = (const ProbDesc *)b->d_probs;
    const uint32_t *cp = (const uint32_t *)b->d_chunk_prob;
    const int *act = (const int *)b->d_active;
    const void *const *Vt = (const void *const *)b->d_V;
    const dim3 gc((unsigned)b->total_chunks);
    const int k = j + 1, hstride = 65 * 2, nr = b->nprob * b->ns;
    // the first pass always; the second only for the problems with a state
    // whose first-pass norm asks for it (the flags are set on the device, so
    // the step stays one fixed kernel sequence for graph capture).  A problem
    // that skips it keeps the first pass's norm partials, so the second
    // norm_final reads the first pass's norm.
    int *need2 = (int *)b->d_need2;
    for (int pass = 0; pass < 2; pass++) {
        const int *need = pass == 0 ? nullptr : need2;
        if (b->prec == RP_C64) {
            RP_LAUNCH(RP_KERNEL_GS, k_batch_dots<float>, gc, dim3(kRedT), 0, Vt, k,
                      (const float2 *)b->V[j + 1], P, cp, act, b->ns, b->chunk_bits,
                      (double *)b->d_part, need);
        } else {
            RP_LAUNCH(RP_KERNEL_GS, k_batch_dots<double>, gc, dim3(kRedT), 0, Vt, k,
                      (const double2 *)b->V[j + 1], P, cp, act, b->ns, b->chunk_bits,
                      (double *)b->d_part, need);
        }
        RP_LAUNCH(RP_KERNEL_GS, k_batch_reduce_dots, dim3((unsigned)b->nprob, (unsigned)(k * b->ns)),
                  dim3(kRedT), 0,
                  (const double *)b->d_part, k, P, act, b->ns, (double *)b->d_h, hstride, H, j,
                  b->kmax, pass == 0 ? 1 : 0, need);
        if (b->prec == RP_C64) {
            RP_LAUNCH(RP_KERNEL_GS, k_batch_update<float>, gc, dim3(kRedT), 0, Vt, k,
                      (float2 *)b->V[j + 1], P, cp, act, b->ns, b->chunk_bits,
                      (const double *)b->d_h, hstride, (double *)b->d_npart, need);
        } else {
            RP_LAUNCH(RP_KERNEL_GS, k_batch_update<double>, gc, dim3(kRedT), 0, Vt, k,
                      (double2 *)b->V[j + 1], P, cp, act, b->ns, b->chunk_bits,
                      (const double *)b->d_h, hstride, (double *)b->d_npart, need);
        }
        // after the first pass: its norm to Im H(j+1, j) and the second-pass
        // flags; after the second: the norm to Re H(j+1, j) and its inverse
        // for the normalization
        RP_LAUNCH(RP_KERNEL_GS, k_batch_norm_final, dim3(nr), dim3(kRedT), 0,
                  (const double *)b->d_npart, P, act, b->nprob, b->ns, (double *)b->d_inv,
                  (double *)nullptr, H, j, b->kmax, pass == 0 ? 1 : 0,
                  pass == 0 ? need2 : (int *)nullptr);
    }
    if (b->prec == RP_C64) {
        RP_LAUNCH(RP_KERNEL_GS, k_batch_scale<float>, gc, dim3(kRedT), 0, (float2 *)b->V[j + 1],
                  P, cp, act, b->ns, b->chunk_bits, (const double *)b->d_inv);
    } else {
        RP_LAUNCH(RP_KERNEL_GS, k_batch_scale<double>, gc, dim3(kRedT), 0,
                  (double2 *)b->V[j + 1], P, cp, act, b->ns, b->chunk_bits,
                  (const double *)b->d_inv);
    }
    return RP_OK;
}

// y += sum_{i<k} c[rhs][i] V_i, c complex double with stride cstride doubles
int rp_batch_lincomb(rp_batch_t b, void *y, int k, const double *c, int cstride)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(b && y && c, "null argument");
    RP_REQUIRE(k >= 1 && k <= 64 && k <= b->kmax + 1, "k must be in [1, min(64, kmax + 1)]");
    RP_REQUIRE(cstride >= 2 * k, "cstride too small");
    for (int i = 0; i < k; i++) RP_REQUIRE(b->V[i], "basis vector not set");
    const dim3 gc((unsigned)b->total_chunks);
    if (b->prec == RP_C64) {
        RP_LAUNCH(RP_KERNEL_OTHER, k_batch_lincomb<float>, gc, dim3(kRedT), 0,
                  (const void *const *)b->d_V, k, (float2 *)y, (const ProbDesc *)b->d_probs,
                  (const uint32_t *)b->d_chunk_prob, (const int *)b->d_active, b->ns,
                  b->chunk_bits, c, cstride);
    } else {
        RP_LAUNCH(RP_KERNEL_OTHER, k_batch_lincomb<double>, gc, dim3(kRedT), 0,
                  (const void *const *)b->d_V, k, (double2 *)y, (const ProbDesc *)b->d_probs,
                  (const uint32_t *)b->d_chunk_prob, (const int *)b->d_active, b->ns,
                  b->chunk_bits, c, cstride);
    }
    return RP_OK;
}

// ---- HIP graphs of Arnoldi steps ---------------------------------------------

struct rp_graph { hipGraph_t g = nullptr; hipGraphExec_t x = nullptr; };

int rp_batch_capture_arnoldi(rp_batch_t b, int j0, int j1, double *H, rp_graph_t *out)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(b && H && out, "null argument");
    RP_REQUIRE(0 <= j0 && j0 < j1 && j1 <= b->kmax, "need 0 <= j0 < j1 <= kmax");
    *out = nullptr;
    Context &c = ctx();
    RP_REQUIRE(!c.prof, "disable profiling (rp_prof_enable(0)) before capturing a graph");
    RP_CHECK_HIP(hipStreamSynchronize(c.stream));
    RP_CHECK_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    for (int j = j0; j < j1 && !st; j++) st = rp_batch_arnoldi_step(b, j, H);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c.stream, &g);
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess)
        return set_error(RP_ERR_RUNTIME, "rp_batch_capture_arnoldi: capture failed: %s",
                         hipGetErrorString(e));
    rp_graph *r = new rp_graph;
    r->g = g;
    e = hipGraphInstantiate(&r->x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        delete r;
        return set_error(RP_ERR_RUNTIME, "rp_batch_capture_arnoldi: instantiate failed: %s",
                         hipGetErrorString(e));
    }
    *out = r;
    return RP_OK;
}

int rp_graph_launch(rp_graph_t g)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(g, "null graph");
    RP_CHECK_HIP(hipGraphLaunch(g->x, ctx().stream));
    return RP_OK;
}

int rp_graph_free(rp_graph_t g)
{
    int st = check_init();
    if (st) return st;
    if (!g) return RP_OK;
    RP_CHECK_HIP(hipStreamSynchronize(ctx().stream));
    if (g->x) (void)hipGraphExecDestroy(g->x);
    if (g->g) (void)hipGraphDestroy(g->g);
    delete g;
    return RP_OK;
}

}  // extern "C"
