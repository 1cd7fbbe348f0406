// rp_apply.hip -- the matrix-free Pauli-sum apply of libreapers_hip.so.
//
// Reference: DevSumOps::upload (inc/gpuimpl.h:1064-1085) copies the raw
// SpinOp array to the device and krnl_apply_ops (src/krnl.cu:80-94) runs one
// thread per output amplitude that walks every term's 2-bit-per-site string
// (SpinOp::apply, inc/ops.h:421-440):
//
//     res[m] += coeff * (-1)^popcount((m ^ x) & z) * vec[m ^ x]
//
// Here the upload does the per-term bit work once on the host:
//   x = flip mask (X bits, even positions of SpinOp::bits),
//   z = phase mask (Z bits, odd positions),
//   c = coeff * (-1)^popcount(x & z), so that
//     res[m] += c * (-1)^popcount(m & z) * vec[m ^ x],
// and sorts the terms by x.  Terms with equal x ("subs") share one gather of
// vec; subs whose x agree above the tile bits b ("groups") share one staged
// input tile.  A Hermitian term has c purely real or purely imaginary, so
// single-term subs are classified real / imaginary / complex on the host and
// the kernel spends 2 FMAs (not 4) per term and amplitude on the first two.
//
// Tiled kernel: one workgroup of NT threads owns the output tile
// [t0, t0 + 2^b); thread tid holds amplitudes t0 + r*NT + tid, r < RP.  For
// each group it stages the input tile t0 ^ (xh << b) in LDS with coalesced
// loads (an XOR with the high part maps an aligned tile onto an aligned
// tile), then applies every sub of the group from LDS: the low bits of x
// permute lanes (tid ^ (x mod NT)) and rows (r ^ (x / NT)).  The sign
// (-1)^popcount(m & z) splits into a per-workgroup part (t0 & z), a per-lane
// part (tid & z) computed once per sub, and a per-row part (r*NT & z) that is
// uniform across the wavefront.  The output accumulates in registers (FP64
// for both state precisions by default) and is written once, with the
// caller's zero_vec fused in when accumulate == 0.
// Workgroups are numbered so that each XCD walks a contiguous range of
// output tiles, whose input tiles overlap in that XCD's L2.
// HBM traffic per apply: (G_b + 1) * 2^len * S bytes, G_b = number of groups.
#include "rp_internal.h"
#include "rp_tables.h"

#include <algorithm>
#include <type_traits>
#include <cstring>
#include <vector>

namespace rp {

constexpr int kSmallThreads = 256;

}  // namespace rp

struct rp_terms {
    int prec = RP_C64;
    int len = 0;
    int tile_bits = 0;       // 0: small kernel (one thread per amplitude)
    int threads = 0;         // workgroup size of the tiled kernel
    int depth = 2;           // LDS tiles in the tiled kernel's pipeline; for the
                             // small kernel the number of term slices (1: none)
    int acc64 = 1;           // accumulate in FP64
    int group_acc = 0;       // FP32 within a group, FP64 across groups (complex<float>)
    uint64_t nterms = 0, nflips = 0, ngroups = 0, nsubs = 0;
    rp::TileMap tmap{};               // output-tile order of the tiled kernels
    void *d_groups = nullptr, *d_subs = nullptr, *d_terms = nullptr, *d_flat = nullptr;
    void *d_cgroups = nullptr, *d_csubs = nullptr, *d_cterms = nullptr;   // compact stream
    void *d_split = nullptr;          // per-slice partial sums of the split small kernel
    size_t split_bytes = 0;
    double apply_passes = 0;          // measured apply cost in state passes (autotune)
    std::vector<uint64_t> raw_bits;   // the uploaded term list, for re-tiling
    std::vector<double> raw_coeff;
};

namespace rp {

// one thread per amplitude; terms in (x-sorted) flat order
template <typename R, typename A>
__global__ void __launch_bounds__(kSmallThreads)
k_apply_small(const typename CxA<R>::T *__restrict__ in, typename CxA<R>::T *__restrict__ out,
              uint32_t dim, const FlatTerm<A> *__restrict__ terms, int nterms, int accumulate)
{
    uint32_t m = blockIdx.x * kSmallThreads + threadIdx.x;
    if (m >= dim) return;
    A ar = 0, ai = 0;
    for (int t = 0; t < nterms; t++) {
        const FlatTerm<A> f = terms[t];
        const auto v = in[m ^ f.x];
        const A vr = v.x, vi = v.y;
        const bool neg = parity32(m & f.z);
        const A cr = neg ? -f.cr : f.cr;
        const A ci = neg ? -f.ci : f.ci;
        ar += cr * vr - ci * vi;
        ai += cr * vi + ci * vr;
    }
    if (accumulate) {
        const auto o = out[m];
        ar += (A)o.x;
        ai += (A)o.y;
    }
    typename CxA<R>::T o;
    o.x = (R)ar;
    o.y = (R)ai;
    out[m] = o;
}

// The small kernel with the terms split into gridDim.y slices: slice s sums
// its contiguous range of the x-sorted terms for one amplitude per thread and
// writes the partial sum; k_split_reduce adds the slices in slice order.  For
// chains of up to 16 sites, where 2^len / 256 workgroups alone cannot fill
// the GPU.
template <typename R, typename A>
__global__ void __launch_bounds__(kSmallThreads)
k_apply_split(const typename CxA<R>::T *__restrict__ in, typename CxA<A>::T *__restrict__ part,
              uint32_t dim, const FlatTerm<A> *__restrict__ terms, int nterms, int per)
{
    uint32_t m = blockIdx.x * kSmallThreads + threadIdx.x;
    if (m >= dim) return;
    const int t0 = blockIdx.y * per;
    const int t1 = min(nterms, t0 + per);
    A ar = 0, ai = 0;
    for (int t = t0; t < t1; t++) {
        const FlatTerm<A> f = terms[t];
        const auto v = in[m ^ f.x];
        const A vr = v.x, vi = v.y;
        const bool neg = parity32(m & f.z);
        const A cr = neg ? -f.cr : f.cr;
        const A ci = neg ? -f.ci : f.ci;
        ar += cr * vr - ci * vi;
        ai += cr * vi + ci * vr;
    }
    typename CxA<A>::T o;
    o.x = ar;
    o.y = ai;
    part[(uint64_t)blockIdx.y * dim + m] = o;
}

template <typename R, typename A>
__global__ void __launch_bounds__(kSmallThreads)
k_split_reduce(const typename CxA<A>::T *__restrict__ part, typename CxA<R>::T *__restrict__ out,
               uint32_t dim, int slices, int accumulate)
{
    for (uint32_t m = blockIdx.x * kSmallThreads + threadIdx.x; m < dim;
         m += gridDim.x * kSmallThreads) {
        A ar = 0, ai = 0;
        for (int sl = 0; sl < slices; sl++) {
            const auto p = part[(uint64_t)sl * dim + m];
            ar += p.x;
            ai += p.y;
        }
        if (accumulate) {
            const auto o = out[m];
            ar += (A)o.x;
            ai += (A)o.y;
        }
        typename CxA<R>::T o;
        o.x = (R)ar;
        o.y = (R)ai;
        out[m] = o;
    }
}

template <typename R, int TB, int NT, int D = 2>
constexpr int apply_waves_per_eu()
{
    constexpr int lds = D * (1 << TB) * 2 * (int)sizeof(R);
    constexpr int wg_by_lds = (160 * 1024) / lds;
    // 1024-thread workgroups: two per CU (32 waves) when their LDS fits
    constexpr int wg_by_waves = (NT == 1024 ? 2048 : 1024) / NT;
    constexpr int wg = wg_by_lds < wg_by_waves ? wg_by_lds : wg_by_waves;
    constexpr int w = (wg < 1 ? 1 : wg) * (NT / 64) / 4;
    return w < 1 ? 1 : w;
}

// Copy the input tile starting at element `src` into LDS buffer `dst` with
// gfx950's asynchronous global->LDS loads: every lane moves 16 bytes, a
// wavefront fills a contiguous 1 KiB run starting at the wave-uniform LDS
// address.  Completion is tracked by the vector memory counter.
template <typename T, int TB, int NT>
__device__ inline void tile_dma(const T *__restrict__ in, uint64_t src, T *dst)
{
    constexpr int TILE_BYTES = (1 << TB) * (int)sizeof(T);
    constexpr int PER_THREAD = TILE_BYTES / (16 * NT);
    static_assert(PER_THREAD >= 1 && TILE_BYTES % (16 * NT) == 0, "tile too small for DMA");
    // uniform tile base plus a 32-bit lane offset (below 64 KiB)
    const char *g = (const char *)(in + src);
    char *l = (char *)dst;
    // wave-uniform: kept in a scalar register, so that the LDS address
    // (M0) of every DMA is scalar arithmetic
    const uint32_t tid = threadIdx.x,
                   wave_base = __builtin_amdgcn_readfirstlane(tid & ~63u);
#pragma unroll
    for (int c = 0; c < PER_THREAD; c++) {
        const uint32_t chunk = c * NT;
        __builtin_amdgcn_global_load_lds(
            (const void *)(g + (uint32_t)((chunk + tid) * 16)),
            (__attribute__((address_space(3))) void *)(l + (size_t)(chunk + wave_base) * 16),
            16, 0, 0);
    }
}

// a sub's coefficient in the partial-sum type P (float copies for FP32
// partial sums of FP64-coefficient layouts: no conversion in the loop)
template <typename P, typename A>
__device__ __forceinline__ P sub_re(const SubRec<A> &S)
{
    if constexpr (std::is_same_v<P, float> && !std::is_same_v<A, float>) return S.fr;
    else return (P)S.cr;
}
template <typename P, typename A>
__device__ __forceinline__ P sub_im(const SubRec<A> &S)
{
    if constexpr (std::is_same_v<P, float> && !std::is_same_v<A, float>) return S.fi;
    else return (P)S.ci;
}

// Sign bit (bit 31) of (-1)^popcount(v), and a coefficient with that sign
// applied by flipping its sign bit
__device__ __forceinline__ uint32_t parity_sign(uint32_t v) { return (uint32_t)__popc(v) << 31; }
// (adding bit 31 flips it like an xor: one shift-add instruction with the
// shift of the sign bit)
__device__ __forceinline__ float flip_sign(float c, uint32_t sb)
{
    return __uint_as_float(__float_as_uint(c) + sb);
}
__device__ __forceinline__ double flip_sign(double c, uint32_t sb)
{
    return __longlong_as_double(__double_as_longlong(c) ^ ((long long)sb << 32));
}
// a * b + c with one rounding
__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }

// the lane-signed coefficient of a sub whose rows flip it again: kept in a
// register, so that each row's flip is one xor with a scalar sign instead
// of the whole popcount per row
template <typename P> __device__ __forceinline__ P opaque(P c)
{
    asm volatile("" : "+v"(c));
    return c;
}

// The LDS address of this lane's input element in tile row 0 for flip mask
// xl: the lane offset is formed once per sub (one xor) and kept opaque, so
// that each row's address is that offset plus a scalar row offset
template <typename T, int NT>
__device__ __forceinline__ const T *sub_lane_base(const T *buf, uint32_t lane8, uint32_t xl)
{
    uint32_t la = lane8 ^ ((xl & (NT - 1)) * (uint32_t)sizeof(T));
    asm volatile("" : "+v"(la));
    return (const T *)((const char *)buf + la);
}

// Apply every sub of group G from the staged input tile `buf` (LDS) to the
// RP accumulated rows of this thread.  Amplitude t0 + r NT + tid of the
// output tile reads the input at (t0 + r NT + tid) ^ x: lane tid ^ (x mod NT)
// of tile row r ^ (x / NT mod RP); lanebits = t0 | tid (the index bits
// outside the row bits) gives the lane's share of the phase sign.
template <typename R, typename A, int TB, int NT, typename P = A>
__device__ __forceinline__ void apply_group(const GroupRuns &G, const typename CxA<R>::T *buf,
                                            const SubRec<A> *__restrict__ subs,
                                            const TermRec<A> *__restrict__ terms,
                                            uint32_t lanebits, uint32_t tid,
                                            P (&accr)[(1 << TB) / NT], P (&acci)[(1 << TB) / NT])
{
    using T = typename CxA<R>::T;
    constexpr int RP = (1 << TB) / NT;
    constexpr int LB = NT == 1024 ? 10 : NT == 512 ? 9 : NT == 256 ? 8 : 7;

    // this lane's byte offset in a tile row
    const uint32_t lane8 = tid * (uint32_t)sizeof(T);
#define RP_SUB_PROLOGUE                                                            \
    const SubRec<A> S = subs[s];                                               \
    const T *row = sub_lane_base<T, NT>(buf, lane8, S.xl);                     \
    const uint32_t xr = S.xl >> LB;                                            \
    const uint32_t lsb = parity_sign(lanebits & S.z);                          \
    const uint32_t zr = (S.z >> LB) & (RP - 1);                                \
    (void)zr;
    // real c, no phase bit among the row bits
    for (uint32_t s = G.b[kRunReal0]; s < G.b[kRunReal0 + 1]; s++) {
        RP_SUB_PROLOGUE
        const P c = flip_sign(sub_re<P>(S), lsb);
#pragma unroll
        for (int r = 0; r < RP; r++) {
            const T v = row[(r ^ xr) * NT];
            accr[r] += c * (P)v.x;
            acci[r] += c * (P)v.y;
        }
    }
    for (uint32_t s = G.b[kRunReal1]; s < G.b[kRunReal1 + 1]; s++) {
        RP_SUB_PROLOGUE
        const P c = opaque(flip_sign(sub_re<P>(S), lsb));
#pragma unroll
        for (int r = 0; r < RP; r++) {
            const T v = row[(r ^ xr) * NT];
            const P cc = flip_sign(c, parity_sign(r & zr));
            accr[r] += cc * (P)v.x;
            acci[r] += cc * (P)v.y;
        }
    }
    // imaginary c = i ci: (i ci)(vr + i vi) = -ci vi + i ci vr
    for (uint32_t s = G.b[kRunImag0]; s < G.b[kRunImag0 + 1]; s++) {
        RP_SUB_PROLOGUE
        const P c = flip_sign(sub_im<P>(S), lsb);
#pragma unroll
        for (int r = 0; r < RP; r++) {
            const T v = row[(r ^ xr) * NT];
            accr[r] -= c * (P)v.y;
            acci[r] += c * (P)v.x;
        }
    }
    for (uint32_t s = G.b[kRunImag1]; s < G.b[kRunImag1 + 1]; s++) {
        RP_SUB_PROLOGUE
        const P c = opaque(flip_sign(sub_im<P>(S), lsb));
#pragma unroll
        for (int r = 0; r < RP; r++) {
            const T v = row[(r ^ xr) * NT];
            const P cc = flip_sign(c, parity_sign(r & zr));
            accr[r] -= cc * (P)v.y;
            acci[r] += cc * (P)v.x;
        }
    }
    for (uint32_t s = G.b[kRunCplx]; s < G.b[kRunCplx + 1]; s++) {
        RP_SUB_PROLOGUE
        const P cr = opaque(flip_sign(sub_re<P>(S), lsb)), ci = opaque(flip_sign(sub_im<P>(S), lsb));
#pragma unroll
        for (int r = 0; r < RP; r++) {
            const T v = row[(r ^ xr) * NT];
            const uint32_t rsb = parity_sign(r & zr);
            const P a = flip_sign(cr, rsb), c = flip_sign(ci, rsb);
            // two nested fused multiply-adds per component (the rounding
            // order of the compact-record kernel)
            accr[r] = fmadd(a, (P)v.x, fmadd(-c, (P)v.y, accr[r]));
            acci[r] = fmadd(a, (P)v.y, fmadd(c, (P)v.x, acci[r]));
        }
    }
#undef RP_SUB_PROLOGUE
    // several terms share x: d(m) = sum_t c_t (-1)^popcount(m & z_t),
    // formed for four rows at a time to bound register use
    for (uint32_t s = G.b[kRunMulti]; s < G.b[kRunMulti + 1]; s++) {
        const SubRec<A> S = subs[s];
        const T *row = sub_lane_base<T, NT>(buf, lane8, S.xl);
        const uint32_t xr = S.xl >> LB;
        constexpr int RC = RP < 4 ? RP : 4;
#pragma unroll
        for (int r0 = 0; r0 < RP; r0 += RC) {
            P dr[RC], di[RC];
#pragma unroll
            for (int q = 0; q < RC; q++) { dr[q] = 0; di[q] = 0; }
            for (uint32_t t = S.term_begin; t < S.term_end; t++) {
                const TermRec<A> C = terms[t];
                const uint32_t lsb = parity_sign(lanebits & C.z);
                const P cr = (P)flip_sign(C.cr, lsb);
                const P ci = (P)flip_sign(C.ci, lsb);
                const uint32_t zr = (C.z >> LB) & (RP - 1);
#pragma unroll
                for (int q = 0; q < RC; q++) {
                    const uint32_t rsb = parity_sign((r0 + q) & zr);
                    dr[q] += flip_sign(cr, rsb);
                    di[q] += flip_sign(ci, rsb);
                }
            }
#pragma unroll
            for (int q = 0; q < RC; q++) {
                const T v = row[((r0 + q) ^ xr) * NT];
                accr[r0 + q] = fmadd(dr[q], (P)v.x, fmadd(-di[q], (P)v.y, accr[r0 + q]));
                acci[r0 + q] = fmadd(dr[q], (P)v.y, fmadd(di[q], (P)v.x, acci[r0 + q]));
            }
        }
    }
}

// s_waitcnt immediate that waits until at most N vector memory operations
// of this wave are outstanding and all of its LDS operations have completed
// (gfx9 encoding: vmcnt in bits 3:0 and 15:14, expcnt 6:4 left at its
// maximum, lgkmcnt 11:8 = 0)
template <int N>
__device__ inline void wait_vmcnt_lds()
{
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4));
}

// Workgroup barrier without the release/acquire fences of __syncthreads(),
// whose release would drain every outstanding global->LDS load.  The caller
// has waited for its own loads into the tile about to be read (vmcnt) and
// for its own LDS reads of the tile about to be overwritten (lgkmcnt 0); the
// compiler barriers keep memory operations on their side.
__device__ inline void tile_barrier()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// Pipeline of D LDS tiles: the input tiles of the next D-1 groups are in
// flight while a group is applied.  Phase K of an unrolled round uses the
// statically named buffer K, so the compiler sees which DMA writes which
// array and does not drain every load before each LDS read.
template <typename R, typename A, int TB, int NT, int D, int K, bool GA>
__device__ __forceinline__ void pipe_phase(
    int g, int ngroups, const typename CxA<R>::T *__restrict__ in,
    const GroupRuns *__restrict__ groups, const SubRec<A> *__restrict__ subs,
    const TermRec<A> *__restrict__ terms, uint32_t t0, uint32_t tid,
    typename CxA<R>::T *b0, typename CxA<R>::T *b1, typename CxA<R>::T *b2,
    typename CxA<R>::T *b3, A (&accr)[(1 << TB) / NT], A (&acci)[(1 << TB) / NT],
    float (&gr)[(1 << TB) / NT], float (&gi)[(1 << TB) / NT])
{
    using T = typename CxA<R>::T;
    constexpr int PT = ((1 << TB) * (int)sizeof(T)) / (16 * NT);   // DMA ops per thread per tile
    T *const bufs[4] = {b0, b1, b2, b3};
    // this thread's share of group g's tile has landed once at most the
    // DMAs issued after it (groups g+1 .. g+D-2) are outstanding; after the
    // barrier every share is visible and group g-1's buffer is no longer read
    if constexpr (D == 2) {
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    } else {
        if (g + D - 2 < ngroups) wait_vmcnt_lds<(D - 2) * PT>();
        else __builtin_amdgcn_s_waitcnt(0);
        tile_barrier();
    }
    if (g + D - 1 < ngroups)
        tile_dma<T, TB, NT>(in, t0 ^ (groups[g + D - 1].xh << TB), bufs[(K + D - 1) % D]);
    if constexpr (GA) {
        // the terms of four consecutive groups summed in FP32 (gr, gi), those
        // sums added to the FP64 accumulators by the caller
        apply_group<R, A, TB, NT, float>(groups[g], bufs[K], subs, terms, t0 | tid, tid, gr, gi);
    } else {
        apply_group<R, A, TB, NT>(groups[g], bufs[K], subs, terms, t0 | tid, tid, accr, acci);
    }
}

// add the FP32 partial sums to the FP64 accumulators and clear them
template <typename A, int RP>
__device__ __forceinline__ void flush_partials(A (&accr)[RP], A (&acci)[RP], float (&gr)[RP],
                                               float (&gi)[RP])
{
#pragma unroll
    for (int r = 0; r < RP; r++) {
        accr[r] += (A)gr[r];
        acci[r] += (A)gi[r];
        gr[r] = 0;
        gi[r] = 0;
        // opaque zeros: the compiler would otherwise propagate the constant
        // into the next round and zero the registers again on every path
        asm volatile("" : "+v"(gr[r]), "+v"(gi[r]));
    }
}

// The body of the tiled apply for the output tile of this workgroup: every
// group's input tile through the LDS pipeline, then the tile written once.
// wv receives the values written (the rounded outputs), for an epilogue.
template <typename R, typename A, int TB, int NT, int D, bool GA>
__device__ __forceinline__ void apply_tile(const typename CxA<R>::T *__restrict__ in,
                                           typename CxA<R>::T *__restrict__ out,
                                           const GroupRuns *__restrict__ groups, int ngroups,
                                           const SubRec<A> *__restrict__ subs,
                                           const TermRec<A> *__restrict__ terms, int accumulate,
                                           typename CxA<R>::T *tile0, typename CxA<R>::T *tile1,
                                           typename CxA<R>::T *tile2, typename CxA<R>::T *tile3,
                                           uint32_t t0, typename CxA<R>::T (&wv)[(1 << TB) / NT])
{
    using T = typename CxA<R>::T;
    constexpr int TILE = 1 << TB;
    constexpr int RP = TILE / NT;          // amplitudes per thread
    const uint32_t tid = threadIdx.x;

    A accr[RP], acci[RP];
    float gr[RP], gi[RP];      // FP32 partial sums (group accumulation)
#pragma unroll
    for (int r = 0; r < RP; r++) {
        accr[r] = 0;
        acci[r] = 0;
        gr[r] = 0;
        gi[r] = 0;
        if constexpr (GA) asm volatile("" : "+v"(gr[r]), "+v"(gi[r]));
    }

    T *const ring[4] = {tile0, tile1, tile2, tile3};
#pragma unroll
    for (int q = 0; q < D - 1; q++)
        if (q < ngroups) tile_dma<T, TB, NT>(in, t0 ^ (groups[q].xh << TB), ring[q]);
    for (int g = 0; g < ngroups; g += D) {
        pipe_phase<R, A, TB, NT, D, 0, GA>(g, ngroups, in, groups, subs, terms, t0, tid, tile0, tile1,
                                       tile2, tile3, accr, acci, gr, gi);
        if (g + 1 >= ngroups) break;
        pipe_phase<R, A, TB, NT, D, 1, GA>(g + 1, ngroups, in, groups, subs, terms, t0, tid, tile0,
                                       tile1, tile2, tile3, accr, acci, gr, gi);
        if constexpr (D > 2) {
            if (g + 2 >= ngroups) break;
            pipe_phase<R, A, TB, NT, D, 2, GA>(g + 2, ngroups, in, groups, subs, terms, t0, tid, tile0,
                                           tile1, tile2, tile3, accr, acci, gr, gi);
        }
        if constexpr (D > 3) {
            if (g + 3 >= ngroups) break;
            pipe_phase<R, A, TB, NT, D, 3, GA>(g + 3, ngroups, in, groups, subs, terms, t0, tid, tile0,
                                           tile1, tile2, tile3, accr, acci, gr, gi);
        }
        // FP32 partial sums of every other round (four groups at D = 2) into
        // the FP64 accumulators
        if constexpr (GA) {
            if ((g / D) & 1) flush_partials(accr, acci, gr, gi);
        }
    }
    if constexpr (GA) flush_partials(accr, acci, gr, gi);
#pragma unroll
    for (int r = 0; r < RP; r++) {
        const uint64_t m = (uint64_t)t0 + r * NT + tid;
        T o;
        if (accumulate) {
            const T p = out[m];
            o.x = (R)(accr[r] + (A)p.x);
            o.y = (R)(acci[r] + (A)p.y);
        } else {
            o.x = (R)accr[r];
            o.y = (R)acci[r];
        }
        out[m] = o;
        wv[r] = o;
    }
}

// XCD-aware numbering: consecutive workgroup ids go to different XCDs, so
// give each XCD a contiguous range of tiles
__device__ __forceinline__ uint32_t xcd_tile_index()
{
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    return (nb & 7u) == 0 ? (b & 7u) * (nb >> 3) + (b >> 3) : b;
}

// the workgroup's output tile under the map tm (rp_tables.h): index bits
// 0..p-1 = position within the XCD's resident workgroups, p..p+2 = XCD,
// above = batch; the grid has 2^n >= 2^(p+3) workgroups
__device__ __forceinline__ uint32_t tile_index(const TileMap &tm)
{
    if (!tm.n) return xcd_tile_index();
    const uint32_t b = blockIdx.x, p = (uint32_t)tm.p, j = b >> 3;
    const uint32_t idx = ((j >> p) << (p + 3)) | ((b & 7u) << p) | (j & ((1u << p) - 1));
    uint32_t t = 0;
    for (int i = 0; i < tm.n; i++)
        if (idx >> i & 1u) t ^= tm.col[i];
    return t;
}

template <typename R, typename A, int TB, int NT, int D, bool GA>
__global__ void __launch_bounds__(NT, (apply_waves_per_eu<R, TB, NT, D>()))
k_apply_tiled(const typename CxA<R>::T *__restrict__ in, typename CxA<R>::T *__restrict__ out,
              const GroupRuns *__restrict__ groups, int ngroups, const SubRec<A> *__restrict__ subs,
              const TermRec<A> *__restrict__ terms, int accumulate, TileMap tm)
{
    using T = typename CxA<R>::T;
    constexpr int TILE = 1 << TB;
    static_assert(NT == 1024 || NT == 512 || NT == 256 || NT == 128,
                  "NT must be a power of two in [128, 1024]");
    static_assert(D >= 2 && D <= 4, "2 to 4 LDS tiles");
    // D separate LDS tiles (the unused ones have one element)
    __shared__ T tile0[TILE];
    __shared__ T tile1[TILE];
    __shared__ T tile2[D > 2 ? TILE : 1];
    __shared__ T tile3[D > 3 ? TILE : 1];
    T wv[TILE / NT];
    apply_tile<R, A, TB, NT, D, GA>(in, out, groups, ngroups, subs, terms, accumulate, tile0, tile1,
                                    tile2, tile3, tile_index(tm) << TB, wv);
}

// ---------------------------------------------------------------------------
// The complex<float> group-accumulating apply over compact sub records.
// The 48-byte SubRec list of a sparse SYK Hamiltonian at 24-30 qubits
// (370-480 subs, 18-23 KiB) does not fit the 16 KiB scalar data cache, so
// every sub record of the kernel above is a dependent load from L2.  Here
// each sub is 16 bytes and each group header 16 bytes (8-12 KiB together);
// the subs of all groups form one list in group order, and the record of
// the next sub, and the header of the group after next, are loaded while
// the current ones are in use.  Each single-term sub branches once on its
// row flip (uniform) to a code path whose LDS row reads are immediate
// offsets from one lane address.
// ---------------------------------------------------------------------------
struct SubC { uint32_t xl, z; float c0, c1; };
struct TermC { uint32_t z; float cr, ci; uint32_t pad; };
// group header: tile index above the tile bits, first sub, subs per class
// run (kRunReal0 .. kRunMulti)
struct GroupC { uint32_t xh, begin; uint8_t n[kRuns]; uint8_t pad[8 - kRuns]; };
static_assert(sizeof(GroupC) == 16, "16-byte group header");

template <int V> struct IC { static constexpr int value = V; };

// f(IC<xr>) for the run-time row flip xr < RP: one code path per value, so
// that the row offsets are immediates
template <int RP, typename F>
__device__ __forceinline__ void with_row_flip(uint32_t xr, F &&f)
{
    if constexpr (RP == 8) {
        switch (xr) {
        case 0: f(IC<0>{}); break;
        case 1: f(IC<1>{}); break;
        case 2: f(IC<2>{}); break;
        case 3: f(IC<3>{}); break;
        case 4: f(IC<4>{}); break;
        case 5: f(IC<5>{}); break;
        case 6: f(IC<6>{}); break;
        default: f(IC<7>{}); break;
        }
    } else if constexpr (RP == 4) {
        switch (xr) {
        case 0: f(IC<0>{}); break;
        case 1: f(IC<1>{}); break;
        case 2: f(IC<2>{}); break;
        default: f(IC<3>{}); break;
        }
    } else if constexpr (RP == 2) {
        if (xr) f(IC<1>{});
        else f(IC<0>{});
    } else {
        f(IC<0>{});
    }
}

// all RP row values of a sub loaded before any is used: the reads go out
// back to back and one wait covers them
template <int RP> __device__ __forceinline__ void pin_rows(float2 (&v)[RP])
{
    if constexpr (RP == 8)
        asm volatile("" : "+v"(v[0].x), "+v"(v[0].y), "+v"(v[1].x), "+v"(v[1].y), "+v"(v[2].x),
                     "+v"(v[2].y), "+v"(v[3].x), "+v"(v[3].y), "+v"(v[4].x), "+v"(v[4].y),
                     "+v"(v[5].x), "+v"(v[5].y), "+v"(v[6].x), "+v"(v[6].y), "+v"(v[7].x),
                     "+v"(v[7].y));
    else if constexpr (RP == 4)
        asm volatile("" : "+v"(v[0].x), "+v"(v[0].y), "+v"(v[1].x), "+v"(v[1].y), "+v"(v[2].x),
                     "+v"(v[2].y), "+v"(v[3].x), "+v"(v[3].y));
    else if constexpr (RP == 2)
        asm volatile("" : "+v"(v[0].x), "+v"(v[0].y), "+v"(v[1].x), "+v"(v[1].y));
    else
        asm volatile("" : "+v"(v[0].x), "+v"(v[0].y));
}

// One single-term sub with the row flip XR fixed: input row j of the
// lane's column (row base `row`, immediate offsets) feeds output row
// j ^ XR.  c (and ci for the complex class) carry the lane's phase sign; zr
// = the phase bits among the row bits (classes 1 and complex).
template <int RP, int NT, uint32_t CLS, int XR>
__device__ __forceinline__ void sub_rows(const float2 *row, float c, float ci, uint32_t zr,
                                         float (&gr)[RP], float (&gi)[RP])
{
    float2 vv[RP];
#pragma unroll
    for (int j = 0; j < RP; j++) vv[j] = row[j * NT];
    pin_rows<RP>(vv);
#pragma unroll
    for (int j = 0; j < RP; j++) {
        const int r = j ^ XR;
        const float2 v = vv[j];
        // explicit fused multiply-adds: one rounding each, and no product
        // that the compiler could share between the row-flip paths
        if constexpr (CLS == kRunReal0) {
            gr[r] = __builtin_fmaf(c, v.x, gr[r]);
            gi[r] = __builtin_fmaf(c, v.y, gi[r]);
        } else if constexpr (CLS == kRunImag0) {
            gr[r] = __builtin_fmaf(-c, v.y, gr[r]);
            gi[r] = __builtin_fmaf(c, v.x, gi[r]);
        } else if constexpr (CLS == kRunReal1) {
            const float cc = flip_sign(c, parity_sign((uint32_t)r & zr));
            gr[r] = __builtin_fmaf(cc, v.x, gr[r]);
            gi[r] = __builtin_fmaf(cc, v.y, gi[r]);
        } else if constexpr (CLS == kRunImag1) {
            const float cc = flip_sign(c, parity_sign((uint32_t)r & zr));
            gr[r] = __builtin_fmaf(-cc, v.y, gr[r]);
            gi[r] = __builtin_fmaf(cc, v.x, gi[r]);
        } else {
            const uint32_t rsb = parity_sign((uint32_t)r & zr);
            const float a = flip_sign(c, rsb), b = flip_sign(ci, rsb);
            gr[r] = __builtin_fmaf(a, v.x, __builtin_fmaf(-b, v.y, gr[r]));
            gi[r] = __builtin_fmaf(a, v.y, __builtin_fmaf(b, v.x, gi[r]));
        }
    }
}

// The subs of one class run [s, end) of a group from the staged tile buf.
// c0 holds the real coefficient (real classes), the imaginary one
// (imaginary classes) or the real part with c1 the imaginary part
// (complex); a multi sub holds its term range [z, bits of c0) in the FP32
// term list.  S is the record at s on entry and at end on exit.
template <int TB, int NT, uint32_t CLS>
__device__ __forceinline__ void stream_run(uint32_t &s, uint32_t end, SubC &S,
                                           const SubC *__restrict__ subs, const float2 *buf,
                                           uint32_t lane8, uint32_t lanebits,
                                           const TermC *__restrict__ terms,
                                           float (&gr)[(1 << TB) / NT], float (&gi)[(1 << TB) / NT])
{
    using T = float2;
    constexpr int RP = (1 << TB) / NT;
    constexpr int LB = NT == 1024 ? 10 : NT == 512 ? 9 : NT == 256 ? 8 : 7;
    for (; s < end; s++) {
        const SubC Sn = subs[s + 1];     // the list has one sentinel record
        const T *row = sub_lane_base<T, NT>(buf, lane8, S.xl);
        const uint32_t xr = S.xl >> LB;
        if constexpr (CLS != kRunMulti) {
            const uint32_t lsb = parity_sign(lanebits & S.z);
            const uint32_t zr = (S.z >> LB) & (RP - 1);
            float c = flip_sign(S.c0, lsb), ci = 0.f;
            // a lane-signed coefficient that each row flips again stays in a
            // register (one xor per row instead of the popcount per row)
            if constexpr (CLS == kRunReal1 || CLS == kRunImag1 || CLS == kRunCplx) c = opaque(c);
            if constexpr (CLS == kRunCplx) ci = opaque(flip_sign(S.c1, lsb));
            with_row_flip<RP>(xr, [&](auto X) {
                sub_rows<RP, NT, CLS, decltype(X)::value>(row, c, ci, zr, gr, gi);
            });
        } else {
            constexpr int RC = RP < 4 ? RP : 4;
            const uint32_t tb = S.z, te = __float_as_uint(S.c0);
#pragma unroll
            for (int r0 = 0; r0 < RP; r0 += RC) {
                float dr[RC], di[RC];
#pragma unroll
                for (int q = 0; q < RC; q++) { dr[q] = 0; di[q] = 0; }
                for (uint32_t t = tb; t < te; t++) {
                    const TermC C = terms[t];
                    const uint32_t lsb = parity_sign(lanebits & C.z);
                    const float cr = flip_sign(C.cr, lsb), ci = flip_sign(C.ci, lsb);
                    const uint32_t zr = (C.z >> LB) & (RP - 1);
#pragma unroll
                    for (int q = 0; q < RC; q++) {
                        const uint32_t rsb = parity_sign((r0 + q) & zr);
                        dr[q] += flip_sign(cr, rsb);
                        di[q] += flip_sign(ci, rsb);
                    }
                }
#pragma unroll
                for (int q = 0; q < RC; q++) {
                    const T v = row[((r0 + q) ^ xr) * NT];
                    gr[r0 + q] = __builtin_fmaf(dr[q], v.x, __builtin_fmaf(-di[q], v.y, gr[r0 + q]));
                    gi[r0 + q] = __builtin_fmaf(dr[q], v.y, __builtin_fmaf(di[q], v.x, gi[r0 + q]));
                }
            }
        }
        S = Sn;
    }
}

// one group of the streamed kernel: wait for its tile (buffer `cur`),
// start the next group's tile into `nxt`, apply the group's subs run by
// run.  Gc / Gn: this and the next group's headers (the header list has two
// sentinel entries); S the record of the group's first sub on entry and of
// the next group's first sub on exit.
template <int TB, int NT>
__device__ __forceinline__ void stream_phase(int g, int ngroups, const float2 *__restrict__ in,
                                             const GroupC *__restrict__ groups,
                                             const SubC *__restrict__ subs,
                                             const TermC *__restrict__ terms, uint32_t t0,
                                             uint32_t tid, const float2 *cur, float2 *nxt,
                                             SubC &S, uint32_t &s, GroupC &Gc, GroupC &Gn,
                                             float (&gr)[(1 << TB) / NT],
                                             float (&gi)[(1 << TB) / NT])
{
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (g + 1 < ngroups) tile_dma<float2, TB, NT>(in, t0 ^ (Gn.xh << TB), nxt);
    const GroupC Gnn = groups[g + 2];
    const uint32_t lane8 = tid * (uint32_t)sizeof(float2), lb = t0 | tid;
    uint32_t e = Gc.begin;
#define RP_STREAM_RUN(CLS)                                                                  \
    e += Gc.n[CLS];                                                                         \
    stream_run<TB, NT, CLS>(s, e, S, subs, cur, lane8, lb, terms, gr, gi);
    RP_STREAM_RUN(kRunReal0)
    RP_STREAM_RUN(kRunReal1)
    RP_STREAM_RUN(kRunImag0)
    RP_STREAM_RUN(kRunImag1)
    RP_STREAM_RUN(kRunCplx)
    RP_STREAM_RUN(kRunMulti)
#undef RP_STREAM_RUN
    Gc = Gn;
    Gn = Gnn;
}

// the whole output tile t0 of the streamed kernel: every group's input tile
// through the two LDS tiles, then the tile written once
template <int TB, int NT>
__device__ __forceinline__ void stream_tile(const float2 *__restrict__ in, float2 *__restrict__ out,
                                            const GroupC *__restrict__ groups, int ngroups,
                                            const SubC *__restrict__ subs,
                                            const TermC *__restrict__ terms, int accumulate,
                                            uint32_t t0, float2 *tile0, float2 *tile1)
{
    using T = float2;
    constexpr int RP = (1 << TB) / NT;
    const uint32_t tid = threadIdx.x;
    double accr[RP], acci[RP];
    float gr[RP], gi[RP];
#pragma unroll
    for (int r = 0; r < RP; r++) {
        accr[r] = 0;
        acci[r] = 0;
        gr[r] = 0;
        gi[r] = 0;
        asm volatile("" : "+v"(gr[r]), "+v"(gi[r]));
    }
    GroupC Gc = groups[0], Gn = groups[1];
    uint32_t s = 0;
    SubC S = subs[0];
    tile_dma<T, TB, NT>(in, t0 ^ (Gc.xh << TB), tile0);
    for (int g = 0; g < ngroups; g += 2) {
        stream_phase<TB, NT>(g, ngroups, in, groups, subs, terms, t0, tid, tile0, tile1, S, s, Gc,
                             Gn, gr, gi);
        if (g + 1 >= ngroups) break;
        stream_phase<TB, NT>(g + 1, ngroups, in, groups, subs, terms, t0, tid, tile1, tile0, S, s,
                             Gc, Gn, gr, gi);
        // FP32 partial sums of every four groups into the FP64 accumulators
        if ((g / 2) & 1) flush_partials(accr, acci, gr, gi);
    }
    flush_partials(accr, acci, gr, gi);
#pragma unroll
    for (int r = 0; r < RP; r++) {
        const uint64_t m = (uint64_t)t0 + r * NT + tid;
        T o;
        if (accumulate) {
            const T p = out[m];
            o.x = (float)(accr[r] + (double)p.x);
            o.y = (float)(acci[r] + (double)p.y);
        } else {
            o.x = (float)accr[r];
            o.y = (float)acci[r];
        }
        out[m] = o;
    }
}

template <int TB, int NT>
__global__ void __launch_bounds__(NT, (apply_waves_per_eu<float, TB, NT, 2>()))
k_apply_stream(const float2 *__restrict__ in, float2 *__restrict__ out,
               const GroupC *__restrict__ groups, int ngroups, const SubC *__restrict__ subs,
               const TermC *__restrict__ terms, int accumulate, TileMap tm)
{
    __shared__ float2 tile0[1 << TB];
    __shared__ float2 tile1[1 << TB];
    stream_tile<TB, NT>(in, out, groups, ngroups, subs, terms, accumulate, tile_index(tm) << TB,
                        tile0, tile1);
}

// ---------------------------------------------------------------------------
// The apply with the first Gram-Schmidt pass's coefficients formed in its
// epilogue: after writing its output tile w, the workgroup reads the same
// rows of V_0..V_{k-1} and writes its FP64 partials of <V_q|w> (columns
// 2q, 2q+1) and of ||w||^2 (column 2 kMaxSweep) to row blockIdx.x of part.
// The w values are the rounded ones just written, so the coefficients are
// those of the stored vector.
// ---------------------------------------------------------------------------
constexpr int kMaxSweep = RP_MAX_SWEEP_VECS;
constexpr int kEpiChunk = 8;                        // vectors per block reduction
struct VecPtrs64 { const void *p[kMaxSweep]; };

// sum NV per-thread doubles over the NT-thread workgroup into sh[(NT/64)*NV + q]
template <int NT, int NV>
__device__ __forceinline__ void wg_reduce(double (&v)[NV], double *sh)
{
    constexpr int W = NT / 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; q++) {
        double x = v[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        v[q] = x;
    }
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) sh[wid * NV + q] = v[q];
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0;
        for (int w = 0; w < W; w++) s += sh[w * NV + threadIdx.x];
        sh[W * NV + threadIdx.x] = s;
    }
    __syncthreads();
}

template <typename R, typename A, int TB, int NT, int D, bool GA>
__global__ void __launch_bounds__(NT, (apply_waves_per_eu<R, TB, NT, D>()))
k_apply_tiled_gs(const typename CxA<R>::T *__restrict__ in, typename CxA<R>::T *__restrict__ out,
                 const GroupRuns *__restrict__ groups, int ngroups,
                 const SubRec<A> *__restrict__ subs, const TermRec<A> *__restrict__ terms,
                 VecPtrs64 V, int k, double *__restrict__ part, int width, TileMap tm)
{
    using T = typename CxA<R>::T;
    constexpr int TILE = 1 << TB;
    constexpr int RP = TILE / NT;
    constexpr int NV = 2 * kEpiChunk + 1;
    __shared__ T tile0[TILE];
    __shared__ T tile1[TILE];
    __shared__ T tile2[D > 2 ? TILE : 1];
    __shared__ T tile3[D > 3 ? TILE : 1];
    __shared__ double sh[(NT / 64 + 1) * NV];
    T wv[RP];
    const uint32_t t0 = tile_index(tm) << TB;
    apply_tile<R, A, TB, NT, D, GA>(in, out, groups, ngroups, subs, terms, 0, tile0, tile1, tile2,
                                    tile3, t0, wv);
    const uint32_t tid = threadIdx.x;
    double wn2 = 0;
#pragma unroll
    for (int r = 0; r < RP; r++) wn2 += (double)wv[r].x * wv[r].x + (double)wv[r].y * wv[r].y;
    double *row = part + (size_t)blockIdx.x * width;
    for (int q0 = 0; q0 == 0 || q0 < k; q0 += kEpiChunk) {
        double acc[NV];
#pragma unroll
        for (int q = 0; q < NV; q++) acc[q] = 0;
#pragma unroll
        for (int q = 0; q < kEpiChunk; q++) {
            if (q0 + q < k) {
                const T *vq = (const T *)V.p[q0 + q];
#pragma unroll
                for (int r = 0; r < RP; r++) {
                    const T x = vq[(uint64_t)t0 + r * NT + tid];
                    const double xr = x.x, xi = x.y, yr = wv[r].x, yi = wv[r].y;
                    acc[2 * q] += xr * yr + xi * yi;
                    acc[2 * q + 1] += xr * yi - xi * yr;
                }
            }
        }
        if (q0 == 0) acc[2 * kEpiChunk] = wn2;
        wg_reduce<NT, NV>(acc, sh);
        constexpr int W = NT / 64;
        if ((int)tid < 2 * kEpiChunk && q0 + (int)tid / 2 < k) row[2 * q0 + tid] = sh[W * NV + tid];
        if (q0 == 0 && tid == 2 * kEpiChunk) row[2 * kMaxSweep] = sh[W * NV + tid];
    }
}

constexpr int tiling_key(int tb, int nt, int d) { return tb * 100000 + nt * 10 + d; }

// the tilings the tiled kernel is instantiated for: {tile bits, threads,
// pipeline depth, element bytes (0: both precisions)}
struct TilingDesc { int tb, nt, depth, esz; };
// Tile bits 0 is the small kernel, its depth the number of term slices.
constexpr int kSplitMaxLen = 16;
constexpr TilingDesc kTilings[] = {
    {0, kSmallThreads, 1, 0}, {0, kSmallThreads, 8, 0}, {0, kSmallThreads, 32, 0},
    {0, kSmallThreads, 128, 0},
    {8, 128, 2, 0}, {10, 128, 2, 0}, {11, 256, 4, 16}, {12, 512, 2, 0},
    {12, 512, 3, 8}, {12, 512, 4, 8}, {12, 1024, 2, 8}, {13, 1024, 2, 8}};

static bool tiling_instantiated(int prec, int tb, int nt, int d)
{
    for (const TilingDesc &t : kTilings)
        if (t.tb == tb && t.nt == nt && t.depth == d &&
            (t.esz == 0 || t.esz == (int)elem_bytes(prec)))
            return true;
    return false;
}

// default tiling (RP_TERMS_AUTOTUNE measures the alternatives).  Up to 16
// sites the small kernel, its terms split into slices when 2^len / 256
// workgroups are too few for the term count: the largest power-of-two slice
// count up to 128 that keeps at most 2048 workgroups and 32 terms per slice.
// Beyond, the tiled kernel: 8 amplitudes per thread, smaller tiles and
// workgroups for shorter chains so that there are enough workgroups.
static void choose_tiling(int prec, int len, uint64_t nterms, int group_acc, int *tb, int *nt,
                          int *depth)
{
    *depth = 2;
    if (len <= kSplitMaxLen) {
        uint64_t blocks = std::max<uint64_t>(1, (1ull << len) / kSmallThreads);
        int sl = 1;
        while (sl < 128 && blocks * (uint64_t)(2 * sl) <= 2048 && nterms >= (uint64_t)(64 * sl))
            sl *= 2;
        if (sl > 1 || len < 10) {
            *tb = 0;
            *nt = kSmallThreads;
            *depth = sl;
            return;
        }
    }
    // 2^12-amplitude tiles with 512 threads: for complex<float> sparse SYK at
    // 24 qubits faster than 2^13 tiles (more groups, but two workgroups per
    // CU instead of one); for complex<double> the largest tile in 128 KiB.
    // For complex<float> with group accumulation from 28 sites (states of
    // 2 GiB and more, beyond every cache) 2^13 tiles with 1024 threads:
    // sparse SYK at 30 qubits takes 197 ms there against 260 ms with 2^12
    // tiles
    if (prec == RP_C64 && group_acc && len >= 28) { *tb = 13; *nt = 1024; return; }
    if (len >= 18) { *tb = 12; *nt = 512; return; }
    if (len >= 14) { *tb = 10; *nt = 128; return; }
    *tb = 8;
    *nt = 128;
}

// the compact lists of the tiled tables (complex<float> states, FP64 term
// coefficients): one 16-byte record per sub in the SubRec order,
// coefficients rounded to float, one sentinel record; one 16-byte header
// per group, two sentinel headers; the multi-sub terms in FP32.  False when
// a class run of some group has more than 255 subs (the header's count
// width): the handle then keeps the SubRec kernel.
static bool compact_stream(const Tables<double> &tab, std::vector<GroupC> &cg,
                           std::vector<SubC> &cs, std::vector<TermC> &ct)
{
    for (const TermRec<double> &t : tab.terms) ct.push_back({t.z, (float)t.cr, (float)t.ci, 0});
    for (const GroupRuns &G : tab.groups) {
        GroupC h{};
        h.xh = G.xh;
        h.begin = G.b[0];
        for (uint32_t k = 0; k < kRuns; k++) {
            if (G.b[k + 1] - G.b[k] > 255) return false;
            h.n[k] = (uint8_t)(G.b[k + 1] - G.b[k]);
            for (uint32_t i = G.b[k]; i < G.b[k + 1]; i++) {
                const SubRec<double> &S = tab.subs[i];
                SubC c{S.xl, S.z, 0.f, 0.f};
                if (k == kRunMulti) {
                    c.z = S.term_begin;
                    memcpy(&c.c0, &S.term_end, sizeof(float));
                } else if (k == kRunImag0 || k == kRunImag1) {
                    c.c0 = S.fi;
                } else {
                    c.c0 = S.fr;
                    c.c1 = S.fi;
                }
                cs.push_back(c);
            }
        }
        cg.push_back(h);
    }
    GroupC end{};
    end.begin = (uint32_t)cs.size();
    cg.push_back(end);
    cg.push_back(end);
    cs.push_back({0, 0, 0.f, 0.f});
    if (ct.empty()) ct.push_back({0, 0.f, 0.f, 0});
    return true;
}

// log2 of the tiled kernel's resident workgroups per XCD (32 CUs) at the
// occupancy apply_waves_per_eu gives it: the p of its output-tile map
static int xcd_resident_bits(int prec, int tb, int nt, int depth)
{
    if (!tb) return 0;
    const int lds = depth * (1 << tb) * (int)elem_bytes(prec);
    int wg = std::min((160 * 1024) / lds, (nt == 1024 ? 2048 : 1024) / nt);
    wg = std::max(wg, 1);
    int p = 5;
    while ((2 << p) <= 32 * wg) p++;
    return p;
}

template <typename R, typename A>
static int upload_impl(rp_terms *h, const uint64_t *bits, const double *coeff, uint64_t n)
{
    Tables<A> tab;
    int st = build_tables<R, A>(h->len, h->tile_bits, h->threads, bits, coeff, n, tab,
                                xcd_resident_bits(h->prec, h->tile_bits, h->threads, h->depth));
    if (st) return st;
    h->tmap = tab.tilemap;
    auto &groups = tab.groups;
    auto &subs = tab.subs;
    auto &terms = tab.terms;
    auto &flat = tab.flat;
    const int tb = h->tile_bits;
    const uint64_t nflips = tab.nflips;
    h->nterms = tab.nterms;
    h->nflips = nflips;
    h->ngroups = tb ? groups.size() : 1;
    h->nsubs = subs.size();
    Context &c = ctx();
    auto up = [&](void **dst, const void *src, size_t bytes) -> int {
        if (!bytes) return RP_OK;
        hipError_t e = dev_malloc((void **)dst, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return set_error(RP_ERR_OUT_OF_MEMORY, "Failed to allocate %zu bytes on GPU device.",
                             bytes);
        }
        RP_CHECK_HIP(hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, c.stream));
        return RP_OK;
    };
    if ((st = up(&h->d_flat, flat.data(), flat.size() * sizeof(FlatTerm<A>)))) return st;
    if (tb == 0 && h->depth > 1) {
        size_t need = (size_t)h->depth * (1ull << h->len) * sizeof(typename CxA<A>::T);
        if (h->split_bytes < need) {
            if (h->d_split) (void)hipFree(h->d_split);
            h->d_split = nullptr;
            h->split_bytes = 0;
            if (dev_malloc((void **)&h->d_split, need) != hipSuccess) {
                (void)hipGetLastError();
                h->d_split = nullptr;
                return set_error(RP_ERR_OUT_OF_MEMORY, "Failed to allocate %zu bytes on GPU device.",
                                 need);
            }
            h->split_bytes = need;
        }
    }
    if (tb) {
        if ((st = up(&h->d_groups, groups.data(), groups.size() * sizeof(GroupRuns)))) return st;
        if ((st = up(&h->d_subs, subs.data(), subs.size() * sizeof(SubRec<A>)))) return st;
        if ((st = up(&h->d_terms, terms.data(), terms.size() * sizeof(TermRec<A>)))) return st;
    }
    if constexpr (sizeof(R) == 4 && sizeof(A) == 8) {
        std::vector<GroupC> cg;
        std::vector<SubC> cs;
        std::vector<TermC> ct;
        if (tb && compact_stream(tab, cg, cs, ct)) {
            if ((st = up(&h->d_cgroups, cg.data(), cg.size() * sizeof(GroupC)))) return st;
            if ((st = up(&h->d_csubs, cs.data(), cs.size() * sizeof(SubC)))) return st;
            if ((st = up(&h->d_cterms, ct.data(), ct.size() * sizeof(TermC)))) return st;
        }
    }
    RP_CHECK_HIP(hipStreamSynchronize(c.stream));
    return RP_OK;
}

static void free_handle(rp_terms *h)
{
    if (!h) return;
    if (h->d_groups) (void)hipFree(h->d_groups);
    if (h->d_subs) (void)hipFree(h->d_subs);
    if (h->d_terms) (void)hipFree(h->d_terms);
    if (h->d_flat) (void)hipFree(h->d_flat);
    if (h->d_split) (void)hipFree(h->d_split);
    for (void *p : {h->d_cgroups, h->d_csubs, h->d_cterms})
        if (p) (void)hipFree(p);
    delete h;
}

template <typename R, typename A, int TB, int NT, int D = 2>
static int launch_tiled(rp_terms *h, void *res, const void *vec, int accumulate)
{
    using T = typename CxA<R>::T;
    const uint64_t grid = (1ull << h->len) >> TB;
    if constexpr (sizeof(R) == 4 && sizeof(A) == 8 && D == 2) {
        if (h->group_acc && h->d_csubs) {
            RP_LAUNCH(RP_KERNEL_APPLY, (k_apply_stream<TB, NT>), dim3((unsigned)grid), dim3(NT), 0,
                      (const T *)vec, (T *)res, (const GroupC *)h->d_cgroups, (int)h->ngroups,
                      (const SubC *)h->d_csubs, (const TermC *)h->d_cterms, accumulate, h->tmap);
            return RP_OK;
        }
        if (h->group_acc) {
            RP_LAUNCH(RP_KERNEL_APPLY, (k_apply_tiled<R, A, TB, NT, D, true>), dim3((unsigned)grid),
                      dim3(NT), 0, (const T *)vec, (T *)res, (const GroupRuns *)h->d_groups,
                      (int)h->ngroups, (const SubRec<A> *)h->d_subs,
                      (const TermRec<A> *)h->d_terms, accumulate, h->tmap);
            return RP_OK;
        }
    }
    RP_LAUNCH(RP_KERNEL_APPLY, (k_apply_tiled<R, A, TB, NT, D, false>), dim3((unsigned)grid),
              dim3(NT), 0, (const T *)vec, (T *)res, (const GroupRuns *)h->d_groups,
              (int)h->ngroups, (const SubRec<A> *)h->d_subs, (const TermRec<A> *)h->d_terms,
              accumulate, h->tmap);
    return RP_OK;
}

template <typename R, typename A>
static int apply_impl(rp_terms *h, void *res, const void *vec, int accumulate)
{
    using T = typename CxA<R>::T;
    const uint64_t dim = 1ull << h->len;
    if (h->nterms == 0) {
        if (!accumulate) RP_CHECK_HIP(hipMemsetAsync(res, 0, dim * sizeof(T), ctx().stream));
        return RP_OK;
    }
    if (h->tile_bits == 0) {
        int grid = (int)((dim + kSmallThreads - 1) / kSmallThreads);
        if (h->depth > 1) {
            using P = typename CxA<A>::T;
            const int slices = h->depth;
            const int per = (int)((h->nterms + slices - 1) / slices);
            RP_LAUNCH(RP_KERNEL_APPLY, (k_apply_split<R, A>), dim3(grid, slices),
                      dim3(kSmallThreads), 0, (const T *)vec, (P *)h->d_split, (uint32_t)dim,
                      (const FlatTerm<A> *)h->d_flat, (int)h->nterms, per);
            RP_LAUNCH(RP_KERNEL_APPLY, (k_split_reduce<R, A>), dim3(grid), dim3(kSmallThreads), 0,
                      (const P *)h->d_split, (T *)res, (uint32_t)dim, slices, accumulate);
            return RP_OK;
        }
        RP_LAUNCH(RP_KERNEL_APPLY, (k_apply_small<R, A>), dim3(grid), dim3(kSmallThreads), 0,
                  (const T *)vec, (T *)res, (uint32_t)dim, (const FlatTerm<A> *)h->d_flat,
                  (int)h->nterms, accumulate);
        return RP_OK;
    }
    switch (tiling_key(h->tile_bits, h->threads, h->depth)) {
    case tiling_key(8, 128, 2): return launch_tiled<R, A, 8, 128>(h, res, vec, accumulate);
    case tiling_key(10, 128, 2): return launch_tiled<R, A, 10, 128>(h, res, vec, accumulate);
    case tiling_key(12, 512, 2): return launch_tiled<R, A, 12, 512>(h, res, vec, accumulate);
    case tiling_key(11, 256, 4):
        if constexpr (sizeof(R) == 8) return launch_tiled<R, A, 11, 256, 4>(h, res, vec, accumulate);
        break;
    case tiling_key(12, 512, 3):
        if constexpr (sizeof(R) == 4) return launch_tiled<R, A, 12, 512, 3>(h, res, vec, accumulate);
        break;
    case tiling_key(12, 512, 4):
        if constexpr (sizeof(R) == 4) return launch_tiled<R, A, 12, 512, 4>(h, res, vec, accumulate);
        break;
    case tiling_key(12, 1024, 2):
        if constexpr (sizeof(R) == 4) return launch_tiled<R, A, 12, 1024>(h, res, vec, accumulate);
        break;
    case tiling_key(13, 1024, 2):
        if constexpr (sizeof(R) == 4) return launch_tiled<R, A, 13, 1024>(h, res, vec, accumulate);
        break;
    default: break;
    }
    return set_error(RP_ERR_RUNTIME, "rp_apply_ops: unsupported tiling b=%d threads=%d depth=%d",
                     h->tile_bits, h->threads, h->depth);
}

// rp_apply_gs's kernel: the tiled apply with the Gram-Schmidt dots in its
// epilogue, for the tilings it is instantiated for (complex<float> states,
// 2^12 tiles, 512 threads, 2 LDS tiles, FP64 or group accumulation) and at
// most max_rows workgroups; *rows = the partial rows written, 0 when the
// handle's layout has no such kernel (the caller then applies and sweeps
// separately)
int apply_dots_launch(rp_terms *h, void *res, const void *vec, const void *const *V, int k,
                      double *part, int width, uint64_t max_rows, uint64_t *rows)
{
    using T = float2;
    *rows = 0;
    if (h->prec != RP_C64 || !h->acc64 || h->nterms == 0 || h->tile_bits != 12 ||
        h->threads != 512 || h->depth != 2 || k < 0 || k > kMaxSweep)
        return RP_OK;
    const uint64_t grid = (1ull << h->len) >> 12;
    if (grid > max_rows) return RP_OK;
    VecPtrs64 P{};
    for (int q = 0; q < k; q++) P.p[q] = V[q];
    if (h->group_acc)
        RP_LAUNCH(RP_KERNEL_APPLY, (k_apply_tiled_gs<float, double, 12, 512, 2, true>),
                  dim3((unsigned)grid), dim3(512), 0, (const T *)vec, (T *)res,
                  (const GroupRuns *)h->d_groups, (int)h->ngroups, (const SubRec<double> *)h->d_subs,
                  (const TermRec<double> *)h->d_terms, P, k, part, width, h->tmap);
    else
        RP_LAUNCH(RP_KERNEL_APPLY, (k_apply_tiled_gs<float, double, 12, 512, 2, false>),
                  dim3((unsigned)grid), dim3(512), 0, (const T *)vec, (T *)res,
                  (const GroupRuns *)h->d_groups, (int)h->ngroups, (const SubRec<double> *)h->d_subs,
                  (const TermRec<double> *)h->d_terms, P, k, part, width, h->tmap);
    *rows = grid;
    return RP_OK;
}

}  // namespace rp

using namespace rp;

static int upload_any(rp_terms *h, const uint64_t *bits, const double *coeff, uint64_t n)
{
    if (h->prec == RP_C128) return upload_impl<double, double>(h, bits, coeff, n);
    if (h->acc64) return upload_impl<float, double>(h, bits, coeff, n);
    return upload_impl<float, float>(h, bits, coeff, n);
}

static void free_tables(rp_terms *h)
{
    for (void **p : {&h->d_groups, &h->d_subs, &h->d_terms, &h->d_flat, &h->d_cgroups, &h->d_csubs,
                     &h->d_cterms})
        if (*p) { (void)hipFree(*p); *p = nullptr; }
}

static int apply_dispatch(rp_terms *h, void *res, const void *vec, int accumulate)
{
    if (h->prec == RP_C128) return apply_impl<double, double>(h, res, vec, accumulate);
    if (h->acc64) return apply_impl<float, double>(h, res, vec, accumulate);
    return apply_impl<float, float>(h, res, vec, accumulate);
}

// rebuild the device tables for another tiling (from the stored term list)
static int retile(rp_terms *h, int tb, int nt, int d)
{
    free_tables(h);
    h->tile_bits = tb;
    h->threads = nt;
    h->depth = d;
    return upload_any(h, h->raw_bits.data(), h->raw_coeff.data(), h->raw_bits.size());
}

// Time every instantiated tiling that fits len on scratch vectors (3
// applies each after a warm-up) and keep the fastest.
static int autotune(rp_terms *h)
{
    Context &c = ctx();
    const size_t bytes = (size_t)elem_bytes(h->prec) << h->len;
    void *x = nullptr, *y = nullptr;
    if (dev_malloc(&x, bytes) != hipSuccess || dev_malloc(&y, bytes) != hipSuccess) {
        (void)hipGetLastError();
        if (x) (void)hipFree(x);
        return set_error(RP_ERR_OUT_OF_MEMORY, "rp_terms_upload_ex: no memory for autotuning");
    }
    int st = RP_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float best = 1e30f;
    TilingDesc pick{h->tile_bits, h->threads, h->depth, 0};
    if (hipMemsetAsync(x, 0, bytes, c.stream) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        st = set_error(RP_ERR_RUNTIME, "rp_terms_upload_ex: autotuning set-up failed");
    for (const TilingDesc &cd : kTilings) {
        if (st) break;
        if (cd.tb > h->len || !tiling_instantiated(h->prec, cd.tb, cd.nt, cd.depth)) continue;
        if (cd.tb == 0 && (h->len > kSplitMaxLen || h->nterms < (uint64_t)(2 * cd.depth))) continue;
        if ((st = retile(h, cd.tb, cd.nt, cd.depth))) break;
        if ((st = apply_dispatch(h, y, x, 0))) break;
        if (hipEventRecord(e0, c.stream) != hipSuccess) { st = RP_ERR_RUNTIME; break; }
        for (int r = 0; r < 3 && !st; r++) st = apply_dispatch(h, y, x, 0);
        float ms = 0;
        if (st || hipEventRecord(e1, c.stream) != hipSuccess ||
            hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            if (!st) st = set_error(RP_ERR_RUNTIME, "rp_terms_upload_ex: autotuning timing failed");
            break;
        }
        if (ms < best) { best = ms; pick = cd; }
    }
    // the kept tiling's apply against one streaming pass of the same vector
    // (3 scalings of x, each reading and writing it: 6 passes)
    // (3 scalings of x, each reading and writing it: 6 passes); a failure
    // here leaves the figure at 0 and the upload as it is
    if (!st && best < 1e30f) {
        float ms = 0;
        int sst = hipEventRecord(e0, c.stream) == hipSuccess ? RP_OK : RP_ERR_RUNTIME;
        for (int r = 0; r < 3 && !sst; r++)
            sst = rp_scale_vec(h->prec, x, 1.0, 0.0, 1ull << h->len);
        if (!sst && hipEventRecord(e1, c.stream) == hipSuccess &&
            hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess &&
            ms > 0)
            h->apply_passes = std::max(1.0, (double)best / 3.0 / ((double)ms / 6.0));
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(c.stream);
    (void)hipFree(x);
    (void)hipFree(y);
    if (st) return st;
    if (pick.tb != h->tile_bits || pick.nt != h->threads || pick.depth != h->depth)
        st = retile(h, pick.tb, pick.nt, pick.depth);
    return st;
}

extern "C" {

int rp_terms_upload_ex(rp_terms_t *out, int prec, int len, const uint64_t *bits,
                       const double *coeff, uint64_t nterms, int flags)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(out, "null output handle");
    *out = nullptr;
    RP_REQUIRE(valid_prec(prec), "bad precision");
    RP_REQUIRE(len >= 1 && len <= 32, "len must be in [1, 32]");
    RP_REQUIRE(nterms == 0 || (bits && coeff), "null term arrays");
    RP_REQUIRE((flags & ~(RP_TERMS_ACC_NATIVE | RP_TERMS_AUTOTUNE | RP_TERMS_ACC_GROUP)) == 0,
               "unknown flags");
    RP_REQUIRE(!((flags & RP_TERMS_ACC_NATIVE) && (flags & RP_TERMS_ACC_GROUP)),
               "RP_TERMS_ACC_NATIVE and RP_TERMS_ACC_GROUP exclude each other");
    rp_terms *h = new rp_terms;
    h->prec = prec;
    h->len = len;
    h->acc64 = (prec == RP_C128 || !(flags & RP_TERMS_ACC_NATIVE)) ? 1 : 0;
    h->group_acc = (prec == RP_C64 && (flags & RP_TERMS_ACC_GROUP)) ? 1 : 0;
    choose_tiling(prec, len, nterms, h->group_acc, &h->tile_bits, &h->threads, &h->depth);
    h->raw_bits.assign(bits, bits + nterms);
    h->raw_coeff.assign(coeff, coeff + 2 * nterms);
    st = upload_any(h, bits, coeff, nterms);
    if (!st && (flags & RP_TERMS_AUTOTUNE) && h->nterms && len <= 31)
        st = autotune(h);
    if (st) {
        free_handle(h);
        return st;
    }
    *out = h;
    return RP_OK;
}

int rp_terms_upload(rp_terms_t *out, int prec, int len, const uint64_t *bits,
                    const double *coeff, uint64_t nterms)
{
    return rp_terms_upload_ex(out, prec, len, bits, coeff, nterms, 0);
}

int rp_terms_free(rp_terms_t h)
{
    int st = check_init();
    if (st) return st;
    if (!h) return RP_OK;
    RP_CHECK_HIP(hipStreamSynchronize(ctx().stream));
    free_handle(h);
    return RP_OK;
}

int rp_terms_set_tiling(rp_terms_t h, int tile_bits, int threads, int depth)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(h, "null term handle");
    RP_REQUIRE(tile_bits <= h->len, "tile bits exceed the chain length");
    // the one-thread-per-amplitude kernels index 32-bit amplitudes and the
    // sliced one keeps depth * 2^len partial sums
    RP_REQUIRE(tile_bits != 0 || h->len <= kSplitMaxLen,
               "tile bits 0 (one thread per amplitude) only up to len 16");
    RP_REQUIRE(tiling_instantiated(h->prec, tile_bits, threads, depth),
               "tiling not instantiated for this precision");
    RP_CHECK_HIP(hipStreamSynchronize(ctx().stream));
    return retile(h, tile_bits, threads, depth);
}

int rp_terms_info(rp_terms_t h, rp_terms_info_t *info)
{
    RP_REQUIRE(h && info, "null argument");
    info->prec = h->prec;
    info->len = h->len;
    info->tile_bits = h->tile_bits;
    info->threads = h->threads;
    info->depth = h->depth;
    info->nterms = h->nterms;
    info->nflips = h->nflips;
    info->ngroups = h->tile_bits ? h->ngroups : h->nflips;
    info->apply_passes = h->apply_passes;
    info->kernel = h->tile_bits == 0 ? RP_APPLY_SMALL
                   : (h->prec == RP_C64 && h->acc64 && h->group_acc && h->depth == 2 && h->d_csubs)
                       ? RP_APPLY_STREAM
                       : RP_APPLY_TILED;
    info->bytes_per_apply = (info->ngroups + 1) * (1ull << h->len) * elem_bytes(h->prec);
    info->tile_map = h->tmap.n;
    return RP_OK;
}

int rp_apply_ops(rp_terms_t h, void *res, const void *vec, int accumulate)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(h, "null term handle");
    RP_REQUIRE(res && vec, "null vector");
    RP_REQUIRE(res != vec, "res and vec must be different buffers");
    return apply_dispatch(h, res, vec, accumulate);
}

}  // extern "C"
