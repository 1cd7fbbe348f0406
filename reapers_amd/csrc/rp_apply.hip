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
// and sorts the terms by x: terms with equal x share one gather of vec, and
// terms whose x agree above the tile bits b share one input tile.
//
// Tiled kernel: one workgroup owns the output tile [T, T + 2^b).  For each
// distinct high part xh = x >> b it stages the input tile T ^ (xh << b) in
// LDS with coalesced loads (an XOR with xh maps an aligned tile onto an
// aligned tile), then applies every term of that group from LDS: the low
// bits of x permute within the tile.  The output accumulates in registers and
// is written once, with the caller's zero_vec fused in when accumulate == 0.
// HBM traffic per apply: (G_b + 1) * 2^len * S bytes, G_b = number of
// distinct xh.
#include "rp_internal.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace rp {

template <typename R> struct CxA;
template <> struct CxA<float> { using T = float2; };
template <> struct CxA<double> { using T = double2; };

struct GroupRec { uint32_t xh, sub_begin, sub_end, pad; };
struct SubRec { uint32_t xl, term_begin, term_end, pad; };
template <typename R> struct TermRec { uint32_t z, pad; R cr, ci; };
template <typename R> struct FlatTerm { uint32_t x, z; R cr, ci; };

constexpr int kApplyThreads = 256;

}  // namespace rp

struct rp_terms {
    int prec = RP_C64;
    int len = 0;
    int tile_bits = 0;       // 0: small kernel (one thread per amplitude)
    uint64_t nterms = 0, nflips = 0, ngroups = 0, nsubs = 0;
    void *d_groups = nullptr, *d_subs = nullptr, *d_terms = nullptr, *d_flat = nullptr;
};

namespace rp {

__device__ inline int parity32(uint32_t v) { return __popc(v) & 1; }

// one thread per amplitude; terms in (x-sorted) flat order
template <typename R>
__global__ void __launch_bounds__(kApplyThreads)
k_apply_small(const typename CxA<R>::T *__restrict__ in, typename CxA<R>::T *__restrict__ out,
              uint32_t dim, const FlatTerm<R> *__restrict__ terms, int nterms, int accumulate)
{
    uint32_t m = blockIdx.x * kApplyThreads + threadIdx.x;
    if (m >= dim) return;
    R ar = 0, ai = 0;
    for (int t = 0; t < nterms; t++) {
        FlatTerm<R> f = terms[t];
        auto v = in[m ^ f.x];
        R cr = parity32(m & f.z) ? -f.cr : f.cr;
        R ci = parity32(m & f.z) ? -f.ci : f.ci;
        ar += cr * v.x - ci * v.y;
        ai += cr * v.y + ci * v.x;
    }
    if (accumulate) {
        auto o = out[m];
        ar += o.x;
        ai += o.y;
    }
    typename CxA<R>::T o;
    o.x = ar;
    o.y = ai;
    out[m] = o;
}

template <typename R, int TB>
__global__ void __launch_bounds__(kApplyThreads)
k_apply_tiled(const typename CxA<R>::T *__restrict__ in, typename CxA<R>::T *__restrict__ out,
              const GroupRec *__restrict__ groups, int ngroups, const SubRec *__restrict__ subs,
              const TermRec<R> *__restrict__ terms, int accumulate)
{
    using T = typename CxA<R>::T;
    constexpr int NT = kApplyThreads;
    constexpr int TILE = 1 << TB;
    constexpr int RP = TILE / NT;          // amplitudes per thread
    __shared__ T tile[TILE];

    const uint32_t t0 = (uint32_t)blockIdx.x << TB;      // output tile base
    const uint32_t tid = threadIdx.x;
    R accr[RP], acci[RP];
#pragma unroll
    for (int r = 0; r < RP; r++) { accr[r] = 0; acci[r] = 0; }

    for (int g = 0; g < ngroups; g++) {
        const GroupRec G = groups[g];
        const uint32_t src = t0 ^ (G.xh << TB);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RP; r++) tile[r * NT + tid] = in[(uint64_t)src + r * NT + tid];
        __syncthreads();
        for (uint32_t s = G.sub_begin; s < G.sub_end; s++) {
            const SubRec S = subs[s];
            const uint32_t lane_src = tid ^ (S.xl & (NT - 1));
            const uint32_t xr = S.xl / NT;
            R dr[RP], di[RP];
#pragma unroll
            for (int r = 0; r < RP; r++) { dr[r] = 0; di[r] = 0; }
            for (uint32_t t = S.term_begin; t < S.term_end; t++) {
                const TermRec<R> C = terms[t];
                const uint32_t hp = parity32(t0 & C.z);
#pragma unroll
                for (int r = 0; r < RP; r++) {
                    const uint32_t p = parity32(((uint32_t)r * NT + tid) & C.z) ^ hp;
                    dr[r] += p ? -C.cr : C.cr;
                    di[r] += p ? -C.ci : C.ci;
                }
            }
#pragma unroll
            for (int r = 0; r < RP; r++) {
                const T v = tile[(r ^ xr) * NT + lane_src];
                accr[r] += dr[r] * v.x - di[r] * v.y;
                acci[r] += dr[r] * v.y + di[r] * v.x;
            }
        }
    }
#pragma unroll
    for (int r = 0; r < RP; r++) {
        const uint64_t m = (uint64_t)t0 + r * NT + tid;
        T o;
        if (accumulate) {
            T p = out[m];
            o.x = p.x + accr[r];
            o.y = p.y + acci[r];
        } else {
            o.x = accr[r];
            o.y = acci[r];
        }
        out[m] = o;
    }
}

static inline uint32_t pext_even(uint64_t b)   // bit 2s -> bit s
{
    uint32_t r = 0;
    for (int s = 0; s < 32; s++) r |= (uint32_t)((b >> (2 * s)) & 1u) << s;
    return r;
}

static inline uint32_t pext_odd(uint64_t b)    // bit 2s+1 -> bit s
{
    uint32_t r = 0;
    for (int s = 0; s < 32; s++) r |= (uint32_t)((b >> (2 * s + 1)) & 1u) << s;
    return r;
}

static int choose_tile_bits(int len)
{
    if (len >= 18) return 12;
    if (len >= 14) return 10;
    if (len >= 8) return 8;
    return 0;
}

template <typename R>
static int upload_impl(rp_terms *h, const uint64_t *bits, const double *coeff, uint64_t n)
{
    struct Tmp { uint32_t x, z; R cr, ci; uint64_t order; };
    std::vector<Tmp> v;
    v.reserve(n);
    const uint64_t sitemask = h->len == 32 ? ~0ull : ((1ull << (2 * h->len)) - 1);
    for (uint64_t t = 0; t < n; t++) {
        R cr = (R)coeff[2 * t], ci = (R)coeff[2 * t + 1];
        if (cr == 0 && ci == 0) continue;
        if (bits[t] & ~sitemask)
            return set_error(RP_ERR_INVALID_ARGUMENT,
                             "rp_terms_upload: term %llu acts on a site >= len=%d",
                             (unsigned long long)t, h->len);
        Tmp e;
        e.x = pext_even(bits[t]);
        e.z = pext_odd(bits[t]);
        if (__builtin_popcount(e.x & e.z) & 1) { cr = -cr; ci = -ci; }
        e.cr = cr;
        e.ci = ci;
        e.order = t;
        v.push_back(e);
    }
    const int tb = h->tile_bits;
    const uint32_t lowmask = tb ? ((1u << tb) - 1) : 0xffffffffu;
    std::stable_sort(v.begin(), v.end(), [&](const Tmp &a, const Tmp &b) {
        uint32_t ah = tb ? a.x >> tb : 0, bh = tb ? b.x >> tb : 0;
        if (ah != bh) return ah < bh;
        return (a.x & lowmask) < (b.x & lowmask);
    });
    h->nterms = v.size();
    std::vector<GroupRec> groups;
    std::vector<SubRec> subs;
    std::vector<TermRec<R>> terms;
    std::vector<FlatTerm<R>> flat;
    uint64_t nflips = 0;
    for (size_t i = 0; i < v.size(); i++) {
        if (i == 0 || v[i].x != v[i - 1].x) nflips++;
        flat.push_back({v[i].x, v[i].z, v[i].cr, v[i].ci});
        if (!tb) continue;
        uint32_t xh = v[i].x >> tb, xl = v[i].x & lowmask;
        if (groups.empty() || groups.back().xh != xh) {
            groups.push_back({xh, (uint32_t)subs.size(), (uint32_t)subs.size(), 0});
        }
        if (subs.size() == groups.back().sub_begin || groups.back().sub_end == groups.back().sub_begin ||
            subs.back().xl != xl) {
            subs.push_back({xl, (uint32_t)terms.size(), (uint32_t)terms.size(), 0});
            groups.back().sub_end = (uint32_t)subs.size();
        }
        terms.push_back({v[i].z, 0, v[i].cr, v[i].ci});
        subs.back().term_end = (uint32_t)terms.size();
    }
    h->nflips = nflips;
    h->ngroups = tb ? groups.size() : 1;
    h->nsubs = subs.size();
    Context &c = ctx();
    auto up = [&](void **dst, const void *src, size_t bytes) -> int {
        if (!bytes) return RP_OK;
        hipError_t e = hipMalloc(dst, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return set_error(RP_ERR_OUT_OF_MEMORY, "Failed to allocate %zu bytes on GPU device.",
                             bytes);
        }
        RP_CHECK_HIP(hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, c.stream));
        return RP_OK;
    };
    int st;
    if ((st = up(&h->d_flat, flat.data(), flat.size() * sizeof(FlatTerm<R>)))) return st;
    if (tb) {
        if ((st = up(&h->d_groups, groups.data(), groups.size() * sizeof(GroupRec)))) return st;
        if ((st = up(&h->d_subs, subs.data(), subs.size() * sizeof(SubRec)))) return st;
        if ((st = up(&h->d_terms, terms.data(), terms.size() * sizeof(TermRec<R>)))) return st;
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
    delete h;
}

template <typename R>
static int apply_impl(rp_terms *h, void *res, const void *vec, int accumulate)
{
    using T = typename CxA<R>::T;
    const uint64_t dim = 1ull << h->len;
    if (h->nterms == 0) {
        if (!accumulate)
            RP_CHECK_HIP(hipMemsetAsync(res, 0, dim * sizeof(T), ctx().stream));
        return RP_OK;
    }
    if (h->tile_bits == 0) {
        int grid = (int)((dim + kApplyThreads - 1) / kApplyThreads);
        RP_LAUNCH(RP_KERNEL_APPLY, k_apply_small<R>, dim3(grid), dim3(kApplyThreads), 0,
                  (const T *)vec, (T *)res, (uint32_t)dim, (const FlatTerm<R> *)h->d_flat,
                  (int)h->nterms, accumulate);
        return RP_OK;
    }
    const uint64_t grid = dim >> h->tile_bits;
    const GroupRec *g = (const GroupRec *)h->d_groups;
    const SubRec *s = (const SubRec *)h->d_subs;
    const TermRec<R> *t = (const TermRec<R> *)h->d_terms;
    const int ng = (int)h->ngroups;
    switch (h->tile_bits) {
    case 8:
        RP_LAUNCH(RP_KERNEL_APPLY, (k_apply_tiled<R, 8>), dim3((unsigned)grid), dim3(kApplyThreads),
                  0, (const T *)vec, (T *)res, g, ng, s, t, accumulate);
        break;
    case 10:
        RP_LAUNCH(RP_KERNEL_APPLY, (k_apply_tiled<R, 10>), dim3((unsigned)grid), dim3(kApplyThreads),
                  0, (const T *)vec, (T *)res, g, ng, s, t, accumulate);
        break;
    case 12:
        RP_LAUNCH(RP_KERNEL_APPLY, (k_apply_tiled<R, 12>), dim3((unsigned)grid), dim3(kApplyThreads),
                  0, (const T *)vec, (T *)res, g, ng, s, t, accumulate);
        break;
    default:
        return set_error(RP_ERR_RUNTIME, "rp_apply_ops: unsupported tile size %d", h->tile_bits);
    }
    return RP_OK;
}

}  // namespace rp

using namespace rp;

extern "C" {

int rp_terms_upload(rp_terms_t *out, int prec, int len, const uint64_t *bits,
                    const double *coeff, uint64_t nterms)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(out, "null output handle");
    *out = nullptr;
    RP_REQUIRE(valid_prec(prec), "bad precision");
    RP_REQUIRE(len >= 1 && len <= 32, "len must be in [1, 32]");
    RP_REQUIRE(nterms == 0 || (bits && coeff), "null term arrays");
    rp_terms *h = new rp_terms;
    h->prec = prec;
    h->len = len;
    h->tile_bits = choose_tile_bits(len);
    st = prec == RP_C64 ? upload_impl<float>(h, bits, coeff, nterms)
                        : upload_impl<double>(h, bits, coeff, nterms);
    if (st) {
        free_handle(h);
        return st;
    }
    *out = h;
    return RP_OK;
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

int rp_terms_info(rp_terms_t h, rp_terms_info_t *info)
{
    RP_REQUIRE(h && info, "null argument");
    info->prec = h->prec;
    info->len = h->len;
    info->tile_bits = h->tile_bits;
    info->nterms = h->nterms;
    info->nflips = h->nflips;
    info->ngroups = h->tile_bits ? h->ngroups : h->nflips;
    info->bytes_per_apply = (info->ngroups + 1) * (1ull << h->len) * elem_bytes(h->prec);
    return RP_OK;
}

int rp_apply_ops(rp_terms_t h, void *res, const void *vec, int accumulate)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(h, "null term handle");
    RP_REQUIRE(res && vec, "null vector");
    RP_REQUIRE(res != vec, "res and vec must be different buffers");
    return h->prec == RP_C64 ? apply_impl<float>(h, res, vec, accumulate)
                             : apply_impl<double>(h, res, vec, accumulate);
}

}  // extern "C"
