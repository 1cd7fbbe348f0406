// rp_tables.h -- device layout of a Pauli-sum term list for the tiled apply
// kernels (single state and batched), built on the host.
//
// Per term: x = flip mask (X bits, even positions of SpinOp::bits), z = phase
// mask (Z bits, odd positions), c = coeff * (-1)^popcount(x & z), so that
//     res[m] += c * (-1)^popcount(m & z) * vec[m ^ x]
// (the reference's SpinOp::apply, inc/ops.h:421-440).  Terms are sorted by x
// (by the classes of the output-tile map first, when there is one); terms
// with equal x form a sub, subs whose x agree above the tile bits b form a
// group; a group stores its subs by class (real / imaginary c, each split by
// whether z has bits among the row bits, complex, several terms).
#pragma once

#include "rp_internal.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace rp {

template <typename R> struct CxA;
template <> struct CxA<float> { using T = float2; };
template <> struct CxA<double> { using T = double2; };

// kinds of subs: one term with real / imaginary / complex c, or several terms
enum : uint32_t { kSubReal = 0, kSubImag = 1, kSubComplex = 2, kSubMulti = 3 };

template <typename A> struct SubRec {
    uint32_t xl, z, kind, term_begin, term_end, pad0;
    float fr, fi;            // cr, ci rounded to float (for FP32 partial sums)
    A cr, ci;
};
template <typename A> struct TermRec { uint32_t z, pad; A cr, ci; };
template <typename A> struct FlatTerm { uint32_t x, z; A cr, ci; };

__device__ inline uint32_t parity32(uint32_t v) { return __popc(v) & 1u; }

// Output-tile order of the single-state tiled apply.  The 2^p workgroups
// resident on one XCD work through the group list roughly in step, so for a
// group with high flip part xh they stage the input tiles t ^ xh of their
// output tiles t.  When those t form a coset of a subspace S (over GF(2)) of
// tile indices, the staged tiles form the coset of xh, and every later group
// whose xh lies in the same coset of S stages the same 2^p tiles again, from
// that XCD's L2.  A TileMap sends a workgroup's index (position within its
// XCD in bits 0..p-1, XCD in bits p..p+2, batch above) to its output tile
// through the columns col[0..n): the first p columns span S, chosen so that
// the groups' xh fall into few classes modulo S, the next three extend S to
// the tiles of all eight XCDs (which share the Infinity Cache), the rest
// complete a basis.  n = 0: the contiguous order of xcd_tile_index().
constexpr int kTileMapMax = 32;
struct TileMap { uint32_t col[kTileMapMax]; int n, p; };

// echelon basis of a GF(2) subspace: piv[i] has leading bit i, or is 0
struct Gf2Span {
    uint32_t piv[32] = {0};
    uint32_t reduce(uint32_t v) const   // the coset's element with no pivot bits
    {
        for (int i = 31; i >= 0; i--)
            if ((v >> i & 1u) && piv[i]) v ^= piv[i];
        return v;
    }
    bool insert(uint32_t v)
    {
        v = reduce(v);
        if (!v) return false;
        piv[31 - __builtin_clz(v)] = v;
        return true;
    }
};

// Greedy choice of the map for the groups' distinct high parts xs (sorted):
// candidate columns are the most frequent differences xh ^ xh' (from at most
// 512 groups), each step takes the one leaving the fewest classes.  xcd and
// chip receive the spans of the first p and p + 3 columns.  Returns the
// number of classes modulo the XCD span.
static inline size_t build_tile_map(const std::vector<uint32_t> &xs, int nbits, int p,
                                    TileMap &m, Gf2Span &xcd, Gf2Span &chip)
{
    std::vector<uint32_t> d;
    const size_t stride = xs.size() > 512 ? (xs.size() + 511) / 512 : 1;
    for (size_t i = 0; i < xs.size(); i += stride)
        for (size_t j = i + stride; j < xs.size(); j += stride) d.push_back(xs[i] ^ xs[j]);
    std::sort(d.begin(), d.end());
    std::vector<std::pair<int, uint32_t>> cand;   // (-count, difference)
    for (size_t i = 0; i < d.size();) {
        size_t j = i;
        while (j < d.size() && d[j] == d[i]) j++;
        cand.push_back({-(int)(j - i), d[i]});
        i = j;
    }
    std::sort(cand.begin(), cand.end());
    if (cand.size() > 128) cand.resize(128);
    for (int i = 0; i < nbits; i++) cand.push_back({0, 1u << i});   // fallback columns
    std::vector<uint32_t> r(xs.size());
    auto classes = [&](const Gf2Span &s) {
        for (size_t i = 0; i < xs.size(); i++) r[i] = s.reduce(xs[i]);
        std::sort(r.begin(), r.end());
        return (size_t)(std::unique(r.begin(), r.end()) - r.begin());
    };
    Gf2Span span;
    size_t nxcd = 0;
    m.n = 0;
    m.p = p;
    for (int c = 0; c < p + 3; c++) {
        size_t best = SIZE_MAX;
        uint32_t col = 0;
        for (const auto &e : cand) {
            Gf2Span s = span;
            if (!s.insert(e.second)) continue;
            const size_t k = classes(s);
            if (k < best) { best = k; col = e.second; }
        }
        span.insert(col);
        m.col[m.n++] = col;
        if (c == p - 1) { xcd = span; nxcd = best; }
    }
    chip = span;
    for (int i = 0; i < nbits; i++)
        if (span.insert(1u << i)) m.col[m.n++] = 1u << i;
    return nxcd;
}

// Sub classes, in the order a group stores its subs: single-term real c with
// no phase bit in the row bits, real with row bits, imaginary (same split),
// complex, and several terms per flip mask.
enum : uint32_t { kRunReal0 = 0, kRunReal1, kRunImag0, kRunImag1, kRunCplx, kRunMulti, kRuns };

struct GroupRuns { uint32_t xh, b[kRuns + 1]; };

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

template <typename A> struct Tables {
    std::vector<GroupRuns> groups;
    std::vector<SubRec<A>> subs;
    std::vector<TermRec<A>> terms;
    std::vector<FlatTerm<A>> flat;   // x-sorted, for the one-thread-per-amplitude kernel
    uint64_t nterms = 0, nflips = 0;
    TileMap tilemap{};   // output-tile order (n = 0: contiguous)
};

// Build the tables of n terms for chain length len at tile bits tb (0: flat
// list only) and workgroup size threads.  Zero coefficients are dropped.
// xcd_bits = p > 0: 2^p workgroups are resident per XCD; build the output-tile
// map and order the groups by its classes (when there are at least p + 3
// tile-index bits).
template <typename R, typename A>
int build_tables(int len, int tb, int threads, const uint64_t *bits, const double *coeff,
                 uint64_t n, Tables<A> &out, int xcd_bits = 0)
{
    struct Tmp { uint32_t x, z; A cr, ci; };
    std::vector<Tmp> v;
    v.reserve(n);
    const uint64_t sitemask = len == 32 ? ~0ull : ((1ull << (2 * len)) - 1);
    for (uint64_t t = 0; t < n; t++) {
        // coefficients arrive as doubles holding values of the SumOps' FpType
        A cr = (A)(R)coeff[2 * t], ci = (A)(R)coeff[2 * t + 1];
        if (cr == 0 && ci == 0) continue;
        if (bits[t] & ~sitemask)
            return set_error(RP_ERR_INVALID_ARGUMENT,
                             "rp_terms_upload: term %llu acts on a site >= len=%d",
                             (unsigned long long)t, len);
        Tmp e;
        e.x = pext_even(bits[t]);
        e.z = pext_odd(bits[t]);
        if (__builtin_popcount(e.x & e.z) & 1) { cr = -cr; ci = -ci; }
        e.cr = cr;
        e.ci = ci;
        v.push_back(e);
    }
    const uint32_t lowmask = tb ? ((1u << tb) - 1) : 0xffffffffu;
    Gf2Span xcd, chip;
    out.tilemap = TileMap{};
    bool mapped = tb && xcd_bits > 0 && len - tb >= xcd_bits + 3 && !v.empty();
    if (mapped) {
        std::vector<uint32_t> xs;
        xs.reserve(v.size());
        for (const Tmp &e : v) xs.push_back(e.x >> tb);
        std::sort(xs.begin(), xs.end());
        xs.erase(std::unique(xs.begin(), xs.end()), xs.end());
        // kept only when it beats the contiguous order, whose XCD classes
        // are the xh >> p
        size_t low = 0;
        for (size_t i = 0; i < xs.size(); i++)
            low += i == 0 || (xs[i] >> xcd_bits) != (xs[i - 1] >> xcd_bits);
        if (build_tile_map(xs, len - tb, xcd_bits, out.tilemap, xcd, chip) >= low) {
            out.tilemap = TileMap{};
            mapped = false;
        }
    }
    // by chip class, XCD class (when mapped), high part, low part
    std::stable_sort(v.begin(), v.end(), [&](const Tmp &a, const Tmp &b) {
        uint32_t ah = tb ? a.x >> tb : 0, bh = tb ? b.x >> tb : 0;
        if (mapped && ah != bh) {
            const uint32_t ac = chip.reduce(ah), bc = chip.reduce(bh);
            if (ac != bc) return ac < bc;
            const uint32_t ax = xcd.reduce(ah), bx = xcd.reduce(bh);
            if (ax != bx) return ax < bx;
        }
        if (ah != bh) return ah < bh;
        return (a.x & lowmask) < (b.x & lowmask);
    });
    out.nterms = v.size();
    auto &groups = out.groups;
    auto &subs = out.subs;
    auto &terms = out.terms;
    auto &flat = out.flat;
    uint64_t nflips = 0;
    for (size_t i = 0; i < v.size(); i++) {
        if (i == 0 || v[i].x != v[i - 1].x) nflips++;
        flat.push_back({v[i].x, v[i].z, v[i].cr, v[i].ci});
    }
    if (tb) {
        const int lb = __builtin_ctz((unsigned)threads);
        const uint32_t rpmask = (1u << (tb - lb)) - 1;
        size_t i = 0;
        while (i < v.size()) {
            const uint32_t xh = v[i].x >> tb;
            // the group's subs, classified, then stored run by run
            std::vector<SubRec<A>> run[kRuns];
            while (i < v.size() && (v[i].x >> tb) == xh) {
                size_t j = i;
                while (j < v.size() && v[j].x == v[i].x) j++;
                SubRec<A> S;
                memset(&S, 0, sizeof(S));
                S.xl = v[i].x & lowmask;
                int cls;
                if (j - i == 1) {
                    S.z = v[i].z;
                    S.cr = v[i].cr;
                    S.ci = v[i].ci;
                    S.fr = (float)v[i].cr;
                    S.fi = (float)v[i].ci;
                    const bool rowz = ((S.z >> lb) & rpmask) != 0;
                    if (v[i].ci == 0) { S.kind = kSubReal; cls = rowz ? kRunReal1 : kRunReal0; }
                    else if (v[i].cr == 0) { S.kind = kSubImag; cls = rowz ? kRunImag1 : kRunImag0; }
                    else { S.kind = kSubComplex; cls = kRunCplx; }
                } else {
                    S.kind = kSubMulti;
                    S.term_begin = (uint32_t)terms.size();
                    for (size_t q = i; q < j; q++) terms.push_back({v[q].z, 0, v[q].cr, v[q].ci});
                    S.term_end = (uint32_t)terms.size();
                    cls = kRunMulti;
                }
                run[cls].push_back(S);
                i = j;
            }
            GroupRuns G;
            G.xh = xh;
            for (int k = 0; k < kRuns; k++) {
                G.b[k] = (uint32_t)subs.size();
                subs.insert(subs.end(), run[k].begin(), run[k].end());
            }
            G.b[kRuns] = (uint32_t)subs.size();
            groups.push_back(G);
        }
    }
    out.nflips = nflips;
    return RP_OK;
}

}  // namespace rp
