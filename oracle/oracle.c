/*
 * oracle.c -- CPU restatement of the REAPERS hot path (test infrastructure
 * only; see oracle.h for the list of reference sources each part follows).
 *
 * Parity status of each part (how it is pinned; see tests/test_oracle_*.py):
 *   mt19937, generate_canonical, normal/bernoulli/uniform_int:
 *       pinned (C++ standard KAT, numpy's independent MT19937, and libstdc++
 *       <random> itself through the shipped C++ host library)
 *   Pauli product/sign/apply/COO: pinned (tests/ops.cc:26-161 tables, closed
 *       forms, dense Kronecker products built with numpy)
 *   Majorana strings: pinned (inc/fermions.h:236-245 worked examples,
 *       anticommutation tests/ops.cc:299-346)
 *   SYK generator: term count/bits/coefficients pinned against the shipped C++
 *       generator that calls libstdc++ <random>; the exact term list of a
 *       reference binary is unpinned (the reference needs Eigen3 to build)
 *   matrix exponential: pinned (tests/ops.cc:198-229 KATs, scipy expm)
 *   Krylov evolve / Lanczos: pinned against exact diagonalization (numpy eigh)
 *       and invariants; the exact restart trajectory of a reference binary is
 *       unpinned (it depends on Eigen's exp)
 */
#include "oracle.h"

/* OR_PLANT selects a deliberately wrong variant for tests that show the
 * pinning checks catch it (0 = the restatement itself):
 *   1 drop the product sign, 2 swap X and Z bits in apply,
 *   3 drop the restart coupling beta in Hfull */
#ifndef OR_PLANT
#define OR_PLANT 0
#endif

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ===================================================================== */
/* libstdc++ <random> restatement                                         */
/* ===================================================================== */

/* std::mt19937 (C++ [rand.eng.mers]): w=32 n=624 m=397 r=31
 * a=0x9908b0df u=11 d=0xffffffff s=7 b=0x9d2c5680 t=15 c=0xefc60000 l=18
 * f=1812433253 */
void or_mt_seed(or_mt19937 *g, uint32_t seed)
{
    g->mt[0] = seed;
    for (int i = 1; i < 624; i++) {
        uint32_t p = g->mt[i - 1];
        g->mt[i] = 1812433253u * (p ^ (p >> 30)) + (uint32_t)i;
    }
    g->idx = 624;
}

static void or_mt_twist(or_mt19937 *g)
{
    for (int i = 0; i < 624; i++) {
        uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
        uint32_t v = g->mt[(i + 397) % 624] ^ (y >> 1);
        if (y & 1u) v ^= 0x9908b0dfu;
        g->mt[i] = v;
    }
    g->idx = 0;
}

uint32_t or_mt_next(or_mt19937 *g)
{
    if (g->idx >= 624) or_mt_twist(g);
    uint32_t y = g->mt[g->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

uint32_t or_mt19937_kat(void)
{
    or_mt19937 g;
    or_mt_seed(&g, 5489u);
    uint32_t v = 0;
    for (int i = 0; i < 10000; i++) v = or_mt_next(&g);
    return v;
}

void or_uniform_int_stream(uint32_t seed, int a, int b, double *out, int64_t n);

void or_mt_stream(uint32_t seed, uint32_t *out, int64_t n)
{
    or_mt19937 g;
    or_mt_seed(&g, seed);
    for (int64_t i = 0; i < n; i++) out[i] = or_mt_next(&g);
}

/* std::generate_canonical<float, 24>(mt19937): one 32-bit draw
 * (GCC 11 bits/random.tcc:3348-3378) */
float or_canonical_f32(or_mt19937 *g)
{
    float sum = (float)or_mt_next(g) * 1.0f;
    float tmp = 4294967296.0f;
    float ret = sum / tmp;
    if (ret >= 1.0f) ret = nextafterf(1.0f, 0.0f);
    return ret;
}

/* std::generate_canonical<double, 53>(mt19937): two 32-bit draws */
double or_canonical_f64(or_mt19937 *g)
{
    double sum = 0.0;
    double tmp = 1.0;
    sum += (double)or_mt_next(g) * tmp;
    tmp *= 4294967296.0;
    sum += (double)or_mt_next(g) * tmp;
    tmp *= 4294967296.0;
    double ret = sum / tmp;
    if (ret >= 1.0) ret = nextafter(1.0, 0.0);
    return ret;
}

/* std::normal_distribution: Marsaglia polar method with one cached value
 * (GCC 11 bits/random.tcc:1802-1836) */
typedef struct { int avail; float saved; } or_normal_f32;
typedef struct { int avail; double saved; } or_normal_f64;

static float or_normal_draw_f32(or_normal_f32 *d, or_mt19937 *g)
{
    float ret;
    if (d->avail) {
        d->avail = 0;
        ret = d->saved;
    } else {
        float x, y, r2;
        do {
            /* result_type(2.0) * aurng() - 1.0: the subtraction is in double */
            x = (float)((double)(2.0f * or_canonical_f32(g)) - 1.0);
            y = (float)((double)(2.0f * or_canonical_f32(g)) - 1.0);
            r2 = x * x + y * y;
        } while ((double)r2 > 1.0 || (double)r2 == 0.0);
        const float mult = sqrtf(-2.0f * logf(r2) / r2);
        d->saved = x * mult;
        d->avail = 1;
        ret = y * mult;
    }
    return ret * 1.0f + 0.0f;
}

static double or_normal_draw_f64(or_normal_f64 *d, or_mt19937 *g)
{
    double ret;
    if (d->avail) {
        d->avail = 0;
        ret = d->saved;
    } else {
        double x, y, r2;
        do {
            x = 2.0 * or_canonical_f64(g) - 1.0;
            y = 2.0 * or_canonical_f64(g) - 1.0;
            r2 = x * x + y * y;
        } while (r2 > 1.0 || r2 == 0.0);
        const double mult = sqrt(-2.0 * log(r2) / r2);
        d->saved = x * mult;
        d->avail = 1;
        ret = y * mult;
    }
    return ret * 1.0 + 0.0;
}

void or_normal_stream_f32(uint32_t seed, float *out, int64_t n)
{
    or_mt19937 g;
    or_normal_f32 d = {0, 0.0f};
    or_mt_seed(&g, seed);
    for (int64_t i = 0; i < n; i++) out[i] = or_normal_draw_f32(&d, &g);
}

/* bernoulli(p), uniform_int(a, b) and generate_canonical streams from one
 * engine, for comparing against libstdc++ */
void or_bernoulli_stream(uint32_t seed, double p, double *out, int64_t n);
void or_uniform_int_stream(uint32_t seed, int a, int b, double *out, int64_t n);
void or_canonical_stream(uint32_t seed, int fp, double *out, int64_t n);

void or_normal_stream_f64(uint32_t seed, double *out, int64_t n)
{
    or_mt19937 g;
    or_normal_f64 d = {0, 0.0};
    or_mt_seed(&g, seed);
    for (int64_t i = 0; i < n; i++) out[i] = or_normal_draw_f64(&d, &g);
}

/* std::bernoulli_distribution(p)(g): generate_canonical<double> < p
 * (GCC 11 bits/random.h:3636-3644) */
static int or_bernoulli(or_mt19937 *g, double p)
{
    return (or_canonical_f64(g) - 0.0) < p * (1.0 - 0.0);
}

void or_bernoulli_stream(uint32_t seed, double p, double *out, int64_t n)
{
    or_mt19937 g;
    or_mt_seed(&g, seed);
    for (int64_t i = 0; i < n; i++) out[i] = or_bernoulli(&g, p) ? 1.0 : 0.0;
}

void or_canonical_stream(uint32_t seed, int fp, double *out, int64_t n)
{
    or_mt19937 g;
    or_mt_seed(&g, seed);
    for (int64_t i = 0; i < n; i++)
        out[i] = fp == 32 ? (double)or_canonical_f32(&g) : or_canonical_f64(&g);
}

/* std::uniform_int_distribution<int>(a, b)(mt19937): Lemire's nearly
 * divisionless reduction with 64-bit products (GCC 11
 * bits/uniform_int_dist.h:246-270, 288-318) */
int or_uniform_int(or_mt19937 *g, int a, int b)
{
    uint64_t urange = (uint64_t)b - (uint64_t)a;
    uint32_t range = (uint32_t)(urange + 1);
    uint64_t product = (uint64_t)or_mt_next(g) * (uint64_t)range;
    uint32_t low = (uint32_t)product;
    if (low < range) {
        uint32_t threshold = (uint32_t)(-range) % range;
        while (low < threshold) {
            product = (uint64_t)or_mt_next(g) * (uint64_t)range;
            low = (uint32_t)product;
        }
    }
    return (int)(uint32_t)(product >> 32) + a;
}

void or_uniform_int_stream(uint32_t seed, int a, int b, double *out, int64_t n)
{
    or_mt19937 g;
    or_mt_seed(&g, seed);
    for (int64_t i = 0; i < n; i++) out[i] = or_uniform_int(&g, a, b);
}

/* ===================================================================== */
/* Pauli strings (inc/ops.h)                                              */
/* ===================================================================== */

/* Sigma codes: id=00 x=01 y=11 z=10 (inc/ops.h:133-143) */
static int or_site(uint64_t bits, int s) { return (int)((bits >> (2 * s)) & 3u); }

/* inc/ops.h:151-161: z.x, z.y, y.x, y.y carry a minus sign */
static int or_prod_has_minus(int op0, int op1)
{
#if OR_PLANT == 1
    /* planted mistake: the product never carries a minus sign */
    (void)op0; (void)op1;
    return 0;
#endif
    switch ((op0 << 2) | op1) {
    case 0x9: case 0xb: case 0xd: case 0xf: return 1;
    default: return 0;
    }
}

/* inc/ops.h:319-327: iterate over the non-identity sites of the LEFT factor */
uint64_t or_pauli_mul(uint64_t b0, uint64_t b1, int *minus)
{
    int m = 0;
    for (int s = 0; s < 32; s++) {
        int o0 = or_site(b0, s);
        if (o0 == 0) continue;
        if (or_prod_has_minus(o0, or_site(b1, s))) m = !m;
    }
    *minus = m;
    return b0 ^ b1;
}

/* inc/ops.h:421-440, bit loop with early exit */
void or_pauli_apply(uint64_t bits, uint32_t n, uint32_t *res, int *minus)
{
    int m = 0;
    uint64_t b = bits;
#if OR_PLANT == 2
    /* planted mistake: X and Z bits swapped */
    b = ((bits & 0x5555555555555555ull) << 1) | ((bits >> 1) & 0x5555555555555555ull);
#endif
    for (int i = 0; i < 32; i++) {
        if (b & 1u) n ^= (uint32_t)1u << i;
        b >>= 1;
        if ((b & 1u) && ((n >> i) & 1u)) m = !m;
        b >>= 1;
        if (!b) break;
    }
    *res = n;
    *minus = m;
}

int or_count_sy(uint64_t bits)
{
    int c = 0;
    for (int s = 0; s < 32; s++)
        if (or_site(bits, s) == 3) c++;
    return c & 3;
}

/* inc/ops.h:553-586: build the COO form of one string by doubling, site 0
 * first; neg[i] marks a -1 entry at row i, column cols[i]. */
void or_pauli_sparse_matrix(uint64_t bits, int len, uint32_t *cols, uint8_t *neg)
{
    if (len == 0) { cols[0] = 0; neg[0] = 0; return; }
    size_t cnt = 0;
    for (int i = 0; i < len; i++) {
        int op = or_site(bits, i);
        int n0 = (op == 3), n1 = (op == 2);       /* y: (true,false); z: (false,true) */
        uint32_t c0 = (op == 1 || op == 3) ? 1u : 0u;
        uint32_t c1 = (op == 1 || op == 3) ? 0u : 1u;
        if (cnt == 0) {
            neg[0] = (uint8_t)n0; neg[1] = (uint8_t)n1;
            cols[0] = c0; cols[1] = c1;
            cnt = 2;
            continue;
        }
        for (size_t k = 0; k < cnt; k++) {
            neg[k + cnt] = (uint8_t)(n1 ? !neg[k] : neg[k]);
            if (n0) neg[k] = (uint8_t)!neg[k];
            cols[k + cnt] = cols[k] | (c1 << i);
            cols[k] |= (c0 << i);
        }
        cnt <<= 1;
    }
}

/* ===================================================================== */
/* precision-generic parts                                                */
/* ===================================================================== */

#define FP float
#define SUF(x) x##_f32
#define FSQRT sqrtf
#define FABS fabsf
#define FEXP expf
#define OR_IS_F32 1
#include "oracle_fp.inc"
#undef FP
#undef SUF
#undef FSQRT
#undef FABS
#undef FEXP
#undef OR_IS_F32

#define FP double
#define SUF(x) x##_f64
#define FSQRT sqrt
#define FABS fabs
#define FEXP exp
#define OR_IS_F32 0
#include "oracle_fp.inc"
#undef FP
#undef SUF
#undef FSQRT
#undef FABS
#undef FEXP
#undef OR_IS_F32

/* dense matrix (inc/cpuimpl.h:372-390): mat(i, cols[i]) += (+/-) coeff,
 * terms in list order, column-major */
void or_dense_matrix_f64(int len, const uint64_t *bits, const double *coeff,
                         int64_t nterms, double *mat)
{
    size_t dim = (size_t)1 << len;
    memset(mat, 0, sizeof(double) * 2 * dim * dim);
    uint32_t *cols = (uint32_t *)malloc(sizeof(uint32_t) * dim);
    uint8_t *neg = (uint8_t *)malloc(dim);
    for (int64_t t = 0; t < nterms; t++) {
        or_pauli_sparse_matrix(bits[t], len, cols, neg);
        for (size_t i = 0; i < dim; i++) {
            double re = neg[i] ? -coeff[2 * t] : coeff[2 * t];
            double im = neg[i] ? -coeff[2 * t + 1] : coeff[2 * t + 1];
            size_t o = 2 * ((size_t)cols[i] * dim + i);
            mat[o] += re;
            mat[o + 1] += im;
        }
    }
    free(cols);
    free(neg);
}
