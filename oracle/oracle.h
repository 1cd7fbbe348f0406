/*
 * oracle.h -- CPU restatement of the REAPERS hot path, used ONLY as test
 * infrastructure (tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py).  Nothing in the shipped library links, loads or calls this code.
 *
 * What is restated, and where it comes from in the reference (cl91/REAPERS):
 *   - Pauli string encoding, product sign rule, apply():   inc/ops.h:133-161,
 *     inc/ops.h:319-327, inc/ops.h:366-377, inc/ops.h:421-440
 *   - COO construction of a Pauli string's matrix:          inc/ops.h:451-586
 *   - term list merge-on-insert, operator products:        inc/cpuimpl.h:233-253,
 *     inc/cpuimpl.h:325-343
 *   - dense matrix of a term list:                         inc/cpuimpl.h:372-390
 *   - Jordan-Wigner Majorana strings (non-block form):     inc/fermions.h:276-296
 *   - SYK Hamiltonian generators:                          inc/syk.h:88-119,
 *     inc/syk.h:54-81, inc/syk.h:152-262
 *   - libstdc++ <random> algorithms those generators call: mt19937,
 *     generate_canonical, normal_distribution (Marsaglia polar with a cached
 *     second value), bernoulli_distribution, uniform_int_distribution
 *     (GCC 11.4 bits/random.tcc:3348, bits/random.tcc:1802, bits/random.h:3636,
 *     bits/uniform_int_dist.h) -- third-party, restated from their source
 *   - CPU matrix-free apply (OpenMP over amplitudes):      inc/cpuimpl.h:737-754
 *   - restarted Krylov exp(-(beta+it)H)psi with iterated Gram-Schmidt and
 *     Arnoldi:                                             inc/algo.h:76-173,
 *     inc/algo.h:266-365
 *   - Lanczos ground state:                                inc/algo.h:180-257
 *   - matrix exponential of the small Hessenberg matrix: the reference calls
 *     Eigen's MatrixFunctions exp (not under /root/reference); restated from
 *     Higham, "The scaling and squaring method for the matrix exponential
 *     revisited", SIAM J. Matrix Anal. Appl. 26 (2005) 1179-1193, with the
 *     degree/threshold table Eigen 3.4 uses for float and double.
 *
 * Floating point: every function is written as the reference's sequence of
 * IEEE operations and this file is compiled with -ffp-contract=off, so that
 * results are reproducible bit for bit (the reference's own unit tests are
 * built the same way, without -march=native: tests/run.sh:7).
 */
#ifndef REAPERS_ORACLE_H
#define REAPERS_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- libstdc++ <random> restatement ---------------------------------- */
typedef struct { uint32_t mt[624]; int idx; } or_mt19937;
void     or_mt_seed(or_mt19937 *g, uint32_t seed);
uint32_t or_mt_next(or_mt19937 *g);
uint32_t or_mt19937_kat(void);            /* 10000th output, default seed */
void     or_mt_stream(uint32_t seed, uint32_t *out, int64_t n);
float    or_canonical_f32(or_mt19937 *g);
double   or_canonical_f64(or_mt19937 *g);
/* normal(0,1) stream exactly as std::normal_distribution<FP> draws it */
void     or_normal_stream_f32(uint32_t seed, float *out, int64_t n);
void     or_normal_stream_f64(uint32_t seed, double *out, int64_t n);
int      or_uniform_int(or_mt19937 *g, int a, int b);
void     or_bernoulli_stream(uint32_t seed, double p, double *out, int64_t n);
void     or_uniform_int_stream(uint32_t seed, int a, int b, double *out, int64_t n);
void     or_canonical_stream(uint32_t seed, int fp, double *out, int64_t n);

/* ---- Pauli string algebra (inc/ops.h) ------------------------------- */
/* product op0*op1: returns the bits, *minus set per inc/ops.h:319-327 */
uint64_t or_pauli_mul(uint64_t b0, uint64_t b1, int *minus);
/* inc/ops.h:421-440: res = n with x flips, *minus = sign of the coefficient */
void     or_pauli_apply(uint64_t bits, uint32_t n, uint32_t *res, int *minus);
/* number of sigma_y sites mod 4 (inc/ops.h:209-217) */
int      or_count_sy(uint64_t bits);
/* COO column list + minus bitmap of one string (inc/ops.h:553-586) */
void     or_pauli_sparse_matrix(uint64_t bits, int len, uint32_t *cols, uint8_t *neg);

/* ---- term generation (inc/fermions.h, inc/syk.h) -------------------- */
/* Majorana chi_n on N/2 sites: one term */
void or_majorana_f32(int N, int n, int std_gamma, uint64_t *bits, float *coeff2);
void or_majorana_f64(int N, int n, int std_gamma, uint64_t *bits, double *coeff2);
void or_majorana_block_f32(int N, int n, int std_gamma, int which, uint64_t *bits, float *coeff2);
void or_majorana_block_f64(int N, int n, int std_gamma, int which, uint64_t *bits, double *coeff2);
int64_t or_syk_block_gen_f32(int N, float sp, float J, int std_gamma, uint32_t seed,
                             int regularize, int which, uint64_t *bits, float *coeff, int64_t cap);
int64_t or_syk_block_gen_f64(int N, double sp, double J, int std_gamma, uint32_t seed,
                             int regularize, int which, uint64_t *bits, double *coeff, int64_t cap);
/* SYK::gen_ham(std::mt19937(seed), regularize); returns the number of terms,
 * or -needed if cap is too small; coeff is interleaved (re, im). */
int64_t or_syk_gen_f32(int N, float sp, float J, int std_gamma, uint32_t seed,
                       int regularize, uint64_t *bits, float *coeff, int64_t cap);
int64_t or_syk_gen_f64(int N, double sp, double J, int std_gamma, uint32_t seed,
                       int regularize, uint64_t *bits, double *coeff, int64_t cap);
/* anticommutator {chi_i, chi_j} as a term list (inc/cpuimpl.h products+merge),
 * used to pin the JW strings; returns number of terms written */
int64_t or_majorana_anticomm_f64(int N, int i, int j, int std_gamma,
                                 uint64_t *bits, double *coeff, int64_t cap);

/* ---- operator application ------------------------------------------ */
/* out += H in, matrix-free, complex interleaved, reference term order */
void or_apply_f64(int len, const uint64_t *bits, const double *coeff, int64_t nterms,
                  const double *in, double *out);
void or_apply_f32(int len, const uint64_t *bits, const float *coeff, int64_t nterms,
                  const float *in, float *out);
/* rows [row_begin, row_end) only, out holding just those rows (out[0] is
   row row_begin): windows of one apply of a state too large to apply whole */
void or_apply_window_f64(int len, const uint64_t *bits, const double *coeff, int64_t nterms,
                         const double *in, double *out, int64_t row_begin, int64_t row_end);
void or_apply_window_f32(int len, const uint64_t *bits, const float *coeff, int64_t nterms,
                         const float *in, float *out, int64_t row_begin, int64_t row_end);
/* rows [row_begin, row_end) only: a bounded sample of one apply, for timing */
void or_apply_rows_f64(int len, const uint64_t *bits, const double *coeff, int64_t nterms,
                       const double *in, double *out, int64_t row_begin, int64_t row_end);
void or_apply_rows_f32(int len, const uint64_t *bits, const float *coeff, int64_t nterms,
                       const float *in, float *out, int64_t row_begin, int64_t row_end);
/* dense matrix, column-major dim x dim complex (inc/cpuimpl.h:372-390) */
void or_dense_matrix_f64(int len, const uint64_t *bits, const double *coeff,
                         int64_t nterms, double *mat);

/* ---- small dense linear algebra ------------------------------------- */
/* Eigen-style scaling and squaring Pade exp of an n x n complex matrix
 * (column-major, interleaved). prec 64 uses the double table (Pade up to 13),
 * prec 32 the single table (Pade up to 7). */
void or_expm_f64(int n, const double *a, double *out);
void or_expm_f32(int n, const float *a, float *out);

/* ---- restarted Krylov (inc/algo.h) ---------------------------------- */
typedef struct {
    int64_t matvecs;     /* apply_ops calls */
    int64_t restarts;    /* outer iterations (niters) */
    int64_t gs_passes;   /* Gram-Schmidt passes over all Arnoldi steps */
    int     breakdown;   /* 1 if Arnoldi broke down */
    double  last_nrm;    /* ||Hexp0|| of the last restart */
} or_krylov_stats;
/* psi <- exp(-(beta + i t) H) psi; psi interleaved complex, dim 2^len */
int or_krylov_evolve_f64(int len, const uint64_t *bits, const double *coeff, int64_t nterms,
                         double *psi, double t, double beta, int krydim, double tol,
                         int max_iters, or_krylov_stats *st);
int or_krylov_evolve_f32(int len, const uint64_t *bits, const float *coeff, int64_t nterms,
                         float *psi, float t, float beta, int krydim, float tol,
                         int max_iters, or_krylov_stats *st);
/* Lanczos ground state (inc/algo.h:180-257), starting vector given in psi
 * instead of a Haar random state; returns the energy, psi <- ground state */
double or_lanczos_ground_state_f64(int len, const uint64_t *bits, const double *coeff,
                                   int64_t nterms, double *psi, int krydim, double eps,
                                   int maxiter, int *iters);

int or_num_threads(void);

/* ---- Krylov-schedule proposer network (oracle_mlp.c) ----------------- */
int    or_mlp_nparams(void);
void   or_mlp_forward_f64(const double *p, const double *X, int64_t n, double *Y);
/* returns the loss; g <- its gradient (W may be null: all ones) */
double or_mlp_loss_grad_f64(const double *p, const double *X, const double *T,
                            const double *W, int64_t n, double *g);
void   or_mlp_adam_f64(double *p, const double *g, double *m, double *v, int64_t np,
                       double lr, double b1, double b2, double eps, double wd,
                       int64_t step);

#ifdef __cplusplus
}
#endif
#endif
