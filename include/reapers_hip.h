/*
 * reapers_hip.h -- C ABI of the gfx950 backend (libreapers_hip.so).
 *
 * This is the drop-in boundary: the header-only C++ API (reapers_amd/include/
 * reapers/) implements the reference's GPUImpl static interface on top of
 * these functions, and no C++ or torch type crosses it.  Every entry point
 * returns an int status (RP_OK on success); rp_last_error() describes the
 * last failure of the calling thread.  The C++ layer maps the status codes to
 * the reference's exception classes (inc/except.h:100-170).
 *
 * Conventions (the reference's, inc/gpuimpl.h:396-1048):
 *   - one device, one library stream with default priority; vector kernels
 *     are ordered on that stream, every function that returns a value to the
 *     host synchronizes first, rp_synchronize() waits for all queued work;
 *   - device pointers are owned by the caller (rp_malloc/rp_free); a term
 *     handle (rp_terms_t) is owned by the library until rp_terms_free();
 *   - complex vectors are interleaved (re, im): RP_C64 = complex<float>
 *     (8 bytes), RP_C128 = complex<double> (16 bytes);
 *   - n is the number of complex elements; no power-of-two restriction.
 *
 * Each declaration names the reference interface it replaces (file:line
 * under the reference tree).
 */
#ifndef REAPERS_HIP_H
#define REAPERS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes; the C++ layer throws the reference's exception for each */
#define RP_OK                   0
#define RP_ERR_INVALID_ARGUMENT 1   /* InvalidArgument (inc/except.h:100-117) */
#define RP_ERR_OUT_OF_MEMORY    2   /* DevOutOfMem (inc/except.h:160-170)     */
#define RP_ERR_RUNTIME          3   /* CudaError/RuntimeError (inc/except.h:119-141) */
#define RP_ERR_NO_DEVICE        4   /* NoGpuDevice (inc/except.h:143-149)     */
#define RP_ERR_NOT_INITIALIZED  5

/* element precisions */
#define RP_C64  0                   /* complex<float>  */
#define RP_C128 1                   /* complex<double> */

/* rounding of a complex<double> -> complex<float> conversion */
#define RP_ROUND_NEAREST      0     /* CPUImpl::copy_vec (inc/cpuimpl.h:620-629) */
#define RP_ROUND_TOWARD_ZERO  1     /* krnl_dtof_vec (src/krnl.cu:33-39)         */

/* memcpy directions */
#define RP_H2D 0
#define RP_D2H 1
#define RP_D2D 2

/* ---------------------------------------------------------------------- */
/* lifecycle -- GPUContext (inc/gpuimpl.h:496-557), sync_devices          */
/* (inc/gpuimpl.h:628-638)                                                 */
/* ---------------------------------------------------------------------- */
int         rp_init(int device);
int         rp_shutdown(void);
const char *rp_last_error(void);
int         rp_synchronize(void);
/* cudaMemGetInfo in GPUContext (inc/gpuimpl.h:553-556) */
int         rp_mem_get_info(size_t *free_bytes, size_t *total_bytes);
/* number of compute units and the device name, for launch sizing/reporting */
int         rp_device_props(int *num_cus, char *name, size_t name_len);

/* ---------------------------------------------------------------------- */
/* memory -- DeviceVec (inc/gpuimpl.h:404-494), cuda_alloc               */
/* (inc/gpuimpl.h:62-66), DeviceVec host<->device conversions              */
/* (inc/gpuimpl.h:438-462)                                                 */
/* ---------------------------------------------------------------------- */
/* rp_free waits for the library stream and keeps the block for the next
 * rp_malloc of the same size (freed device memory is cleared by the driver
 * before it can be allocated again, seconds for the 8.6 GB vectors of a
 * 30-qubit evolution); kept blocks count as free in rp_mem_get_info and go
 * back to HIP when any allocation of the library runs out of memory and at
 * rp_shutdown. */
int rp_malloc(void **ptr, size_t bytes);
int rp_free(void *ptr);
int rp_memcpy(void *dst, const void *src, size_t bytes, int direction);

/* ---------------------------------------------------------------------- */
/* BLAS-1 -- GPUImpl vector operations                                     */
/* ---------------------------------------------------------------------- */
/* zero_vec (inc/gpuimpl.h:855-862) */
int rp_zero_vec(int prec, void *v, uint64_t n);
/* copy_vec, same or converting precision (inc/gpuimpl.h:869-889,
 * src/krnl.cu:25-39, src/krnl.cu:157-184); rounding applies to C128->C64 */
int rp_copy_vec(int dst_prec, void *dst, int src_prec, const void *src, uint64_t n,
                int rounding);
/* add_vec(v0, v1): v0 += v1 (src/krnl.cu:41-48, inc/gpuimpl.h:908-918) */
int rp_add_vec(int prec, void *v0, const void *v1, uint64_t n);
/* add_vec(res, v0, v1): res = v0 + v1 (src/krnl.cu:50-58, inc/gpuimpl.h:944-955) */
int rp_add_vec3(int prec, void *res, const void *v0, const void *v1, uint64_t n);
/* add_vec(v0, s, v1): v0 += s*v1, s complex (inc/gpuimpl.h:806-818,
 * inc/gpuimpl.h:920-942; cublasXaxpy) */
int rp_axpy(int prec, void *v0, double s_re, double s_im, const void *v1, uint64_t n);
/* scale_vec(v, s): v *= s (inc/gpuimpl.h:822-830, inc/gpuimpl.h:957-976) */
int rp_scale_vec(int prec, void *v, double s_re, double s_im, uint64_t n);
/* vec_prod: <v0|v1> = sum conj(v0[i]) v1[i] (inc/gpuimpl.h:980-995;
 * cublasXdotc), accumulated in FP64 for both precisions */
int rp_vec_prod(int prec, const void *v0, const void *v1, uint64_t n, double *re, double *im);
/* vec_norm = sqrt(Re <v|v>) (inc/gpuimpl.h:999-1002), FP64 accumulation */
int rp_vec_norm(int prec, const void *v, uint64_t n, double *nrm);
/* init_random (inc/gpuimpl.h:1050-1062): Re, Im ~ N(0,1) from a seeded
 * Philox4x32-10 stream (the reference seeds cuRAND from random_device), then
 * normalized with an FP64 norm */
int rp_init_random(int prec, void *v, uint64_t n, uint64_t seed);

/* ---------------------------------------------------------------------- */
/* Pauli-sum operator -- DevSumOps (inc/gpuimpl.h:125-393)                 */
/* ---------------------------------------------------------------------- */
typedef struct rp_terms *rp_terms_t;

/* DevSumOps::upload (inc/gpuimpl.h:1064-1085): takes the host term list of a
 * SumOps -- SpinOp bits (2 bits per site, X at bit 2s, Z at bit 2s+1,
 * inc/ops.h:133-143) and the coefficient (the SpinOp coeff member, in which
 * sigma_y carries its i: inc/ops.h:285-288) as interleaved doubles -- and
 * builds the flip/phase masks and the grouped device layout the apply kernel
 * uses.  len is the spin chain length (1..32). */
int rp_terms_upload(rp_terms_t *out, int prec, int len, const uint64_t *bits,
                    const double *coeff, uint64_t nterms);
/* rp_terms_upload with options: RP_TERMS_ACC_NATIVE accumulates a complex<float>
 * apply in FP32 instead of FP64 (faster, one FP32 rounding per term instead
 * of one per output element). */
#define RP_TERMS_ACC_NATIVE 1
/* RP_TERMS_AUTOTUNE times the instantiated apply tilings (tile bits and
 * workgroup size) on scratch vectors at upload and keeps the fastest */
#define RP_TERMS_AUTOTUNE 2
/* RP_TERMS_ACC_GROUP (complex<float> only): the terms of every four
 * consecutive tile groups are summed in FP32 and those sums in FP64 -- FP32
 * arithmetic in the inner loop, the running sum never rounded to FP32 (used
 * by the 2-stage pipelines; others accumulate in FP64) */
#define RP_TERMS_ACC_GROUP 4
int rp_terms_upload_ex(rp_terms_t *out, int prec, int len, const uint64_t *bits,
                       const double *coeff, uint64_t nterms, int flags);
/* DevSumOps::release (inc/gpuimpl.h:1087-1105) */
int rp_terms_free(rp_terms_t h);

typedef struct {
    int      prec;          /* RP_C64 / RP_C128 */
    int      len;           /* spin chain length */
    int      tile_bits;     /* b: log2 of the amplitude tile one workgroup owns
                               (0: one thread per amplitude, no tiles) */
    uint64_t nterms;        /* terms after dropping zero coefficients */
    uint64_t nflips;        /* distinct flip masks x (G_0) */
    uint64_t ngroups;       /* distinct x >> b: input tile passes (G_b) */
    uint64_t bytes_per_apply; /* algorithmic HBM bytes (G_b + 1) * 2^len * S */
    int      threads;       /* workgroup size of the tiled kernel */
    int      depth;         /* LDS tiles in its pipeline (input tiles in flight + 1);
                               for tile_bits 0 the number of term slices */
    int      kernel;        /* RP_APPLY_SMALL, RP_APPLY_TILED or RP_APPLY_STREAM */
    double   apply_passes;  /* the kept tiling's apply time over one streaming pass
                               of the state (half an rp_scale_vec), measured by the
                               autotuner at upload; 0 without RP_TERMS_AUTOTUNE */
    int      tile_map;      /* tiled kernels: columns of the output-tile map that
                               orders the groups by XCD class (0: contiguous tiles,
                               groups by flip mask) */
} rp_terms_info_t;
/* the apply kernel family a term handle launches: one thread per amplitude
   (tile_bits 0); tiles with per-sub records; tiles with the compact sub
   stream (complex<float>, group accumulation, two LDS tiles) */
#define RP_APPLY_SMALL  0
#define RP_APPLY_TILED  1
#define RP_APPLY_STREAM 2
int rp_terms_info(rp_terms_t h, rp_terms_info_t *info);
/* Rebuild the layout for a given tiling of the tiled kernel (no counterpart
 * in the reference): tile bits / threads / pipeline depth must be one of the
 * instantiated combinations -- (8,128,2), (10,128,2), (12,512,2) for both
 * precisions; (12,512,3), (12,512,4), (12,1024,2), (13,1024,2) for RP_C64; (11,256,4) for
 * RP_C128 -- or tile bits 0 with 256 threads: the one-thread-per-amplitude
 * kernel, its depth the number of term slices (1, 8, 32, 128) that
 * separate workgroups sum before a second pass adds them in slice order.
 * The autotuner (RP_TERMS_AUTOTUNE) chooses among the same set, the sliced
 * kernel only up to len 16. */
int rp_terms_set_tiling(rp_terms_t h, int tile_bits, int threads, int depth);

/* GPUImpl::apply_ops + krnl_apply_ops (src/krnl.cu:220-257, 80-94):
 *   accumulate != 0: res += H vec   (the reference's semantics)
 *   accumulate == 0: res  = H vec   (zero_vec fused into the apply)
 * res and vec must be different buffers of 2^len elements of the handle's
 * precision. */
int rp_apply_ops(rp_terms_t h, void *res, const void *vec, int accumulate);

/* The Arnoldi step's apply fused with its first Gram-Schmidt pass (no
 * counterpart in the reference, which calls apply_ops and then one vec_prod
 * per basis vector, inc/algo.h:86-101 and 160): res = H vec, then the pass of
 * rp_gs_pass_fused over V[0..k) (k <= 64) on res.  Where the handle's layout
 * allows (complex<float> states with FP64 or group accumulation, 2^12-
 * amplitude tiles of 512 threads, up to 2^28 amplitudes) the coefficients
 * <V_q|res> and ||res||^2 are formed in the apply kernel from the values it
 * writes, saving the pass's separate read of res; otherwise this is
 * rp_apply_ops followed by rp_gs_pass_fused.  res must differ from vec and
 * from every V[q]. */
int rp_apply_gs(rp_terms_t h, void *res, const void *vec, const void *const *V, int k,
                double *hcoef, double *norm, double fuse_min, int *normalized);

/* ---------------------------------------------------------------------- */
/* fused Krylov sweeps -- replace the per-vector dot/axpy/norm calls of    */
/* Krylov::orthogonalize (inc/algo.h:76-107) and the answer update of      */
/* Krylov::evolve (inc/algo.h:336-345)                                     */
/* ---------------------------------------------------------------------- */
/* the largest number of basis vectors one fused sweep takes (k below); it
 * bounds the Krylov dimension of an evolution */
#define RP_MAX_SWEEP_VECS 64
/* One classical Gram-Schmidt pass of w against V[0..k-1] (device pointers):
 *   h[i] = <V[i]|w> (FP64), w -= sum_i h[i] V[i], *norm = ||w|| (FP64).
 * h receives 2k doubles (re, im).  k may be 0 (then only the norm). */
int rp_gs_pass(int prec, const void *const *V, int k, void *w, uint64_t n,
               double *h, double *norm);
/* rp_gs_pass with the normalization fused into the update sweep: when the
 * coefficients give beta^2 = ||w||^2 - sum |h_q|^2 >= ||w||^2 / 2 (the
 * Pythagorean norm of the result for an orthonormal basis, free of
 * cancellation) and beta >= fuse_min > 0, w is written as (w - V h) / beta,
 * *normalized = 1 and *norm = beta; otherwise the plain pass (*normalized =
 * 0, *norm = the computed norm).  Saves the separate scale sweep of the
 * Arnoldi step (inc/algo.h:170). */
int rp_gs_pass_fused(int prec, const void *const *V, int k, void *w, uint64_t n, double *h,
                     double *norm, double fuse_min, int *normalized);
/* rp_gs_pass_fused that allocates nothing, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid it); no counterpart in the reference.  Same arguments and
 * results as rp_gs_pass_fused; its reduction partials and results live in a
 * fixed 1.06 MB device array of the library (one per GPU, taken at rp_init)
 * instead of the library workspace, which limits k to 64 (larger k is
 * rejected with RP_ERR_INVALID_ARGUMENT; n is unlimited).  Every call on a
 * GPU shares that GPU's array and leaves its last contents there: calls that
 * overlap on one GPU (from several host threads or streams) race. */
int rp_gs_pass_fused_noalloc(int prec, const void *const *V, int k, void *w, uint64_t n,
                             double *h, double *norm, double fuse_min, int *normalized);
/* y += sum_i c[i] V[i] for k vectors (any k; one sweep per 16 of them), c
   given as 2k doubles (re, im) */
int rp_lincomb(int prec, void *y, const void *const *V, int k, const double *c, uint64_t n);

/* ---------------------------------------------------------------------- */
/* dense matrices and exact diagonalization -- DevSumOps::MatrixType       */
/* (inc/gpuimpl.h:139-300), get_matrix / get_eigensystem / _matexp        */
/* (inc/gpuimpl.h:370-388, 1179-1263), krnl_eye / krnl_vec_exp /          */
/* krnl_scale_vecs (src/krnl.cu:113-154), mat_mul / mat_tr                 */
/* (src/krnl.cu:260-300).  Matrices are column-major, leading dimension =  */
/* rows, complex of the given precision, in caller-owned device memory;    */
/* n = 2^len <= 2^15.                                                       */
/* ---------------------------------------------------------------------- */
/* the dense form of a term list (bits and interleaved coefficients as for
 * rp_terms_upload): mat[r + c n] = <r|H|c>, n = 2^len */
int rp_dense_build(int prec, int len, const uint64_t *bits, const double *coeff, uint64_t nterms,
                   void *mat);
/* Hermitian eigensystem (cusolverDnXsyevd in get_eigensystem,
 * inc/gpuimpl.h:1179-1233): the lower triangle of mat is read, mat is
 * overwritten by the column eigenvectors and w (device, n reals of the
 * precision) receives the eigenvalues in ascending order.  rocSOLVER's
 * divide-and-conquer heevd runs first; if its eigenvectors are not finite
 * (it fails so on matrices made of decoupled blocks) it runs again on
 * U^H A U, U a Walsh-Hadamard matrix (n a power of two) times random phases,
 * and the eigenvectors are rotated back; heev (QR iteration) is the last
 * resort. */
int rp_dense_eigh(int prec, uint64_t n, void *mat, void *w);
/* exp(c H) = V diag(exp(c w)) V^H from H's eigensystem (_matexp,
 * inc/gpuimpl.h:1235-1252); work is n*n scratch elements; res, work and V
 * must be distinct */
int rp_dense_expm_eig(int prec, uint64_t n, const void *V, const void *w, double c_re,
                      double c_im, void *res, void *work);
/* C = alpha op(A) op(B) + beta C with op 0 = as is, 1 = conjugate transpose
 * (mat_mul, inc/gpuimpl.h:1024-1036); m x k times k x n */
int rp_dense_gemm(int prec, uint64_t m, uint64_t n, uint64_t k, int op_a, int op_b, double a_re,
                  double a_im, const void *A, uint64_t lda, const void *B, uint64_t ldb,
                  double b_re, double b_im, void *C, uint64_t ldc);
/* trace summed in FP64 (mat_tr, inc/gpuimpl.h:1040-1042) */
int rp_dense_trace(int prec, uint64_t n, const void *mat, double *re, double *im);
/* identity (krnl_eye via MatrixType::Identity, inc/gpuimpl.h:1107-1113) */
int rp_dense_eye(int prec, uint64_t n, void *mat);

/* ---------------------------------------------------------------------- */
/* kernel timing (HIP events on the library stream)                       */
/* ---------------------------------------------------------------------- */
#define RP_KERNEL_APPLY 0           /* rp_apply_ops launches  */
#define RP_KERNEL_GS    1           /* rp_gs_pass launches    */
#define RP_KERNEL_OTHER 2           /* every other launch     */
/* when enabled, each launch is bracketed by events and its duration added to
 * a per-class total at the next synchronization; on = RP_PROF_CLASSES(mask)
 * brackets only the classes whose bit (1 << class) is set, any other
 * nonzero value all of them */
#define RP_PROF_SELECT 0x100
#define RP_PROF_CLASSES(mask) (RP_PROF_SELECT | ((mask) & 0xff))
int rp_prof_enable(int on);
int rp_prof_reset(void);
int rp_prof_query(int kernel_class, uint64_t *launches, double *total_ms);

/* --- Batched right-hand sides (no counterpart in the reference, which
 * evolves one State at a time, inc/state.h:413-424 and inc/algo.h:266-365;
 * the per-problem arithmetic is the same as rp_apply_ops, rp_gs_pass and
 * rp_lincomb).  A batch holds nprob Hamiltonians on len[p] qubits (12..30,
 * uneven) with nrhs states each (1, 2 or 4).  A batch vector is one device
 * allocation of info.vec_bytes bytes; problem p's block starts at basis
 * offset off[p] (rp_batch_offsets) and state s of amplitude m is element
 * (off[p] + m) * nrhs + s.  Every operation is one launch (or a fixed
 * sequence of launches) over all problems whose active flag is set. */
typedef struct rp_batch *rp_batch_t;
typedef struct rp_graph *rp_graph_t;
#define RP_BATCH_AUTOTUNE 1   /* time the instantiated tilings, keep the fastest */
/* complex<float> batches: the terms of each pair of tile groups are summed in
 * FP32 and those sums in FP64 (as RP_TERMS_ACC_GROUP; others sum in FP64) */
#define RP_BATCH_ACC_GROUP 2
/* with the 32 KiB tile: 512 threads (2 rows each) instead of 256 (4 rows);
 * the autotuner also times this variant */
#define RP_BATCH_WIDE 4
/* force the apply's tile bits b: the apply is instantiated for tiles of
 * 32 KiB and 64 KiB, b = log2(32 KiB / (nrhs * element bytes)) or b + 1 */
#define RP_BATCH_TILE_BITS(b) ((b) << 8)
/* one problem of at least 20 qubits: the apply gathers each state into a
 * plane of its own, applies the single-state kernel (rp_apply_ops, the
 * problem's terms uploaded with its own autotuner) to every plane and
 * scatters the results back (2 * nrhs * 2^len * S bytes of scratch).  Taken
 * without this flag when RP_BATCH_AUTOTUNE times it faster than the batched
 * kernel, or from 22 qubits without the autotuner; RP_BATCH_TILE_BITS, and a
 * term list with no nonzero term, keep the batched kernel */
#define RP_BATCH_PLANES 8
typedef struct {
    int prec, nprob, nrhs, kmax;
    int tile_bits, threads, grid;   /* apply tiling and persistent grid (with
                                       planes: the single-state kernel's tiling
                                       and workgroups per plane) */
    int planes;                     /* 1: the plane apply (RP_BATCH_PLANES) */
    uint64_t nbasis;                /* sum of 2^len[p] */
    uint64_t vec_bytes;             /* bytes of one batch vector */
    uint64_t nterms, ngroups;       /* over all problems */
    uint64_t bytes_per_apply;       /* sum_p (G_b,p + 1) 2^len[p] nrhs S */
} rp_batch_info_t;
/* terms of problem p: bits[term_off[p] .. term_off[p+1]), coeff interleaved;
 * kmax bounds the Krylov dimension of rp_batch_arnoldi_step */
int rp_batch_create(rp_batch_t *out, int prec, int nprob, const int *len, const uint64_t *bits,
                    const double *coeff, const int64_t *term_off, int nrhs, int kmax, int flags);
int rp_batch_free(rp_batch_t b);
int rp_batch_info(rp_batch_t b, rp_batch_info_t *info);
int rp_batch_offsets(rp_batch_t b, uint64_t *off);              /* nprob + 1 entries */
int rp_batch_groups(rp_batch_t b, uint64_t *groups);            /* G_b of each problem */
int rp_batch_set_active(rp_batch_t b, const int *active_host);  /* nprob flags */
/* the basis table V_0..V_{n-1} read by the Arnoldi step and lincomb */
int rp_batch_set_basis(rp_batch_t b, void *const *V, int n);
/* res (+)= H_p vec_p for every active problem */
int rp_batch_apply(rp_batch_t b, void *res, const void *vec, int accumulate);
/* V_i /= ||V_i|| per state; nrm_dev[p * nrhs + s] = the norm (device) */
int rp_batch_normalize(rp_batch_t b, int i, double *nrm_dev);
/* Arnoldi step j: V_{j+1} = A V_j, one Gram-Schmidt pass (FP64 dots) and a
 * second for the problems with a state whose first-pass norm is below
 * eta^2 / 2 = 1/4 (the reference's selective criterion from its initial
 * norms 1 and eta / 2, decided on the device),
 * H(0..j+1, j) into H_dev (per state a column-major (kmax+1) x kmax complex
 * double matrix, states consecutive), V_{j+1} normalized; kernels only.
 * H(j+1, j) is real: its real part is the norm after the last pass (the
 * first, when the second did not run), its imaginary slot holds the norm
 * after the first pass (for the reference's breakdown test, norm < norm_1 /
 * sqrt(2): inc/algo.h:79-106). */
int rp_batch_arnoldi_step(rp_batch_t b, int j, double *H_dev);
/* y += sum_{i<k} c[rhs][i] V_i, c complex double (device), cstride doubles per state */
int rp_batch_lincomb(rp_batch_t b, void *y, int k, const double *c_dev, int cstride);
/* Arnoldi steps j0..j1-1 captured as one HIP graph (no memset or copy nodes) */
int rp_batch_capture_arnoldi(rp_batch_t b, int j0, int j1, double *H_dev, rp_graph_t *out);
int rp_graph_launch(rp_graph_t g);
int rp_graph_free(rp_graph_t g);

/* --- Krylov-schedule proposer network (no counterpart in the reference,
 * whose Krylov::evolve uses one fixed dimension for all restarts,
 * inc/algo.h:266-269).  A 12 -> 64 -> 64 -> 2 FP32 perceptron with tanh
 * hidden layers; parameters are one array of RP_MLP_NPARAMS floats laid out
 * W1[64][12], b1[64], W2[64][64], b2[64], W3[2][64], b3[2].
 * Training and inference run on the library stream; all pointers except
 * the infer call's x_host / y_host and loss are device pointers. */
#define RP_MLP_IN 12
#define RP_MLP_HIDDEN 64
#define RP_MLP_OUT 2
#define RP_MLP_NPARAMS 5122
#define RP_MLP_MAX_INFER 64
/* Y[n][2] = net(X[n][12]) */
int rp_mlp_forward(const float *params, const float *X, int64_t n, float *Y);
/* loss = sum_{s,o} W[s][o] (net(X[s])_o - T[s][o])^2 / n (W may be null:
 * all ones) and grad = d loss / d params; deterministic for a given n */
int rp_mlp_loss_grad(const float *params, const float *X, const float *T, const float *W,
                     int64_t n, float *grad, double *loss);
/* one AdamW step (decoupled weight decay), step counts from 1 */
int rp_mlp_adam(float *params, const float *grad, float *m, float *v, int64_t nparams,
                double lr, double beta1, double beta2, double eps, double weight_decay,
                int64_t step);
/* forward of n <= RP_MLP_MAX_INFER host samples; blocks until y_host is set */
int rp_mlp_infer(const float *params, const float *x_host, int n, float *y_host);

#ifdef __cplusplus
}
#endif
#endif
