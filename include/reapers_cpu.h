/*
 * reapers_cpu.h -- C ABI of the CPU backend (libreapers_cpu.so).
 *
 * The header-only C++ API built with REAPERS_NOGPU, where DefImpl is
 * CPUImpl (the reference's backend switch, inc/reapers.h:96-100, and its
 * CPUImpl, inc/cpuimpl.h:513-755).  The library links neither HIP nor any
 * GPU runtime; it exists so that programs without a C++ compiler of their
 * own (Python's ctypes) can run the host backend.  Vectors are interleaved
 * complex numbers in host memory (fp 32: complex<float>, fp 64:
 * complex<double>); term lists are SpinOp bits and interleaved
 * coefficients, as for rp_terms_upload.  Every function returns 0 or a
 * status of include/reapers_hip.h's RP_ERR_* family; rpc_last_error() gives
 * the exception's message.
 */
#ifndef REAPERS_CPU_H
#define REAPERS_CPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char *rpc_last_error(void);

/* CPUImpl::apply_ops (inc/cpuimpl.h:737-754): out (+)= H in */
int rpc_apply(int fp, int len, const uint64_t *bits, const double *coeff, uint64_t nterms,
              const void *in, void *out, int accumulate);

/* State<Fp, CPUImpl>::evolve with Krylov<CPUImpl> (inc/state.h:413-424,
 * inc/algo.h:266-365): psi = exp(-(beta + i t) H) psi in place */
int rpc_evolve(int fp, int len, const uint64_t *bits, const double *coeff, uint64_t nterms,
               void *psi, double t, double beta, int krydim, double tol, int64_t *matvecs);

/* ExactDiagonalization<CPUImpl>::evolve (inc/algo.h:28-54) through the host
 * dense forms (get_eigensystem / matexp, inc/cpuimpl.h:336-420) */
int rpc_ed_evolve(int fp, int len, const uint64_t *bits, const double *coeff, uint64_t nterms,
                  void *psi, double t, double beta);

/* Lanczos ground state (inc/algo.h:180-257) from psi; returns the energy,
 * psi becomes the ground-state estimate */
int rpc_ground_state(int fp, int len, const uint64_t *bits, const double *coeff,
                     uint64_t nterms, void *psi, int krydim, double eps, double *energy);

/* eigensystem of the dense form (HostSumOps::get_eigensystem): w ascending
 * (2^len doubles), V column-major eigenvectors (4^len complex<double>) */
int rpc_eigensystem(int len, const uint64_t *bits, const double *coeff, uint64_t nterms,
                    double *w, void *V);

/* CPUImpl::init_random: the Philox4x32-10 stream of rp_init_random */
int rpc_random_state(int fp, int len, uint64_t seed, void *out);

#ifdef __cplusplus
}
#endif
#endif
