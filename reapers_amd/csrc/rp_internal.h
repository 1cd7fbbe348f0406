// rp_internal.h -- shared state of libreapers_hip.so (not installed).
//
// One process-wide context on one device, like the reference's GPUContext
// singleton (inc/gpuimpl.h:496-559), but with an explicit stream of default
// priority, a device workspace for reduction partials, pinned host memory for
// scalar results, and optional per-launch event timing.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>
#include <vector>

#include "../../include/reapers_hip.h"

namespace rp {

struct float2c { float x, y; };
struct double2c { double x, y; };

struct ProfPending {
    int cls;
    hipEvent_t start, stop;
};

struct Context {
    bool initialized = false;
    int device = 0;
    int num_cus = 0;
    std::string name;
    hipStream_t stream = nullptr;
    // device workspace: reduction partials and small device-side results
    void *work = nullptr;
    size_t work_bytes = 0;
    // pinned host buffer for scalars copied back to the host
    double *pinned = nullptr;
    size_t pinned_doubles = 0;
    // partial rows of the apply's Gram-Schmidt epilogue (rp_apply_gs),
    // allocated on first use
    void *epi_work = nullptr;
    size_t epi_bytes = 0;
    // profiling
    bool prof = false;
    int prof_mask = 7;                // kernel classes bracketed (bit = class)
    uint64_t prof_launches[3] = {0, 0, 0};
    double prof_ms[3] = {0, 0, 0};
    std::vector<ProfPending> pending;
    std::vector<hipEvent_t> event_pool;
};

Context &ctx();

// record the last error of this thread and return the status code
int set_error(int code, const char *fmt, ...);

// Device allocation for every buffer of the library: hipMalloc, and when
// that runs out of memory, the blocks kept by rp_free are returned to HIP
// and the allocation is tried once more.
hipError_t dev_malloc(void **ptr, size_t bytes);
// bytes held in freed blocks kept for reuse
size_t pooled_bytes();
// Return the kept blocks to HIP when less than `bytes` of device memory is
// free: for libraries that allocate on their own (rocBLAS, rocSOLVER) and
// cannot fall back on the pool.
void ensure_device_headroom(size_t bytes);
int check_init();
// take the address of rp_vec.hip's file-scope scratch array on the current
// device (once per device, at rp_init)
int vec_scratch_setup(int device);
// the apply with the Gram-Schmidt dots in its epilogue (rp_apply.hip); rows
// = partial rows written to part (row stride width doubles), 0 if the term
// handle's layout has no such kernel
int apply_dots_launch(rp_terms_t h, void *res, const void *vec, const void *const *V, int k,
                      double *part, int width, uint64_t max_rows, uint64_t *rows);

// launch bracketing for the optional per-class timing
void prof_begin(int cls, hipEvent_t *start, hipEvent_t *stop);
void prof_end(int cls, hipEvent_t start, hipEvent_t stop);
int prof_collect();   // must be called after a stream synchronization

// blocks of the persistent-style reduction kernels
constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;

inline size_t elem_bytes(int prec) { return prec == RP_C64 ? 8 : 16; }
inline bool valid_prec(int prec) { return prec == RP_C64 || prec == RP_C128; }

}  // namespace rp

#define RP_CHECK_HIP(expr)                                                        \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess)                                                     \
            return ::rp::set_error(RP_ERR_RUNTIME, "%s:%d: %s failed: %s",       \
                                   __FILE__, __LINE__, #expr,                     \
                                   hipGetErrorString(e_));                        \
    } while (0)

#define RP_CHECK_LAUNCH()                                                         \
    do {                                                                          \
        hipError_t e_ = hipGetLastError();                                        \
        if (e_ != hipSuccess)                                                     \
            return ::rp::set_error(RP_ERR_RUNTIME, "%s:%d: kernel launch failed: %s", \
                                   __FILE__, __LINE__, hipGetErrorString(e_));    \
    } while (0)

#define RP_REQUIRE(cond, msg)                                                     \
    do {                                                                          \
        if (!(cond))                                                              \
            return ::rp::set_error(RP_ERR_INVALID_ARGUMENT, "%s: %s", __func__, msg); \
    } while (0)

// launch a kernel on the library stream, bracketed by timing events when
// profiling is on
#define RP_LAUNCH(cls, kernel, grid, block, shmem, ...)                           \
    do {                                                                          \
        ::rp::Context &c_ = ::rp::ctx();                                          \
        hipEvent_t s_ = nullptr, t_ = nullptr;                                    \
        const bool p_ = c_.prof && ((c_.prof_mask >> (cls)) & 1);                 \
        if (p_) ::rp::prof_begin((cls), &s_, &t_);                                \
        hipLaunchKernelGGL(kernel, (grid), (block), (shmem), c_.stream, __VA_ARGS__); \
        RP_CHECK_LAUNCH();                                                        \
        if (p_) ::rp::prof_end((cls), s_, t_);                                    \
    } while (0)
