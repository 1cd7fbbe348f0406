// rp_runtime.hip -- context, errors, memory and kernel timing of
// libreapers_hip.so.  Replaces the reference's GPUContext
// (inc/gpuimpl.h:496-559) and the CUDA_CALL / cuda_alloc error handling
// (inc/gpuimpl.h:26-66).
#include "rp_internal.h"

#include <cstdarg>
#include <map>
#include <mutex>
#include <unordered_map>
#include <cstdio>
#include <cstring>

namespace rp {

static Context g_ctx;
static thread_local char g_err[1024] = "";

Context &ctx() { return g_ctx; }

int set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int check_init()
{
    if (!g_ctx.initialized)
        return set_error(RP_ERR_NOT_INITIALIZED, "rp_init() has not been called");
    return RP_OK;
}

static hipEvent_t pool_get()
{
    if (!g_ctx.event_pool.empty()) {
        hipEvent_t e = g_ctx.event_pool.back();
        g_ctx.event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void prof_begin(int cls, hipEvent_t *start, hipEvent_t *stop)
{
    (void)cls;
    *start = pool_get();
    *stop = pool_get();
    if (*start) (void)hipEventRecord(*start, g_ctx.stream);
}

void prof_end(int cls, hipEvent_t start, hipEvent_t stop)
{
    if (!start || !stop) return;
    (void)hipEventRecord(stop, g_ctx.stream);
    g_ctx.pending.push_back({cls, start, stop});
}

int prof_collect()
{
    for (auto &p : g_ctx.pending) {
        float ms = 0.f;
        hipError_t e = hipEventElapsedTime(&ms, p.start, p.stop);
        if (e == hipSuccess) {
            g_ctx.prof_launches[p.cls] += 1;
            g_ctx.prof_ms[p.cls] += (double)ms;
        }
        g_ctx.event_pool.push_back(p.start);
        g_ctx.event_pool.push_back(p.stop);
    }
    g_ctx.pending.clear();
    return RP_OK;
}

}  // namespace rp

using namespace rp;

extern "C" {

const char *rp_last_error(void) { return g_err; }

int rp_init(int device)
{
    Context &c = ctx();
    if (c.initialized) {
        if (c.device == device) return RP_OK;
        return set_error(RP_ERR_INVALID_ARGUMENT,
                         "rp_init: already initialized on device %d", c.device);
    }
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        return set_error(RP_ERR_NO_DEVICE, "rp_init: no GPU device available (%s)",
                         hipGetErrorString(e));
    if (device < 0 || device >= count)
        return set_error(RP_ERR_INVALID_ARGUMENT, "rp_init: device %d out of range [0,%d)",
                         device, count);
    RP_CHECK_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    RP_CHECK_HIP(hipGetDeviceProperties(&prop, device));
    c.device = device;
    c.num_cus = prop.multiProcessorCount;
    c.name = prop.gcnArchName;
    // default-priority stream, not synchronizing with the legacy null stream
    RP_CHECK_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    // partials for up to 64 complex FP64 values and one norm per reduction
    // block, plus room for the reduced results
    c.work_bytes = (size_t)kRedBlocks * (2 * 64 + 1) * sizeof(double) + 64 * 1024;
    RP_CHECK_HIP(dev_malloc(&c.work, c.work_bytes));
    c.pinned_doubles = 4096;
    RP_CHECK_HIP(hipHostMalloc((void **)&c.pinned, c.pinned_doubles * sizeof(double),
                               hipHostMallocDefault));
    int st = vec_scratch_setup(device);
    if (st) return st;
    c.initialized = true;
    return RP_OK;
}

}  // extern "C"

namespace {
size_t release_pool();
}

extern "C" {

int rp_shutdown(void)
{
    Context &c = ctx();
    if (!c.initialized) return RP_OK;
    (void)hipStreamSynchronize(c.stream);
    prof_collect();
    for (auto e : c.event_pool) (void)hipEventDestroy(e);
    c.event_pool.clear();
    (void)release_pool();
    if (c.work) (void)hipFree(c.work);
    if (c.epi_work) (void)hipFree(c.epi_work);
    if (c.pinned) (void)hipHostFree(c.pinned);
    if (c.stream) (void)hipStreamDestroy(c.stream);
    c.work = nullptr;
    c.epi_work = nullptr;
    c.epi_bytes = 0;
    c.pinned = nullptr;
    c.stream = nullptr;
    c.initialized = false;
    return RP_OK;
}

int rp_synchronize(void)
{
    int st = check_init();
    if (st) return st;
    RP_CHECK_HIP(hipStreamSynchronize(ctx().stream));
    return prof_collect();
}

int rp_mem_get_info(size_t *free_bytes, size_t *total_bytes)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(free_bytes && total_bytes, "null output pointer");
    RP_CHECK_HIP(hipMemGetInfo(free_bytes, total_bytes));
    // blocks kept for reuse count as free: an allocation gets them back
    *free_bytes += pooled_bytes();
    return RP_OK;
}

int rp_device_props(int *num_cus, char *name, size_t name_len)
{
    int st = check_init();
    if (st) return st;
    if (num_cus) *num_cus = ctx().num_cus;
    if (name && name_len) {
        strncpy(name, ctx().name.c_str(), name_len - 1);
        name[name_len - 1] = 0;
    }
    return RP_OK;
}

}  // extern "C"

// Blocks of rp_malloc are kept on rp_free, by size, and handed out again to
// an rp_malloc of the same size: freeing device memory makes the driver
// clear it before it can be allocated again, which for the tens of 8.6 GB
// Krylov vectors of one 30-qubit evolution costs seconds per call.  The
// pool is returned to HIP whenever an allocation runs out of memory, and
// at rp_shutdown.
namespace {
std::mutex g_mem_mu;
std::unordered_map<void *, size_t> g_live;        // rp_malloc blocks in use
std::multimap<size_t, void *> g_pool;             // freed blocks by size
size_t g_pool_bytes = 0;

size_t release_pool()
{
    std::lock_guard<std::mutex> lk(g_mem_mu);
    const size_t n = g_pool_bytes;
    for (auto &kv : g_pool) (void)hipFree(kv.second);
    g_pool.clear();
    g_pool_bytes = 0;
    return n;
}
}  // namespace

namespace rp {
hipError_t dev_malloc(void **ptr, size_t bytes)
{
    hipError_t e = hipMalloc(ptr, bytes);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        if (release_pool() > 0) e = hipMalloc(ptr, bytes);
    }
    return e;
}

size_t pooled_bytes()
{
    std::lock_guard<std::mutex> lk(g_mem_mu);
    return g_pool_bytes;
}

void ensure_device_headroom(size_t bytes)
{
    if (pooled_bytes() == 0) return;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        (void)hipGetLastError();
        return;
    }
    if (free_b < bytes) (void)release_pool();
}
}  // namespace rp

extern "C" {

int rp_malloc(void **ptr, size_t bytes)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(ptr, "null output pointer");
    *ptr = nullptr;
    if (bytes == 0) return RP_OK;
    {
        std::lock_guard<std::mutex> lk(g_mem_mu);
        auto it = g_pool.find(bytes);
        if (it != g_pool.end()) {
            *ptr = it->second;
            g_pool.erase(it);
            g_pool_bytes -= bytes;
            g_live[*ptr] = bytes;
            return RP_OK;
        }
    }
    hipError_t e = dev_malloc(ptr, bytes);
    if (e == hipErrorOutOfMemory || (e == hipSuccess && *ptr == nullptr)) {
        (void)hipGetLastError();
        *ptr = nullptr;
        return set_error(RP_ERR_OUT_OF_MEMORY,
                         "Failed to allocate %zu bytes on GPU device.", bytes);
    }
    RP_CHECK_HIP(e);
    std::lock_guard<std::mutex> lk(g_mem_mu);
    g_live[*ptr] = bytes;
    return RP_OK;
}

int rp_free(void *ptr)
{
    int st = check_init();
    if (st) return st;
    if (!ptr) return RP_OK;
    // work queued on the library stream may still read the buffer; after
    // this wait a reuse of the block is ordered after it
    RP_CHECK_HIP(hipStreamSynchronize(ctx().stream));
    {
        std::lock_guard<std::mutex> lk(g_mem_mu);
        auto it = g_live.find(ptr);
        if (it != g_live.end()) {
            g_pool.emplace(it->second, ptr);
            g_pool_bytes += it->second;
            g_live.erase(it);
            return RP_OK;
        }
    }
    RP_CHECK_HIP(hipFree(ptr));
    return RP_OK;
}

int rp_memcpy(void *dst, const void *src, size_t bytes, int direction)
{
    int st = check_init();
    if (st) return st;
    if (bytes == 0) return RP_OK;
    RP_REQUIRE(dst && src, "null pointer");
    hipMemcpyKind kind;
    switch (direction) {
    case RP_H2D: kind = hipMemcpyHostToDevice; break;
    case RP_D2H: kind = hipMemcpyDeviceToHost; break;
    case RP_D2D: kind = hipMemcpyDeviceToDevice; break;
    default: return set_error(RP_ERR_INVALID_ARGUMENT, "rp_memcpy: bad direction %d", direction);
    }
    RP_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, kind, ctx().stream));
    if (direction != RP_D2D) RP_CHECK_HIP(hipStreamSynchronize(ctx().stream));
    return RP_OK;
}

int rp_prof_enable(int on)
{
    int st = check_init();
    if (st) return st;
    ctx().prof = on != 0;
    // RP_PROF_CLASSES(mask) selects kernel classes; any other nonzero value all
    ctx().prof_mask = (on & RP_PROF_SELECT) ? (on & 0xff) : 0xff;
    return RP_OK;
}

int rp_prof_reset(void)
{
    int st = check_init();
    if (st) return st;
    RP_CHECK_HIP(hipStreamSynchronize(ctx().stream));
    prof_collect();
    for (int i = 0; i < 3; i++) {
        ctx().prof_launches[i] = 0;
        ctx().prof_ms[i] = 0;
    }
    return RP_OK;
}

int rp_prof_query(int kernel_class, uint64_t *launches, double *total_ms)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(kernel_class >= 0 && kernel_class < 3, "bad kernel class");
    RP_CHECK_HIP(hipStreamSynchronize(ctx().stream));
    prof_collect();
    if (launches) *launches = ctx().prof_launches[kernel_class];
    if (total_ms) *total_ms = ctx().prof_ms[kernel_class];
    return RP_OK;
}

}  // extern "C"
