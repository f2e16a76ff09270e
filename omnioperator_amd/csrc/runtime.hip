// Runtime part of libomni_gpu.so: error reporting, device memory helpers and kernel timing.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "omni_gpu_internal.h"

namespace omni_gpu {

namespace {
thread_local std::string g_last_error;

struct PendingKernel {
    std::string name;
    hipEvent_t start;
    hipEvent_t stop;
};

std::mutex g_prof_mu;
bool g_prof_on = false;
std::vector<PendingKernel> g_pending;
std::map<std::string, std::pair<int64_t, double>> g_prof_totals;

// Resolves finished event pairs into the per-name totals.  Caller holds g_prof_mu.
void DrainPendingLocked()
{
    for (auto &p : g_pending) {
        (void)hipEventSynchronize(p.stop);
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
            auto &t = g_prof_totals[p.name];
            t.first += 1;
            t.second += ms;
        }
        (void)hipEventDestroy(p.start);
        (void)hipEventDestroy(p.stop);
    }
    g_pending.clear();
}
}  // namespace

int SetError(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int HipFail(hipError_t e, const char *what, const char *file, int line)
{
    int code = (e == hipErrorOutOfMemory) ? OMNI_GPU_ERR_NOMEM : OMNI_GPU_ERR_HIP;
    return SetError(code, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
}

int CheckLaunch(const char *kernel)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return SetError(OMNI_GPU_ERR_HIP, "launch of %s failed: %s", kernel, hipGetErrorString(e));
    return OMNI_GPU_OK;
}

int DevAlloc(void **p, size_t bytes, hipStream_t s)
{
    hipError_t e = hipMallocAsync(p, bytes, s);
    if (e != hipSuccess) {
        *p = nullptr;
        return HipFail(e, "hipMallocAsync", __FILE__, __LINE__);
    }
    return OMNI_GPU_OK;
}

void DevFree(void *p, hipStream_t s)
{
    if (p) (void)hipFreeAsync(p, s);
}

namespace {
std::mutex g_stream_mu;
std::map<int, std::vector<hipStream_t>> g_idle_streams;
}  // namespace

int AcquireStream(hipStream_t *s)
{
    int dev = CurrentDevice();
    {
        std::lock_guard<std::mutex> g(g_stream_mu);
        auto &idle = g_idle_streams[dev];
        if (!idle.empty()) {
            *s = idle.back();
            idle.pop_back();
            return OMNI_GPU_OK;
        }
    }
    OG_HIP(hipStreamCreateWithFlags(s, hipStreamNonBlocking));
    return OMNI_GPU_OK;
}

void ReleaseStream(hipStream_t s)
{
    if (!s) return;
    if (hipStreamSynchronize(s) != hipSuccess) {
        (void)hipStreamDestroy(s);  // a stream in an error state is not reused
        return;
    }
    int dev = 0;
    if (hipStreamGetDevice(s, &dev) != hipSuccess) {
        (void)hipStreamDestroy(s);
        return;
    }
    std::lock_guard<std::mutex> g(g_stream_mu);
    g_idle_streams[dev].push_back(s);
}

int DevBuf::Grow(size_t need, size_t keep, hipStream_t s)
{
    if (need <= bytes && p) return OMNI_GPU_OK;
    void *np = nullptr;
    size_t sz = need < 256 ? 256 : need;
    OG_TRY(DevAlloc(&np, sz, s));
    if (p && keep) OG_HIP(hipMemcpyAsync(np, p, keep < bytes ? keep : bytes, hipMemcpyDeviceToDevice, s));
    if (p) DevFree(p, s);
    p = np;
    bytes = sz;
    return OMNI_GPU_OK;
}

KernelScope::KernelScope(hipStream_t s, const char *name) : s_(s), name_(name)
{
    if (!g_prof_on) return;
    if (hipEventCreate(&start_) != hipSuccess) {
        start_ = nullptr;
        return;
    }
    (void)hipEventRecord(start_, s_);
}

KernelScope::~KernelScope()
{
    if (!start_) return;
    hipEvent_t stop = nullptr;
    if (hipEventCreate(&stop) != hipSuccess) {
        (void)hipEventDestroy(start_);
        return;
    }
    (void)hipEventRecord(stop, s_);
    std::lock_guard<std::mutex> g(g_prof_mu);
    g_pending.push_back({name_, start_, stop});
    if (g_pending.size() > 4096) DrainPendingLocked();
}

int CurrentDevice()
{
    int d = 0;
    (void)hipGetDevice(&d);
    return d;
}

}  // namespace omni_gpu

using namespace omni_gpu;

extern "C" {

int omni_gpu_init(int device)
{
    int n = 0;
    OG_HIP(hipGetDeviceCount(&n));
    OG_CHECK(device >= 0 && device < n, OMNI_GPU_ERR_INVALID, "device %d out of range (%d devices)", device, n);
    OG_HIP(hipSetDevice(device));
    hipMemPool_t pool;
    OG_HIP(hipDeviceGetDefaultMemPool(&pool, device));
    uint64_t threshold = UINT64_MAX;
    OG_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &threshold));
    return OMNI_GPU_OK;
}

const char *omni_gpu_last_error(void) { return g_last_error.c_str(); }

int omni_gpu_malloc(void **ptr, size_t bytes)
{
    OG_CHECK(ptr != nullptr, OMNI_GPU_ERR_INVALID, "null output pointer");
    OG_HIP(hipMalloc(ptr, bytes == 0 ? 1 : bytes));
    return OMNI_GPU_OK;
}

int omni_gpu_free(void *ptr)
{
    if (ptr) OG_HIP(hipFree(ptr));
    return OMNI_GPU_OK;
}

int omni_gpu_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    if (bytes) OG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return OMNI_GPU_OK;
}

int omni_gpu_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    if (bytes) OG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return OMNI_GPU_OK;
}

int omni_gpu_memcpy_d2d(void *dst, const void *src, size_t bytes)
{
    if (bytes) OG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return OMNI_GPU_OK;
}

int omni_gpu_memset(void *dst, int value, size_t bytes)
{
    if (bytes) OG_HIP(hipMemset(dst, value, bytes));
    return OMNI_GPU_OK;
}

int omni_gpu_device_synchronize(void)
{
    OG_HIP(hipDeviceSynchronize());
    return OMNI_GPU_OK;
}

int omni_gpu_profile_enable(int on)
{
    std::lock_guard<std::mutex> g(g_prof_mu);
    g_prof_on = on != 0;
    return OMNI_GPU_OK;
}

int omni_gpu_profile_reset(void)
{
    std::lock_guard<std::mutex> g(g_prof_mu);
    DrainPendingLocked();
    g_prof_totals.clear();
    return OMNI_GPU_OK;
}

int omni_gpu_profile_get(const char *name_prefix, int64_t *launches, double *total_ms)
{
    OG_CHECK(name_prefix && launches && total_ms, OMNI_GPU_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(g_prof_mu);
    DrainPendingLocked();
    int64_t n = 0;
    double ms = 0;
    size_t plen = strlen(name_prefix);
    for (auto &kv : g_prof_totals) {
        if (kv.first.compare(0, plen, name_prefix) == 0) {
            n += kv.second.first;
            ms += kv.second.second;
        }
    }
    *launches = n;
    *total_ms = ms;
    return OMNI_GPU_OK;
}

}  // extern "C"
