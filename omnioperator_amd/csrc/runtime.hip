// Runtime part of libomni_gpu.so: error reporting, device memory helpers and kernel timing.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <unordered_map>
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

// Timing events are recycled (created once per device, never destroyed): creating two events per launch
// would add host time to every timed launch.  Caller of the Locked forms holds g_prof_mu.
std::map<int, std::vector<hipEvent_t>> g_free_events;

void GiveEventLocked(hipEvent_t e)
{
    g_free_events[CurrentDevice()].push_back(e);
}

void GiveEvent(hipEvent_t e)
{
    std::lock_guard<std::mutex> g(g_prof_mu);
    GiveEventLocked(e);
}

hipEvent_t TakeEvent()
{
    {
        std::lock_guard<std::mutex> g(g_prof_mu);
        auto &v = g_free_events[CurrentDevice()];
        if (!v.empty()) {
            hipEvent_t e = v.back();
            v.pop_back();
            return e;
        }
    }
    hipEvent_t e = nullptr;
    return hipEventCreate(&e) == hipSuccess ? e : nullptr;
}

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
        GiveEventLocked(p.start);
        GiveEventLocked(p.stop);
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

namespace {
// Device block cache.  hipFreeAsync of a large block costs hundreds of microseconds on this runtime even
// with the pool's release threshold raised (the pool checks and trims physical memory on the call), which
// is more than a whole operator's kernels on mid-sized inputs.  Freed blocks are therefore kept here and
// handed out again by size class:
//  - a block freed on stream s is reusable at once by later work on s (stream order protects it);
//  - once s has been synchronized (ReleaseStream), it is reusable on every stream of its device.
// Blocks beyond kCacheLimit cached bytes are returned to the pool instead.
struct BlockCache {
    std::mutex mu;
    std::unordered_map<void *, std::pair<size_t, int>> live;           // block -> (size class, device)
    std::map<hipStream_t, std::vector<std::pair<size_t, void *>>> on_stream;  // freed, stream-ordered
    std::map<int, std::multimap<size_t, void *>> ready;                // freed and synchronized
    size_t cached = 0;
};
BlockCache g_cache;
constexpr size_t kCacheLimit = static_cast<size_t>(16) << 30;

size_t SizeClass(size_t n)
{
    if (n <= (static_cast<size_t>(1) << 20)) {
        size_t c = 512;
        while (c < n) c <<= 1;
        return c;
    }
    const size_t g = static_cast<size_t>(2) << 20;
    return (n + g - 1) / g * g;
}

bool TakeCached(size_t cls, int dev, hipStream_t s, void **p)
{
    std::lock_guard<std::mutex> g(g_cache.mu);
    auto it = g_cache.on_stream.find(s);
    if (it != g_cache.on_stream.end()) {
        auto &v = it->second;
        for (size_t i = v.size(); i-- > 0;) {
            if (v[i].first == cls) {
                *p = v[i].second;
                v.erase(v.begin() + static_cast<std::ptrdiff_t>(i));
                g_cache.cached -= cls;
                return true;
            }
        }
    }
    auto &r = g_cache.ready[dev];
    auto f = r.find(cls);
    if (f == r.end()) return false;
    *p = f->second;
    r.erase(f);
    g_cache.cached -= cls;
    return true;
}

// Returns every cached block of `dev` to the runtime (after the device has drained, so that no queued work
// still uses a block freed on some stream).
void TrimDevice(int dev)
{
    (void)hipDeviceSynchronize();
    std::lock_guard<std::mutex> g(g_cache.mu);
    for (auto &kv : g_cache.ready[dev]) {
        (void)hipFree(kv.second);
        g_cache.live.erase(kv.second);
        g_cache.cached -= kv.first;
    }
    g_cache.ready[dev].clear();
    for (auto &sv : g_cache.on_stream) {
        auto &v = sv.second;
        for (size_t i = v.size(); i-- > 0;) {
            auto lv = g_cache.live.find(v[i].second);
            if (lv == g_cache.live.end() || lv->second.second != dev) continue;
            (void)hipFree(v[i].second);
            g_cache.live.erase(lv);
            g_cache.cached -= v[i].first;
            v.erase(v.begin() + static_cast<std::ptrdiff_t>(i));
        }
    }
}
}  // namespace

int DevAlloc(void **p, size_t bytes, hipStream_t s)
{
    const int dev = CurrentDevice();
    const size_t cls = SizeClass(bytes);
    if (TakeCached(cls, dev, s, p)) return OMNI_GPU_OK;
    hipError_t e = hipMallocAsync(p, cls, s);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        TrimDevice(dev);
        e = hipMallocAsync(p, cls, s);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        return HipFail(e, "hipMallocAsync", __FILE__, __LINE__);
    }
    std::lock_guard<std::mutex> g(g_cache.mu);
    g_cache.live[*p] = {cls, dev};
    return OMNI_GPU_OK;
}

void DevFree(void *p, hipStream_t s)
{
    if (!p) return;
    {
        std::lock_guard<std::mutex> g(g_cache.mu);
        auto it = g_cache.live.find(p);
        if (it != g_cache.live.end() && g_cache.cached + it->second.first <= kCacheLimit) {
            g_cache.on_stream[s].push_back({it->second.first, p});
            g_cache.cached += it->second.first;
            return;
        }
        if (it != g_cache.live.end()) g_cache.live.erase(it);
    }
    (void)hipFreeAsync(p, s);
}

void MemoryStats(int64_t *in_use, int64_t *cached)
{
    std::lock_guard<std::mutex> g(g_cache.mu);
    size_t live = 0;
    for (const auto &kv : g_cache.live) live += kv.second.first;
    *in_use = static_cast<int64_t>(live - g_cache.cached);
    *cached = static_cast<int64_t>(g_cache.cached);
}

namespace {
// Pinned staging block per stream for small host<->device transfers: a copy from pageable memory costs
// tens of microseconds of host time on this runtime, one from pinned memory is an ordinary async copy.
// Space is handed out in order and recycled whenever the stream is known to have drained.
struct Staging {
    uint8_t *base = nullptr;
    size_t used = 0;
};
constexpr size_t kStagingBytes = static_cast<size_t>(1) << 20;
std::mutex g_staging_mu;
std::map<hipStream_t, Staging> g_staging;

Staging *StagingOf(hipStream_t s)
{
    std::lock_guard<std::mutex> g(g_staging_mu);
    Staging &st = g_staging[s];
    if (!st.base) {
        void *p = nullptr;
        if (hipHostMalloc(&p, kStagingBytes, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            g_staging.erase(s);
            return nullptr;
        }
        st.base = static_cast<uint8_t *>(p);
        st.used = 0;
    }
    return &st;
}

void ResetStaging(hipStream_t s)
{
    std::lock_guard<std::mutex> g(g_staging_mu);
    auto it = g_staging.find(s);
    if (it != g_staging.end()) it->second.used = 0;
}

void DropStaging(hipStream_t s)
{
    std::lock_guard<std::mutex> g(g_staging_mu);
    auto it = g_staging.find(s);
    if (it == g_staging.end()) return;
    if (it->second.base) (void)hipHostFree(it->second.base);
    g_staging.erase(it);
}

constexpr size_t Align64(size_t n) { return (n + 63) & ~static_cast<size_t>(63); }
}  // namespace

int SyncStream(hipStream_t s)
{
    OG_HIP(hipStreamSynchronize(s));
    ResetStaging(s);
    return OMNI_GPU_OK;
}

int CopyToDevice(void *dst, const void *src, size_t n, hipStream_t s)
{
    if (n == 0) return OMNI_GPU_OK;
    Staging *st = StagingOf(s);
    if (!st || Align64(n) > kStagingBytes) {
        OG_HIP(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, s));
        return OMNI_GPU_OK;
    }
    if (st->used + Align64(n) > kStagingBytes) OG_TRY(SyncStream(s));
    uint8_t *slot = st->base + st->used;
    st->used += Align64(n);
    std::memcpy(slot, src, n);
    OG_HIP(hipMemcpyAsync(dst, slot, n, hipMemcpyHostToDevice, s));
    return OMNI_GPU_OK;
}

int CopyToHost(hipStream_t s, std::initializer_list<HostRead> reads)
{
    size_t total = 0;
    for (const HostRead &r : reads) total += Align64(r.n);
    Staging *st = StagingOf(s);
    if (!st || total > kStagingBytes) {
        for (const HostRead &r : reads) {
            if (r.n) OG_HIP(hipMemcpyAsync(r.dst, r.src, r.n, hipMemcpyDeviceToHost, s));
        }
        return SyncStream(s);
    }
    if (st->used + total > kStagingBytes) OG_TRY(SyncStream(s));
    size_t off = st->used;
    for (const HostRead &r : reads) {
        if (r.n) OG_HIP(hipMemcpyAsync(st->base + off, r.src, r.n, hipMemcpyDeviceToHost, s));
        off += Align64(r.n);
    }
    OG_HIP(hipStreamSynchronize(s));
    off = st->used;
    for (const HostRead &r : reads) {
        if (r.n) std::memcpy(r.dst, st->base + off, r.n);
        off += Align64(r.n);
    }
    st->used = 0;
    return OMNI_GPU_OK;
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
        // a stream in an error state is not reused; the blocks freed on it are dropped (leaked), since
        // whether its work still touches them is unknown
        {
            std::lock_guard<std::mutex> g(g_cache.mu);
            auto it = g_cache.on_stream.find(s);
            if (it != g_cache.on_stream.end()) {
                for (auto &b : it->second) {
                    g_cache.live.erase(b.second);
                    g_cache.cached -= b.first;
                }
                g_cache.on_stream.erase(it);
            }
        }
        DropStaging(s);
        (void)hipStreamDestroy(s);
        return;
    }
    ResetStaging(s);
    int dev = 0;
    if (hipStreamGetDevice(s, &dev) != hipSuccess) {
        DropStaging(s);
        (void)hipStreamDestroy(s);
        return;
    }
    {
        // the stream has drained: blocks freed on it become reusable on any stream of the device
        std::lock_guard<std::mutex> g(g_cache.mu);
        auto it = g_cache.on_stream.find(s);
        if (it != g_cache.on_stream.end()) {
            auto &r = g_cache.ready[dev];
            for (auto &b : it->second) r.insert({b.first, b.second});
            g_cache.on_stream.erase(it);
        }
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
    start_ = TakeEvent();
    if (start_) (void)hipEventRecord(start_, s_);
}

KernelScope::~KernelScope()
{
    if (!start_) return;
    hipEvent_t stop = TakeEvent();
    if (!stop) {
        GiveEvent(start_);
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

// The copies and the memset run on the null stream, which does not order the library's own streams (they are
// non-blocking), and hipMemset / pageable hipMemcpy may return before the device has finished.  Each call
// therefore waits for the null stream, so that work a handle queues afterwards never races with it.
int omni_gpu_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    if (bytes) {
        OG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
        OG_HIP(hipStreamSynchronize(nullptr));
    }
    return OMNI_GPU_OK;
}

int omni_gpu_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    if (bytes) {
        OG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        OG_HIP(hipStreamSynchronize(nullptr));
    }
    return OMNI_GPU_OK;
}

int omni_gpu_memcpy_d2d(void *dst, const void *src, size_t bytes)
{
    if (bytes) {
        OG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
        OG_HIP(hipStreamSynchronize(nullptr));
    }
    return OMNI_GPU_OK;
}

int omni_gpu_memset(void *dst, int value, size_t bytes)
{
    if (bytes) {
        OG_HIP(hipMemset(dst, value, bytes));
        OG_HIP(hipStreamSynchronize(nullptr));
    }
    return OMNI_GPU_OK;
}

int omni_gpu_host_alloc(void **ptr, size_t bytes)
{
    OG_CHECK(ptr != nullptr, OMNI_GPU_ERR_INVALID, "null output pointer");
    OG_HIP(hipHostMalloc(ptr, bytes == 0 ? 1 : bytes, hipHostMallocDefault));
    return OMNI_GPU_OK;
}

int omni_gpu_host_free(void *ptr)
{
    if (ptr) OG_HIP(hipHostFree(ptr));
    return OMNI_GPU_OK;
}

int omni_gpu_device_synchronize(void)
{
    OG_HIP(hipDeviceSynchronize());
    return OMNI_GPU_OK;
}

int omni_gpu_memory_stats(int64_t *in_use_bytes, int64_t *cached_bytes)
{
    OG_CHECK(in_use_bytes && cached_bytes, OMNI_GPU_ERR_INVALID, "null argument");
    MemoryStats(in_use_bytes, cached_bytes);
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
