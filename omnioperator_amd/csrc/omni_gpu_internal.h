// Internal helpers shared by the HIP translation units of libomni_gpu.so (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <initializer_list>
#include <string>

#include "../../include/omni_gpu.h"

// Global (HBM) address space qualifier for pointers the compiler cannot otherwise prove global, so that loads
// through them are global rather than flat loads.
#ifndef OG_GLOBAL
#define OG_GLOBAL __attribute__((address_space(1)))
#endif

// Lanes per wavefront (CDNA: 64).  Overridable only so that a CPU debugging build can run each thread as its
// own wave.
#ifndef OG_WAVE_SIZE
#define OG_WAVE_SIZE 64
#endif

// Dynamic LDS array of a kernel (its size is the launch's dynamic shared-memory bytes).
#ifndef OG_DYNAMIC_LDS
#define OG_DYNAMIC_LDS(T, name) extern __shared__ T name[]
#endif

namespace omni_gpu {

constexpr int kWaveSize = OG_WAVE_SIZE;
constexpr int kXcds = 8;  // MI355X: workgroup b of a launch runs on XCD b mod 8, each XCD with its own L2

// ---- errors ---------------------------------------------------------------------------------------------
int SetError(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int HipFail(hipError_t e, const char *what, const char *file, int line);

#define OG_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t og_e_ = (call);                                                      \
        if (og_e_ != hipSuccess) return ::omni_gpu::HipFail(og_e_, #call, __FILE__, __LINE__); \
    } while (0)

#define OG_CHECK(cond, code, ...)                                 \
    do {                                                          \
        if (!(cond)) return ::omni_gpu::SetError(code, __VA_ARGS__); \
    } while (0)

#define OG_TRY(expr)              \
    do {                          \
        int og_rc_ = (expr);      \
        if (og_rc_ != OMNI_GPU_OK) return og_rc_; \
    } while (0)

// Clears the error state left by a failed launch so the next check reports the right call.
int CheckLaunch(const char *kernel);

// ---- device memory ---------------------------------------------------------------------------------------
// Stream-ordered device blocks.  Allocation rounds up to a size class and first reuses a cached block (one
// freed earlier on the same stream, or on any stream that has since been synchronized); only a miss reaches
// hipMallocAsync.  DevFree keeps the block for reuse instead of returning it to the runtime.
int DevAlloc(void **p, size_t bytes, hipStream_t s);
void DevFree(void *p, hipStream_t s);
// Bytes of device blocks the library holds for live buffers, and bytes of freed blocks kept for reuse.
void MemoryStats(int64_t *in_use, int64_t *cached);

// Per-device pool of non-blocking streams (default priority).  Each handle owns one stream while it
// lives; creating and destroying a HIP stream per operator would cost more host time than the operator's
// own work on small inputs.  ReleaseStream synchronizes the stream before returning it to the pool.
int AcquireStream(hipStream_t *s);
void ReleaseStream(hipStream_t s);

// Small transfers through the stream's pinned staging block.  CopyToDevice is stream-ordered and the host
// source may be reused at once; CopyToHost enqueues every read, synchronizes the stream once and then fills
// the destinations.  SyncStream is hipStreamSynchronize that also recycles the staging space.
struct HostRead {
    void *dst;
    const void *src;
    size_t n;
};
int CopyToDevice(void *dst, const void *src, size_t n, hipStream_t s);
int CopyToHost(hipStream_t s, std::initializer_list<HostRead> reads);
int SyncStream(hipStream_t s);

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    // Grows to at least `need` bytes; contents are NOT preserved.
    int Reserve(size_t need, hipStream_t s)
    {
        if (need <= bytes && p) return OMNI_GPU_OK;
        if (p) DevFree(p, s);
        p = nullptr;
        bytes = 0;
        size_t sz = need < 256 ? 256 : need;
        OG_TRY(DevAlloc(&p, sz, s));
        bytes = sz;
        return OMNI_GPU_OK;
    }
    // Grows to at least `need` bytes keeping the first `keep` bytes.
    int Grow(size_t need, size_t keep, hipStream_t s);
    void Release(hipStream_t s)
    {
        if (p) DevFree(p, s);
        p = nullptr;
        bytes = 0;
    }
    template <typename T> T *as() const { return static_cast<T *>(p); }
};

// ---- batched-stream advisor (learned.hip) ----------------------------------------------------------------------
// On stream s: per-batch sketch features of the key column of a batched stream (batch b: rows
// [offsets[b], offsets[b + 1]), device offsets) into X[nb][16]; when `mlp` (an omni_gpu_mlp_create handle) is given,
// the network's outputs into Y[nb][4] and, when `modes` is given, each batch's first phase (0 = per-batch LDS
// table when at most `limit` groups are predicted, else 2 = hash-partitioned passes).  Batches of at most
// `min_rows` rows get zero features and no outputs (their callers decide them without the network).  The sketch
// and the network's inference of a batch run in one workgroup of one kernel.  Kernels only (no memset,
// copy or synchronisation): capturable in a graph.
int LaunchBatchAdvisor(hipStream_t s, void *mlp, const void *keys, const uint8_t *nulls, int width,
                       const int64_t *offsets, int nb, float *X, float *Y, int32_t *modes, float limit,
                       int64_t min_rows);

// ---- kernel timing ----------------------------------------------------------------------------------------
// When profiling is on, a KernelScope brackets one launch with HIP events on the launch stream.
class KernelScope {
public:
    KernelScope(hipStream_t s, const char *name);
    ~KernelScope();

private:
    hipStream_t s_;
    const char *name_;
    hipEvent_t start_ = nullptr;
};

int CurrentDevice();

// ---- type helpers -----------------------------------------------------------------------------------------
inline int TypeWidth(int32_t t)
{
    switch (t) {
        case OMNI_TYPE_SHORT: return 2;
        case OMNI_TYPE_INT:
        case OMNI_TYPE_DATE32: return 4;
        case OMNI_TYPE_LONG:
        case OMNI_TYPE_TIMESTAMP:
        case OMNI_TYPE_DECIMAL64:
        case OMNI_TYPE_DOUBLE: return 8;
        case OMNI_TYPE_DECIMAL128: return 16;
        case OMNI_TYPE_BOOLEAN:
        case OMNI_TYPE_BYTE: return 1;
        default: return 0;
    }
}

inline bool IsIntegerKeyType(int32_t t)
{
    return t == OMNI_TYPE_SHORT || t == OMNI_TYPE_INT || t == OMNI_TYPE_DATE32 || t == OMNI_TYPE_LONG ||
           t == OMNI_TYPE_TIMESTAMP || t == OMNI_TYPE_DECIMAL64;
}

// ---- device helpers -----------------------------------------------------------------------------------------
constexpr unsigned long long kEmptyKey = 0x8000000000000000ULL;  // slot sentinel; a real key equal to it
                                                                  // is routed to a side group/chain

__device__ __forceinline__ bool NullBit(const uint8_t *nulls, int64_t i)
{
    return nulls != nullptr && ((nulls[i >> 3] >> (i & 7)) & 1);
}

// Loads element i of an integer-like column of width 1/2/4/8 as a sign-extended 64-bit word.  Width 1 is BYTE
// (int8) or BOOLEAN, whose values are 0 or 1 (the reference's Vector<bool>), so both read as int8.
__device__ __forceinline__ unsigned long long LoadWord(const void *v, int width, int64_t i)
{
    switch (width) {
        case 8: return static_cast<const unsigned long long *>(v)[i];
        case 4: return static_cast<unsigned long long>(static_cast<long long>(static_cast<const int32_t *>(v)[i]));
        case 2: return static_cast<unsigned long long>(static_cast<long long>(static_cast<const int16_t *>(v)[i]));
        default: return static_cast<unsigned long long>(static_cast<long long>(static_cast<const int8_t *>(v)[i]));
    }
}

// 64-bit finaliser (murmur3 fmix64); the table hash of this library.  Hash values are internal: no
// output depends on them, so the GPU uses a mixer that spreads dense keys across slots instead of the
// reference's CRC32C / identity hashes.
__device__ __forceinline__ unsigned long long Mix64(unsigned long long k)
{
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ULL;
    k ^= k >> 33;
    return k;
}

__device__ __forceinline__ int LaneId() { return static_cast<int>(__lane_id()); }
// Wave of the calling thread within its workgroup.
__device__ __forceinline__ int WaveInBlock() { return static_cast<int>(threadIdx.x) / kWaveSize; }

__device__ __forceinline__ unsigned long long LaneMaskLt()
{
    int l = LaneId();
    return l == 0 ? 0ULL : (~0ULL >> (64 - l));
}

// Wave-aggregated reservation of one slot per lane that wants one.  Safe under divergence: only the
// lanes executing the call take part (ballot counts active lanes; the leader is one of them).
__device__ __forceinline__ unsigned long long WaveReserveOne(unsigned long long *counter, bool want)
{
    unsigned long long mask = __ballot(want);
    if (!want) return 0;
    int leader = __ffsll(static_cast<long long>(mask)) - 1;
    unsigned long long base = 0;
    if (LaneId() == leader) base = atomicAdd(counter, static_cast<unsigned long long>(__popcll(mask)));
    unsigned int lo = __builtin_amdgcn_readlane(static_cast<unsigned int>(base), leader);
    unsigned int hi = __builtin_amdgcn_readlane(static_cast<unsigned int>(base >> 32), leader);
    return ((static_cast<unsigned long long>(hi) << 32) | lo) + __popcll(mask & LaneMaskLt());
}

// Wave-aggregated reservation of `want` consecutive slots per lane.  MUST be called by all 64 lanes of a
// converged wave (full-wave inclusive scan); lanes with nothing to reserve pass 0.
__device__ __forceinline__ unsigned long long WaveReserve(unsigned long long *counter, unsigned int want)
{
    int lane = LaneId();
    unsigned int incl = want;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        unsigned int up = __shfl_up(incl, d, 64);
        if (lane >= d) incl += up;
    }
    unsigned int total = __builtin_amdgcn_readlane(incl, 63);
    unsigned long long base = 0;
    if (lane == 0 && total) base = atomicAdd(counter, static_cast<unsigned long long>(total));
    unsigned int lo = __builtin_amdgcn_readlane(static_cast<unsigned int>(base), 0);
    unsigned int hi = __builtin_amdgcn_readlane(static_cast<unsigned int>(base >> 32), 0);
    return ((static_cast<unsigned long long>(hi) << 32) | lo) + (incl - want);
}

}  // namespace omni_gpu
