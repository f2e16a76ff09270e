// Batched group-by over a stream of many VectorBatches of uneven size, each an independent group-by.
//
// The reference runs one HashAggregationOperator::AddInput per batch (core/src/operator/aggregation/
// group_aggregation.cpp:320-406).  Here the whole stream is resident in HBM and ONE persistent launch drains a
// device queue of batches: every workgroup takes the next batch index with an atomic, aggregates that batch
// in an LDS table, and writes the batch's groups straight into its output range.  A batch whose distinct keys
// do not fit the LDS table spills, exactly, into the workgroup's private scratch table in HBM (the rows that
// found no LDS slot go there, then the LDS partials are merged in); results never depend on which path ran.
//
// Layout: the stream's columns are concatenated; batch b owns input rows [off[b], off[b+1]) and output rows
// [off[b], off[b] + groups[b]).  LDS table: S probe slots + 2 side slots (NULL key, key equal to the slot
// sentinel).  Scratch: one region per resident workgroup, kept all-empty between batches (compaction resets
// the slots it reads).
#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "agg_device.h"

namespace omni_gpu {

struct StreamArgs {
    KeyDev kd;
    AggArgs args;
    LdsLayout L;          // LDS table (S = probe slots; the layout holds S + 2 slots)
    LdsLayout SL;         // scratch table layout for one workgroup region (S = scratch probe slots)
    uint8_t *scratch;     // base of the scratch regions
    int64_t scratch_stride;
    const int64_t *offsets;
    const int32_t *order;   // processing order of the batches
    const int32_t *modes;   // per batch: 0 = LDS first (spill if full), 2 = scratch table directly; nullptr = all 0
    int32_t nbatches;
    int32_t max_probe;
    unsigned long long *queue;
    OutArgs out;
    unsigned long long *groups;  // per batch
    int overflow_as_null;
    unsigned int *error_flag;
    unsigned long long *stats;   // [0] LDS-only batches, [1] scratch batches
};

__device__ __forceinline__ void InitSlots(uint8_t *base, const LdsLayout &L, const AggArgs &args, int from, int to,
                                          int step)
{
    unsigned long long *keys = reinterpret_cast<unsigned long long *>(base + L.keys_off);
    for (int i = from; i < to; i += step) {
        keys[i] = kEmptyKey;
        for (int a = 0; a < args.n; ++a) {
            int kind = args.a[a].kind;
            if (Is32(kind)) reinterpret_cast<int *>(base + L.off[a][0])[i] = static_cast<int>(InitWord(kind));
            else reinterpret_cast<unsigned long long *>(base + L.off[a][0])[i] = InitWord(kind);
            if (L.off[a][1] >= 0) reinterpret_cast<unsigned long long *>(base + L.off[a][1])[i] = 0;
            if (L.off[a][2] >= 0) reinterpret_cast<unsigned long long *>(base + L.off[a][2])[i] = 0;
            if (L.flag_off[a] >= 0) base[L.flag_off[a] + i] = 0;
        }
    }
}

// Adds the partial state `src` into `dst` (the two are different tables; dst may be updated concurrently).
template <typename D, typename S> __device__ __forceinline__ void MergeT(const AggArgs &args, const D &dst, const S &src)
{
    typedef unsigned long long u64;
    for (int i = 0; i < args.n; ++i) {
        const AggDev &d = args.a[i];
        switch (d.kind) {
            case K_COUNT_ALL:
            case K_COUNT_COL: atomicAdd(dst.template p<u64>(i, 0), *src.template p<u64>(i, 0)); break;
            case K_AVG:
                atomicAdd(dst.template p<double>(i, 0), *src.template p<double>(i, 0));
                atomicAdd(dst.template p<u64>(i, 1), *src.template p<u64>(i, 1));
                break;
            default:
                if (!*src.flag(i)) break;
                *dst.flag(i) = 1;
                switch (d.kind) {
                    case K_SUM_I64: atomicAdd(dst.template p<u64>(i, 0), *src.template p<u64>(i, 0)); break;
                    case K_SUM_F64: atomicAdd(dst.template p<double>(i, 0), *src.template p<double>(i, 0)); break;
                    case K_SUM_DEC:
                        AddDec3(dst.template p<u64>(i, 0), dst.template p<u64>(i, 1), dst.template p<long long>(i, 2),
                                *src.template p<u64>(i, 0), *src.template p<u64>(i, 1), *src.template p<long long>(i, 2));
                        break;
                    case K_MIN_I64: atomicMin(dst.template p<long long>(i, 0), *src.template p<long long>(i, 0)); break;
                    case K_MAX_I64: atomicMax(dst.template p<long long>(i, 0), *src.template p<long long>(i, 0)); break;
                    case K_MIN_I32: atomicMin(dst.template p<int>(i, 0), *src.template p<int>(i, 0)); break;
                    case K_MAX_I32: atomicMax(dst.template p<int>(i, 0), *src.template p<int>(i, 0)); break;
                    case K_MIN_F64: atomicMin(dst.template p<double>(i, 0), *src.template p<double>(i, 0)); break;
                    case K_MAX_F64: atomicMax(dst.template p<double>(i, 0), *src.template p<double>(i, 0)); break;
                    default: break;
                }
        }
    }
}

// Writes one group (key + aggregates) to output row r.  NULL bits are OR-ed into pre-cleared bitmaps.
template <typename Ref>
__device__ __forceinline__ void EmitRow(const StreamArgs &a, const Ref &s, unsigned long long key, bool keyNull, int64_t r)
{
    const OutDev &ko = a.out.o[0];
    StoreWidth(ko.values, ko.width, r, keyNull ? 0ULL : key);
    if (keyNull) atomicOr(reinterpret_cast<unsigned int *>(ko.nulls) + (r >> 5), 1u << (r & 31));
    for (int i = 0; i < a.args.n; ++i) {
        const OutDev &o = a.out.o[1 + i];
        bool isNull = FinalizeAgg(a.args.a[i], i, s, o, r, a.overflow_as_null, a.error_flag);
        if (isNull) atomicOr(reinterpret_cast<unsigned int *>(o.nulls) + (r >> 5), 1u << (r & 31));
    }
}

// Linear-probing find-or-insert into a table of `S` probe slots (keys at base + L.keys_off); unbounded probe.
__device__ __forceinline__ int ScratchSlot(uint8_t *base, const LdsLayout &L, unsigned long long key, unsigned int mask)
{
    unsigned long long *keys = reinterpret_cast<unsigned long long *>(base + L.keys_off);
    unsigned int pos = static_cast<unsigned int>(Mix64(key)) & mask;
    while (true) {
        unsigned long long k = keys[pos];
        if (k == kEmptyKey) {
            k = atomicCAS(&keys[pos], kEmptyKey, key);
            if (k == kEmptyKey) return static_cast<int>(pos);
        }
        if (k == key) return static_cast<int>(pos);
        pos = (pos + 1) & mask;
    }
}

template <int BLOCK>
__device__ __forceinline__ void ProcessBatch(const StreamArgs &a, int bi, uint8_t *lds, unsigned int *sh)
{
    const LdsLayout &L = a.L;
    const int S = L.S;
    unsigned long long *lkeys = reinterpret_cast<unsigned long long *>(lds + L.keys_off);
    const int64_t start = a.offsets[bi], end = a.offsets[bi + 1];
    const int64_t rows = end - start;
    const bool direct = a.modes != nullptr && a.modes[bi] == 2;
    uint8_t *scr = a.scratch + static_cast<int64_t>(blockIdx.x) * a.scratch_stride;
    unsigned int smask = 1;
    while (static_cast<int64_t>(smask) < 2 * rows) smask <<= 1;
    smask -= 1;  // scratch probe range for this batch: next power of two >= 2 * rows

    InitSlots(lds, L, a.args, threadIdx.x, S + 2, BLOCK);
    if (threadIdx.x == 0) {
        sh[0] = direct ? 1u : 0u;  // rows went to the scratch table
        sh[1] = 0;                 // output rows of this batch
    }
    __syncthreads();
    for (int64_t row = start + threadIdx.x; row < end; row += BLOCK) {
        bool isNull = NullBit(a.kd.nulls[0], row);
        unsigned long long key = isNull ? 0ULL : LoadWord(a.kd.values[0], a.kd.width[0], row);
        int slot = -1;
        if (isNull) {
            slot = S;
            lkeys[S] = 0;  // marks the side slot used
        } else if (key == kEmptyKey) {
            slot = S + 1;
            lkeys[S + 1] = 0;
        } else if (!direct) {
            unsigned long long h = Mix64(key);
            unsigned int pos = static_cast<unsigned int>(((h >> 32) * static_cast<unsigned long long>(S)) >> 32);
            for (int step = 0; step < a.max_probe; ++step) {
                unsigned long long k = lkeys[pos];
                if (k == kEmptyKey) {
                    k = atomicCAS(&lkeys[pos], kEmptyKey, key);
                    if (k == kEmptyKey) k = key;
                }
                if (k == key) {
                    slot = static_cast<int>(pos);
                    break;
                }
                pos = pos + 1 == static_cast<unsigned int>(S) ? 0 : pos + 1;
            }
        }
        if (slot >= 0) {
            UpdateAggsT(a.args, LdsRef{lds, L, slot}, row);
        } else {
            if (!direct) sh[0] = 1u;
            int ss = ScratchSlot(scr, a.SL, key, smask);
            UpdateAggsT(a.args, LdsRef{scr, a.SL, ss}, row);
        }
    }
    __syncthreads();
    const bool spilled = sh[0] != 0;
    if (spilled) {
        // merge the LDS partials into the scratch table
        for (int i = threadIdx.x; i < S; i += BLOCK) {
            unsigned long long k = lkeys[i];
            if (k == kEmptyKey) continue;
            int ss = ScratchSlot(scr, a.SL, k, smask);
            MergeT(a.args, LdsRef{scr, a.SL, ss}, LdsRef{lds, L, i});
        }
        __syncthreads();
        unsigned long long *skeys = reinterpret_cast<unsigned long long *>(scr + a.SL.keys_off);
        for (unsigned int i = threadIdx.x; i <= smask; i += BLOCK) {
            unsigned long long k = skeys[i];
            if (k == kEmptyKey) continue;
            unsigned int r = atomicAdd(&sh[1], 1u);
            EmitRow(a, LdsRef{scr, a.SL, static_cast<int>(i)}, k, false, start + r);
            InitSlots(scr, a.SL, a.args, static_cast<int>(i), static_cast<int>(i) + 1, 1);  // leave it empty
        }
    } else {
        for (int i = threadIdx.x; i < S; i += BLOCK) {
            unsigned long long k = lkeys[i];
            if (k == kEmptyKey) continue;
            unsigned int r = atomicAdd(&sh[1], 1u);
            EmitRow(a, LdsRef{lds, L, i}, k, false, start + r);
        }
    }
    if (threadIdx.x < 2 && lkeys[S + threadIdx.x] != kEmptyKey) {
        unsigned int r = atomicAdd(&sh[1], 1u);
        EmitRow(a, LdsRef{lds, L, S + static_cast<int>(threadIdx.x)}, threadIdx.x == 0 ? 0ULL : kEmptyKey,
                threadIdx.x == 0, start + r);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a.groups[bi] = sh[1];
        atomicAdd(&a.stats[spilled ? 1 : 0], 1ULL);
    }
}

// single_batch >= 0: process only that batch (one workgroup, one launch per batch).
// single_batch < 0: persistent -- every workgroup drains the shared queue until it is empty.
template <int BLOCK> __global__ void __launch_bounds__(BLOCK) agg_stream(StreamArgs a, int single_batch)
{
    extern __shared__ unsigned long long lds_words[];
    uint8_t *lds = reinterpret_cast<uint8_t *>(lds_words);
    __shared__ unsigned int sh[2];
    __shared__ int next;
    if (single_batch >= 0) {
        ProcessBatch<BLOCK>(a, single_batch, lds, sh);
        return;
    }
    while (true) {
        if (threadIdx.x == 0) {
            unsigned long long q = atomicAdd(a.queue, 1ULL);
            next = q < static_cast<unsigned long long>(a.nbatches) ? a.order[q] : -1;
        }
        __syncthreads();
        int bi = next;
        __syncthreads();
        if (bi < 0) break;
        ProcessBatch<BLOCK>(a, bi, lds, sh);
    }
}

__global__ void stream_reset(unsigned long long *queue, unsigned int *words, int64_t nwords)
{
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i == 0) *queue = 0;
    if (i < nwords) words[i] = 0;
}

__global__ void stream_init_scratch(uint8_t *scratch, int64_t stride, int regions, LdsLayout SL, AggArgs args)
{
    int region = blockIdx.y;
    if (region >= regions) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < SL.S) InitSlots(scratch + region * stride, SL, args, i, i + 1, 1);
}

// ---- host side ----------------------------------------------------------------------------------------------

int ClassifyAgg(const omni_gpu_agg_spec &sp, int i, AggDev &d);  // agg.hip
LdsLayout MakeLdsLayout(const AggArgs &args, int budget);      // agg.hip

struct StreamHandle {
    omni_gpu_agg_spec spec{};
    hipStream_t stream = nullptr;
    KeyDev kd{};
    AggArgs args{};
    int block = 256;
    int lds_budget = 32 * 1024;
    int max_probe = 32;
    int wgs_per_cu = 0;  // 0 = occupancy limit
    int64_t max_batch_rows = 0;
    LdsLayout L{}, SL{};
    int regions = 0;
    DevBuf scratch, offsets, order, modes, groups, counters;  // counters: [0] queue, [8] error u32, [16] stats u64[2]
    std::vector<int64_t> h_offsets;
    std::vector<int32_t> h_order, h_modes;
    int32_t dev_nb = -1;
    bool dev_has_modes = false;
    // captured steady state
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<uint64_t> graph_key;
};

int Prepare(StreamHandle &h)
{
    // LDS table: the layout's arrays hold L.S slots; the probe range is L.S - 64, leaving room for the
    // 2 side slots.  Scratch: the next power of two >= 2 * max batch rows probe slots per region.
    h.L = MakeLdsLayout(h.args, h.lds_budget);
    h.L.S -= 64;
    int sslots = 1;
    while (sslots < 2 * h.max_batch_rows) sslots <<= 1;
    {
        LdsLayout SL{};
        SL.S = sslots;
        int64_t off = 0;
        auto take = [&](int64_t bytes) {
            int64_t o = off;
            off += (bytes + 255) & ~255LL;
            return static_cast<int>(o);
        };
        SL.keys_off = take(8LL * sslots);
        for (int i = 0; i < h.args.n; ++i) {
            int k = h.args.a[i].kind;
            SL.off[i][0] = take((Is32(k) ? 4LL : 8LL) * sslots);
            SL.off[i][1] = (k == K_AVG || k == K_SUM_DEC) ? take(8LL * sslots) : -1;
            SL.off[i][2] = k == K_SUM_DEC ? take(8LL * sslots) : -1;
            SL.flag_off[i] = (k != K_COUNT_ALL && k != K_COUNT_COL && k != K_AVG) ? take(sslots) : -1;
        }
        SL.bytes = static_cast<int>(off);
        h.SL = SL;
    }
    int dev = CurrentDevice();
    int cus = 0;
    OG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    int occ = 0;
    if (h.block == 512) {
        OG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, agg_stream<512>, 512, h.L.bytes));
    } else {
        OG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, agg_stream<256>, 256, h.L.bytes));
    }
    occ = std::max(occ, 1);
    if (h.wgs_per_cu > 0) occ = std::min(occ, h.wgs_per_cu);
    h.regions = cus * occ;
    OG_TRY(h.scratch.Reserve(static_cast<size_t>(h.SL.bytes) * h.regions, h.stream));
    dim3 grid((h.SL.S + 255) / 256, h.regions);
    {
        KernelScope ks(h.stream, "stream_init_scratch");
        stream_init_scratch<<<grid, 256, 0, h.stream>>>(h.scratch.as<uint8_t>(), h.SL.bytes, h.regions, h.SL, h.args);
        OG_TRY(CheckLaunch("stream_init_scratch"));
    }
    if (h.graph_exec) {
        (void)hipGraphExecDestroy(h.graph_exec);
        (void)hipGraphDestroy(h.graph);
        h.graph_exec = nullptr;
        h.graph = nullptr;
        h.graph_key.clear();
    }
    return OMNI_GPU_OK;
}

// Clears the output NULL bitmaps and the queue, then launches the batch kernel(s).  Inside a graph capture
// no timing events are recorded (capturing == true).
int LaunchStream(StreamHandle &h, const StreamArgs &a, int64_t total_rows, int plain, int nb, bool capturing)
{
    hipStream_t s = h.stream;
    for (int c = 0; c < 1 + h.args.n; ++c) {
        int64_t words = (total_rows + 31) / 32;
        stream_reset<<<static_cast<unsigned int>((words + 255) / 256), 256, 0, s>>>(
            h.counters.as<unsigned long long>(), reinterpret_cast<unsigned int *>(a.out.o[c].nulls), words);
        OG_TRY(CheckLaunch("stream_reset"));
    }
    std::unique_ptr<KernelScope> ks;
    if (!capturing) ks.reset(new KernelScope(s, plain ? "agg_stream_per_batch" : "agg_stream_persistent"));
    if (plain) {
        for (int b = 0; b < nb; ++b) {
            if (h.block == 512) agg_stream<512><<<1, 512, h.L.bytes, s>>>(a, b);
            else agg_stream<256><<<1, 256, h.L.bytes, s>>>(a, b);
        }
    } else {
        if (h.block == 512) agg_stream<512><<<h.regions, 512, h.L.bytes, s>>>(a, -1);
        else agg_stream<256><<<h.regions, 256, h.L.bytes, s>>>(a, -1);
    }
    return CheckLaunch("agg_stream");
}

}  // namespace omni_gpu

using namespace omni_gpu;

extern "C" {

int omni_gpu_agg_stream_create(const omni_gpu_agg_spec *spec, int64_t max_batch_rows, void **handle)
{
    OG_CHECK(spec && handle, OMNI_GPU_ERR_INVALID, "null argument");
    OG_CHECK(spec->num_keys == 1 && TypeWidth(spec->key_types[0]) > 0 && TypeWidth(spec->key_types[0]) <= 8,
             OMNI_GPU_ERR_INVALID, "the batched stream supports one key column of <= 8 bytes");
    OG_CHECK(max_batch_rows > 0 && max_batch_rows <= (1 << 20), OMNI_GPU_ERR_INVALID, "max_batch_rows out of range");
    std::unique_ptr<StreamHandle> h(new StreamHandle());
    h->spec = *spec;
    h->kd.num = 1;
    h->kd.W = 1;
    h->kd.width[0] = TypeWidth(spec->key_types[0]);
    h->args.n = spec->num_aggs;
    for (int i = 0; i < spec->num_aggs; ++i) OG_TRY(ClassifyAgg(*spec, i, h->args.a[i]));
    h->max_batch_rows = max_batch_rows;
    OG_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    OG_TRY(h->counters.Reserve(64, h->stream));
    OG_HIP(hipMemsetAsync(h->counters.p, 0, 64, h->stream));
    OG_TRY(Prepare(*h));
    OG_HIP(hipStreamSynchronize(h->stream));
    *handle = h.release();
    return OMNI_GPU_OK;
}

int omni_gpu_agg_stream_tune(void *handle, int32_t block_threads, int32_t lds_bytes, int32_t max_probe,
                             int32_t wgs_per_cu)
{
    OG_CHECK(handle, OMNI_GPU_ERR_INVALID, "null handle");
    StreamHandle &h = *static_cast<StreamHandle *>(handle);
    OG_CHECK(block_threads == 256 || block_threads == 512, OMNI_GPU_ERR_INVALID, "block_threads must be 256 or 512");
    OG_CHECK(lds_bytes >= 8 * 1024 && lds_bytes <= 64 * 1024, OMNI_GPU_ERR_INVALID, "lds_bytes must be in [8K, 64K]");
    OG_CHECK(max_probe >= 1 && max_probe <= 256, OMNI_GPU_ERR_INVALID, "max_probe must be in [1, 256]");
    h.block = block_threads;
    h.lds_budget = lds_bytes;
    h.max_probe = max_probe;
    h.wgs_per_cu = std::max(0, wgs_per_cu);
    OG_TRY(Prepare(h));
    OG_HIP(hipStreamSynchronize(h.stream));
    return OMNI_GPU_OK;
}

int omni_gpu_agg_stream_info(void *handle, int32_t *lds_slots, int32_t *workgroups, int64_t *scratch_bytes)
{
    OG_CHECK(handle, OMNI_GPU_ERR_INVALID, "null handle");
    StreamHandle &h = *static_cast<StreamHandle *>(handle);
    if (lds_slots) *lds_slots = h.L.S;
    if (workgroups) *workgroups = h.regions;
    if (scratch_bytes) *scratch_bytes = static_cast<int64_t>(h.SL.bytes) * h.regions;
    return OMNI_GPU_OK;
}

int omni_gpu_agg_stream_run(void *handle, const omni_gpu_column *keys, const omni_gpu_column *aggs,
                            const int64_t *batch_offsets, int32_t num_batches, const int32_t *batch_modes,
                            int32_t flags, omni_gpu_out_column *outs, int64_t *groups_per_batch)
{
    OG_CHECK(handle && keys && batch_offsets && outs && groups_per_batch, OMNI_GPU_ERR_INVALID, "null argument");
    OG_CHECK(num_batches > 0, OMNI_GPU_ERR_INVALID, "no batches");
    StreamHandle &h = *static_cast<StreamHandle *>(handle);
    hipStream_t s = h.stream;
    OG_CHECK(keys[0].type_id == h.spec.key_types[0] && keys[0].values, OMNI_GPU_ERR_INVALID, "key column mismatch");
    int64_t total = batch_offsets[num_batches];
    OG_CHECK(batch_offsets[0] == 0, OMNI_GPU_ERR_INVALID, "batch_offsets[0] must be 0");
    for (int b = 0; b < num_batches; ++b) {
        int64_t r = batch_offsets[b + 1] - batch_offsets[b];
        OG_CHECK(r >= 0 && r <= h.max_batch_rows, OMNI_GPU_ERR_INVALID, "batch %d has %lld rows (max %lld)", b,
                 static_cast<long long>(r), static_cast<long long>(h.max_batch_rows));
    }
    StreamArgs a{};
    a.kd = h.kd;
    a.kd.values[0] = keys[0].values;
    a.kd.nulls[0] = keys[0].nulls;
    a.args = h.args;
    for (int i = 0; i < h.args.n; ++i) {
        if (h.args.a[i].kind == K_COUNT_ALL) continue;
        OG_CHECK(aggs && aggs[i].type_id == h.spec.agg_in_types[i] && aggs[i].values, OMNI_GPU_ERR_INVALID,
                 "aggregate %d input mismatch", i);
        a.args.a[i].values = aggs[i].values;
        a.args.a[i].nulls = aggs[i].nulls;
    }
    for (int c = 0; c < 1 + h.args.n; ++c) {
        int expect = c == 0 ? h.spec.key_types[0] : h.spec.agg_out_types[c - 1];
        OG_CHECK(outs[c].type_id == expect && outs[c].values && outs[c].nulls, OMNI_GPU_ERR_INVALID,
                 "output column %d mismatch", c);
        a.out.o[c].width = TypeWidth(expect);
        a.out.o[c].values = outs[c].values;
        a.out.o[c].nulls = outs[c].nulls;
    }
    a.out.nkeys = 1;
    a.out.naggs = h.args.n;
    // batch descriptors: copied only when they change (a replayed stream keeps them resident)
    bool changed = h.dev_nb != num_batches || h.dev_has_modes != (batch_modes != nullptr) ||
                   !std::equal(batch_offsets, batch_offsets + num_batches + 1, h.h_offsets.begin(), h.h_offsets.end()) ||
                   (batch_modes && !std::equal(batch_modes, batch_modes + num_batches, h.h_modes.begin(), h.h_modes.end()));
    if (changed) {
        h.h_offsets.assign(batch_offsets, batch_offsets + num_batches + 1);
        if (batch_modes) h.h_modes.assign(batch_modes, batch_modes + num_batches);
        else h.h_modes.clear();
        // largest batches first: the longest tasks start early, so the queue drains evenly
        h.h_order.resize(num_batches);
        for (int b = 0; b < num_batches; ++b) h.h_order[b] = b;
        std::stable_sort(h.h_order.begin(), h.h_order.end(), [&](int x, int y) {
            return batch_offsets[x + 1] - batch_offsets[x] > batch_offsets[y + 1] - batch_offsets[y];
        });
        OG_TRY(h.offsets.Reserve(sizeof(int64_t) * (num_batches + 1), s));
        OG_TRY(h.order.Reserve(sizeof(int32_t) * num_batches, s));
        OG_TRY(h.groups.Reserve(sizeof(unsigned long long) * num_batches, s));
        OG_HIP(hipMemcpyAsync(h.offsets.p, batch_offsets, sizeof(int64_t) * (num_batches + 1), hipMemcpyHostToDevice, s));
        OG_HIP(hipMemcpyAsync(h.order.p, h.h_order.data(), sizeof(int32_t) * num_batches, hipMemcpyHostToDevice, s));
        if (batch_modes) {
            OG_TRY(h.modes.Reserve(sizeof(int32_t) * num_batches, s));
            OG_HIP(hipMemcpyAsync(h.modes.p, batch_modes, sizeof(int32_t) * num_batches, hipMemcpyHostToDevice, s));
        }
        h.dev_nb = num_batches;
        h.dev_has_modes = batch_modes != nullptr;
    }
    a.scratch = h.scratch.as<uint8_t>();
    a.scratch_stride = h.SL.bytes;
    a.L = h.L;
    a.SL = h.SL;
    a.offsets = h.offsets.as<int64_t>();
    a.order = h.order.as<int32_t>();
    a.modes = batch_modes ? h.modes.as<int32_t>() : nullptr;
    a.nbatches = num_batches;
    a.max_probe = h.max_probe;
    a.queue = h.counters.as<unsigned long long>();
    a.error_flag = reinterpret_cast<unsigned int *>(h.counters.as<uint8_t>() + 8);
    a.stats = reinterpret_cast<unsigned long long *>(h.counters.as<uint8_t>() + 16);
    a.groups = h.groups.as<unsigned long long>();
    a.overflow_as_null = h.spec.overflow_as_null;
    int plain = flags & 1;
    int use_graph = (flags & 2) && !plain;
    OG_HIP(hipMemsetAsync(h.counters.as<uint8_t>() + 8, 0, 4, s));
    if (use_graph) {
        // the captured launch sequence is keyed by every pointer and size it bakes in
        std::vector<uint64_t> key = {reinterpret_cast<uint64_t>(keys[0].values), reinterpret_cast<uint64_t>(keys[0].nulls),
                                     static_cast<uint64_t>(num_batches), static_cast<uint64_t>(total),
                                     static_cast<uint64_t>(batch_modes != nullptr), static_cast<uint64_t>(h.block),
                                     static_cast<uint64_t>(h.L.bytes)};
        for (int i = 0; i < h.args.n; ++i) {
            key.push_back(reinterpret_cast<uint64_t>(a.args.a[i].values));
            key.push_back(reinterpret_cast<uint64_t>(a.args.a[i].nulls));
        }
        for (int c = 0; c < 1 + h.args.n; ++c) {
            key.push_back(reinterpret_cast<uint64_t>(outs[c].values));
            key.push_back(reinterpret_cast<uint64_t>(outs[c].nulls));
        }
        if (!h.graph_exec || key != h.graph_key) {
            if (h.graph_exec) {
                (void)hipGraphExecDestroy(h.graph_exec);
                (void)hipGraphDestroy(h.graph);
                h.graph_exec = nullptr;
                h.graph = nullptr;
            }
            OG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            int rc = LaunchStream(h, a, total, 0, num_batches, true);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(s, &g);
            OG_TRY(rc);
            if (e != hipSuccess) return HipFail(e, "hipStreamEndCapture", __FILE__, __LINE__);
            h.graph = g;
            OG_HIP(hipGraphInstantiate(&h.graph_exec, g, nullptr, nullptr, 0));
            h.graph_key = key;
        }
        KernelScope ks(s, "agg_stream_graph");
        OG_HIP(hipGraphLaunch(h.graph_exec, s));
    } else {
        OG_TRY(LaunchStream(h, a, total, plain, num_batches, false));
    }
    std::vector<unsigned long long> g(num_batches);
    unsigned int err = 0;
    OG_HIP(hipMemcpyAsync(g.data(), h.groups.p, sizeof(unsigned long long) * num_batches, hipMemcpyDeviceToHost, s));
    OG_HIP(hipMemcpyAsync(&err, a.error_flag, 4, hipMemcpyDeviceToHost, s));
    OG_HIP(hipStreamSynchronize(s));
    for (int b = 0; b < num_batches; ++b) groups_per_batch[b] = static_cast<int64_t>(g[b]);
    if (err) return SetError(OMNI_GPU_ERR_OVERFLOW, "decimal overflow in SUM (overflowAsNull is false)");
    return OMNI_GPU_OK;
}

int omni_gpu_agg_stream_stats(void *handle, int64_t *lds_batches, int64_t *scratch_batches)
{
    OG_CHECK(handle, OMNI_GPU_ERR_INVALID, "null handle");
    StreamHandle &h = *static_cast<StreamHandle *>(handle);
    unsigned long long st[2] = {0, 0};
    OG_HIP(hipMemcpyAsync(st, h.counters.as<uint8_t>() + 16, 16, hipMemcpyDeviceToHost, h.stream));
    OG_HIP(hipMemsetAsync(h.counters.as<uint8_t>() + 16, 0, 16, h.stream));
    OG_HIP(hipStreamSynchronize(h.stream));
    if (lds_batches) *lds_batches = static_cast<int64_t>(st[0]);
    if (scratch_batches) *scratch_batches = static_cast<int64_t>(st[1]);
    return OMNI_GPU_OK;
}

int omni_gpu_agg_stream_destroy(void *handle)
{
    if (!handle) return OMNI_GPU_OK;
    StreamHandle *h = static_cast<StreamHandle *>(handle);
    if (h->graph_exec) {
        (void)hipGraphExecDestroy(h->graph_exec);
        (void)hipGraphDestroy(h->graph);
    }
    hipStream_t s = h->stream;
    h->scratch.Release(s);
    h->offsets.Release(s);
    h->order.Release(s);
    h->modes.Release(s);
    h->groups.Release(s);
    h->counters.Release(s);
    if (s) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
    delete h;
    return OMNI_GPU_OK;
}

}  // extern "C"
