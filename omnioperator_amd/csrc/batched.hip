// Batched group-by over a stream of many VectorBatches of uneven size, each an independent group-by.
//
// The reference runs one HashAggregationOperator::AddInput + GetOutput per batch (core/src/operator/
// aggregation/group_aggregation.cpp:320-406, 454-474).  Here the whole stream is resident in HBM and two
// launches cover every batch (phase 1: LDS table per batch, abandoned when the batch's keys do not fit;
// phase 2: exact per-batch HBM table for the abandoned ones), optionally replayed from a captured hipGraph.
// A per-batch mode array (from a predicted group count) can send a batch straight to phase 2.
#include <algorithm>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "agg_layout.h"
#include "batch_fused.h"
#include "batches_device.h"

namespace omni_gpu {

constexpr int kScanBlock = 1024;
constexpr float kAdvisorSlotShare = 0.5f;  // the advisor sends a batch past phase 1 above this share of its slots
constexpr int kXcdCount = 8;  // XCDs of an MI355X; workgroup w runs on XCD w mod 8

// order (optional): workgroup i takes batch order[i] -- the batches by decreasing rows, so that the longest
// workgroups start first and the launch ends with short ones (longest-processing-time-first).
template <int BLOCK> __global__ void __launch_bounds__(BLOCK) batch_lds(const BatchConfig *__restrict__ cfg, int first,
                                                                         const int32_t *__restrict__ order)
{
    extern __shared__ unsigned long long lds_words[];
    __shared__ unsigned int sh[2];
    const int i = first + static_cast<int>(blockIdx.x);
    BatchLds<BLOCK>(*cfg, order ? order[i] : i, reinterpret_cast<uint8_t *>(lds_words), sh);
}

template <int BLOCK> __global__ void __launch_bounds__(BLOCK) batch_sketch(const BatchConfig *__restrict__ cfg)
{
    __shared__ unsigned int bits[kSketchWords];
    __shared__ unsigned int sh[1];
    BatchSketch<BLOCK>(*cfg, static_cast<int>(blockIdx.x), bits, sh);
}

template <int BLOCK> __global__ void __launch_bounds__(BLOCK) batch_split(const BatchConfig *__restrict__ cfg,
                                                                           const int32_t *__restrict__ order)
{
    __shared__ unsigned int hist[kMaxParts];
    __shared__ unsigned int bits[kLcWords];
    __shared__ unsigned int sh[2];
    BatchSplit<BLOCK>(*cfg, order[blockIdx.x], hist, bits, sh);
}

template <int BLOCK> __global__ void __launch_bounds__(BLOCK) batch_part(const BatchConfig *__restrict__ cfg)
{
    extern __shared__ unsigned long long lds_words[];
    __shared__ unsigned int sh[4];
    BatchPart<BLOCK>(*cfg, static_cast<int64_t>(blockIdx.x), reinterpret_cast<uint8_t *>(lds_words), sh);
}

// One workgroup per batch does the batch's whole group-by (BatchFused), with the generic table (NA = 0) or the lite
// table of NA aggregates over KW-byte keys.  The grid is persistent (about the workgroups the chip holds at once):
// each workgroup takes the next batch from the device queue head (*queue, zero at launch), in the order of
// `order` (batches by decreasing rows), until every batch is taken.  512-lane workgroups are held to 80 VGPRs
// (six waves per SIMD: three workgroups per CU, as many as their LDS tables allow).
template <int BLOCK, int NA, int KW>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(BLOCK == 512 ? 6 : 1)))
batch_fused(const BatchConfig *__restrict__ cfg, const int32_t *__restrict__ order, unsigned int *__restrict__ queue)
{
    extern __shared__ unsigned long long lds_words[];
    __shared__ unsigned int cnt[kFusedParts + 1];
    __shared__ unsigned int aux[kFusedParts + 1];
    __shared__ unsigned int ccnt[kFusedParts + 1];
    __shared__ unsigned int wsum[BLOCK / kWaveSize];
    __shared__ unsigned int sh[4];
    __shared__ unsigned int head;
    const unsigned int nb = static_cast<unsigned int>(cfg->nbatches);
    uint8_t *lds = reinterpret_cast<uint8_t *>(lds_words);
    while (true) {
        if (threadIdx.x == 0) head = atomicAdd(queue, 1u);
        __syncthreads();
        const unsigned int i = head;
        if (i >= nb) break;  // the same value in every thread: the whole workgroup leaves
        // the configuration is read again for each batch: an index the compiler cannot see is zero keeps it from
        // holding every field in registers across batches
        unsigned int zero;
        __asm__ volatile("s_mov_b32 %0, 0" : "=s"(zero));
        const BatchConfig &c = cfg[zero];
        if constexpr (NA == 0) {
            BatchFused<BLOCK>(c, GenericTable{c, lds, LoadFusedIn(c)}, order[i], lds, c.L2.S, c.L2.bytes, cnt, aux,
                              ccnt, wsum, sh);
        } else {
            BatchFused<BLOCK>(c, MakeLiteTable<NA, KW>(c, lds), order[i], lds, c.lite_smax, c.lite_bytes, cnt, aux,
                              ccnt, wsum, sh);
        }
        __syncthreads();  // the batch's LDS and every read of `head` are done before the next batch
    }
}

typedef void (*FusedKernelFn)(const BatchConfig *, const int32_t *, unsigned int *);

template <int BLOCK, int KW>
FusedKernelFn FusedLiteKernel(int na)
{
    switch (na) {
        case 1: return batch_fused<BLOCK, 1, KW>;
        case 2: return batch_fused<BLOCK, 2, KW>;
        case 3: return batch_fused<BLOCK, 3, KW>;
        default: return batch_fused<BLOCK, 4, KW>;
    }
}

// The fused kernel for this configuration and its dynamic LDS bytes.
template <int BLOCK>
FusedKernelFn FusedKernel(const BatchConfig &hc, int *lds)
{
    if (hc.lite_na == 0) {
        *lds = hc.L2.bytes;
        return batch_fused<BLOCK, 0, 8>;
    }
    *lds = hc.lite_bytes;
    return hc.kd.width[0] == 8 ? FusedLiteKernel<BLOCK, 8>(hc.lite_na) : FusedLiteKernel<BLOCK, 4>(hc.lite_na);
}

// Persistent grid of the fused kernel: the workgroups the device holds at once (at most one per batch).
int FusedGrid(const BatchConfig &hc, int block, int nb, int *grid)
{
    int lds = 0;
    const FusedKernelFn k = block == 512 ? FusedKernel<512>(hc, &lds) : FusedKernel<256>(hc, &lds);
    int per_cu = 0, dev = 0, cus = 0;
    OG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(k), block,
                                                        static_cast<size_t>(lds)));
    OG_HIP(hipGetDevice(&dev));
    OG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    *grid = std::max(1, std::min(nb, std::max(1, per_cu) * std::max(1, cus)));
    return OMNI_GPU_OK;
}

// Phase 3 launches are grid-stride over their tiles / batches (grid capped at kPhase3Grid) and return at once
// when phase 2 ran and no partition overflowed, so an idle phase 3 costs a few microseconds.
constexpr int64_t kPhase3Grid = 2048;

__device__ __forceinline__ bool Phase3Idle(const BatchConfig &c) { return c.hbm_flag == 2 && *c.redo == 0u; }

template <int BLOCK> __global__ void __launch_bounds__(BLOCK) batch_hbm_init(const BatchConfig *__restrict__ cfg, int64_t first,
                                                                              int64_t count)
{
    if (Phase3Idle(*cfg)) return;
    for (int64_t t = blockIdx.x; t < count; t += gridDim.x) BatchHbmInit<BLOCK>(*cfg, first + t);
}

template <int BLOCK> __global__ void __launch_bounds__(BLOCK) batch_hbm_insert(const BatchConfig *__restrict__ cfg, int64_t first,
                                                                                int64_t count)
{
    if (Phase3Idle(*cfg)) return;
    for (int64_t t = blockIdx.x; t < count; t += gridDim.x) BatchHbmInsert<BLOCK>(*cfg, first + t);
}

template <int BLOCK> __global__ void __launch_bounds__(BLOCK) batch_hbm_emit(const BatchConfig *__restrict__ cfg, int first, int count)
{
    __shared__ unsigned int sh[2];
    if (Phase3Idle(*cfg)) return;
    for (int b = blockIdx.x; b < count; b += gridDim.x) BatchHbmEmit<BLOCK>(*cfg, first + b, sh);
}

// HBM-table region offsets of the batches flagged for phase 2 / 3.  only_redo: the regions are needed only when a
// batch was flagged for the exact HBM pass (the one-workgroup-per-batch form), so the kernel returns at once
// otherwise.
__global__ void __launch_bounds__(kScanBlock) batch_scan(const BatchConfig *__restrict__ cfg, int only_redo)
{
    __shared__ unsigned long long wsum[kScanBlock / kWaveSize];
    if (only_redo && *cfg->redo == 0u) return;
    ScanRegions<kScanBlock>(*cfg, wsum);
}

__global__ void batch_clear(unsigned int *words, int64_t nwords)
{
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < nwords) words[i] = 0;
}

struct BatchesHandle {
    omni_gpu_agg_spec spec{};
    hipStream_t stream = nullptr;
    BatchConfig cfg{};       // host copy of the device-resident configuration
    int block = 256;
    int lds_budget = 32 * 1024;
    int max_probe = 32;
    int64_t max_batch_rows = 0;
    int part_budget = 64 * 1024;  // LDS bytes of a phase-2 partition table
    int part_rows = 0;            // rows per phase-2 partition (0: half the partition table's slots)
    DevBuf dcfg, offsets, tiles, modes, spilled, groups, region_off, pool, counters;  // counters: [0] error, [4] redo
    DevBuf items, parts, parts_used;  // phase-2 work items, partitions launched / used per batch
    DevBuf order;                     // batches by decreasing rows (phase 1 and 2a workgroup order)
    DevBuf sketch_modes;              // per-batch first phase from the sketch launch or the advisor
    DevBuf perm_keys, perm_vals;      // partition-ordered rows of the one-workgroup-per-batch form
    bool fused = false;               // this run takes the one-workgroup-per-batch form (flag bit 4)
    int fused_grid = 1;               // its persistent grid
    void *advisor = nullptr;          // MLP handle (omni_gpu_agg_batches_set_advisor)
    bool use_advisor = false;         // this run's first phases come from the advisor
    DevBuf adv_x, adv_y;              // the advisor's per-batch features and outputs
    int64_t nitems = 0;
    bool last_per_batch = false;
    std::vector<int64_t> h_tiles;  // phase-3 tile prefix over the batches
    size_t pool_bytes = 0;
    std::vector<int64_t> h_offsets;
    std::vector<int32_t> h_modes;
    bool have_desc = false;
    bool has_modes = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<uint64_t> graph_key;
};

void DropGraph(BatchesHandle &h)
{
    if (h.graph_exec) (void)hipGraphExecDestroy(h.graph_exec);
    if (h.graph) (void)hipGraphDestroy(h.graph);
    h.graph_exec = nullptr;
    h.graph = nullptr;
    h.graph_key.clear();
}

int Layout(BatchesHandle &h)
{
    h.cfg.L = MakeLdsLayout(h.cfg.args, h.lds_budget);
    h.cfg.L.S -= 2;  // probe slots [0, S); the last two slots of the arrays are the side slots
    OG_CHECK(h.cfg.L.S >= 62, OMNI_GPU_ERR_INVALID, "LDS budget %d bytes holds only %d slots for these aggregates",
             h.lds_budget, h.cfg.L.S);
    h.cfg.max_probe = h.max_probe;
    h.cfg.L2 = MakeLdsLayout(h.cfg.args, h.part_budget);
    h.cfg.L2.S -= 2;
    OG_CHECK(h.cfg.L2.S >= 62, OMNI_GPU_ERR_INVALID, "LDS budget %d bytes holds only %d slots for these aggregates",
             h.part_budget, h.cfg.L2.S);
    h.have_desc = false;  // the phase-2 items depend on the partition table size
    DropGraph(h);
    return OMNI_GPU_OK;
}

// Phase-2 work items for batches of the given row counts: batch b gets P_b partitions, the smallest power of
// two with rows / P_b <= half the partition table (so a partition overflows only if its hash share is far
// above the mean).  Items of batch b go to workgroup ids == b (mod 8), i.e. to one XCD, consecutively, so the
// batch's rows stay in that XCD's L2 while its partitions re-read them; the per-XCD lists are interleaved
// and padded with empty items.
std::vector<int4> PartItems(const int64_t *offsets, int nb, int table_slots, int part_rows, std::vector<int32_t> &parts,
                            const std::vector<int32_t> &order)
{
    const int64_t cpart = part_rows > 0 ? part_rows : std::max(1, table_slots / 2);
    std::vector<std::vector<int4>> lists(kXcdCount);
    parts.assign(nb, 1);
    for (int i = 0; i < nb; ++i) {  // batch order[i]'s split workgroup is workgroup i, on XCD i mod 8
        const int b = order[i];
        const int64_t rows = offsets[b + 1] - offsets[b];
        int P = 1;
        while (P < kMaxParts && static_cast<int64_t>(P) * cpart < rows) P <<= 1;
        parts[b] = P;
        for (int p = 0; p < P; ++p) lists[i % kXcdCount].push_back(make_int4(b, p, P, 0));
    }
    size_t longest = 0;
    for (auto &l : lists) longest = std::max(longest, l.size());
    std::vector<int4> items(longest * kXcdCount, make_int4(-1, 0, 1, 0));
    for (int x = 0; x < kXcdCount; ++x) {
        for (size_t k = 0; k < lists[x].size(); ++k) items[k * kXcdCount + x] = lists[x][k];
    }
    return items;
}

// Launches the whole sequence on the handle's stream (no timing events: it may be under graph capture).
int LaunchBatches(BatchesHandle &h, int64_t total_rows, int nb, bool per_batch)
{
    hipStream_t s = h.stream;
    const BatchConfig *dc = h.dcfg.as<BatchConfig>();
    int64_t words = (total_rows + 31) / 32;
    for (int c = 0; c < 1 + h.cfg.args.n; ++c) {
        batch_clear<<<static_cast<unsigned int>((words + 255) / 256), 256, 0, s>>>(
            reinterpret_cast<unsigned int *>(h.cfg.out.o[c].nulls), words);
        OG_TRY(CheckLaunch("batch_clear"));
    }
    int lds_bytes = h.cfg.L.bytes;
    if (h.fused && !per_batch) {
        // the advisor's predictions (no per-batch plan: the fused kernel reads the predicted group counts)
        if (h.use_advisor) {
            // batches that go direct on their row count alone are not sketched
            const int smax = h.cfg.lite_na ? h.cfg.lite_smax : h.cfg.L2.S;
            // the batch_fused kernel's rows-only direct rule: max(partition fit, kDirectLoad of the table's slots)
            const int64_t fit = h.cfg.part_fit > 0
                                    ? h.cfg.part_fit
                                    : std::max(static_cast<int64_t>(kFusedLoad * static_cast<float>(smax)),
                                               static_cast<int64_t>(kDirectLoad * static_cast<float>(smax)));
            OG_TRY(LaunchBatchAdvisor(s, h.advisor, h.cfg.kd.values[0], h.cfg.kd.nulls[0], h.cfg.kd.width[0],
                                      h.offsets.as<int64_t>(), nb, h.adv_x.as<float>(), h.adv_y.as<float>(), nullptr,
                                      0.f, fit));
        }
        const int32_t *order = h.order.as<int32_t>();
        unsigned int *queue = reinterpret_cast<unsigned int *>(h.counters.as<uint8_t>() + 8);
        batch_clear<<<1, 64, 0, s>>>(queue, 1);  // a kernel, not a memset: the sequence may be captured
        OG_TRY(CheckLaunch("batch_clear"));
        int lds = 0;
        const FusedKernelFn k = h.block == 512 ? FusedKernel<512>(h.cfg, &lds) : FusedKernel<256>(h.cfg, &lds);
        k<<<h.fused_grid, h.block, lds, s>>>(dc, order, queue);
        OG_TRY(CheckLaunch("batch_fused"));
        batch_scan<<<1, kScanBlock, 0, s>>>(dc, 1);
        OG_TRY(CheckLaunch("batch_scan"));
    } else if (per_batch) {
        for (int b = 0; b < nb; ++b) {
            if (h.block == 512) batch_lds<512><<<1, 512, lds_bytes, s>>>(dc, b, nullptr);
            else batch_lds<256><<<1, 256, lds_bytes, s>>>(dc, b, nullptr);
        }
    } else {
        const int32_t *order = h.order.as<int32_t>();
        if (h.use_advisor) {
            OG_TRY(LaunchBatchAdvisor(s, h.advisor, h.cfg.kd.values[0], h.cfg.kd.nulls[0], h.cfg.kd.width[0],
                                      h.offsets.as<int64_t>(), nb, h.adv_x.as<float>(), h.adv_y.as<float>(),
                                      h.cfg.sketch_modes, kAdvisorSlotShare * static_cast<float>(h.cfg.L.S), -1));
        } else if (h.cfg.sketch_modes) {
            batch_sketch<256><<<nb, 256, 0, s>>>(dc);
            OG_TRY(CheckLaunch("batch_sketch"));
        }
        if (h.block == 512) batch_lds<512><<<nb, 512, lds_bytes, s>>>(dc, 0, order);
        else batch_lds<256><<<nb, 256, lds_bytes, s>>>(dc, 0, order);
    }
    if (!h.fused || per_batch) {
        OG_TRY(CheckLaunch("batch_lds"));
        batch_scan<<<1, kScanBlock, 0, s>>>(dc, 0);
        OG_TRY(CheckLaunch("batch_scan"));
    }
    if (!per_batch && !h.fused && h.nitems > 0) {
        const unsigned int g = static_cast<unsigned int>(h.nitems);
        if (h.block == 512) {
            batch_split<512><<<nb, 512, 0, s>>>(dc, h.order.as<int32_t>());
            batch_part<512><<<g, 512, h.cfg.L2.bytes, s>>>(dc);
        } else {
            batch_split<256><<<nb, 256, 0, s>>>(dc, h.order.as<int32_t>());
            batch_part<256><<<g, 256, h.cfg.L2.bytes, s>>>(dc);
        }
        OG_TRY(CheckLaunch("batch_part"));
    }
    const int64_t tiles = h.cfg.total_tiles;
    auto grid = [](int64_t n) { return static_cast<unsigned int>(std::max<int64_t>(1, std::min(n, kPhase3Grid))); };
    if (per_batch) {
        for (int b = 0; b < nb; ++b) {
            const int64_t t0 = h.h_tiles[b], nt = h.h_tiles[b + 1] - h.h_tiles[b];
            if (nt == 0) continue;
            if (h.block == 512) {
                batch_hbm_init<512><<<grid(nt), 512, 0, s>>>(dc, t0, nt);
                batch_hbm_insert<512><<<grid(nt), 512, 0, s>>>(dc, t0, nt);
                batch_hbm_emit<512><<<1, 512, 0, s>>>(dc, b, 1);
            } else {
                batch_hbm_init<256><<<grid(nt), 256, 0, s>>>(dc, t0, nt);
                batch_hbm_insert<256><<<grid(nt), 256, 0, s>>>(dc, t0, nt);
                batch_hbm_emit<256><<<1, 256, 0, s>>>(dc, b, 1);
            }
        }
    } else if (tiles > 0) {
        if (h.block == 512) {
            batch_hbm_init<512><<<grid(tiles), 512, 0, s>>>(dc, 0, tiles);
            batch_hbm_insert<512><<<grid(tiles), 512, 0, s>>>(dc, 0, tiles);
            batch_hbm_emit<512><<<grid(nb), 512, 0, s>>>(dc, 0, nb);
        } else {
            batch_hbm_init<256><<<grid(tiles), 256, 0, s>>>(dc, 0, tiles);
            batch_hbm_insert<256><<<grid(tiles), 256, 0, s>>>(dc, 0, tiles);
            batch_hbm_emit<256><<<grid(nb), 256, 0, s>>>(dc, 0, nb);
        }
    }
    return CheckLaunch("batch_hbm");
}

}  // namespace omni_gpu

using namespace omni_gpu;

extern "C" {

int omni_gpu_agg_batches_create(const omni_gpu_agg_spec *spec, int64_t max_batch_rows, void **handle)
{
    OG_CHECK(spec && handle, OMNI_GPU_ERR_INVALID, "null argument");
    OG_CHECK(spec->num_keys == 1 && TypeWidth(spec->key_types[0]) > 0 && TypeWidth(spec->key_types[0]) <= 8,
             OMNI_GPU_ERR_INVALID, "batched group-by supports one key column of <= 8 bytes");
    OG_CHECK(max_batch_rows > 0 && max_batch_rows <= (1 << 20), OMNI_GPU_ERR_INVALID, "max_batch_rows out of range");
    std::unique_ptr<BatchesHandle> h(new BatchesHandle());
    h->spec = *spec;
    h->max_batch_rows = max_batch_rows;
    BatchConfig &c = h->cfg;
    c.kd.num = 1;
    c.kd.W = 1;
    c.kd.width[0] = TypeWidth(spec->key_types[0]);
    c.args.n = spec->num_aggs;
    for (int i = 0; i < spec->num_aggs; ++i) {
        OG_TRY(ClassifyAgg(*spec, i, c.args.a[i]));
        // the row-order check of a DECIMAL64 -> DECIMAL64 or DECIMAL128 SUM needs the per-batch path
        OG_CHECK(c.args.a[i].kind != K_SUM_DEC_CHK, OMNI_GPU_ERR_INVALID,
                 "the batched stream sums decimals into DECIMAL128 from DECIMAL64 only (aggregate %d); other decimal "
                 "sums can overflow part-way and take omni_gpu_agg_create",
                 i);
    }
    c.args.row_base = 0;
    c.R = UnitLayout(c.args);
    c.overflow_as_null = spec->overflow_as_null;
    OG_TRY(AcquireStream(&h->stream));
    OG_TRY(h->counters.Reserve(64, h->stream));
    OG_HIP(hipMemsetAsync(h->counters.p, 0, 64, h->stream));
    OG_TRY(h->dcfg.Reserve(sizeof(BatchConfig), h->stream));
    OG_TRY(Layout(*h));
    OG_TRY(SyncStream(h->stream));
    *handle = h.release();
    return OMNI_GPU_OK;
}

int omni_gpu_agg_batches_tune(void *handle, int32_t block_threads, int32_t lds_bytes, int32_t max_probe)
{
    OG_CHECK(handle, OMNI_GPU_ERR_INVALID, "null handle");
    BatchesHandle &h = *static_cast<BatchesHandle *>(handle);
    OG_CHECK(block_threads == 256 || block_threads == 512, OMNI_GPU_ERR_INVALID, "block_threads must be 256 or 512");
    OG_CHECK(lds_bytes >= 8 * 1024 && lds_bytes <= 64 * 1024, OMNI_GPU_ERR_INVALID, "lds_bytes must be in [8K, 64K]");
    OG_CHECK(max_probe >= 1 && max_probe <= 256, OMNI_GPU_ERR_INVALID, "max_probe must be in [1, 256]");
    h.block = block_threads;
    h.lds_budget = lds_bytes;
    h.max_probe = max_probe;
    return Layout(h);
}

int omni_gpu_agg_batches_tune_partitions(void *handle, int32_t part_lds_bytes, int32_t part_rows)
{
    OG_CHECK(handle, OMNI_GPU_ERR_INVALID, "null handle");
    BatchesHandle &h = *static_cast<BatchesHandle *>(handle);
    OG_CHECK(part_lds_bytes >= 8 * 1024 && part_lds_bytes <= 64 * 1024, OMNI_GPU_ERR_INVALID,
             "part_lds_bytes must be in [8K, 64K]");
    OG_CHECK(part_rows >= 0, OMNI_GPU_ERR_INVALID, "part_rows must be >= 0");
    h.part_budget = part_lds_bytes;
    h.part_rows = part_rows;
    return Layout(h);
}

int omni_gpu_agg_batches_info(void *handle, int32_t *lds_slots, int64_t *pool_bytes)
{
    OG_CHECK(handle, OMNI_GPU_ERR_INVALID, "null handle");
    BatchesHandle &h = *static_cast<BatchesHandle *>(handle);
    if (lds_slots) *lds_slots = h.cfg.L.S;
    if (pool_bytes) *pool_bytes = static_cast<int64_t>(h.pool_bytes);
    return OMNI_GPU_OK;
}

int omni_gpu_agg_batches_run(void *handle, const omni_gpu_column *keys, const omni_gpu_column *aggs,
                             const int64_t *batch_offsets, int32_t num_batches, const int32_t *batch_modes,
                             int32_t flags, omni_gpu_out_column *outs, int64_t *groups_per_batch)
{
    OG_CHECK(handle && keys && batch_offsets && outs && groups_per_batch, OMNI_GPU_ERR_INVALID, "null argument");
    OG_CHECK(num_batches > 0, OMNI_GPU_ERR_INVALID, "no batches");
    BatchesHandle &h = *static_cast<BatchesHandle *>(handle);
    hipStream_t s = h.stream;
    BatchConfig &c = h.cfg;
    OG_CHECK(keys[0].type_id == h.spec.key_types[0] && keys[0].values, OMNI_GPU_ERR_INVALID, "key column mismatch");
    OG_CHECK(batch_offsets[0] == 0, OMNI_GPU_ERR_INVALID, "batch_offsets[0] must be 0");
    for (int b = 0; b < num_batches; ++b) {
        int64_t r = batch_offsets[b + 1] - batch_offsets[b];
        OG_CHECK(r >= 0 && r <= h.max_batch_rows, OMNI_GPU_ERR_INVALID, "batch %d has %lld rows (max %lld)", b,
                 static_cast<long long>(r), static_cast<long long>(h.max_batch_rows));
    }
    if (batch_modes) {
        for (int b = 0; b < num_batches; ++b) {
            OG_CHECK(batch_modes[b] == 0 || batch_modes[b] == 2, OMNI_GPU_ERR_INVALID, "batch mode must be 0 or 2");
        }
    }
    const int64_t total = batch_offsets[num_batches];
    c.kd.values[0] = keys[0].values;
    c.kd.nulls[0] = keys[0].nulls;
    for (int i = 0; i < c.args.n; ++i) {
        if (c.args.a[i].kind == K_COUNT_ALL) continue;
        OG_CHECK(aggs && aggs[i].type_id == h.spec.agg_in_types[i] && aggs[i].values, OMNI_GPU_ERR_INVALID,
                 "aggregate %d input mismatch", i);
        c.args.a[i].values = aggs[i].values;
        c.args.a[i].nulls = aggs[i].nulls;
    }
    for (int k = 0; k < 1 + c.args.n; ++k) {
        int expect = k == 0 ? h.spec.key_types[0] : h.spec.agg_out_types[k - 1];
        OG_CHECK(outs[k].type_id == expect && outs[k].values && outs[k].nulls, OMNI_GPU_ERR_INVALID,
                 "output column %d mismatch", k);
        c.out.o[k].width = TypeWidth(expect);
        c.out.o[k].values = outs[k].values;
        c.out.o[k].nulls = outs[k].nulls;
    }
    c.out.nkeys = 1;
    c.out.naggs = c.args.n;
    // batch descriptors and per-batch arrays: rebuilt only when the stream's shape changes
    bool changed = !h.have_desc || h.has_modes != (batch_modes != nullptr) ||
                   h.h_offsets.size() != static_cast<size_t>(num_batches) + 1 ||
                   !std::equal(batch_offsets, batch_offsets + num_batches + 1, h.h_offsets.begin()) ||
                   (batch_modes && !std::equal(batch_modes, batch_modes + num_batches, h.h_modes.begin()));
    if (changed) {
        DropGraph(h);
        h.h_offsets.assign(batch_offsets, batch_offsets + num_batches + 1);
        if (batch_modes) h.h_modes.assign(batch_modes, batch_modes + num_batches);
        else h.h_modes.clear();
        h.has_modes = batch_modes != nullptr;
        OG_TRY(h.offsets.Reserve(sizeof(int64_t) * (num_batches + 1), s));
        OG_TRY(h.spilled.Reserve(sizeof(unsigned int) * num_batches, s));
        OG_TRY(h.groups.Reserve(sizeof(unsigned long long) * num_batches, s));
        OG_TRY(h.region_off.Reserve(sizeof(unsigned long long) * num_batches, s));
        OG_TRY(CopyToDevice(h.offsets.p, batch_offsets, sizeof(int64_t) * (num_batches + 1), s));
        h.h_tiles.assign(num_batches + 1, 0);
        for (int b = 0; b < num_batches; ++b) {
            // at least one tile per batch: an empty batch sent to phase 2 still initialises its table
            h.h_tiles[b + 1] = h.h_tiles[b] +
                               std::max<int64_t>(1, (batch_offsets[b + 1] - batch_offsets[b] + kHbmTile - 1) / kHbmTile);
        }
        OG_CHECK(h.h_tiles[num_batches] < 0x7FFFFFFF, OMNI_GPU_ERR_INVALID, "too many rows for one run");
        OG_TRY(h.tiles.Reserve(sizeof(int64_t) * (num_batches + 1), s));
        OG_TRY(CopyToDevice(h.tiles.p, h.h_tiles.data(), sizeof(int64_t) * (num_batches + 1), s));
        if (batch_modes) {
            OG_TRY(h.modes.Reserve(sizeof(int32_t) * num_batches, s));
            OG_TRY(CopyToDevice(h.modes.p, batch_modes, sizeof(int32_t) * num_batches, s));
        }
        // region pool: room for every batch to need its HBM table
        size_t pool = 0;
        for (int b = 0; b < num_batches; ++b) {
            pool += static_cast<size_t>(RegionBytes(c.R, batch_offsets[b + 1] - batch_offsets[b]));
        }
        OG_TRY(h.pool.Reserve(pool, s));
        h.pool_bytes = pool;
        std::vector<int32_t> order(num_batches);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return batch_offsets[a + 1] - batch_offsets[a] > batch_offsets[b + 1] - batch_offsets[b];
        });
        OG_TRY(h.order.Reserve(sizeof(int32_t) * order.size(), s));
        OG_TRY(CopyToDevice(h.order.p, order.data(), sizeof(int32_t) * order.size(), s));
        std::vector<int32_t> parts;
        const std::vector<int4> items = PartItems(batch_offsets, num_batches, c.L2.S, h.part_rows, parts, order);
        h.nitems = static_cast<int64_t>(items.size());
        OG_CHECK(h.nitems < 0x7FFFFFFF, OMNI_GPU_ERR_INVALID, "too many phase-2 items for one run");
        OG_TRY(h.items.Reserve(sizeof(int4) * items.size(), s));
        OG_TRY(CopyToDevice(h.items.p, items.data(), sizeof(int4) * items.size(), s));
        OG_TRY(h.parts.Reserve(sizeof(int32_t) * parts.size(), s));
        OG_TRY(h.parts_used.Reserve(sizeof(int32_t) * parts.size(), s));
        OG_TRY(CopyToDevice(h.parts.p, parts.data(), sizeof(int32_t) * parts.size(), s));
        h.have_desc = true;
    }
    c.offsets = h.offsets.as<int64_t>();
    c.tile_start = h.tiles.as<int64_t>();
    c.total_tiles = h.h_tiles.back();
    c.modes = batch_modes ? h.modes.as<int32_t>() : nullptr;
    c.nbatches = num_batches;
    c.spilled = h.spilled.as<unsigned int>();
    c.groups = h.groups.as<unsigned long long>();
    c.region_off = h.region_off.as<unsigned long long>();
    c.region_pool = h.pool.as<uint8_t>();
    c.items = h.items.as<int4>();
    c.parts = h.parts.as<int32_t>();
    c.parts_used = h.parts_used.as<int32_t>();
    c.redo = reinterpret_cast<unsigned int *>(h.counters.as<uint8_t>() + 4);
    c.nitems = h.nitems;
    c.hbm_flag = (flags & 1) ? 1 : 2;  // the per-batch form has no phase 2
    // flag bit 4: one workgroup per batch does the whole group-by (BatchFused)
    h.fused = !(flags & 1) && (flags & 16);
    c.carry = 0;
    c.carry_agg = -1;
    c.part_fit = h.part_rows;
    c.lite_na = 0;
    c.perm_keys = nullptr;
    c.perm_vals = nullptr;
    c.perm_rows = nullptr;
    if (h.fused) {
        // the rows carry their value word when every aggregate is COUNT(*) or reads one column of at most 8 bytes
        // without NULLs (and the key has no NULLs); otherwise they carry their row number
        bool carry = c.kd.nulls[0] == nullptr;
        for (int i = 0; i < c.args.n && carry; ++i) {
            const AggDev &d = c.args.a[i];
            if (d.kind == K_COUNT_ALL) continue;
            const bool kindOk = d.kind == K_COUNT_COL || d.kind == K_SUM_I64 || d.kind == K_SUM_F64 || d.kind == K_AVG ||
                                d.kind == K_SUM_DEC || d.kind == K_MIN_I64 || d.kind == K_MAX_I64 ||
                                d.kind == K_MIN_I32 || d.kind == K_MAX_I32;
            carry = kindOk && d.nulls == nullptr && d.in_width <= 8;
            if (carry && c.carry_agg >= 0) {
                const AggDev &f = c.args.a[c.carry_agg];
                carry = f.values == d.values && f.in_width == d.in_width;
            }
            if (carry && c.carry_agg < 0) c.carry_agg = i;
        }
        if (!carry) c.carry_agg = -1;
        c.carry = carry ? 1 : 0;
        // the lite table: at most kLiteAggs aggregates, each COUNT or an int64 SUM / MIN / MAX with an 8-byte output
        bool lite = carry && c.args.n >= 1 && c.args.n <= kLiteAggs && (c.kd.width[0] == 4 || c.kd.width[0] == 8) &&
                    (c.carry_agg < 0 || c.args.a[c.carry_agg].in_width == 8);
        for (int i = 0; i < c.args.n && lite; ++i) {
            const int k = c.args.a[i].kind;
            lite = (k == K_COUNT_ALL || k == K_COUNT_COL || k == K_SUM_I64 || k == K_MIN_I64 || k == K_MAX_I64) &&
                   c.out.o[1 + i].width == 8;
            if (lite) {
                c.lite_op[i] = k == K_SUM_I64 ? kLiteSum : k == K_MIN_I64 ? kLiteMin : k == K_MAX_I64 ? kLiteMax
                                                                                                   : kLiteCount;
            }
        }
        c.lite_na = lite ? c.args.n : 0;
        if (lite) {
            const int slots = (h.part_budget / (8 * (1 + c.args.n))) & ~63;  // keys + one word per aggregate
            c.lite_smax = slots - 2;
            c.lite_bytes = 8 * (1 + c.args.n) * slots;
        }
        const size_t n = static_cast<size_t>(std::max<int64_t>(1, total));
        OG_TRY(h.perm_keys.Reserve(sizeof(unsigned long long) * n, s));
        OG_TRY(h.perm_vals.Reserve((carry ? sizeof(unsigned long long) : sizeof(unsigned int)) * n, s));
        c.perm_keys = h.perm_keys.as<unsigned long long>();
        if (carry) c.perm_vals = h.perm_vals.as<unsigned long long>();
        else c.perm_rows = h.perm_vals.as<unsigned int>();
    }
    // flag bit 3: the advisor network picks each batch's first phase from a per-batch sketch; flag bit 2: a
    // linear-counting estimate does (neither with caller modes or the per-batch form; bit 3 wins)
    h.use_advisor = !(flags & 1) && (flags & 8) && batch_modes == nullptr;
    OG_CHECK(!h.use_advisor || h.advisor, OMNI_GPU_ERR_STATE, "flag bit 3 needs omni_gpu_agg_batches_set_advisor");
    const bool sketch = !(flags & 1) && ((flags & 4) || h.use_advisor) && batch_modes == nullptr;
    if (sketch && !h.fused) OG_TRY(h.sketch_modes.Reserve(sizeof(int32_t) * num_batches, s));
    if (h.use_advisor) {
        OG_TRY(h.adv_x.Reserve(sizeof(float) * 16 * num_batches, s));
        OG_TRY(h.adv_y.Reserve(sizeof(float) * 4 * num_batches, s));
    }
    c.sketch_modes = sketch && !h.fused ? h.sketch_modes.as<int32_t>() : nullptr;
    h.cfg.sketch_modes = c.sketch_modes;
    c.pred = h.fused && h.use_advisor ? h.adv_y.as<float>() : nullptr;
    if (h.fused) OG_TRY(FusedGrid(c, h.block, num_batches, &h.fused_grid));
    h.last_per_batch = (flags & 1) != 0;
    c.error_flag = reinterpret_cast<unsigned int *>(h.counters.as<uint8_t>());
    OG_TRY(CopyToDevice(h.dcfg.p, &c, sizeof(BatchConfig), s));
    OG_HIP(hipMemsetAsync(c.error_flag, 0, 8, s));  // error flag and phase-3 redo flag
    const bool per_batch = flags & 1;
    const bool use_graph = (flags & 2) && !per_batch;
    if (use_graph) {
        // the captured sequence bakes in the configuration buffer (whose contents are refreshed above) and
        // the launch shapes; it is re-captured when any of them changes
        std::vector<uint64_t> key = {static_cast<uint64_t>(num_batches), static_cast<uint64_t>(total),
                                     static_cast<uint64_t>(h.block), static_cast<uint64_t>(c.L.bytes),
                                     reinterpret_cast<uint64_t>(h.dcfg.p), c.sketch_modes ? 1ULL : 0ULL,
                                     h.use_advisor ? reinterpret_cast<uint64_t>(h.advisor) : 0ULL,
                                     reinterpret_cast<uint64_t>(h.adv_x.p), reinterpret_cast<uint64_t>(h.adv_y.p),
                                     reinterpret_cast<uint64_t>(h.sketch_modes.p), reinterpret_cast<uint64_t>(keys[0].values),
                                     reinterpret_cast<uint64_t>(keys[0].nulls), h.fused ? 1ULL : 0ULL,
                                     static_cast<uint64_t>(c.L2.bytes), static_cast<uint64_t>(c.lite_na),
                                     static_cast<uint64_t>(h.fused_grid)};
        for (int k = 0; k < 1 + c.args.n; ++k) key.push_back(reinterpret_cast<uint64_t>(outs[k].nulls));
        if (!h.graph_exec || key != h.graph_key) {
            DropGraph(h);
            OG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            int rc = LaunchBatches(h, total, num_batches, false);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(s, &g);
            OG_TRY(rc);
            if (e != hipSuccess) return HipFail(e, "hipStreamEndCapture", __FILE__, __LINE__);
            h.graph = g;
            OG_HIP(hipGraphInstantiate(&h.graph_exec, g, nullptr, nullptr, 0));
            h.graph_key = key;
        }
        KernelScope ks(s, "agg_batches_graph");
        OG_HIP(hipGraphLaunch(h.graph_exec, s));
    } else {
        KernelScope ks(s, per_batch ? "agg_batches_per_batch" : "agg_batches");
        OG_TRY(LaunchBatches(h, total, num_batches, per_batch));
    }
    std::vector<unsigned long long> g(num_batches);
    unsigned int err = 0;
    OG_TRY(CopyToHost(s, {{g.data(), h.groups.p, sizeof(unsigned long long) * num_batches}, {&err, c.error_flag, 4}}));
    for (int b = 0; b < num_batches; ++b) groups_per_batch[b] = static_cast<int64_t>(g[b]);
    if (err) return SetError(OMNI_GPU_ERR_OVERFLOW, "decimal overflow in SUM (overflowAsNull is false)");
    return OMNI_GPU_OK;
}

int omni_gpu_agg_batches_set_advisor(void *handle, void *mlp)
{
    OG_CHECK(handle, OMNI_GPU_ERR_INVALID, "null handle");
    BatchesHandle &h = *static_cast<BatchesHandle *>(handle);
    if (h.advisor != mlp) DropGraph(h);
    h.advisor = mlp;
    return OMNI_GPU_OK;
}

int omni_gpu_agg_batches_phase_of(void *handle, int32_t *phase)
{
    OG_CHECK(handle && phase, OMNI_GPU_ERR_INVALID, "null argument");
    BatchesHandle &h = *static_cast<BatchesHandle *>(handle);
    int64_t nb = h.have_desc ? static_cast<int64_t>(h.h_offsets.size()) - 1 : 0;
    std::vector<unsigned int> sp(static_cast<size_t>(std::max<int64_t>(nb, 1)));
    if (nb > 0) OG_TRY(CopyToHost(h.stream, {{sp.data(), h.spilled.p, sizeof(unsigned int) * nb}}));
    for (int64_t b = 0; b < nb; ++b) {
        const unsigned int f = sp[b];
        phase[b] = f == 0 ? 1 : (f == 1 && !h.last_per_batch ? 2 : 3);
    }
    return OMNI_GPU_OK;
}

int omni_gpu_agg_batches_stats(void *handle, int64_t *lds_batches, int64_t *hbm_batches)
{
    OG_CHECK(handle, OMNI_GPU_ERR_INVALID, "null handle");
    BatchesHandle &h = *static_cast<BatchesHandle *>(handle);
    // derived from the per-batch flags of the last run (no counters on the hot path)
    int64_t nb = h.have_desc ? static_cast<int64_t>(h.h_offsets.size()) - 1 : 0;
    std::vector<unsigned int> sp(static_cast<size_t>(std::max<int64_t>(nb, 1)));
    if (nb > 0) {
        OG_TRY(CopyToHost(h.stream, {{sp.data(), h.spilled.p, sizeof(unsigned int) * nb}}));
    }
    int64_t hbm = 0;
    for (int64_t b = 0; b < nb; ++b) hbm += sp[b] ? 1 : 0;
    if (lds_batches) *lds_batches = nb - hbm;
    if (hbm_batches) *hbm_batches = hbm;
    return OMNI_GPU_OK;
}

int omni_gpu_agg_batches_phases(void *handle, int64_t *phase1, int64_t *phase2, int64_t *phase3)
{
    OG_CHECK(handle, OMNI_GPU_ERR_INVALID, "null handle");
    BatchesHandle &h = *static_cast<BatchesHandle *>(handle);
    int64_t nb = h.have_desc ? static_cast<int64_t>(h.h_offsets.size()) - 1 : 0;
    std::vector<unsigned int> sp(static_cast<size_t>(std::max<int64_t>(nb, 1)));
    if (nb > 0) {
        OG_TRY(CopyToHost(h.stream, {{sp.data(), h.spilled.p, sizeof(unsigned int) * nb}}));
    }
    // flags of the last run: 0 finished in phase 1; 1 finished in phase 2 (or, in the per-batch form, in
    // HBM tables); 2 finished in HBM tables after a phase-2 partition overflowed
    int64_t n[3] = {0, 0, 0};
    for (int64_t b = 0; b < nb; ++b) n[sp[b] < 3 ? sp[b] : 2] += 1;
    if (h.last_per_batch) {
        n[2] += n[1];
        n[1] = 0;
    }
    if (phase1) *phase1 = n[0];
    if (phase2) *phase2 = n[1];
    if (phase3) *phase3 = n[2];
    return OMNI_GPU_OK;
}

int omni_gpu_agg_batches_destroy(void *handle)
{
    if (!handle) return OMNI_GPU_OK;
    BatchesHandle *h = static_cast<BatchesHandle *>(handle);
    DropGraph(*h);
    hipStream_t s = h->stream;
    h->dcfg.Release(s);
    h->items.Release(s);
    h->parts.Release(s);
    h->parts_used.Release(s);
    h->order.Release(s);
    h->sketch_modes.Release(s);
    h->adv_x.Release(s);
    h->adv_y.Release(s);
    h->offsets.Release(s);
    h->tiles.Release(s);
    h->modes.Release(s);
    h->spilled.Release(s);
    h->groups.Release(s);
    h->region_off.Release(s);
    h->pool.Release(s);
    h->counters.Release(s);
    ReleaseStream(s);
    delete h;
    return OMNI_GPU_OK;
}

}  // extern "C"
