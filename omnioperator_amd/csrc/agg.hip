// Group-by hash aggregation on gfx950.
//
// Replaces the per-row Emplace loop + per-aggregator ProcessGroup scatter of HashAggregationOperator
// (core/src/operator/aggregation/group_aggregation.cpp:811-890, aggregator/operations_hash_aggregator.h:16-32)
// and its output traversal (group_aggregation.cpp:892-955, 1486-1538).
//
// Device layout (all in HBM, sized by the host):
//   single-word keys (one key column of width <= 8):
//     slots[cap]      {u64 key; u32 gid+1; u32 pad}   open addressing, linear probing, 64-bit CAS on key
//     side cells      special[0] = NULL-key group, special[1] = group of the key equal to the slot sentinel
//   multi-word keys (several key columns or a DECIMAL128 key), W words per key:
//     ctrl[cap]       u32: 0 empty, 1 being written, gid+2 published
//     skeys[cap * W]  key words (word 0 carries the NULL bits of every key column and a <=4-byte column)
//   per group, indexed by a group id:
//     gkeys[G * W], gnull[G] (single-word only), gvalid[G], and one state row per group (array of
//     structures: every aggregate's words of a group share one row, so a row update touches one cache line)
//   group ids: each wave takes chunks of ids from one counter and hands them out to its lanes; ids a wave
//     reserved but never used stay invalid and the output skips them (prefix-scan compaction of gvalid)
//
// Results are independent of the hash function and of insertion order for every integer aggregate
// (wrapping SUM, COUNT, MIN, MAX and exact 192-bit decimal accumulation); double SUM/AVG depend on the
// order of atomic additions and are held to a relative tolerance.
#include <algorithm>
#include <limits>
#include <cstring>
#include <memory>
#include <vector>

#include "agg_kernels.h"
#include "agg_layout.h"
#include "scan_device.h"

namespace omni_gpu {

// ---- host side ----------------------------------------------------------------------------------------------

constexpr double kMaxLoad = 0.5;

inline unsigned int Blocks(int64_t n) { return static_cast<unsigned int>((n + kBlock - 1) / kBlock); }

inline uint64_t NextPow2(uint64_t x)
{
    uint64_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

struct AggHandle {
    omni_gpu_agg_spec spec{};
    int device = 0;
    hipStream_t stream = nullptr;
    bool single = true;
    KeyDev kd{};
    AggArgs args{};  // value pointers filled per batch
    uint64_t cap = 0;
    int64_t max_groups = 0;
    DevBuf slots, ctrl, skeys, gkeys, gnull, gvalid, states, row_tmpl, cursors;
    DevBuf index, tile_counts;      // output compaction: live group ids in ascending order
    DevBuf stage_vals[kStageCols], stage_nulls[kStageCols], stage_src, stage_rows;  // omni_gpu_agg_add_batches
    DevBuf counters;  // [0] ids handed out u64, [8] overflow u32, [12] error u32, [16..24) special u32[2],
                      // [24] spills u64, [32] live groups u32, [36] checked decimal groups crossed u32,
                      // [64] per-workgroup key min / max pairs i64[2 * kMinMaxMaxBlocks] (ARRAY rule)
    bool checked = false;           // some aggregate is a checked decimal SUM (K_SUM_DEC_CHK): global table only,
                                    // each launch records its rows' group ids and is settled and replayed
    DevBuf row_gid;                 // group id of each row of the current launch (checked decimal SUM)
    int64_t rows_added = 0;         // input rows before the current launch (positions for double MIN/MAX ties)
    int64_t known_ids = 0;          // ids handed out at the last synchronisation
    int64_t pending_rows = 0;       // rows added since then
    int64_t live_groups = -1;       // live groups after the last compaction (-1: stale)
    bool use_lds = false;
    LdsLayout lds{};
    int64_t lds_rows = 0;    // rows sent through the LDS kernel since the last spill check
    bool map_planned = false;       // the direct-index map was considered (first insert)
    DevBuf amap;                    // direct-index map cells (ARRAY rule)
    ArrayMapDev am{};
    long long map_min = 0;          // the map's key range [map_min, map_min + am.range)
    int64_t spills_seen = 0;

    unsigned long long *ngroups() const { return reinterpret_cast<unsigned long long *>(counters.as<uint8_t>()); }
    unsigned int *overflow() const { return reinterpret_cast<unsigned int *>(counters.as<uint8_t>() + 8); }
    unsigned int *error_flag() const { return reinterpret_cast<unsigned int *>(counters.as<uint8_t>() + 12); }
    unsigned int *special() const { return reinterpret_cast<unsigned int *>(counters.as<uint8_t>() + 16); }
    unsigned long long *spills() const { return reinterpret_cast<unsigned long long *>(counters.as<uint8_t>() + 24); }
    unsigned int *live() const { return reinterpret_cast<unsigned int *>(counters.as<uint8_t>() + 32); }
    unsigned int *crossed() const { return reinterpret_cast<unsigned int *>(counters.as<uint8_t>() + 36); }

    // Ids a wave takes at a time: about the groups expected (seen so far, or the capacity hint if larger) over
    // the wave slots, a power of two in [1, 512], so that small group counts leave no unused reserved ids and
    // large ones need few counter atomics -- also before the first synchronisation of a presized table.
    unsigned long long Chunk() const
    {
        const int64_t expect = std::max<int64_t>(known_ids, spec.capacity_hint);
        unsigned long long c = 1;
        while (c < 512 && static_cast<int64_t>(c) * 2 * kCursorSlots <= expect) c <<= 1;
        return c;
    }

    GroupCtx Ctx() const
    {
        GroupCtx g{};
        g.ngroups = ngroups();
        g.max_groups = static_cast<unsigned long long>(max_groups);
        g.overflow = overflow();
        g.gkeys = gkeys.as<unsigned long long>();
        g.gnull = single ? gnull.as<uint8_t>() : nullptr;
        g.gvalid = gvalid.as<uint8_t>();
        g.special = special();
        g.wave = nullptr;
        g.chunk = Chunk();
        return g;
    }
};

int Fill32(hipStream_t s, void *p, unsigned int v, int64_t n)
{
    if (n <= 0) return OMNI_GPU_OK;
    KernelScope ks(s, "fill_u32");
    fill_u32<<<Blocks(n), kBlock, 0, s>>>(static_cast<unsigned int *>(p), v, n);
    return CheckLaunch("fill_u32");
}

int Fill8(hipStream_t s, void *p, uint8_t v, int64_t n)
{
    if (n <= 0) return OMNI_GPU_OK;
    KernelScope ks(s, "fill_u8");
    fill_u8<<<Blocks(n), kBlock, 0, s>>>(static_cast<uint8_t *>(p), v, n);
    return CheckLaunch("fill_u8");
}

// Allocates a table of `cap` slots and group arrays for cap * kMaxLoad groups, keeping group ids [0, keep).
int Resize(AggHandle &h, uint64_t cap, int64_t keep)
{
    hipStream_t s = h.stream;
    int64_t newMax = static_cast<int64_t>(static_cast<double>(cap) * kMaxLoad);
    int W = h.single ? 1 : h.kd.W;
    const size_t R = static_cast<size_t>(h.args.row_bytes);
    // group arrays (contents of the first `keep` ids kept)
    OG_TRY(h.gkeys.Grow(sizeof(unsigned long long) * W * newMax, sizeof(unsigned long long) * W * keep, s));
    if (h.single) OG_TRY(h.gnull.Grow(newMax, keep, s));
    OG_TRY(h.gvalid.Grow(newMax, keep, s));
    OG_TRY(h.states.Grow(R * newMax, R * keep, s));
    h.args.states = h.states.as<uint8_t>();
    OG_TRY(Fill8(s, h.gvalid.as<uint8_t>() + keep, 0, newMax - keep));
    {
        const int64_t words = (newMax - keep) * static_cast<int64_t>(R / 8);
        if (words > 0) {
            KernelScope ks(s, "agg_init_states");
            init_state_rows<<<Blocks(words), kBlock, 0, s>>>(h.states.as<unsigned long long>(),
                                                             h.row_tmpl.as<unsigned long long>(),
                                                             static_cast<int>(R / 8), keep, newMax - keep);
            OG_TRY(CheckLaunch("init_state_rows"));
        }
    }
    // fresh table, then re-insert the kept groups under their ids
    unsigned long long mask = cap - 1;
    if (h.single) {
        OG_TRY(h.slots.Reserve(sizeof(Slot1) * cap, s));
        {
            KernelScope ks(s, "agg_init_slots");
            init_slots1<<<Blocks(static_cast<int64_t>(cap)), kBlock, 0, s>>>(h.slots.as<Slot1>(), static_cast<int64_t>(cap));
            OG_TRY(CheckLaunch("init_slots1"));
        }
        if (keep > 0) {
            KernelScope ks(s, "agg_rehash");
            agg_rehash_single<<<Blocks(keep), kBlock, 0, s>>>(h.gkeys.as<unsigned long long>(), h.gvalid.as<uint8_t>(),
                                                               keep, h.slots.as<Slot1>(), mask, h.special());
            OG_TRY(CheckLaunch("agg_rehash_single"));
        }
    } else {
        OG_TRY(h.ctrl.Reserve(sizeof(unsigned int) * cap, s));
        OG_TRY(h.skeys.Reserve(sizeof(unsigned long long) * W * cap, s));
        OG_TRY(Fill32(s, h.ctrl.p, 0, static_cast<int64_t>(cap)));
        if (keep > 0) {
            KernelScope ks(s, "agg_rehash");
            agg_rehash_multi<<<Blocks(keep), kBlock, 0, s>>>(h.gkeys.as<unsigned long long>(), h.gvalid.as<uint8_t>(),
                                                              keep, W, h.ctrl.as<unsigned int>(),
                                                              h.skeys.as<unsigned long long>(), mask);
            OG_TRY(CheckLaunch("agg_rehash_multi"));
        }
    }
    h.cap = cap;
    h.max_groups = newMax;
    return OMNI_GPU_OK;
}

int SyncGroups(AggHandle &h)
{
    unsigned long long g = 0, sp = 0;
    unsigned int ovf = 0;
    OG_TRY(CopyToHost(h.stream, {{&g, h.ngroups(), 8}, {&ovf, h.overflow(), 4}, {&sp, h.spills(), 8}}));
    OG_CHECK(ovf == 0, OMNI_GPU_ERR_STATE, "aggregation table overflowed its group arrays (internal sizing error)");
    h.known_ids = static_cast<int64_t>(g);
    h.pending_rows = 0;
    if (h.use_lds && h.spec.strategy == OMNI_GPU_STRATEGY_AUTO && h.lds_rows > 0) {
        // Adaptive fallback: when most rows miss the LDS table (too many groups per workgroup chunk),
        // later batches go straight to the global table.
        int64_t delta = static_cast<int64_t>(sp) - h.spills_seen;
        h.spills_seen = static_cast<int64_t>(sp);
        if (delta * 2 > h.lds_rows) h.use_lds = false;
        h.lds_rows = 0;
    }
    return OMNI_GPU_OK;
}

// Live group ids in ascending order into h.index (prefix-scan compaction of the gvalid flags); sets
// h.live_groups.  Ids reserved by waves but never used are skipped.
int Compact(AggHandle &h)
{
    OG_TRY(SyncGroups(h));
    hipStream_t s = h.stream;
    const int64_t n = h.known_ids;
    if (n == 0) {
        h.live_groups = 0;
        return OMNI_GPU_OK;
    }
    const int64_t tiles = (n + kFlagTile - 1) / kFlagTile;
    OG_TRY(h.tile_counts.Reserve(sizeof(unsigned int) * tiles, s));
    OG_TRY(h.index.Reserve(sizeof(unsigned int) * n, s));
    {
        KernelScope ks(s, "agg_compact");
        flag_tile_counts<<<static_cast<unsigned int>(tiles), kScanBlock, 0, s>>>(h.gvalid.as<uint8_t>(), n,
                                                                                 h.tile_counts.as<unsigned int>());
        scan_sums<<<1, kScanBlock, 0, s>>>(h.tile_counts.as<unsigned int>(), static_cast<int>(tiles), h.live());
        flag_compact<<<static_cast<unsigned int>(tiles), kScanBlock, 0, s>>>(
            h.gvalid.as<uint8_t>(), n, h.tile_counts.as<unsigned int>(), h.index.as<unsigned int>());
        OG_TRY(CheckLaunch("agg_compact"));
    }
    unsigned int live = 0;
    OG_TRY(CopyToHost(s, {{&live, h.live(), 4}}));
    h.live_groups = live;
    return OMNI_GPU_OK;
}

constexpr int kLdsBudget = 64 * 1024;

}  // namespace omni_gpu

using namespace omni_gpu;

extern "C" {

int omni_gpu_agg_create(const omni_gpu_agg_spec *spec, void **handle)
{
    OG_CHECK(spec && handle, OMNI_GPU_ERR_INVALID, "null argument");
    OG_CHECK(spec->num_keys >= 1 && spec->num_keys <= OMNI_GPU_MAX_KEYS, OMNI_GPU_ERR_INVALID,
             "num_keys must be in [1, %d]", OMNI_GPU_MAX_KEYS);
    OG_CHECK(spec->num_aggs >= 0 && spec->num_aggs <= OMNI_GPU_MAX_AGGS, OMNI_GPU_ERR_INVALID,
             "num_aggs must be in [0, %d]", OMNI_GPU_MAX_AGGS);
    OG_CHECK(spec->strategy >= OMNI_GPU_STRATEGY_AUTO && spec->strategy <= OMNI_GPU_STRATEGY_GLOBAL,
             OMNI_GPU_ERR_INVALID, "strategy %d is not available on this path", spec->strategy);
    std::unique_ptr<AggHandle> h(new AggHandle());
    h->spec = *spec;
    h->device = CurrentDevice();
    // key layout
    KeyDev &kd = h->kd;
    kd.num = spec->num_keys;
    for (int c = 0; c < kd.num; ++c) {
        kd.width[c] = TypeWidth(spec->key_types[c]);
        OG_CHECK(kd.width[c] > 0, OMNI_GPU_ERR_INVALID, "key type %d is not supported", spec->key_types[c]);
    }
    h->single = kd.num == 1 && kd.width[0] <= 8;
    if (h->single) {
        kd.W = 1;
        kd.word[0] = 0;
    } else {
        int w = 1;  // word 0: NULL bits (low 32) + first <= 4-byte column (high 32)
        bool word0Free = true;
        for (int c = 0; c < kd.num; ++c) {
            if (kd.width[c] <= 4 && word0Free) {
                kd.in_word0[c] = 1;
                kd.word[c] = 0;
                word0Free = false;
            } else {
                kd.word[c] = w;
                w += kd.width[c] == 16 ? 2 : 1;
            }
        }
        kd.W = w;
        OG_CHECK(kd.W <= kMaxWords, OMNI_GPU_ERR_INVALID, "composite key too wide (%d words)", kd.W);
    }
    h->args.n = spec->num_aggs;
    for (int i = 0; i < spec->num_aggs; ++i) {
        OG_TRY(ClassifyAgg(*spec, i, h->args.a[i]));
        h->checked |= h->args.a[i].kind == K_SUM_DEC_CHK;
    }
    AssignRowLayout(h->args);
    OG_TRY(AcquireStream(&h->stream));
    // 64 bytes of counters, then the per-workgroup key min / max pairs the ARRAY rule plans its map from
    OG_TRY(h->counters.Reserve(64 + 16 * kMinMaxMaxBlocks, h->stream));
    OG_HIP(hipMemsetAsync(h->counters.p, 0, 80, h->stream));
    OG_TRY(h->cursors.Reserve(sizeof(unsigned long long) * 2 * kCursorSlots, h->stream));
    OG_HIP(hipMemsetAsync(h->cursors.p, 0, sizeof(unsigned long long) * 2 * kCursorSlots, h->stream));
    std::vector<unsigned long long> tmpl = RowTemplate(h->args);
    OG_TRY(h->row_tmpl.Reserve(sizeof(unsigned long long) * tmpl.size(), h->stream));
    OG_TRY(CopyToDevice(h->row_tmpl.p, tmpl.data(), sizeof(unsigned long long) * tmpl.size(), h->stream));
    uint64_t want = spec->capacity_hint > 0 ? static_cast<uint64_t>(spec->capacity_hint) : 16384;
    uint64_t cap = NextPow2(static_cast<uint64_t>(static_cast<double>(want + 2) / kMaxLoad) + 1);
    cap = std::max<uint64_t>(cap, 1024);
    OG_TRY(Resize(*h, cap, 0));
    // Strategy: the LDS-local table serves single-word keys whose group count fits it (a capacity hint
    // above half the LDS slots, or a multi-word key, selects the global table directly).
    h->lds = MakeLdsLayout(h->args, kLdsBudget, h->single ? 1 : h->kd.W);
    bool ldsFits = h->lds.S >= 256;
    if (spec->strategy == OMNI_GPU_STRATEGY_LDS) {
        OG_CHECK(ldsFits, OMNI_GPU_ERR_INVALID, "the aggregates leave too few LDS slots for the LDS strategy");
        h->use_lds = true;
    } else if (spec->strategy == OMNI_GPU_STRATEGY_AUTO) {
        h->use_lds = ldsFits && (spec->capacity_hint <= 0 || spec->capacity_hint * 2 <= h->lds.S);
    }
    // A checked decimal SUM is replayed in row order from each row's group id, which the global kernels record;
    // the strategy only moves time, so such a handle takes the global table whatever was asked.
    if (h->checked) h->use_lds = false;
    *handle = h.release();
    return OMNI_GPU_OK;
}

// The ARRAY hash-table rule (AggHashTableRule::ARRAY) for one integer key: VectorAnalyzer::CheckArrayMap
// (core/src/operator/hashmap/vector_analyzer.cpp:10-80) over the first inserted rows -- no array when every key is
// NULL, when max - min could overflow or exceeds 2^21 - 2; otherwise the range is widened by max(50, max - min)
// on each side, saturating at the type's limits, and declined when the widened range and its NULL slot exceed
// 2^21 entries.  The reference resizes once and then moves to its hash map
// when later keys leave the range; here such keys go to the hash table directly (results are the same).
int PlanArrayMap(AggHandle &h, int64_t rows)
{
    h.map_planned = true;
    const int t = h.spec.key_types[0];
    if (h.spec.hash_table_rule != OMNI_GPU_HASH_TABLE_ARRAY || !h.single || h.kd.num != 1 ||
        !(t == OMNI_TYPE_SHORT || t == OMNI_TYPE_INT || t == OMNI_TYPE_DATE32 || t == OMNI_TYPE_LONG ||
          t == OMNI_TYPE_TIMESTAMP || t == OMNI_TYPE_DECIMAL64)) {
        return OMNI_GPU_OK;
    }
    hipStream_t s = h.stream;
    uint8_t *sc = h.counters.as<uint8_t>() + 64;  // room for kMinMaxMaxBlocks {min, max} pairs
    const unsigned int grid = static_cast<unsigned int>(
        std::max<int64_t>(1, std::min<int64_t>((rows + kBlock * kMinMaxUnroll - 1) / (kBlock * kMinMaxUnroll),
                                               kMinMaxMaxBlocks)));
    {
        KernelScope ks(s, "agg_key_minmax");
        agg_key_minmax<<<grid, kBlock, 0, s>>>(h.kd, rows, reinterpret_cast<long long *>(sc));
        OG_TRY(CheckLaunch("agg_key_minmax"));
    }
    std::vector<long long> pairs(2 * static_cast<size_t>(grid));
    OG_TRY(CopyToHost(s, {{pairs.data(), sc, pairs.size() * sizeof(long long)}}));
    long long mm[2] = {std::numeric_limits<long long>::max(), std::numeric_limits<long long>::min()};
    for (unsigned int b = 0; b < grid; ++b) {
        mm[0] = std::min(mm[0], pairs[2 * b]);
        mm[1] = std::max(mm[1], pairs[2 * b + 1]);
    }
    const long long mn = mm[0], mx = mm[1];
    if (mx < mn) return OMNI_GPU_OK;                                             // every key NULL
    if (mx >= 0 && mn < mx - std::numeric_limits<long long>::max()) return OMNI_GPU_OK;  // max - min overflows
    if (mx - mn > (2LL << 20) - 2) return OMNI_GPU_OK;                           // range above 2M entries
    long long tmin = std::numeric_limits<long long>::min(), tmax = std::numeric_limits<long long>::max();
    if (t == OMNI_TYPE_SHORT) {
        tmin = std::numeric_limits<int16_t>::min();
        tmax = std::numeric_limits<int16_t>::max();
    } else if (t == OMNI_TYPE_INT || t == OMNI_TYPE_DATE32) {
        tmin = std::numeric_limits<int32_t>::min();
        tmax = std::numeric_limits<int32_t>::max();
    }
    const long long reserve = std::max(50LL, mx - mn);
    const long long hi = mx <= tmax - reserve ? mx + reserve : tmax;
    const long long lo = mn >= tmin + reserve ? mn - reserve : tmin;
    const uint64_t cells = static_cast<uint64_t>(hi) - static_cast<uint64_t>(lo) + 1;
    // the widened range plus the NULL slot must stay within RANGE_UPPER_BOUND = 2^21 entries (GetRange() check,
    // vector_analyzer.cpp:74-77)
    if (cells + 1 > (2ULL << 20)) return OMNI_GPU_OK;
    OG_TRY(h.amap.Reserve(sizeof(unsigned int) * cells, s));
    OG_TRY(Fill32(s, h.amap.p, 0u, static_cast<int64_t>(cells)));
    h.am.cells = h.amap.as<unsigned int>();
    h.am.min = static_cast<unsigned long long>(lo);
    h.am.range = cells;
    h.map_min = lo;
    return OMNI_GPU_OK;
}

int omni_gpu_agg_add_batch(void *handle, const omni_gpu_column *keys, const omni_gpu_column *aggs, int64_t rows)
{
    OG_CHECK(handle, OMNI_GPU_ERR_INVALID, "null handle");
    AggHandle &h = *static_cast<AggHandle *>(handle);
    OG_CHECK(rows >= 0, OMNI_GPU_ERR_INVALID, "negative row count");
    if (rows == 0) return OMNI_GPU_OK;
    OG_CHECK(keys, OMNI_GPU_ERR_INVALID, "null key columns");
    OG_CHECK(h.args.n == 0 || aggs, OMNI_GPU_ERR_INVALID, "null aggregate columns");
    for (int c = 0; c < h.kd.num; ++c) {
        OG_CHECK(keys[c].type_id == h.spec.key_types[c], OMNI_GPU_ERR_INVALID, "key %d type %d != spec %d", c,
                 keys[c].type_id, h.spec.key_types[c]);
        OG_CHECK(keys[c].values, OMNI_GPU_ERR_INVALID, "key %d has no values", c);
        h.kd.values[c] = keys[c].values;
        h.kd.nulls[c] = keys[c].nulls;
    }
    for (int i = 0; i < h.args.n; ++i) {
        AggDev &d = h.args.a[i];
        if (d.kind == K_COUNT_ALL) {
            d.values = nullptr;
            d.nulls = nullptr;
            continue;
        }
        OG_CHECK(aggs[i].type_id == h.spec.agg_in_types[i], OMNI_GPU_ERR_INVALID, "aggregate %d input type %d != spec %d",
                 i, aggs[i].type_id, h.spec.agg_in_types[i]);
        OG_CHECK(aggs[i].values, OMNI_GPU_ERR_INVALID, "aggregate %d has no values", i);
        d.values = aggs[i].values;
        d.nulls = aggs[i].nulls;
    }
    h.live_groups = -1;
    // Capacity: ids after this batch <= ids now + one per new row + one chunk per wave slot in reserve.
    auto need = [&]() {
        return h.known_ids + h.pending_rows + rows + static_cast<int64_t>(h.Chunk()) * kCursorSlots + 2;
    };
    if (need() > h.max_groups) {
        OG_TRY(SyncGroups(h));
        if (need() > h.max_groups) {
            uint64_t cap = h.cap;
            while (static_cast<int64_t>(static_cast<double>(cap) * kMaxLoad) < need()) cap *= 4;
            OG_TRY(Resize(h, cap, h.known_ids));
        }
    }
    if (!h.map_planned) OG_TRY(PlanArrayMap(h, rows));
    h.args.row_base = h.rows_added;
    h.rows_added += rows;
    GroupCtx g = h.Ctx();
    unsigned long long *cur = h.cursors.as<unsigned long long>();
    if (h.use_lds) {
        // Rows per thread: each workgroup merges its distinct keys into the global table once, and those merges
        // contend on the same state rows, so few groups favour long chunks (16 rows per thread halves the
        // merges: 1M rows over 1,000 groups 0.175 -> 0.125 ms); from about 4096 groups on, a 16384-row chunk
        // holds most of them, the merge no longer shrinks and 8 rows per thread keeps twice the workgroups
        // busy (1M rows over 10,000 groups: 0.40 ms at 8, 0.46 ms at 16).
        const int64_t expect = std::max<int64_t>(h.known_ids, h.spec.capacity_hint);
        const int64_t rpt = expect >= 4096 ? 8 : 16;
        int64_t blocks = (rows + kLdsBlock * rpt - 1) / (kLdsBlock * rpt);
        blocks = std::max<int64_t>(1, std::min<int64_t>(blocks, kMaxLdsBlocks));
        int64_t per = (rows + blocks - 1) / blocks;
        KernelScope ks(h.stream, "agg_insert_lds");
        if (h.single) {
            agg_insert_lds<<<static_cast<unsigned int>(blocks), kLdsBlock, h.lds.bytes, h.stream>>>(
                h.kd, rows, per, h.slots.as<Slot1>(), h.cap - 1, g, h.args, h.lds, h.spills(), cur, h.am);
        } else {
            agg_insert_lds_multi<<<static_cast<unsigned int>(blocks), kLdsBlock, h.lds.bytes, h.stream>>>(
                h.kd, rows, per, h.ctrl.as<unsigned int>(), h.skeys.as<unsigned long long>(), h.cap - 1, g, h.args,
                h.lds, h.spills(), cur);
        }
        OG_TRY(CheckLaunch("agg_insert_lds"));
        h.lds_rows += rows;
        h.pending_rows += rows;
        return OMNI_GPU_OK;
    }
    unsigned int *rowGid = nullptr;
    if (h.checked) {
        OG_TRY(h.row_gid.Reserve(sizeof(unsigned int) * static_cast<size_t>(rows), h.stream));
        rowGid = h.row_gid.as<unsigned int>();
    }
    const unsigned int grid = static_cast<unsigned int>(std::min<int64_t>(Blocks(rows), kMaxInsertBlocks));
    {
        KernelScope ks(h.stream, h.single ? "agg_insert_single" : "agg_insert_multi");
        if (h.single) {
            agg_insert_single<<<grid, kBlock, 0, h.stream>>>(h.kd, rows, h.slots.as<Slot1>(), h.cap - 1, g, h.args, cur,
                                                             h.am, rowGid);
        } else {
            agg_insert_multi<<<grid, kBlock, 0, h.stream>>>(h.kd, rows, h.ctrl.as<unsigned int>(),
                                                            h.skeys.as<unsigned long long>(), h.cap - 1, g, h.args, cur,
                                                            rowGid);
        }
        OG_TRY(CheckLaunch("agg_insert"));
    }
    if (h.checked) {
        // checked decimal SUM: mark the groups whose sums passed their limits, then replay their rows in order
        // (the replay launch returns at once when no group was marked)
        KernelScope ks(h.stream, "agg_dec_check");
        agg_dec_settle<<<grid, kBlock, 0, h.stream>>>(rows, rowGid, h.args, h.crossed());
        agg_dec_replay<<<1, kReplayBlock, 0, h.stream>>>(rows, rowGid, h.args, h.crossed());
        OG_TRY(CheckLaunch("agg_dec_check"));
    }
    h.pending_rows += rows;
    return OMNI_GPU_OK;
}

int omni_gpu_agg_add_batches(void *handle, const omni_gpu_column *keys, const omni_gpu_column *aggs,
                             const int64_t *rows, int32_t num_batches)
{
    OG_CHECK(handle && rows && num_batches >= 0, OMNI_GPU_ERR_INVALID, "null argument");
    AggHandle &h = *static_cast<AggHandle *>(handle);
    if (num_batches == 0) return OMNI_GPU_OK;
    OG_CHECK(keys, OMNI_GPU_ERR_INVALID, "null key columns");
    OG_CHECK(h.args.n == 0 || aggs, OMNI_GPU_ERR_INVALID, "null aggregate columns");
    const int nk = h.kd.num, na = h.args.n;
    if (num_batches == 1) return omni_gpu_agg_add_batch(handle, keys, aggs, rows[0]);
    std::vector<long long> start(static_cast<size_t>(num_batches) + 1, 0);
    for (int b = 0; b < num_batches; ++b) {
        OG_CHECK(rows[b] >= 0, OMNI_GPU_ERR_INVALID, "negative row count in batch %d", b);
        start[b + 1] = start[b] + rows[b];
    }
    const long long total = start[num_batches];
    if (total == 0) return OMNI_GPU_OK;
    // staged columns: the keys, then each aggregate input that is not COUNT(*) and is not the same column
    // (same pointers in every batch) as an earlier one
    std::vector<const omni_gpu_column *> col0;      // the column in batch 0 (type, pointers)
    std::vector<std::pair<const omni_gpu_column *, int>> cols;  // (array, stride) per staged column
    int keyCol[OMNI_GPU_MAX_KEYS], aggCol[OMNI_GPU_MAX_AGGS];
    for (int c = 0; c < nk; ++c) {
        OG_CHECK(keys[c].type_id == h.spec.key_types[c], OMNI_GPU_ERR_INVALID, "key %d type %d != spec %d", c,
                 keys[c].type_id, h.spec.key_types[c]);
        keyCol[c] = static_cast<int>(cols.size());
        cols.push_back({keys + c, nk});
    }
    for (int i = 0; i < na; ++i) {
        aggCol[i] = -1;
        if (h.args.a[i].kind == K_COUNT_ALL) continue;
        OG_CHECK(aggs[i].type_id == h.spec.agg_in_types[i], OMNI_GPU_ERR_INVALID,
                 "aggregate %d input type %d != spec %d", i, aggs[i].type_id, h.spec.agg_in_types[i]);
        for (int j = 0; j < static_cast<int>(cols.size()) && aggCol[i] < 0; ++j) {
            bool same = TypeWidth(cols[j].first[0].type_id) == TypeWidth(aggs[i].type_id);
            for (int b = 0; same && b < num_batches; ++b) {
                const omni_gpu_column &x = cols[j].first[static_cast<size_t>(b) * cols[j].second];
                const omni_gpu_column &y = aggs[static_cast<size_t>(b) * na + i];
                same = x.values == y.values && x.nulls == y.nulls;
            }
            if (same) aggCol[i] = j;
        }
        if (aggCol[i] < 0) {
            aggCol[i] = static_cast<int>(cols.size());
            cols.push_back({aggs + i, na});
        }
    }
    const int nc = static_cast<int>(cols.size());
    hipStream_t s = h.stream;
    StageArgs a{};
    a.nb = num_batches;
    a.ncols = nc;
    a.rows = total;
    std::vector<StageSrc> src(static_cast<size_t>(num_batches) * nc);
    for (int c = 0; c < nc; ++c) {
        const int w = TypeWidth(cols[c].first[0].type_id);
        bool anyNull = false;
        for (int b = 0; b < num_batches; ++b) {
            const omni_gpu_column &x = cols[c].first[static_cast<size_t>(b) * cols[c].second];
            OG_CHECK(rows[b] == 0 || x.values, OMNI_GPU_ERR_INVALID, "batch %d column %d has no values", b, c);
            src[static_cast<size_t>(b) * nc + c] = {x.values, x.nulls};
            anyNull = anyNull || (x.nulls != nullptr && rows[b] > 0);
        }
        OG_TRY(h.stage_vals[c].Reserve(static_cast<size_t>(w) * total, s));
        a.width[c] = w;
        a.dst[c] = h.stage_vals[c].p;
        if (anyNull) {
            OG_TRY(h.stage_nulls[c].Reserve(sizeof(unsigned int) * ((total + 31) / 32), s));
            a.dst_nulls[c] = h.stage_nulls[c].as<unsigned int>();
        }
    }
    OG_TRY(h.stage_src.Reserve(sizeof(StageSrc) * src.size(), s));
    OG_TRY(h.stage_rows.Reserve(sizeof(long long) * start.size(), s));
    OG_TRY(CopyToDevice(h.stage_src.p, src.data(), sizeof(StageSrc) * src.size(), s));
    OG_TRY(CopyToDevice(h.stage_rows.p, start.data(), sizeof(long long) * start.size(), s));
    a.src = h.stage_src.as<StageSrc>();
    a.row_start = h.stage_rows.as<long long>();
    {
        KernelScope ks(s, "agg_stage");
        const long long wblocks = ((total + 31) / 32 + kBlock - 1) / kBlock;
        agg_stage_values<<<dim3(static_cast<unsigned int>(Blocks(total)), nc), kBlock, 0, s>>>(a);
        agg_stage_nulls<<<dim3(static_cast<unsigned int>(wblocks), nc), kBlock, 0, s>>>(a);
        OG_TRY(CheckLaunch("agg_stage"));
    }
    omni_gpu_column sk[OMNI_GPU_MAX_KEYS] = {}, sa[OMNI_GPU_MAX_AGGS] = {};
    for (int c = 0; c < nk; ++c) {
        sk[c].type_id = keys[c].type_id;
        sk[c].values = a.dst[keyCol[c]];
        sk[c].nulls = reinterpret_cast<const uint8_t *>(a.dst_nulls[keyCol[c]]);
    }
    for (int i = 0; i < na; ++i) {
        sa[i].type_id = h.spec.agg_in_types[i];
        if (aggCol[i] < 0) continue;
        sa[i].values = a.dst[aggCol[i]];
        sa[i].nulls = reinterpret_cast<const uint8_t *>(a.dst_nulls[aggCol[i]]);
    }
    return omni_gpu_agg_add_batch(handle, sk, sa, total);
}

int omni_gpu_agg_wait(void *handle)
{
    OG_CHECK(handle, OMNI_GPU_ERR_INVALID, "null handle");
    OG_TRY(SyncStream(static_cast<AggHandle *>(handle)->stream));
    return OMNI_GPU_OK;
}

int omni_gpu_agg_num_groups(void *handle, int64_t *groups)
{
    OG_CHECK(handle && groups, OMNI_GPU_ERR_INVALID, "null argument");
    AggHandle &h = *static_cast<AggHandle *>(handle);
    if (h.live_groups < 0) OG_TRY(Compact(h));
    *groups = h.live_groups;
    return OMNI_GPU_OK;
}

int omni_gpu_agg_get_output(void *handle, int64_t start, int64_t count, omni_gpu_out_column *outs)
{
    OG_CHECK(handle && outs, OMNI_GPU_ERR_INVALID, "null argument");
    AggHandle &h = *static_cast<AggHandle *>(handle);
    if (h.live_groups < 0) OG_TRY(Compact(h));
    OG_CHECK(start >= 0 && count >= 0 && start + count <= h.live_groups, OMNI_GPU_ERR_INVALID,
             "output range [%lld, %lld) outside [0, %lld)", static_cast<long long>(start),
             static_cast<long long>(start + count), static_cast<long long>(h.live_groups));
    if (count == 0) return OMNI_GPU_OK;
    OutArgs oa{};
    oa.nkeys = h.kd.num;
    oa.naggs = h.args.n;
    for (int c = 0; c < h.kd.num + h.args.n; ++c) {
        int expect = c < h.kd.num ? h.spec.key_types[c] : h.spec.agg_out_types[c - h.kd.num];
        OG_CHECK(outs[c].type_id == expect, OMNI_GPU_ERR_INVALID, "output column %d type %d != %d", c, outs[c].type_id, expect);
        OG_CHECK(outs[c].values && outs[c].nulls, OMNI_GPU_ERR_INVALID, "output column %d lacks buffers", c);
        oa.o[c].width = TypeWidth(outs[c].type_id);
        oa.o[c].values = outs[c].values;
        oa.o[c].nulls = outs[c].nulls;
    }
    OG_HIP(hipMemsetAsync(h.error_flag(), 0, 4, h.stream));
    {
        KernelScope ks(h.stream, "agg_output");
        agg_output<<<Blocks(count), kBlock, 0, h.stream>>>(start, count, h.kd, h.single ? 1 : 0,
                                                          h.index.as<unsigned int>(), h.gkeys.as<unsigned long long>(),
                                                          h.gnull.as<uint8_t>(), h.args, oa, h.spec.overflow_as_null,
                                                          h.error_flag());
        OG_TRY(CheckLaunch("agg_output"));
    }
    unsigned int err = 0;
    OG_TRY(CopyToHost(h.stream, {{&err, h.error_flag(), 4}}));
    if (err) return SetError(OMNI_GPU_ERR_OVERFLOW, "decimal overflow in SUM (overflowAsNull is false)");
    return OMNI_GPU_OK;
}

int omni_gpu_agg_table_info(void *handle, int32_t *array_map, int64_t *range_min, int64_t *range_cells)
{
    OG_CHECK(handle && array_map && range_min && range_cells, OMNI_GPU_ERR_INVALID, "null argument");
    const AggHandle &h = *static_cast<const AggHandle *>(handle);
    *array_map = h.am.cells != nullptr ? 1 : 0;
    *range_min = h.am.cells != nullptr ? h.map_min : 0;
    *range_cells = h.am.cells != nullptr ? static_cast<int64_t>(h.am.range) : 0;
    return OMNI_GPU_OK;
}

int omni_gpu_agg_destroy(void *handle)
{
    if (!handle) return OMNI_GPU_OK;
    AggHandle *h = static_cast<AggHandle *>(handle);
    hipStream_t s = h->stream;
    for (DevBuf *b : {&h->slots, &h->ctrl, &h->skeys, &h->gkeys, &h->gnull, &h->gvalid, &h->states, &h->row_tmpl,
                      &h->cursors, &h->index, &h->tile_counts, &h->counters, &h->stage_src, &h->stage_rows,
                      &h->amap, &h->row_gid}) {
        b->Release(s);
    }
    for (int c = 0; c < kStageCols; ++c) {
        h->stage_vals[c].Release(s);
        h->stage_nulls[c].Release(s);
    }
    ReleaseStream(s);
    delete h;
    return OMNI_GPU_OK;
}

}  // extern "C"
