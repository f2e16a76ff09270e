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
#include <cmath>
#include <limits>
#include <cstring>
#include <memory>
#include <vector>

#include "agg_kernels.h"
#include "agg_layout.h"
#include "agg_radix.h"
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

// A batch recorded by the radix strategy (its columns are read when the groups are first asked for).
struct RadixBatch {
    omni_gpu_column keys[OMNI_GPU_MAX_KEYS];
    omni_gpu_column aggs[OMNI_GPU_MAX_AGGS];
    int64_t rows;
};

constexpr int kRadixStores = 3;

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
    // radix strategy (agg_radix.h): batches are recorded and partitioned when the groups are first asked for
    bool radix = false;
    bool radix_done = false;        // a radix aggregation ran: later batches go to the hash table (after a rehash)
    KeyDev rkd{};                   // multi-word key layout of the radix encoding (also for single-word handles)
    std::vector<RadixBatch> rbatches;
    int64_t radix_rows = 0;
    DevBuf rvals[kRadixStores][kRadixCols], rnulls[kRadixStores][kRadixCols];
    DevBuf rsegs, rtiles, rcols, rcounts, rsums, rstarts, rfstart, rparents, rovf, rmisc, rhll;

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
int RadixFinish(AggHandle &h);

int Compact(AggHandle &h)
{
    if (!h.rbatches.empty()) OG_TRY(RadixFinish(h));
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

// ---- radix strategy (agg_radix.h) -------------------------------------------------------------------------------

constexpr int kRadixLdsBudget = 32 * 1024;  // LDS table of a final piece: five workgroups share a CU

inline int Log2(uint64_t p)
{
    int k = 0;
    while ((1ULL << k) < p) ++k;
    return k;
}

// Device-wide exclusive scan of n u32 values in place; the total goes to *total (device).
int ScanU32(AggHandle &h, unsigned int *v, long long n, unsigned int *total)
{
    hipStream_t s = h.stream;
    const long long tiles = (n + kScanTile - 1) / kScanTile;
    OG_CHECK(tiles < (1LL << 31), OMNI_GPU_ERR_INVALID, "scan too long");
    OG_TRY(h.rsums.Reserve(sizeof(unsigned int) * std::max<long long>(1, tiles), s));
    KernelScope ks(s, "agg_radix_scan");
    scan_tiles<<<static_cast<unsigned int>(tiles), kScanBlock, 0, s>>>(v, n, h.rsums.as<unsigned int>());
    scan_sums<<<1, kScanBlock, 0, s>>>(h.rsums.as<unsigned int>(), static_cast<int>(tiles), total);
    scan_add<<<static_cast<unsigned int>(tiles), kScanBlock, 0, s>>>(v, n, h.rsums.as<unsigned int>());
    return CheckLaunch("agg_radix_scan");
}

// Stored columns of a radix aggregation: the key columns, then each aggregate input that is not COUNT(*) and
// not the same column (same pointers in every batch) as an earlier one.
struct RadixPlan {
    int nc = 0;
    int width[kRadixCols];
    bool nullable[kRadixCols];
    int aggCol[OMNI_GPU_MAX_AGGS];
    std::vector<RadixSeg> segs;  // one per recorded batch
};

int PlanRadixColumns(const AggHandle &h, RadixPlan &rp)
{
    const int nk = h.kd.num, na = h.args.n;
    const size_t nb = h.rbatches.size();
    rp.segs.assign(nb, RadixSeg{});
    auto colOf = [&](size_t b, int c, int src) -> const omni_gpu_column & {
        return src < nk ? h.rbatches[b].keys[src] : h.rbatches[b].aggs[src - nk];
    };
    std::vector<int> source;  // per stored column: key c, or nk + aggregate i
    for (int c = 0; c < nk; ++c) source.push_back(c);
    for (int i = 0; i < na; ++i) {
        rp.aggCol[i] = -1;
        if (h.args.a[i].kind == K_COUNT_ALL) continue;
        for (int j = 0; j < static_cast<int>(source.size()) && rp.aggCol[i] < 0; ++j) {
            bool same = TypeWidth(colOf(0, j, source[j]).type_id) == TypeWidth(h.spec.agg_in_types[i]);
            for (size_t b = 0; same && b < nb; ++b) {
                const omni_gpu_column &x = colOf(b, j, source[j]), &y = h.rbatches[b].aggs[i];
                same = x.values == y.values && x.nulls == y.nulls;
            }
            if (same) rp.aggCol[i] = j;
        }
        if (rp.aggCol[i] < 0) {
            rp.aggCol[i] = static_cast<int>(source.size());
            source.push_back(nk + i);
        }
    }
    rp.nc = static_cast<int>(source.size());
    OG_CHECK(rp.nc <= kRadixCols, OMNI_GPU_ERR_INVALID, "too many stored columns");
    for (int c = 0; c < rp.nc; ++c) {
        rp.width[c] = TypeWidth(colOf(0, c, source[c]).type_id);
        rp.nullable[c] = false;
        for (size_t b = 0; b < nb; ++b) {
            const omni_gpu_column &x = colOf(b, c, source[c]);
            rp.segs[b].values[c] = x.values;
            rp.segs[b].nulls[c] = x.nulls;
            rp.nullable[c] = rp.nullable[c] || x.nulls != nullptr;
        }
    }
    return OMNI_GPU_OK;
}

// The columns of store k as a segment (and as the scatter's destination).
RadixSeg StoreSeg(const AggHandle &h, const RadixPlan &rp, int k)
{
    RadixSeg g{};
    for (int c = 0; c < rp.nc; ++c) {
        g.values[c] = h.rvals[k][c].p;
        g.nulls[c] = rp.nullable[c] ? h.rnulls[k][c].as<uint8_t>() : nullptr;
    }
    return g;
}

int ReserveStore(AggHandle &h, const RadixPlan &rp, int k, int64_t rows)
{
    hipStream_t s = h.stream;
    for (int c = 0; c < rp.nc; ++c) {
        OG_TRY(h.rvals[k][c].Reserve(static_cast<size_t>(rp.width[c]) * std::max<int64_t>(1, rows), s));
        if (rp.nullable[c]) {
            const int64_t words = std::max<int64_t>(1, (rows + 31) / 32);
            OG_TRY(h.rnulls[k][c].Reserve(sizeof(unsigned int) * words, s));
            OG_TRY(Fill32(s, h.rnulls[k][c].p, 0, words));
        }
    }
    return OMNI_GPU_OK;
}

// One partitioning level: the tiles of `segs` (parent-major count matrix of `entries` entries) into store `dst`
// in P children per parent by hash bits [shift, shift + log2 P).
int RadixLevel(AggHandle &h, const RadixPlan &rp, const std::vector<RadixSeg> &segs,
               const std::vector<RadixTile> &tiles, long long entries, int P, int shift, int dst, int64_t rows,
               uint8_t *hll = nullptr)
{
    hipStream_t s = h.stream;
    if (tiles.empty()) return OMNI_GPU_OK;
    OG_TRY(ReserveStore(h, rp, dst, rows));
    std::vector<RadixCol> cols(rp.nc);
    for (int c = 0; c < rp.nc; ++c) {
        cols[c].width = rp.width[c];
        cols[c].values = h.rvals[dst][c].p;
        cols[c].nulls = rp.nullable[c] ? h.rnulls[dst][c].as<unsigned int>() : nullptr;
    }
    OG_TRY(h.rsegs.Reserve(sizeof(RadixSeg) * segs.size(), s));
    OG_TRY(CopyToDevice(h.rsegs.p, segs.data(), sizeof(RadixSeg) * segs.size(), s));
    OG_TRY(h.rtiles.Reserve(sizeof(RadixTile) * tiles.size(), s));
    OG_TRY(CopyToDevice(h.rtiles.p, tiles.data(), sizeof(RadixTile) * tiles.size(), s));
    OG_TRY(h.rcols.Reserve(sizeof(RadixCol) * cols.size(), s));
    OG_TRY(CopyToDevice(h.rcols.p, cols.data(), sizeof(RadixCol) * cols.size(), s));
    OG_TRY(h.rcounts.Reserve(sizeof(unsigned int) * std::max<long long>(1, entries), s));
    RadixPass a{};
    a.segs = h.rsegs.as<RadixSeg>();
    a.tiles = h.rtiles.as<RadixTile>();
    a.cols = h.rcols.as<RadixCol>();
    a.nc = rp.nc;
    a.shift = shift;
    a.P = P;
    a.counts = h.rcounts.as<unsigned int>();
    a.kd = h.rkd;
    const unsigned int grid = static_cast<unsigned int>(tiles.size());
    {
        KernelScope ks(s, "agg_radix_hist");
        agg_radix_hist<<<grid, kRadixThreads, 0, s>>>(a, hll);
        OG_TRY(CheckLaunch("agg_radix_hist"));
    }
    OG_TRY(ScanU32(h, a.counts, entries, reinterpret_cast<unsigned int *>(h.rmisc.as<uint8_t>())));
    {
        KernelScope ks(s, "agg_radix_scatter");
        if (P <= kRadixFirstFanout) agg_radix_scatter<kRadixFirstFanout><<<grid, kRadixThreads, 0, s>>>(a);
        else agg_radix_scatter<kRadixMaxFanout><<<grid, kRadixThreads, 0, s>>>(a);
        OG_TRY(CheckLaunch("agg_radix_scatter"));
    }
    return OMNI_GPU_OK;
}

// Piece starts of the last level into h.rfstart: parents x P pieces (+ the end), from the scanned count matrix.
int RadixPieceStarts(AggHandle &h, const std::vector<long long> &cbase, const std::vector<int> &tpar,
                     const unsigned int *pstart, int P, unsigned int total)
{
    hipStream_t s = h.stream;
    const int parents = static_cast<int>(cbase.size());
    OG_TRY(h.rparents.Reserve(sizeof(long long) * parents + sizeof(int) * parents, s));
    std::vector<uint8_t> buf(sizeof(long long) * parents + sizeof(int) * parents);
    std::memcpy(buf.data(), cbase.data(), sizeof(long long) * parents);
    std::memcpy(buf.data() + sizeof(long long) * parents, tpar.data(), sizeof(int) * parents);
    OG_TRY(CopyToDevice(h.rparents.p, buf.data(), buf.size(), s));
    const long long n = static_cast<long long>(parents) * P;
    OG_TRY(h.rfstart.Reserve(sizeof(unsigned int) * (n + 1), s));
    KernelScope ks(s, "agg_radix_pieces");
    agg_radix_piece_starts<<<static_cast<unsigned int>((n + 1 + kBlock - 1) / kBlock), kBlock, 0, s>>>(
        h.rcounts.as<unsigned int>(), h.rparents.as<long long>(),
        reinterpret_cast<const int *>(h.rparents.as<uint8_t>() + sizeof(long long) * parents), pstart, parents, P,
        total, h.rfstart.as<unsigned int>());
    return CheckLaunch("agg_radix_piece_starts");
}

// Group arrays for `want` ids keeping [0, keep); the rows past keep are not initialised (the radix path writes
// every word of the groups it creates, and hands ids out exactly; Resize initialises the rest before a hash-table
// insert).
int GrowGroupsRaw(AggHandle &h, int64_t want, int64_t keep)
{
    hipStream_t s = h.stream;
    const int W = h.single ? 1 : h.kd.W;
    const size_t R = static_cast<size_t>(h.args.row_bytes);
    OG_TRY(h.gkeys.Grow(sizeof(unsigned long long) * W * want, sizeof(unsigned long long) * W * keep, s));
    if (h.single) OG_TRY(h.gnull.Grow(want, keep, s));
    OG_TRY(h.gvalid.Grow(want, keep, s));
    OG_TRY(h.states.Grow(R * want, R * keep, s));
    h.args.states = h.states.as<uint8_t>();
    h.max_groups = want;
    return OMNI_GPU_OK;
}

// Count-matrix columns of a parent's tiles tiles[first, first + T) (workgroup b runs tile b): the workgroups of
// one XCD (b mod 8) take adjacent columns, so within a child the runs written by one XCD's workgroups are adjacent
// and their partial lines meet in that XCD's L2 before being written back.
void XcdColumns(std::vector<RadixTile> &tiles, size_t first, int T)
{
    int count[kXcds] = {0};
    for (int t = 0; t < T; ++t) ++count[(first + t) % kXcds];
    int next[kXcds];
    for (int x = 0, run = 0; x < kXcds; ++x) {
        next[x] = run;
        run += count[x];
    }
    for (int t = 0; t < T; ++t) tiles[first + t].tcol = next[(first + t) % kXcds]++;
}

// HyperLogLog estimate of the distinct hashes from kHllRegs registers (small-range correction by linear counting).
double HllEstimate(const uint8_t *M)
{
    const double m = kHllRegs;
    double z = 0;
    int zeros = 0;
    for (int j = 0; j < kHllRegs; ++j) {
        z += std::ldexp(1.0, -static_cast<int>(M[j]));
        zeros += M[j] == 0;
    }
    const double e = 0.7213 / (1.0 + 1.079 / m) * m * m / z;
    if (e <= 2.5 * m && zeros > 0) return m * std::log(m / zeros);
    return e;
}

// Partitions `rows` rows of `segs` (tiles of at most kRadixTile rows per segment) in two levels: PA partitions
// (store 0) by the hash bits just below `shift`, then PB sub-partitions of each (store 1) by the next bits; shift
// drops by the bits used.  PB and the final step's passes (*subBits) follow from the groups expected, a HyperLogLog
// estimate the first level's histogram takes: pieces of about half an LDS table of groups, and more passes only
// when even 2^11 sub-partitions would leave pieces over 85% of a table.  Leaves the
// piece starts in h.rfstart and the store holding the pieces in *store.
int RadixPartition(AggHandle &h, const RadixPlan &rp, const LdsLayout &L, const std::vector<RadixSeg> &segs,
                   const std::vector<int64_t> &segRows, int64_t rows, int &shift, int &store, long long &pieces,
                   int &subBits)
{
    hipStream_t s = h.stream;
    std::vector<RadixTile> tiles;
    for (size_t b = 0; b < segs.size(); ++b) {
        for (int64_t r = 0; r < segRows[b]; r += kRadixTile) {
            RadixTile t{};
            t.row0 = r;
            t.seg = static_cast<int>(b);
            t.rows = static_cast<int>(std::min<int64_t>(kRadixTile, segRows[b] - r));
            t.tcol = static_cast<int>(tiles.size());
            tiles.push_back(t);
        }
    }
    const int TA = static_cast<int>(tiles.size());
    for (RadixTile &t : tiles) t.tpar = TA;
    XcdColumns(tiles, 0, TA);
    const int PA = static_cast<int>(std::min<uint64_t>(
        kRadixFirstFanout, NextPow2(std::max<uint64_t>(2, static_cast<uint64_t>((rows + kRadixTile - 1) / kRadixTile)))));
    const int ka = Log2(static_cast<uint64_t>(PA));
    OG_CHECK(shift - ka >= 32, OMNI_GPU_ERR_STATE, "radix aggregation ran out of hash bits (%lld rows left)",
             static_cast<long long>(rows));
    shift -= ka;
    OG_TRY(h.rhll.Reserve(static_cast<size_t>(TA) * kHllRegs + sizeof(unsigned int) * kHllRegs, s));
    OG_TRY(RadixLevel(h, rp, segs, tiles, static_cast<long long>(PA) * TA, PA, shift, 0, rows, h.rhll.as<uint8_t>()));
    OG_TRY(RadixPieceStarts(h, {0LL}, {TA}, nullptr, PA, static_cast<unsigned int>(rows)));
    unsigned int *regs = reinterpret_cast<unsigned int *>(h.rhll.as<uint8_t>() + static_cast<size_t>(TA) * kHllRegs);
    OG_TRY(Fill32(s, regs, 0, kHllRegs));
    {
        KernelScope ks(s, "agg_radix_hll");
        agg_radix_hll_reduce<<<(TA + kHllTilesPerBlock - 1) / kHllTilesPerBlock, 256, 0, s>>>(h.rhll.as<uint8_t>(),
                                                                                               TA, regs);
        OG_TRY(CheckLaunch("agg_radix_hll_reduce"));
    }
    std::vector<unsigned int> R(kHllRegs);
    std::vector<uint8_t> M(kHllRegs);
    std::vector<unsigned int> pstart(PA + 1);
    OG_TRY(CopyToHost(s, {{R.data(), regs, sizeof(unsigned int) * kHllRegs},
                          {pstart.data(), h.rfstart.p, sizeof(unsigned int) * (PA + 1)}}));
    for (int j = 0; j < kHllRegs; ++j) M[j] = static_cast<uint8_t>(R[j]);
    const double est = std::min(static_cast<double>(rows), std::max(1.0, HllEstimate(M.data()) * 1.05));
    const double perPiece = 0.5 * L.S;
    const uint64_t wantB = NextPow2(std::max<uint64_t>(1, static_cast<uint64_t>(std::ceil(est / (PA * perPiece)))));
    const int PB = static_cast<int>(std::min<uint64_t>(kRadixMaxFanout, wantB));
    const int kb = Log2(static_cast<uint64_t>(PB));
    // passes only when the pieces would fill their tables past 85% (a pass visits its rows in hash order, a gather)
    const uint64_t passes = NextPow2(std::max<uint64_t>(
        1, static_cast<uint64_t>(std::ceil(est / (static_cast<double>(PA) * PB * 0.85 * L.S)))));
    subBits = std::min(Log2(passes), 6);
    OG_CHECK(shift - kb - subBits >= 32, OMNI_GPU_ERR_STATE, "radix aggregation ran out of hash bits (%lld rows left)",
             static_cast<long long>(rows));
    store = 0;
    pieces = PA;
    if (PB <= 1) return OMNI_GPU_OK;
    OG_TRY(h.rstarts.Reserve(sizeof(unsigned int) * (PA + 1), s));
    OG_TRY(CopyToDevice(h.rstarts.p, pstart.data(), sizeof(unsigned int) * (PA + 1), s));
    // level B: each level-A partition of store 0 is a parent
    tiles.clear();
    std::vector<long long> cbase(PA);
    std::vector<int> tpar(PA);
    long long entries = 0;
    for (int p = 0; p < PA; ++p) {
        const int64_t n = static_cast<int64_t>(pstart[p + 1]) - pstart[p];
        const int T = static_cast<int>((n + kRadixTile - 1) / kRadixTile);
        cbase[p] = entries;
        tpar[p] = T;
        const size_t first = tiles.size();
        for (int t = 0; t < T; ++t) {
            RadixTile x{};
            x.row0 = static_cast<int64_t>(pstart[p]) + static_cast<int64_t>(t) * kRadixTile;
            x.cbase = entries;
            x.seg = 0;
            x.rows = static_cast<int>(std::min<int64_t>(kRadixTile, n - static_cast<int64_t>(t) * kRadixTile));
            x.tpar = T;
            tiles.push_back(x);
        }
        XcdColumns(tiles, first, T);
        entries += static_cast<long long>(PB) * T;
    }
    shift -= kb;
    OG_TRY(RadixLevel(h, rp, {StoreSeg(h, rp, 0)}, tiles, entries, PB, shift, 1, rows));
    OG_TRY(RadixPieceStarts(h, cbase, tpar, h.rstarts.as<unsigned int>(), PB, static_cast<unsigned int>(rows)));
    store = 1;
    pieces = static_cast<long long>(PA) * PB;
    return OMNI_GPU_OK;
}

// Aggregates every recorded batch by partitioning (agg_radix.h) and creates its groups.
int RadixFinish(AggHandle &h)
{
    if (h.rbatches.empty()) return OMNI_GPU_OK;
    hipStream_t s = h.stream;
    const int64_t total = h.radix_rows;
    OG_CHECK(total < 0xFFFFFFF0LL, OMNI_GPU_ERR_INVALID, "the radix strategy takes fewer than 2^32 rows");
    RadixPlan rp;
    OG_TRY(PlanRadixColumns(h, rp));
    OG_TRY(h.rmisc.Reserve(64, s));
    OG_TRY(SyncGroups(h));
    // the LDS table of a final piece; at least two key words so that it has slot states
    const LdsLayout L = MakeLdsLayout(h.args, kRadixLdsBudget, std::max(2, h.rkd.W));
    OG_CHECK(L.S >= 64, OMNI_GPU_ERR_INVALID, "the aggregates leave too few LDS slots for the radix strategy");
    std::vector<int64_t> segRows(h.rbatches.size());
    for (size_t b = 0; b < h.rbatches.size(); ++b) segRows[b] = h.rbatches[b].rows;
    int shift = 64, store = 0, subBits = 0;
    long long pieces = 0;
    OG_TRY(RadixPartition(h, rp, L, rp.segs, segRows, total, shift, store, pieces, subBits));
    // ids: a round takes at most one per row; room for an eighth more rows of overflow rounds before the arrays
    // must grow
    int64_t idsEnd = h.known_ids;
    OG_TRY(GrowGroupsRaw(h, idsEnd + total + total / 8 + 1024, idsEnd));
    int64_t roundRows = total;
    OG_TRY(h.rovf.Reserve(sizeof(unsigned int) * std::max<int64_t>(1, total), s));
    unsigned int *ovfCount = reinterpret_cast<unsigned int *>(h.rmisc.as<uint8_t>() + 8);
    while (true) {
        // the final pieces of `store`
        RadixFinal f{};
        f.kd = h.rkd;
        AggArgs args = h.args;
        for (int c = 0; c < h.rkd.num; ++c) {
            f.kd.values[c] = h.rvals[store][c].p;
            f.kd.nulls[c] = rp.nullable[c] ? h.rnulls[store][c].as<uint8_t>() : nullptr;
        }
        for (int i = 0; i < args.n; ++i) {
            if (rp.aggCol[i] < 0) continue;
            args.a[i].values = h.rvals[store][rp.aggCol[i]].p;
            args.a[i].nulls = rp.nullable[rp.aggCol[i]] ? h.rnulls[store][rp.aggCol[i]].as<uint8_t>() : nullptr;
        }
        f.fstart = h.rfstart.as<unsigned int>();
        f.L = L;
        f.ngroups = h.ngroups();
        f.max_groups = static_cast<unsigned long long>(h.max_groups);
        f.overflow = h.overflow();
        f.gkeys = h.gkeys.as<unsigned long long>();
        f.gnull = h.single ? h.gnull.as<uint8_t>() : nullptr;
        f.gvalid = h.gvalid.as<uint8_t>();
        f.special = h.special();
        f.single = h.single ? 1 : 0;
        f.single_hi = h.rkd.in_word0[0];
        f.ovf_rows = h.rovf.as<unsigned int>();
        f.ovf_count = ovfCount;
        f.sub_bits = subBits;
        f.sub_shift = shift - subBits;
        OG_HIP(hipMemsetAsync(ovfCount, 0, 4, s));
        if (idsEnd + roundRows + 2 > h.max_groups) {
            OG_TRY(GrowGroupsRaw(h, idsEnd + 2 * roundRows + 1024, idsEnd));
            f.max_groups = static_cast<unsigned long long>(h.max_groups);
            f.gkeys = h.gkeys.as<unsigned long long>();
            f.gnull = h.single ? h.gnull.as<uint8_t>() : nullptr;
            f.gvalid = h.gvalid.as<uint8_t>();
            args.states = h.args.states;
        }
        idsEnd += roundRows;
        {
            KernelScope ks(s, "agg_radix_final");
            agg_radix_final<<<static_cast<unsigned int>(pieces), kRadixFinalThreads, RadixFinalLds(L, 1 << subBits), s>>>(
                f, args);
            OG_TRY(CheckLaunch("agg_radix_final"));
        }
        unsigned int ovf = 0;
        OG_TRY(CopyToHost(s, {{&ovf, ovfCount, 4}}));
        if (ovf == 0) break;
        // rows whose key found no slot (every row of such a key): gathered, then partitioned again on lower bits
        OG_TRY(ReserveStore(h, rp, 2, ovf));
        {
            std::vector<RadixCol> cols(rp.nc);
            for (int c = 0; c < rp.nc; ++c) {
                cols[c].width = rp.width[c];
                cols[c].values = h.rvals[2][c].p;
                cols[c].nulls = rp.nullable[c] ? h.rnulls[2][c].as<unsigned int>() : nullptr;
            }
            const RadixSeg src = StoreSeg(h, rp, store);
            OG_TRY(h.rsegs.Reserve(sizeof(RadixSeg), s));
            OG_TRY(CopyToDevice(h.rsegs.p, &src, sizeof(RadixSeg), s));
            OG_TRY(h.rcols.Reserve(sizeof(RadixCol) * cols.size(), s));
            OG_TRY(CopyToDevice(h.rcols.p, cols.data(), sizeof(RadixCol) * cols.size(), s));
            KernelScope ks(s, "agg_radix_gather");
            agg_radix_gather<<<(ovf + kBlock - 1) / kBlock, kBlock, 0, s>>>(h.rsegs.as<RadixSeg>(), h.rcols.as<RadixCol>(),
                                                                            rp.nc, h.rovf.as<unsigned int>(), ovf);
            OG_TRY(CheckLaunch("agg_radix_gather"));
        }
        shift -= subBits;  // the passes' bits are spent too
        OG_TRY(RadixPartition(h, rp, L, {StoreSeg(h, rp, 2)}, {static_cast<int64_t>(ovf)}, ovf, shift, store, pieces,
                              subBits));
        roundRows = ovf;
    }
    h.rbatches.clear();
    h.radix_rows = 0;
    h.radix_done = true;
    OG_TRY(SyncGroups(h));
    return OMNI_GPU_OK;
}

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
    OG_CHECK(spec->strategy >= OMNI_GPU_STRATEGY_AUTO && spec->strategy <= OMNI_GPU_STRATEGY_RADIX,
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
    // the radix strategy's key encoding: the multi-word layout, also for a single-word key (NULL bit in word 0)
    {
        KeyDev &r = h->rkd;
        r = kd;
        int w = 1;
        bool word0Free = true;
        for (int c = 0; c < r.num; ++c) {
            r.in_word0[c] = 0;
            if (r.width[c] <= 4 && word0Free) {
                r.in_word0[c] = 1;
                r.word[c] = 0;
                word0Free = false;
            } else {
                r.word[c] = w;
                w += r.width[c] == 16 ? 2 : 1;
            }
        }
        r.W = w;
    }
    h->args.n = spec->num_aggs;
    for (int i = 0; i < spec->num_aggs; ++i) {
        OG_TRY(ClassifyAgg(*spec, i, h->args.a[i]));
        h->checked |= h->args.a[i].kind == K_SUM_DEC_CHK;
    }
    AssignRowLayout(h->args);
    // The radix strategy reorders rows, so it serves every aggregate whose result does not depend on the row order
    // (the checked decimal SUM and double MIN/MAX do); AUTO takes it for a capacity hint of 2^21 groups or more.
    bool radixOk = true;
    for (int i = 0; i < spec->num_aggs; ++i) {
        const int k = h->args.a[i].kind;
        radixOk = radixOk && k != K_SUM_DEC_CHK && k != K_MIN_F64 && k != K_MAX_F64;
    }
    OG_CHECK(spec->strategy != OMNI_GPU_STRATEGY_RADIX || radixOk, OMNI_GPU_ERR_INVALID,
             "the radix strategy does not keep the row order that a checked decimal SUM or a double MIN / MAX needs");
    h->radix = spec->strategy == OMNI_GPU_STRATEGY_RADIX ||
               (spec->strategy == OMNI_GPU_STRATEGY_AUTO && radixOk && spec->capacity_hint >= (1LL << 21));
    OG_TRY(AcquireStream(&h->stream));
    // 64 bytes of counters, then the per-workgroup key min / max pairs the ARRAY rule plans its map from
    OG_TRY(h->counters.Reserve(64 + 16 * kMinMaxMaxBlocks, h->stream));
    OG_HIP(hipMemsetAsync(h->counters.p, 0, 80, h->stream));
    OG_TRY(h->cursors.Reserve(sizeof(unsigned long long) * 2 * kCursorSlots, h->stream));
    OG_HIP(hipMemsetAsync(h->cursors.p, 0, sizeof(unsigned long long) * 2 * kCursorSlots, h->stream));
    std::vector<unsigned long long> tmpl = RowTemplate(h->args);
    OG_TRY(h->row_tmpl.Reserve(sizeof(unsigned long long) * tmpl.size(), h->stream));
    OG_TRY(CopyToDevice(h->row_tmpl.p, tmpl.data(), sizeof(unsigned long long) * tmpl.size(), h->stream));
    uint64_t want = spec->capacity_hint > 0 && !h->radix ? static_cast<uint64_t>(spec->capacity_hint) : 16384;
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
    if (h->checked || h->radix) h->use_lds = false;
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
    if (h.radix && !h.radix_done) {  // recorded; aggregated when the groups are first asked for
        RadixBatch rb{};
        for (int c = 0; c < h.kd.num; ++c) rb.keys[c] = keys[c];
        for (int i = 0; i < h.args.n; ++i) {
            if (h.args.a[i].kind != K_COUNT_ALL) rb.aggs[i] = aggs[i];
        }
        rb.rows = rows;
        h.rbatches.push_back(rb);
        h.radix_rows += rows;
        return OMNI_GPU_OK;
    }
    // Capacity: ids after this batch <= ids now + one per new row + one chunk per wave slot in reserve.
    auto need = [&]() {
        return h.known_ids + h.pending_rows + rows + static_cast<int64_t>(h.Chunk()) * kCursorSlots + 2;
    };
    if (h.radix) {
        // batches after a radix aggregation go to the hash table, which first takes every existing group
        OG_TRY(SyncGroups(h));
        uint64_t cap = h.cap;
        while (static_cast<int64_t>(static_cast<double>(cap) * kMaxLoad) < need()) cap *= 4;
        OG_TRY(Resize(h, cap, h.known_ids));
        h.radix = false;
    }
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
    if (h.radix && !h.radix_done) {  // each batch is recorded as it is: nothing to stage
        for (int b = 0; b < num_batches; ++b) {
            OG_TRY(omni_gpu_agg_add_batch(handle, keys + static_cast<size_t>(b) * nk,
                                          na ? aggs + static_cast<size_t>(b) * na : nullptr, rows[b]));
        }
        return OMNI_GPU_OK;
    }
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
    AggHandle &h = *static_cast<AggHandle *>(handle);
    if (!h.rbatches.empty()) OG_TRY(RadixFinish(h));
    OG_TRY(SyncStream(h.stream));
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
    for (int k = 0; k < kRadixStores; ++k) {
        for (int c = 0; c < kRadixCols; ++c) {
            h->rvals[k][c].Release(s);
            h->rnulls[k][c].Release(s);
        }
    }
    for (DevBuf *b : {&h->rsegs, &h->rtiles, &h->rcols, &h->rcounts, &h->rsums, &h->rstarts, &h->rfstart,
                      &h->rparents, &h->rovf, &h->rmisc, &h->rhll}) {
        b->Release(s);
    }
    ReleaseStream(s);
    delete h;
    return OMNI_GPU_OK;
}

}  // extern "C"
