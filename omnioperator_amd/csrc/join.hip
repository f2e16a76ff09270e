// Inner hash join (single equality key) on gfx950.
//
// Build replaces HashBuilderOperator's end-of-input table build (core/src/operator/join/hash_builder.cpp:182-213,
// join_hash_table_variants.cpp:254-283, 695-808): the same choice between a direct-index array table (dense
// integer keys: range <= 8 * 2^ceil(log2(n / 0.75)) and <= UINT32_MAX) and a hash table, NULL build keys skipped.
// Probe replaces LookupJoinOperator::ProbeBatchForInnerJoin + the build-column gather
// (core/src/operator/join/lookup_join.cpp:924-1035, 1657-1779).
//
// Device layout:
//   build store   keys[n] (int64, sign-extended), key_null[n] (u8), payload values / NULL bytes per column,
//                 gathered from every retained build batch in one launch at finish
//   hash table    slots[cap] {i64 key; i32 head; u32 inline payload} (16 B), linear probing at load <= 0.25
//                 from a multiplicative home slot; duplicate keys chain through next[n].  Tables of 8 to
//                 4,096 regions of 8,192 slots are built region by region in LDS (partitioned build).
//   array table   heads[range] (i32 build row, -1 none) + next[n]
// Probe: 16 probe rows per lane held in registers; matches are compacted with wave ballots (unique keys) or a
// wave-level exclusive scan (duplicate keys) and one atomic per workgroup on the output counter, so each
// workgroup writes its output rows contiguously.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <vector>

#include "join_device.h"
#include "join_keys.h"

namespace omni_gpu {

// ---- host side ----------------------------------------------------------------------------------------------

struct JoinHandle {
    omni_gpu_join_spec spec{};
    hipStream_t stream = nullptr;
    int64_t n = 0;             // build rows
    std::vector<BuildBatch> pending;  // retained batches, gathered at finish
    DevBuf keys, knull;
    DevBuf mm_parts;           // per-workgroup key min / max pairs of the build gather
    bool key_nulls = false;    // some build batch has a NULL bitmap on its key column
    DevBuf pay[OMNI_GPU_MAX_AGGS], pay_null[OMNI_GPU_MAX_AGGS];
    bool pay_nullable[OMNI_GPU_MAX_AGGS] = {};
    bool finished = false;
    bool is_array = false;
    bool inline_pay = false;
    bool has_dup = false;
    int64_t max_dup = 0;
    // The duplicate-key flag of a finished build is read back without a stream synchronisation: copied into
    // flag_cell (pinned) before flags_ev; has_dup / inline_pay / max_dup are valid once flags_pending is false.
    bool flags_pending = false;
    unsigned int *flag_cell = nullptr;
    hipEvent_t flags_ev = nullptr;
    int probe_group = 1;       // probe width G of join_probe
    int partitions = -1;       // radix-partitioned probe: -1 auto, 0 off, else the partition count
    int last_partitions = 0;   // partition count the last probe used (1 = direct probe)
    int mode = kModeInner;     // probe-side emission rule of the join type (omni_gpu_join_set_type)
    bool preserve_build = false;  // build-preserving outer join: probes mark matched build rows in visited
    bool existence = false;    // existence join: probed with omni_gpu_join_probe_exists only
    DevBuf visited;            // one byte per build row (preserve_build)
    DevBuf part_keys, part_rows, part_counts, part_meta;
    DevBuf ovf_keys, ovf_rows;  // partitioned build: keys whose probe run leaves their region
    uint64_t slots_n = 0;
    long long min_key = 0;
    unsigned long long range = 0;
    // scratch: [0] sentinel head i32, [4] has_dup u32, [8] max_len u32, [12] sentinel inline payload u32,
    //          [16] min/max i64[2], [32] match counter u64, [40] build overflow count u32
    DevBuf slots, heads, next, scratch;
    DevBuf batch_desc, build_desc;
    // keys other than one fixed-width column (join_keys.h): their translation into int64 join keys
    JoinKeyCols kc{};
    int32_t jkey_type = 0;             // type of the join key the kernels see (LONG for translated keys)
    std::vector<KeyBatch> pending_keys;  // the key columns of each retained build batch
    DevBuf dict;                       // dictionary of the build side's wide keys (KeyDict records)
    uint64_t dict_slots = 0;           // its slots (a power of two)
    DevBuf bkeys, bnulls, pkeys, pnulls, kdesc;  // translated build / probe keys and their NULL bitmaps

    // Every device buffer of the handle (destroy releases exactly these; a new member is added here).
    template <typename F> void ForEachBuffer(F f)
    {
        for (DevBuf *b : {&keys, &knull, &mm_parts, &visited, &part_keys, &part_rows, &part_counts, &part_meta,
                          &ovf_keys, &ovf_rows, &slots, &heads, &next, &scratch, &batch_desc, &build_desc,
                          &dict, &bkeys, &bnulls, &pkeys, &pnulls, &kdesc}) {
            f(*b);
        }
        for (int p = 0; p < OMNI_GPU_MAX_AGGS; ++p) {
            f(pay[p]);
            f(pay_null[p]);
        }
    }
};

namespace {
// Host-visible words for the duplicate-key flag: pinned pages cut into 64-byte cells, shared by every join
// handle of the process (pages are kept until the process ends).
std::mutex g_cell_mu;
std::vector<unsigned int *> g_free_cells;

int TakeFlagCell(unsigned int **cell)
{
    std::lock_guard<std::mutex> g(g_cell_mu);
    if (g_free_cells.empty()) {
        void *page = nullptr;
        OG_HIP(hipHostMalloc(&page, 4096, hipHostMallocDefault));
        for (int i = 0; i < 4096 / 64; ++i) {
            g_free_cells.push_back(reinterpret_cast<unsigned int *>(static_cast<uint8_t *>(page) + 64 * i));
        }
    }
    *cell = g_free_cells.back();
    g_free_cells.pop_back();
    return OMNI_GPU_OK;
}

void ReturnFlagCell(unsigned int *cell)
{
    std::lock_guard<std::mutex> g(g_cell_mu);
    g_free_cells.push_back(cell);
}
}  // namespace

// Waits for the finished build's duplicate-key flag (only for the work queued before it, not the whole
// stream) and sets the facts that pick probe kernel variants.
int ResolveBuildFlags(JoinHandle &h)
{
    if (!h.flags_pending) return OMNI_GPU_OK;
    OG_HIP(hipEventSynchronize(h.flags_ev));
    h.flags_pending = false;
    h.has_dup = *h.flag_cell != 0;
    if (h.has_dup) h.inline_pay = false;
    h.max_dup = h.has_dup ? -1 : 1;  // longest chain: measured on first request
    return OMNI_GPU_OK;
}

// Partition count of the radix-partitioned probe, 1 for the direct probe.  Automatic mode partitions tables
// above kAutoPartBytes when the probe side has at least half as many rows as the table has slots (so the
// partitioning pays for itself), with regions of kWinSlots slots for the LDS window probe; tables too big for
// kPartMax such regions use the L2 probe with kPartMax partitions.
constexpr uint64_t kAutoPartBytes = 8ull << 20;

int ChoosePartitions(const JoinHandle &h, int64_t rows)
{
    if (h.is_array || h.partitions == 0 || h.slots_n == 0) return 1;
    if (rows >= (1LL << 32) - 1) return 1;  // u32 entry indices
    int P;
    if (h.partitions > 0) {
        P = h.partitions;
    } else {
        // -1 (automatic): only tables that pay for it; -2: always, with the automatic region size
        if (h.partitions == -1 &&
            (h.slots_n * sizeof(JoinSlot) < kAutoPartBytes || static_cast<uint64_t>(rows) * 2 < h.slots_n)) {
            return 1;
        }
        P = 8;
        while (P < kPartMax && h.slots_n / static_cast<uint64_t>(P) > static_cast<uint64_t>(kWinSlots)) P <<= 1;
    }
    if (static_cast<uint64_t>(P) * 64 > h.slots_n) return 1;  // regions of at least 64 slots
    return P;
}

// Histogram, scan, plan and scatter of `rows` rows (the batches of `a`, nwg tiles of kPartTile) into P
// partitions of the table; fills pa.  `tag` names the stage for the kernel timers.
int Partition(JoinHandle &h, const ProbeArgs &a, const TableDev &t, int P, int64_t nwg, int64_t rows, PartArgs &pa,
              const char *hist_tag, const char *scan_tag, const char *scatter_tag)
{
    hipStream_t s = h.stream;
    pa = PartArgs{};
    pa.nparts = P;
    int lgS = 0, lgP = 0;
    while ((1ULL << lgS) < h.slots_n) ++lgS;
    while ((1 << lgP) < P) ++lgP;
    pa.shift = lgS - lgP;
    pa.ntiles = nwg;
    pa.xcd_cols = (nwg + kXcds - 1) / kXcds;
    pa.ncol = pa.xcd_cols * kXcds;
    const int64_t L = static_cast<int64_t>(P) * pa.ncol;
    OG_CHECK(L < 0x7FFFFFFFLL * kScanTile, OMNI_GPU_ERR_INVALID, "partition count matrix too large");
    const int ntile = static_cast<int>((L + kScanTile - 1) / kScanTile);
    const int per = P / kXcds;
    OG_TRY(h.part_keys.Reserve(sizeof(long long) * rows, s));
    OG_TRY(h.part_rows.Reserve(sizeof(unsigned int) * rows, s));
    OG_TRY(h.part_counts.Reserve(sizeof(unsigned int) * L, s));
    const size_t metaWords = static_cast<size_t>(ntile) + 1 + (P + 1) + kXcds * (per + 1) + kXcds;
    OG_TRY(h.part_meta.Reserve(sizeof(unsigned int) * metaWords, s));
    unsigned int *meta = h.part_meta.as<unsigned int>();
    unsigned int *sums = meta;
    pa.total = meta + ntile;
    pa.part_start = pa.total + 1;
    pa.xcd_prefix = pa.part_start + P + 1;
    pa.xcd_next = pa.xcd_prefix + kXcds * (per + 1);
    pa.counts = h.part_counts.as<unsigned int>();
    pa.pkeys = h.part_keys.as<long long>();
    pa.prows = h.part_rows.as<unsigned int>();
    {
        KernelScope ks(s, hist_tag);
        join_part_hist<<<static_cast<unsigned int>(pa.ncol), kBlock, 0, s>>>(a, t, pa);
        OG_TRY(CheckLaunch("join_part_hist"));
    }
    {
        KernelScope ks(s, scan_tag);
        scan_tiles<<<ntile, kScanBlock, 0, s>>>(pa.counts, L, sums);
        scan_sums<<<1, kScanBlock, 0, s>>>(sums, ntile, pa.total);
        scan_add<<<ntile, kScanBlock, 0, s>>>(pa.counts, L, sums);
        join_part_plan<<<1, kBlock, 0, s>>>(pa);
        OG_TRY(CheckLaunch("join_part_scan"));
    }
    {
        KernelScope ks(s, scatter_tag);
        const size_t lds = ScatterLdsBytes(P);
        OG_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&join_part_scatter),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
        join_part_scatter<<<static_cast<unsigned int>(nwg), kScatterBlock, lds, s>>>(a, t, pa);
        OG_TRY(CheckLaunch("join_part_scatter"));
    }
    return OMNI_GPU_OK;
}

// Radix-partitioned probe; `ty` (join types other than inner) needs regions of at most kWinSlots slots.
int LaunchPartitioned(JoinHandle &h, const ProbeArgs &a, const TableDev &t, int P, int64_t nwg, int64_t rows,
                      const TypedArgs *ty = nullptr)
{
    hipStream_t s = h.stream;
    PartArgs pa;
    OG_TRY(Partition(h, a, t, P, nwg, rows, pa, "join_probe_hist", "join_probe_scan", "join_probe_scatter"));
    OG_TRY(ResolveBuildFlags(h));  // the probe kernel variant; the partitioning above did not need it
    int lgP = 0;
    while ((1 << lgP) < P) ++lgP;
    int cus = 0;
    OG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, CurrentDevice()));
    const uint64_t region = h.slots_n >> lgP;
    if (region <= static_cast<uint64_t>(kWinSlots)) {
        // window probe: about 4 workgroups per CU in all, each partition split evenly
        const int splits = std::max(1, std::min(64, (cus * 4 + P - 1) / P));
        const size_t lds = kWinLdsBytes;
        const unsigned int wgrid = static_cast<unsigned int>(P) * static_cast<unsigned int>(splits);
        KernelScope ks(s, "join_probe_window");
        // template arguments: inline payload, duplicate keys, join-type emission
        auto launch = [&](auto kernel) -> int {
            OG_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
            kernel<<<wgrid, kWinBlock, lds, s>>>(a, t, pa, splits, ty ? *ty : TypedArgs{});
            return OMNI_GPU_OK;
        };
        if (ty && h.has_dup) OG_TRY(launch(&join_window_probe<false, true, true>));
        else if (ty && h.inline_pay) OG_TRY(launch(&join_window_probe<true, false, true>));
        else if (ty) OG_TRY(launch(&join_window_probe<false, false, true>));
        else if (h.has_dup) OG_TRY(launch(&join_window_probe<false, true, false>));
        else if (h.inline_pay) OG_TRY(launch(&join_window_probe<true, false, false>));
        else OG_TRY(launch(&join_window_probe<false, false, false>));
        return CheckLaunch("join_window_probe");
    }
    OG_CHECK(ty == nullptr, OMNI_GPU_ERR_INVALID, "join types other than inner need LDS window regions");
    const unsigned int pgrid = static_cast<unsigned int>(std::max(kXcds, (cus * 4 / kXcds) * kXcds));
    KernelScope ks(s, "join_probe_part");
    if (h.has_dup) join_part_probe<false, true><<<pgrid, kBlock, 0, s>>>(a, t, pa);
    else if (h.inline_pay) join_part_probe<true, false><<<pgrid, kBlock, 0, s>>>(a, t, pa);
    else join_part_probe<false, false><<<pgrid, kBlock, 0, s>>>(a, t, pa);
    return CheckLaunch("join_part_probe");
}

// Partitioned build of the hash table (P regions of kWinSlots slots): the build store's rows are partitioned
// like probe rows, each region is built in LDS by one workgroup and written out whole, then the keys whose
// probe run left their region are inserted into the global table.
// Partitions the build store's rows into the P regions of a table of h.slots_n slots (only the table's size is
// needed, not its memory).
int PartitionBuildRows(JoinHandle &h, int P, PartArgs &pa)
{
    ProbeArgs a{};
    a.nb = 1;
    a.single.keys = h.keys.p;
    a.single.nulls = nullptr;  // NULL build keys are partitioned and skipped by the region build
    a.single.rows = h.n;
    a.single.width = 8;
    TableDev t{};
    t.mask = h.slots_n - 1;
    const int64_t nwg = (h.n + kPartTile - 1) / kPartTile;
    return Partition(h, a, t, P, nwg, h.n, pa, "join_build_hist", "join_build_scan", "join_build_scatter");
}

// `pre`: the rows were already partitioned by PartitionBuildRows (nullptr: partition them here).
int BuildPartitioned(JoinHandle &h, int P, unsigned int *has_dup, const PartArgs *pre)
{
    hipStream_t s = h.stream;
    uint8_t *sc = h.scratch.as<uint8_t>();
    TableDev t{};
    t.slots = h.slots.as<JoinSlot>();
    t.mask = h.slots_n - 1;
    PartArgs pa;
    if (pre) pa = *pre;
    else OG_TRY(PartitionBuildRows(h, P, pa));
    OG_TRY(h.ovf_keys.Reserve(sizeof(long long) * h.n, s));
    OG_TRY(h.ovf_rows.Reserve(sizeof(unsigned int) * h.n, s));
    BuildArgs b{};
    b.knull = h.key_nulls ? h.knull.as<uint8_t>() : nullptr;  // no NULL bitmap in any batch: nothing to skip
    b.next = h.next.as<int>();
    b.sentinel_head = reinterpret_cast<int *>(sc);
    b.sentinel_pay = reinterpret_cast<unsigned int *>(sc + 12);
    b.has_dup = has_dup;
    b.inline_pay = h.inline_pay ? h.pay[0].p : nullptr;
    b.inline_width = h.inline_pay ? TypeWidth(h.spec.payload_types[0]) : 0;
    b.ovf_keys = h.ovf_keys.as<long long>();
    b.ovf_rows = h.ovf_rows.as<unsigned int>();
    b.ovf_count = reinterpret_cast<unsigned int *>(sc + 40);
    const size_t lds = sizeof(JoinSlot) * (static_cast<size_t>(1) << pa.shift);
    KernelScope ks(s, "join_build_window");
    OG_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&join_window_build),
                               hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    join_window_build<<<static_cast<unsigned int>(P), kWinBlock, lds, s>>>(t, pa, b, h.slots.as<JoinSlot>());
    OG_TRY(CheckLaunch("join_window_build"));
    join_build_overflow<<<256, kBlock, 0, s>>>(t, b, h.slots.as<JoinSlot>());
    return CheckLaunch("join_build_overflow");
}

template <int G>
void LaunchProbe(const JoinHandle &h, dim3 grid, hipStream_t s, const ProbeArgs &a, const TableDev &t)
{
    if (h.is_array) {
        if (h.has_dup) join_probe<true, false, true, G><<<grid, kBlock, 0, s>>>(a, t);
        else join_probe<true, false, false, G><<<grid, kBlock, 0, s>>>(a, t);
    } else if (h.has_dup) {
        join_probe<false, false, true, G><<<grid, kBlock, 0, s>>>(a, t);
    } else if (h.inline_pay) {
        join_probe<false, true, false, G><<<grid, kBlock, 0, s>>>(a, t);
    } else {
        join_probe<false, false, false, G><<<grid, kBlock, 0, s>>>(a, t);
    }
}

// Payload output descriptors for `cap` output rows; clears the output NULL bitmaps.
int SetupPayloadOutputs(JoinHandle &h, omni_gpu_out_column *outs, int64_t cap, PayloadDev *pay)
{
    hipStream_t s = h.stream;
    for (int p = 0; p < h.spec.num_payloads; ++p) {
        OG_CHECK(outs && outs[p].type_id == h.spec.payload_types[p], OMNI_GPU_ERR_INVALID, "payload output %d type mismatch", p);
        OG_CHECK(!h.pay_nullable[p] || cap == 0 || !outs[p].values || outs[p].nulls, OMNI_GPU_ERR_INVALID,
                 "payload %d has NULLs: its output needs a null bitmap", p);
        pay[p].width = TypeWidth(h.spec.payload_types[p]);
        pay[p].values = h.pay[p].p;
        pay[p].nulls = h.pay_nullable[p] ? h.pay_null[p].as<uint8_t>() : nullptr;
        pay[p].out = outs[p].values;
        pay[p].out_nulls = outs[p].nulls;
        if (outs[p].values && outs[p].nulls && cap > 0) {
            int64_t words = (cap + 31) / 32;
            KernelScope ks(s, "join_clear_nulls");
            join_fill_i32<<<Blocks(words), kBlock, 0, s>>>(reinterpret_cast<int *>(outs[p].nulls), 0, words);
            OG_TRY(CheckLaunch("join_fill_i32"));
        }
    }
    return OMNI_GPU_OK;
}

TableDev MakeTable(const JoinHandle &h)
{
    TableDev t{};
    t.slots = h.slots.as<JoinSlot>();
    t.mask = h.slots_n ? h.slots_n - 1 : 0;
    t.heads = h.heads.as<int>();
    t.min_key = h.min_key;
    t.range = h.range;
    t.next = h.next.as<int>();
    t.sentinel_head = reinterpret_cast<const int *>(h.scratch.as<uint8_t>());
    t.sentinel_pay = reinterpret_cast<const unsigned int *>(h.scratch.as<uint8_t>() + 12);
    return t;
}

int RunProbe(JoinHandle &h, const ProbeBatch *batches, int nb, int64_t total_blocks, int64_t part_blocks,
             long long *out_rows,
             omni_gpu_out_column *outs, int64_t cap, int64_t *num_matches)
{
    OG_CHECK(h.finished, OMNI_GPU_ERR_STATE, "probe before the build side finished");
    OG_CHECK(num_matches, OMNI_GPU_ERR_INVALID, "null num_matches");
    OG_CHECK(cap == 0 || out_rows, OMNI_GPU_ERR_INVALID, "null out_probe_rows");
    hipStream_t s = h.stream;
    ProbeArgs a{};
    a.nb = nb;
    a.npay = h.spec.num_payloads;
    a.out_rows = out_rows;
    a.cap = cap;
    a.counter = reinterpret_cast<unsigned long long *>(h.scratch.as<uint8_t>() + 32);
    OG_CHECK(!h.existence, OMNI_GPU_ERR_STATE, "an existence join is probed with omni_gpu_join_probe_exists");
    if (h.mode == kModeLeft) {  // unmatched probe rows get NULL payloads
        for (int p = 0; p < a.npay; ++p) {
            OG_CHECK(cap == 0 || !outs || !outs[p].values || outs[p].nulls, OMNI_GPU_ERR_INVALID,
                     "outer join: payload output %d needs a null bitmap", p);
        }
    }
    if (h.mode == kModeSemi || h.mode == kModeAnti) a.npay = 0;  // the probe rows alone are output
    else OG_TRY(SetupPayloadOutputs(h, outs, cap, a.pay));
    if (nb == 1) {
        a.single = batches[0];
        a.batches = nullptr;
    } else {
        OG_TRY(h.batch_desc.Reserve(sizeof(ProbeBatch) * nb, s));
        OG_TRY(CopyToDevice(h.batch_desc.p, batches, sizeof(ProbeBatch) * nb, s));
        a.batches = h.batch_desc.as<ProbeBatch>();
    }
    OG_HIP(hipMemsetAsync(a.counter, 0, 8, s));
    const TableDev t = MakeTable(h);
    int64_t total_rows = 0;
    for (int i = 0; i < nb; ++i) total_rows += batches[i].rows;
    const bool typed = h.mode != kModeInner || h.preserve_build;
    int P = (total_blocks > 0 && h.n > 0) ? ChoosePartitions(h, total_rows) : 1;
    if (typed && P > 1) {  // the join types take the window probe only (regions of at most kWinSlots slots)
        int lgP = 0;
        while ((1 << lgP) < P) ++lgP;
        if ((h.slots_n >> lgP) > static_cast<uint64_t>(kWinSlots)) P = 1;
    }
    h.last_partitions = P;
    if (P <= 1) OG_TRY(ResolveBuildFlags(h));  // LaunchPartitioned resolves after queueing the partitioning
    if (typed && P > 1) {
        TypedArgs ty{h.mode, 0, h.preserve_build ? h.visited.as<uint8_t>() : nullptr};
        OG_TRY(LaunchPartitioned(h, a, t, P, part_blocks, total_rows, &ty));
        bool nullable = false;
        for (int i = 0; i < nb; ++i) nullable = nullable || batches[i].nulls != nullptr;
        if (nullable && (h.mode == kModeLeft || h.mode == kModeAnti)) {
            KernelScope ks(s, "join_probe_null_keys");
            join_emit_null_keys<<<static_cast<unsigned int>(total_blocks), kBlock, 0, s>>>(a);
            OG_TRY(CheckLaunch("join_emit_null_keys"));
        }
    } else if (typed) {
        if (total_blocks > 0) {  // also with no build rows: outer and anti joins still emit every probe row
            TypedArgs ty{h.mode, h.n == 0 ? 1 : 0, h.preserve_build ? h.visited.as<uint8_t>() : nullptr};
            dim3 grid(static_cast<unsigned int>(total_blocks));
            KernelScope ks(s, "join_probe_typed");
            // template arguments: array table, duplicate keys, inline payload
            if (h.n == 0) join_probe_typed<false, false, false><<<grid, kBlock, 0, s>>>(a, t, ty);
            else if (h.is_array && h.has_dup) join_probe_typed<true, true, false><<<grid, kBlock, 0, s>>>(a, t, ty);
            else if (h.is_array) join_probe_typed<true, false, false><<<grid, kBlock, 0, s>>>(a, t, ty);
            else if (h.has_dup) join_probe_typed<false, true, false><<<grid, kBlock, 0, s>>>(a, t, ty);
            else if (h.inline_pay) join_probe_typed<false, false, true><<<grid, kBlock, 0, s>>>(a, t, ty);
            else join_probe_typed<false, false, false><<<grid, kBlock, 0, s>>>(a, t, ty);
            OG_TRY(CheckLaunch("join_probe_typed"));
        }
    } else if (P > 1) {
        OG_TRY(LaunchPartitioned(h, a, t, P, part_blocks, total_rows));
    } else if (total_blocks > 0 && h.n > 0) {
        KernelScope ks(s, "join_probe");
        dim3 grid(static_cast<unsigned int>(total_blocks));
        switch (h.probe_group) {
            case 1: LaunchProbe<1>(h, grid, s, a, t); break;
            case 2: LaunchProbe<2>(h, grid, s, a, t); break;
            case 4: LaunchProbe<4>(h, grid, s, a, t); break;
            case 16: LaunchProbe<16>(h, grid, s, a, t); break;
            default: LaunchProbe<8>(h, grid, s, a, t); break;
        }
        OG_TRY(CheckLaunch("join_probe"));
    }
    unsigned long long m = 0;
    OG_TRY(CopyToHost(s, {{&m, a.counter, 8}}));
    *num_matches = static_cast<int64_t>(m);
    if (static_cast<int64_t>(m) > cap) {
        return SetError(OMNI_GPU_ERR_OUTPUT_FULL, "join produced %llu matches for an output capacity of %lld", m,
                        static_cast<long long>(cap));
    }
    return OMNI_GPU_OK;
}

int FillBatch(ProbeBatch &b, const omni_gpu_column &c, int64_t rows, int64_t row_base, int64_t block_start,
              int64_t pblock_start, int32_t key_type)
{
    OG_CHECK(c.type_id == key_type, OMNI_GPU_ERR_INVALID, "probe key type %d != build key type %d", c.type_id, key_type);
    OG_CHECK(rows == 0 || c.values, OMNI_GPU_ERR_INVALID, "probe key has no values");
    OG_CHECK(rows >= 0, OMNI_GPU_ERR_INVALID, "negative row count");
    b.keys = c.values;
    b.nulls = c.nulls;
    b.rows = rows;
    b.row_base = row_base;
    b.block_start = block_start;
    b.pblock_start = pblock_start;
    b.width = TypeWidth(c.type_id);
    return OMNI_GPU_OK;
}

// The reference's key variant for the spec's key columns (HashBuilderOperatorFactory::InitVariant,
// hash_builder.cpp:44-119) as a translation mode of join_keys.h.
int PlanJoinKeys(const omni_gpu_join_spec &sp, JoinKeyCols &kc)
{
    const int n = sp.num_keys <= 1 ? 1 : sp.num_keys;
    OG_CHECK(n <= OMNI_GPU_MAX_KEYS, OMNI_GPU_ERR_INVALID, "num_keys must be at most %d", OMNI_GPU_MAX_KEYS);
    kc = JoinKeyCols{};
    kc.n = n;
    for (int c = 0; c < n; ++c) {
        kc.type[c] = n == 1 ? sp.key_type : sp.key_types[c];
        kc.width[c] = TypeWidth(kc.type[c]);
        OG_CHECK(kc.width[c] > 0, OMNI_GPU_ERR_INVALID, "join key type %d is not a fixed-width type", kc.type[c]);
    }
    if (n == 1) {
        const int t = kc.type[0];
        if (IsIntegerKeyType(t) || t == OMNI_TYPE_DOUBLE || t == OMNI_TYPE_BOOLEAN || t == OMNI_TYPE_BYTE) {
            kc.mode = kKeySingle;
        } else {
            OG_CHECK(t == OMNI_TYPE_DECIMAL128, OMNI_GPU_ERR_INVALID, "join key type %d is not supported", t);
            kc.mode = kKeyDecimal128;
            kc.W = 2;
        }
        return OMNI_GPU_OK;
    }
    // GetTypeLength: BYTE 8, SHORT 16, INT 32, LONG and DECIMAL64 64 bits; any other type serialises the key
    int bits = 0;
    bool packable = true;
    for (int c = 0; c < n; ++c) {
        switch (kc.type[c]) {
            case OMNI_TYPE_BYTE: kc.bits[c] = 8; break;
            case OMNI_TYPE_SHORT: kc.bits[c] = 16; break;
            case OMNI_TYPE_INT: kc.bits[c] = 32; break;
            case OMNI_TYPE_LONG:
            case OMNI_TYPE_DECIMAL64: kc.bits[c] = 64; break;
            default: packable = false; break;
        }
        bits += kc.bits[c];
    }
    if (packable && bits <= 32) {
        kc.mode = kKeyPacked32;
    } else if (packable && bits <= 64) {
        kc.mode = kKeyPacked64;
    } else if (packable && bits <= 128) {
        kc.mode = kKeyPacked128;
        kc.W = 2;
    } else {
        kc.mode = kKeySerialized;
        kc.W = 0;
        for (int c = 0; c < n; ++c) kc.W += kc.width[c] == 16 ? 2 : 1;
    }
    return OMNI_GPU_OK;
}

// Translates nb batches of key columns (cols[b * kc.n + c]) into int64 join keys: out_keys gets the keys of all
// batches in order, out_nulls a NULL bitmap per batch (starting at its row 0); `cols` returns each batch's
// translated key column.  INSERT adds the keys to the dictionary (build side).
template <bool INSERT>
int TranslateKeys(JoinHandle &h, const omni_gpu_column *in, const int64_t *rows, int nb, DevBuf &out_keys,
                  DevBuf &out_nulls, std::vector<omni_gpu_column> &cols)
{
    hipStream_t s = h.stream;
    std::vector<KeyBatch> kb(nb);
    int64_t total = 0, words = 0, blocks = 0;
    for (int b = 0; b < nb; ++b) {
        OG_CHECK(rows[b] >= 0, OMNI_GPU_ERR_INVALID, "negative row count");
        for (int c = 0; c < h.kc.n; ++c) {
            const omni_gpu_column &col = in[static_cast<size_t>(b) * h.kc.n + c];
            OG_CHECK(col.type_id == h.kc.type[c], OMNI_GPU_ERR_INVALID, "key column %d type %d != %d", c, col.type_id,
                     h.kc.type[c]);
            OG_CHECK(rows[b] == 0 || col.values, OMNI_GPU_ERR_INVALID, "key column %d has no values", c);
            kb[b].values[c] = col.values;
            kb[b].nulls[c] = col.nulls;
        }
        kb[b].rows = rows[b];
        kb[b].out_base = total;
        kb[b].null_word_base = words;
        kb[b].block_start = blocks;
        total += rows[b];
        words += (rows[b] + 31) / 32;
        blocks += (rows[b] + kKeyTile - 1) / kKeyTile;
    }
    OG_CHECK(blocks < 0x7FFFFFFF, OMNI_GPU_ERR_INVALID, "too many key rows for one launch");
    OG_TRY(out_keys.Reserve(sizeof(long long) * std::max<int64_t>(1, total), s));
    OG_TRY(out_nulls.Reserve(sizeof(unsigned int) * std::max<int64_t>(1, words), s));
    // descriptors, then the compact array of their first workgroups that the kernel searches
    std::vector<uint8_t> desc(sizeof(KeyBatch) * nb + sizeof(int64_t) * nb);
    std::memcpy(desc.data(), kb.data(), sizeof(KeyBatch) * nb);
    for (int b = 0; b < nb; ++b) {
        std::memcpy(desc.data() + sizeof(KeyBatch) * nb + sizeof(int64_t) * b, &kb[b].block_start, sizeof(int64_t));
    }
    OG_TRY(h.kdesc.Reserve(desc.size(), s));
    OG_TRY(CopyToDevice(h.kdesc.p, desc.data(), desc.size(), s));
    KeyDict d{};
    if (h.kc.mode != kKeyPacked32 && h.kc.mode != kKeyPacked64) {
        d.recs = h.dict.as<unsigned long long>();
        OG_CHECK(h.dict_slots > 0, OMNI_GPU_ERR_STATE, "probe before the build side finished");
        d.mask = h.dict_slots - 1;
        d.W = h.kc.W;
        d.S = DictStride(h.kc.W);
    }
    if (blocks > 0) {
        KernelScope ks(s, INSERT ? "join_translate_build" : "join_translate_probe");
        const int64_t *starts = reinterpret_cast<const int64_t *>(h.kdesc.as<uint8_t>() + sizeof(KeyBatch) * nb);
        // equal batches (the last may be shorter) let the kernel find its batch without a search
        int64_t uniform = nb > 0 && rows[0] > 0 ? (rows[0] + kKeyTile - 1) / kKeyTile : 0;
        for (int b = 0; b < nb && uniform > 0; ++b) {
            if (b + 1 < nb ? rows[b] != rows[0] : rows[b] > rows[0] || rows[b] == 0) uniform = 0;
        }
        if (h.kc.W <= 2) {
            join_translate_keys<INSERT, 2><<<static_cast<unsigned int>(blocks), kKeyThreads, 0, s>>>(
                h.kc, h.kdesc.as<KeyBatch>(), starts, nb, uniform, d, out_keys.as<long long>(),
                out_nulls.as<unsigned int>());
        } else {
            join_translate_keys<INSERT, kJoinKeyMaxWords><<<static_cast<unsigned int>(blocks), kKeyThreads, 0, s>>>(
                h.kc, h.kdesc.as<KeyBatch>(), starts, nb, uniform, d, out_keys.as<long long>(),
                out_nulls.as<unsigned int>());
        }
        OG_TRY(CheckLaunch("join_translate_keys"));
    }
    cols.resize(nb);
    for (int b = 0; b < nb; ++b) {
        cols[b] = omni_gpu_column{};
        cols[b].type_id = OMNI_TYPE_LONG;
        cols[b].values = out_keys.as<long long>() + kb[b].out_base;
        cols[b].nulls = reinterpret_cast<const uint8_t *>(out_nulls.as<unsigned int>() + kb[b].null_word_base);
    }
    return OMNI_GPU_OK;
}

// Probe-side output columns of join output rows (LookupJoinOutputBuilder::ConstructProbeColumns,
// lookup_join.cpp:1853-1877), gathered on the device: column c of output row k is probe column c's row rows[k].
constexpr int kRowGatherCols = 8;  // columns per launch
struct RowGatherCol {
    const void *src;
    const uint8_t *src_nulls;  // nullptr: no NULLs
    void *dst;
    unsigned int *dst_nulls;   // 32-bit words of the output rows' bits, nullptr: not written
    int width;
};
struct RowGatherArgs {
    RowGatherCol c[kRowGatherCols];
    int ncols;
    const long long *rows;
    int64_t n;
};

__global__ void __launch_bounds__(256) join_gather_rows(RowGatherArgs g)
{
    const int64_t k = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;  // rows of a wave: 64 aligned ones
    const bool live = k < g.n;
    const int64_t r = live ? g.rows[k] : 0;
    for (int c = 0; c < g.ncols; ++c) {
        const RowGatherCol &gc = g.c[c];
        if (live) {
            switch (gc.width) {
                case 16: {
                    const auto *s2 = static_cast<const unsigned long long *>(gc.src) + 2 * r;
                    auto *d2 = static_cast<unsigned long long *>(gc.dst) + 2 * k;
                    d2[0] = s2[0];
                    d2[1] = s2[1];
                    break;
                }
                case 8: static_cast<unsigned long long *>(gc.dst)[k] = static_cast<const unsigned long long *>(gc.src)[r]; break;
                case 4: static_cast<unsigned int *>(gc.dst)[k] = static_cast<const unsigned int *>(gc.src)[r]; break;
                case 2: static_cast<unsigned short *>(gc.dst)[k] = static_cast<const unsigned short *>(gc.src)[r]; break;
                default: static_cast<uint8_t *>(gc.dst)[k] = static_cast<const uint8_t *>(gc.src)[r]; break;
            }
        }
        if (gc.dst_nulls) {
            const bool isNull = live && gc.src_nulls && ((gc.src_nulls[r >> 3] >> (r & 7)) & 1);
            const unsigned long long m = __ballot(isNull);
            const int lane = LaneId();
            const int64_t word = ((k - lane) >> 5) + (lane >> 5);  // lanes 0 and 32 write the wave's two words
            if ((lane & 31) == 0 && word * 32 < g.n) gc.dst_nulls[word] = static_cast<unsigned int>(m >> lane);
        }
    }
}

// Probe rows' hash-table partitions of a multi-partition join (operatorCount > 1): LookupJoinOperator's
// PopulateProbeHashes (lookup_join.cpp:1572-1612, CalculateColHashes et al.) -- per row h = 0, then per non-NULL
// key column h = 31 * h + HashValue(value) -- and HashUtil::GetRawHashPartition(h, mask) (hash_util.h:146-154).
// HashValue: int16 / int32 / int64 sign-extended, rotl(v * P2, 31) * P1 (hash_util.h:74-82); DOUBLE through
// DoubleToLongBits (NaNs as 0x7ff8000000000000, :84-87, 270-279); BOOLEAN 1231 / 1237 (:89-92); DECIMAL64 the
// value (:94-97); DECIMAL128 the xxhash of the low word XOR that of the high word without its sign bit (:99-103).
constexpr unsigned long long kXxP1 = 0x9E3779B185EBCA87ULL, kXxP2 = 0xC2B2AE3D27D4EB4FULL,
                             kXxP3 = 0x165667B19E3779F9ULL, kXxP4 = 0x85EBCA77C2B2AE63ULL,
                             kXxP5 = 0x27D4EB2F165667C5ULL;

__device__ __forceinline__ unsigned long long Rotl64(unsigned long long v, int r) { return (v << r) | (v >> (64 - r)); }

// xxhash64 of one 64-bit word with seed 0 (XxHash64UpdateTail + XxHash64FinalShuffle from P5 + 8)
__device__ __forceinline__ unsigned long long XxWord(unsigned long long v)
{
    unsigned long long h = kXxP5 + 8;
    h = Rotl64(h ^ (Rotl64(v * kXxP2, 31) * kXxP1), 27) * kXxP1 + kXxP4;
    h ^= h >> 33;
    h *= kXxP2;
    h ^= h >> 29;
    h *= kXxP3;
    h ^= h >> 32;
    return h;
}

__device__ __forceinline__ unsigned long long HashUtilValue(unsigned long long v) { return Rotl64(v * kXxP2, 31) * kXxP1; }

struct PartKeyCols {
    int n;
    int types[OMNI_GPU_MAX_KEYS];
    const void *values[OMNI_GPU_MAX_KEYS];
    const uint8_t *nulls[OMNI_GPU_MAX_KEYS];
};

__global__ void __launch_bounds__(256) join_partition_ids(PartKeyCols kc, int64_t rows, unsigned int mask,
                                                          unsigned int *__restrict__ out)
{
    const int64_t r = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (r >= rows) return;
    unsigned long long h = 0;
    for (int c = 0; c < kc.n; ++c) {
        if (NullBit(kc.nulls[c], r)) continue;
        unsigned long long hv;
        switch (kc.types[c]) {
            case OMNI_TYPE_SHORT: hv = HashUtilValue(static_cast<long long>(static_cast<const short *>(kc.values[c])[r])); break;
            case OMNI_TYPE_INT:
            case OMNI_TYPE_DATE32: hv = HashUtilValue(static_cast<long long>(static_cast<const int *>(kc.values[c])[r])); break;
            case OMNI_TYPE_BOOLEAN: hv = static_cast<const uint8_t *>(kc.values[c])[r] ? 1231ULL : 1237ULL; break;
            case OMNI_TYPE_DECIMAL64: hv = static_cast<const unsigned long long *>(kc.values[c])[r]; break;
            case OMNI_TYPE_DOUBLE: {
                unsigned long long bits = static_cast<const unsigned long long *>(kc.values[c])[r];
                if ((bits & 0x7FF0000000000000ULL) == 0x7FF0000000000000ULL && (bits & 0x000FFFFFFFFFFFFFULL) != 0) {
                    bits = 0x7FF8000000000000ULL;
                }
                hv = HashUtilValue(bits);
                break;
            }
            case OMNI_TYPE_DECIMAL128: {
                const auto *w = static_cast<const unsigned long long *>(kc.values[c]) + 2 * r;
                hv = XxWord(w[0]) ^ XxWord(w[1] & 0x7FFFFFFFFFFFFFFFULL);
                break;
            }
            default: hv = HashUtilValue(static_cast<const unsigned long long *>(kc.values[c])[r]); break;
        }
        h = 31ULL * h + hv;
    }
    out[r] = static_cast<unsigned int>(XxWord(__brevll(h))) & mask;
}

}  // namespace omni_gpu

using namespace omni_gpu;

extern "C" {

int omni_gpu_join_build_create(const omni_gpu_join_spec *spec, void **handle)
{
    OG_CHECK(spec && handle, OMNI_GPU_ERR_INVALID, "null argument");
    // single keys the reference keeps in a fixed-width table (HashBuilderOperatorFactory::InitVariant,
    // hash_builder.cpp:45-80) are joined directly -- DOUBLE on its 64-bit pattern (the int64 table), BOOLEAN on
    // 0 / 1 --, the other key variants through their translation (join_keys.h)
    JoinKeyCols kc{};
    OG_TRY(PlanJoinKeys(*spec, kc));
    OG_CHECK(spec->num_payloads >= 0 && spec->num_payloads <= OMNI_GPU_MAX_AGGS, OMNI_GPU_ERR_INVALID,
             "num_payloads out of range");
    for (int p = 0; p < spec->num_payloads; ++p) {
        OG_CHECK(TypeWidth(spec->payload_types[p]) > 0, OMNI_GPU_ERR_INVALID, "payload type %d not supported",
                 spec->payload_types[p]);
    }
    std::unique_ptr<JoinHandle> h(new JoinHandle());
    h->spec = *spec;
    h->kc = kc;
    h->jkey_type = kc.mode == kKeySingle ? spec->key_type : OMNI_TYPE_LONG;
    h->spec.key_type = h->jkey_type;
    OG_TRY(AcquireStream(&h->stream));
    OG_TRY(h->scratch.Reserve(64, h->stream));
    *handle = h.release();
    return OMNI_GPU_OK;
}

int omni_gpu_join_build_add(void *handle, const omni_gpu_column *key, const omni_gpu_column *payloads, int64_t rows)
{
    OG_CHECK(handle && key, OMNI_GPU_ERR_INVALID, "null argument");
    JoinHandle &h = *static_cast<JoinHandle *>(handle);
    OG_CHECK(!h.finished, OMNI_GPU_ERR_STATE, "build side already finished");
    OG_CHECK(rows >= 0, OMNI_GPU_ERR_INVALID, "negative row count");
    if (rows == 0) return OMNI_GPU_OK;
    OG_CHECK(h.n + rows < 0x7FFFFFFF, OMNI_GPU_ERR_INVALID, "build side exceeds 2^31-1 rows");
    if (h.kc.mode == kKeySingle) {
        OG_CHECK(key->type_id == h.spec.key_type, OMNI_GPU_ERR_INVALID, "build key type mismatch");
        OG_CHECK(key->values, OMNI_GPU_ERR_INVALID, "build key has no values");
    } else {  // `key` holds kc.n columns, translated at finish
        KeyBatch kb{};
        for (int c = 0; c < h.kc.n; ++c) {
            OG_CHECK(key[c].type_id == h.kc.type[c] && key[c].values, OMNI_GPU_ERR_INVALID,
                     "build key column %d: type %d (expected %d) or no values", c, key[c].type_id, h.kc.type[c]);
            kb.values[c] = key[c].values;
            kb.nulls[c] = key[c].nulls;
        }
        kb.rows = rows;
        h.pending_keys.push_back(kb);
    }
    OG_CHECK(h.spec.num_payloads == 0 || payloads, OMNI_GPU_ERR_INVALID, "null payload columns");
    BuildBatch bb{};
    bb.key = key->values;
    bb.knull = key->nulls;
    bb.rows = rows;
    bb.row_base = h.n;
    for (int p = 0; p < h.spec.num_payloads; ++p) {
        OG_CHECK(payloads[p].type_id == h.spec.payload_types[p], OMNI_GPU_ERR_INVALID, "payload %d type mismatch", p);
        OG_CHECK(payloads[p].values, OMNI_GPU_ERR_INVALID, "payload %d has no values", p);
        bb.pay[p] = payloads[p].values;
        bb.pnull[p] = payloads[p].nulls;
        if (payloads[p].nulls) h.pay_nullable[p] = true;
    }
    h.pending.push_back(bb);
    h.n += rows;
    return OMNI_GPU_OK;
}

int omni_gpu_join_build_finish(void *handle)
{
    OG_CHECK(handle, OMNI_GPU_ERR_INVALID, "null handle");
    JoinHandle &h = *static_cast<JoinHandle *>(handle);
    OG_CHECK(!h.finished, OMNI_GPU_ERR_STATE, "build side already finished");
    hipStream_t s = h.stream;
    uint8_t *sc = h.scratch.as<uint8_t>();
    // scratch: sentinel head -1, has_dup 0, max chain 0, sentinel inline payload, then min/max seeds
    struct {
        int head, has_dup;
        unsigned int max_len, sentinel_pay;
        long long mm[2];
        unsigned long long counter;
        unsigned int ovf, pad;
    } init = {-1, 0, 0, 0, {std::numeric_limits<long long>::max(), std::numeric_limits<long long>::min()}, 0, 0, 0};
    static_assert(sizeof(init) == 48, "scratch layout");
    OG_TRY(CopyToDevice(sc, &init, sizeof(init), s));
    h.finished = true;
    if (h.preserve_build && h.n > 0) {
        OG_TRY(h.visited.Reserve(static_cast<size_t>(h.n), s));
        OG_HIP(hipMemsetAsync(h.visited.p, 0, static_cast<size_t>(h.n), s));
    }
    if (h.kc.mode != kKeySingle) {
        // every build batch's key columns into int64 join keys (dictionary ids for wide keys); the gather below
        // reads them as the batches' key columns
        if (h.kc.mode != kKeyPacked32 && h.kc.mode != kKeyPacked64) {
            uint64_t cap = 1024;
            while (cap < static_cast<uint64_t>(2 * h.n)) cap <<= 1;
            const int S = DictStride(h.kc.W);
            OG_TRY(h.dict.Reserve(sizeof(unsigned long long) * S * cap, s));
            h.dict_slots = cap;
            KernelScope ks(s, "join_init_table");
            join_dict_clear<<<Blocks(static_cast<int64_t>(cap)), kBlock, 0, s>>>(h.dict.as<unsigned long long>(), S,
                                                                                  h.kc.W, static_cast<int64_t>(cap));
            OG_TRY(CheckLaunch("join_dict_clear"));
        }
        const int nb = static_cast<int>(h.pending.size());
        std::vector<int64_t> rows(nb);
        std::vector<omni_gpu_column> in(static_cast<size_t>(nb) * h.kc.n);
        for (int b = 0; b < nb; ++b) {
            rows[b] = h.pending[b].rows;
            for (int c = 0; c < h.kc.n; ++c) {
                omni_gpu_column &col = in[static_cast<size_t>(b) * h.kc.n + c];
                col.type_id = h.kc.type[c];
                col.values = h.pending_keys[b].values[c];
                col.nulls = h.pending_keys[b].nulls[c];
            }
        }
        std::vector<omni_gpu_column> cols;
        if (nb > 0) OG_TRY(TranslateKeys<true>(h, in.data(), rows.data(), nb, h.bkeys, h.bnulls, cols));
        for (int b = 0; b < nb; ++b) {
            h.pending[b].key = cols[b].values;
            h.pending[b].knull = cols[b].nulls;
        }
        h.pending_keys.clear();
    }
    if (h.n == 0) {
        h.max_dup = 0;
        h.pending.clear();
        return OMNI_GPU_OK;
    }
    // build store: every retained batch gathered in one launch
    OG_TRY(h.keys.Reserve(sizeof(long long) * h.n, s));
    OG_TRY(h.knull.Reserve(h.n, s));
    GatherArgs g{};
    g.nb = static_cast<int>(h.pending.size());
    g.key_width = TypeWidth(h.spec.key_type);
    g.npay = h.spec.num_payloads;
    g.keys = h.keys.as<long long>();
    g.knull = h.knull.as<uint8_t>();
    for (int p = 0; p < g.npay; ++p) {
        g.pay_width[p] = TypeWidth(h.spec.payload_types[p]);
        OG_TRY(h.pay[p].Reserve(static_cast<size_t>(g.pay_width[p]) * h.n, s));
        g.pay[p] = h.pay[p].p;
        if (h.pay_nullable[p]) {
            OG_TRY(h.pay_null[p].Reserve(h.n, s));
            g.pnull[p] = h.pay_null[p].as<uint8_t>();
        }
    }
    // JoinHashTableVariants::BuildHashTable tries the array table for INT / DATE32 / SHORT / LONG / TIMESTAMP
    // keys only (join_hash_table_variants.cpp:700-736); DECIMAL64, DOUBLE and BOOLEAN keys always hash.
    const int kt = h.spec.key_type;
    const bool arrayCandidate = h.spec.strategy == 0 && h.kc.mode == kKeySingle &&
                                (kt == OMNI_TYPE_INT || kt == OMNI_TYPE_DATE32 || kt == OMNI_TYPE_SHORT ||
                                 kt == OMNI_TYPE_LONG || kt == OMNI_TYPE_TIMESTAMP);
    int64_t blocks = 0;
    h.key_nulls = false;
    for (BuildBatch &bb : h.pending) {
        h.key_nulls |= bb.knull != nullptr;
        bb.block_start = blocks;
        blocks += (bb.rows + kGatherTile - 1) / kGatherTile;
    }
    OG_CHECK(blocks < 0x7FFFFFFF, OMNI_GPU_ERR_INVALID, "build side too large for one gather launch");
    OG_TRY(h.build_desc.Reserve(sizeof(BuildBatch) * h.pending.size(), s));
    OG_TRY(CopyToDevice(h.build_desc.p, h.pending.data(), sizeof(BuildBatch) * h.pending.size(), s));
    g.batches = h.build_desc.as<BuildBatch>();
    long long *dmm = reinterpret_cast<long long *>(sc + 16);
    if (arrayCandidate) {
        OG_TRY(h.mm_parts.Reserve(2 * sizeof(long long) * static_cast<size_t>(blocks), s));
        g.mm = h.mm_parts.as<long long>();
    }
    // Once the gather is queued, every return waits for it: the caller may release its build buffers as soon
    // as finish returns, also after an error.
    struct GatherWait {
        hipStream_t s;
        bool armed = false;
        ~GatherWait()
        {
            if (armed) (void)hipStreamSynchronize(s);
        }
    } gatherWait{s};
    {
        KernelScope ks(s, "join_gather_build");
        join_gather_build<<<static_cast<unsigned int>(blocks), 256, 0, s>>>(g);
        gatherWait.armed = true;
        OG_TRY(CheckLaunch("join_gather_build"));
        if (g.mm) {
            join_minmax_reduce<<<1, 1024, 0, s>>>(g.mm, blocks, dmm);
            OG_TRY(CheckLaunch("join_minmax_reduce"));
        }
    }
    // The handle's pinned cell: [0] the duplicate-key flag, [16] the key min / max. Each read-back is an async
    // copy before flags_ev, so a wait covers the work queued before it and not the whole stream.
    if (!h.flag_cell) OG_TRY(TakeFlagCell(&h.flag_cell));
    if (!h.flags_ev) OG_HIP(hipEventCreateWithFlags(&h.flags_ev, hipEventDisableTiming));
    const long long *cell_mm = reinterpret_cast<const long long *>(reinterpret_cast<uint8_t *>(h.flag_cell) + 16);
    if (g.mm) {
        OG_HIP(hipMemcpyAsync(const_cast<long long *>(cell_mm), dmm, 2 * sizeof(long long), hipMemcpyDeviceToHost,
                              s));
    }
    OG_HIP(hipEventRecord(h.flags_ev, s));  // after the gather: the caller's build buffers are free after it
    OG_TRY(h.next.Reserve(sizeof(int) * h.n, s));
    // The hash table's size depends on n only (load <= 1/4: linear-probing runs stay short, and the window
    // probe's LDS rounds and the direct probe's line crossings both scale with the longest runs a wave meets).
    // Its partitioned build starts by partitioning the build rows: queued here, ahead of the min/max read-back
    // below, so the GPU works while the host waits for the array-table decision (an array table leaves the
    // partitioning unused).
    uint64_t hashSlots = 1024;
    while (hashSlots < static_cast<uint64_t>(4 * h.n)) hashSlots <<= 1;
    const uint64_t bP = hashSlots / kWinSlots;
    const bool partBuild = bP >= static_cast<uint64_t>(kXcds) && bP <= static_cast<uint64_t>(kPartMax) &&
                           h.n < (1LL << 32) - 1;
    PartArgs bpa;
    bool prePartitioned = false;
    if (partBuild) {
        h.slots_n = hashSlots;
        OG_TRY(PartitionBuildRows(h, static_cast<int>(bP), bpa));
        prePartitioned = true;
    }
    // initDegree and the array-table rule of JoinHashTableVariants::BuildHashTable / TryToBuildArrayTable.
    bool arrayOk = false;
    if (arrayCandidate) {
        OG_HIP(hipEventSynchronize(h.flags_ev));
        const long long mm[2] = {cell_mm[0], cell_mm[1]};
        if (mm[1] >= mm[0]) {
            double degree = std::ceil(std::log2(static_cast<double>(h.n) / 0.75));
            long double len = std::pow(2.0L, static_cast<long double>(degree));
            long double range = static_cast<long double>(mm[1]) - static_cast<long double>(mm[0]);
            // max - min must not overflow the key type (TryToBuildArrayTable's `min < 0 && T_MAX + min < max`)
            long double typeMax = kt == OMNI_TYPE_SHORT ? 32767.0L
                                  : (kt == OMNI_TYPE_INT || kt == OMNI_TYPE_DATE32) ? 2147483647.0L
                                                                                    : 9223372036854775807.0L;
            arrayOk = range <= typeMax && range <= 4294967295.0L && range <= 8.0L * len;
            if (arrayOk) {
                h.min_key = mm[0];
                h.range = static_cast<unsigned long long>(mm[1] - mm[0]) + 1;
            }
        }
    }
    h.pending.clear();
    unsigned int *has_dup = reinterpret_cast<unsigned int *>(sc + 4);
    if (arrayOk) {
        h.is_array = true;
        h.slots_n = 0;
        OG_TRY(h.heads.Reserve(sizeof(int) * h.range, s));
        {
            KernelScope ks(s, "join_init_table");
            join_fill_i32<<<Blocks(static_cast<int64_t>(h.range)), kBlock, 0, s>>>(h.heads.as<int>(), -1,
                                                                                    static_cast<int64_t>(h.range));
            OG_TRY(CheckLaunch("join_fill_i32"));
        }
        KernelScope ks(s, "join_build");
        join_build_array<<<Blocks(h.n), kBlock, 0, s>>>(h.keys.as<long long>(), h.knull.as<uint8_t>(), h.n,
                                                        h.heads.as<int>(), h.min_key, h.next.as<int>(), has_dup);
        OG_TRY(CheckLaunch("join_build_array"));
    } else {
        h.slots_n = hashSlots;
        OG_TRY(h.slots.Reserve(sizeof(JoinSlot) * h.slots_n, s));
        h.inline_pay = h.spec.num_payloads == 1 && TypeWidth(h.spec.payload_types[0]) <= 4 && !h.pay_nullable[0];
        if (partBuild) {
            OG_TRY(BuildPartitioned(h, static_cast<int>(bP), has_dup, prePartitioned ? &bpa : nullptr));
        } else {
            {
                KernelScope ks(s, "join_init_table");
                join_init_slots<<<Blocks(static_cast<int64_t>(h.slots_n)), kBlock, 0, s>>>(
                    h.slots.as<JoinSlot>(), static_cast<int64_t>(h.slots_n));
                OG_TRY(CheckLaunch("join_init_slots"));
            }
            KernelScope ks(s, "join_build");
            join_build_hash<<<Blocks(h.n), kBlock, 0, s>>>(h.keys.as<long long>(), h.knull.as<uint8_t>(), h.n,
                                                           h.slots.as<JoinSlot>(), h.slots_n - 1, h.next.as<int>(),
                                                           reinterpret_cast<int *>(sc),
                                                           reinterpret_cast<unsigned int *>(sc + 12), has_dup,
                                                           h.inline_pay ? h.pay[0].p : nullptr,
                                                           h.inline_pay ? TypeWidth(h.spec.payload_types[0]) : 0);
            OG_TRY(CheckLaunch("join_build_hash"));
        }
    }
    // the probe kernel variant depends on whether any key repeats: read back without a stream
    // synchronisation, so the build kernels overlap the host's set-up of the probe (ResolveBuildFlags)
    // the gather must have read the caller's build buffers before finish returns (the array-table decision
    // above already waited for it)
    if (!g.mm) OG_HIP(hipEventSynchronize(h.flags_ev));
    gatherWait.armed = false;
    OG_HIP(hipMemcpyAsync(h.flag_cell, has_dup, 4, hipMemcpyDeviceToHost, s));
    OG_HIP(hipEventRecord(h.flags_ev, s));
    h.flags_pending = true;
    return OMNI_GPU_OK;
}

// Longest build-row chain of one key (the largest number of matches one probe row can produce).
static int MaxChain(JoinHandle &h)
{
    hipStream_t s = h.stream;
    uint8_t *sc = h.scratch.as<uint8_t>();
    unsigned int *max_len = reinterpret_cast<unsigned int *>(sc + 8);
    OG_HIP(hipMemsetAsync(max_len, 0, 4, s));
    {
        KernelScope ks(s, "join_max_chain");
        if (h.is_array) {
            join_max_chain<<<Blocks(static_cast<int64_t>(h.range)), kBlock, 0, s>>>(h.heads.as<int>(),
                                                                                    static_cast<int64_t>(h.range), 1,
                                                                                    h.next.as<int>(), max_len);
        } else {
            join_max_chain<<<Blocks(static_cast<int64_t>(h.slots_n)), kBlock, 0, s>>>(
                reinterpret_cast<const int *>(h.slots.as<uint8_t>() + 8), static_cast<int64_t>(h.slots_n), 4,
                h.next.as<int>(), max_len);
        }
        OG_TRY(CheckLaunch("join_max_chain"));
    }
    if (!h.is_array) {  // the sentinel-key chain is outside the table
        KernelScope ks(s, "join_max_chain");
        join_max_chain<<<1, kBlock, 0, s>>>(reinterpret_cast<const int *>(sc), 1, 1, h.next.as<int>(), max_len);
        OG_TRY(CheckLaunch("join_max_chain"));
    }
    unsigned int len = 0;
    OG_TRY(CopyToHost(s, {{&len, max_len, 4}}));
    h.max_dup = len;
    return OMNI_GPU_OK;
}

int omni_gpu_join_build_info(void *handle, int32_t *is_array, int64_t *max_dup, int64_t *table_slots)
{
    OG_CHECK(handle, OMNI_GPU_ERR_INVALID, "null handle");
    JoinHandle &h = *static_cast<JoinHandle *>(handle);
    OG_CHECK(h.finished, OMNI_GPU_ERR_STATE, "build side not finished");
    OG_TRY(ResolveBuildFlags(h));
    if (max_dup && h.max_dup < 0) OG_TRY(MaxChain(h));
    if (is_array) *is_array = h.is_array ? 1 : 0;
    if (max_dup) *max_dup = h.max_dup;
    if (table_slots) *table_slots = h.is_array ? static_cast<int64_t>(h.range) : static_cast<int64_t>(h.slots_n);
    return OMNI_GPU_OK;
}

int omni_gpu_join_tune(void *handle, int32_t probe_group, int32_t partitions)
{
    OG_CHECK(handle, OMNI_GPU_ERR_INVALID, "null handle");
    OG_CHECK(probe_group == 1 || probe_group == 2 || probe_group == 4 || probe_group == 8 || probe_group == 16,
             OMNI_GPU_ERR_INVALID, "probe_group must be 1, 2, 4, 8 or 16");
    OG_CHECK(partitions == -2 || partitions == -1 || partitions == 0 ||
                 (partitions >= kXcds && partitions <= kPartMax && (partitions & (partitions - 1)) == 0),
             OMNI_GPU_ERR_INVALID,
             "partitions must be -1 (auto), -2 (always), 0 (off) or a power of two in [%d, %d]", kXcds, kPartMax);
    JoinHandle &h = *static_cast<JoinHandle *>(handle);
    h.probe_group = probe_group;
    h.partitions = partitions;
    return OMNI_GPU_OK;
}

int omni_gpu_join_probe_info(void *handle, int32_t *partitions)
{
    OG_CHECK(handle && partitions, OMNI_GPU_ERR_INVALID, "null argument");
    *partitions = static_cast<JoinHandle *>(handle)->last_partitions;
    return OMNI_GPU_OK;
}

int omni_gpu_join_probe(void *handle, const omni_gpu_column *probe_key, int64_t rows, int64_t probe_row_base,
                        int64_t *out_probe_rows, omni_gpu_out_column *out_payloads, int64_t out_capacity,
                        int64_t *num_matches)
{
    OG_CHECK(handle && probe_key, OMNI_GPU_ERR_INVALID, "null argument");
    JoinHandle &h = *static_cast<JoinHandle *>(handle);
    std::vector<omni_gpu_column> tk;
    if (h.kc.mode != kKeySingle) {  // the key columns into int64 join keys (a key not in the build side: NULL)
        OG_CHECK(h.finished, OMNI_GPU_ERR_STATE, "probe before the build side finished");
        OG_TRY(TranslateKeys<false>(h, probe_key, &rows, 1, h.pkeys, h.pnulls, tk));
        probe_key = tk.data();
    }
    ProbeBatch b{};
    OG_TRY(FillBatch(b, *probe_key, rows, probe_row_base, 0, 0, h.spec.key_type));
    return RunProbe(h, &b, 1, (rows + kProbeTile - 1) / kProbeTile, (rows + kPartTile - 1) / kPartTile,
                    reinterpret_cast<long long *>(out_probe_rows), out_payloads, out_capacity, num_matches);
}

int omni_gpu_join_probe_batches(void *handle, const omni_gpu_column *keys, const int64_t *rows, int32_t num_batches,
                                int64_t *out_probe_rows, omni_gpu_out_column *out_payloads, int64_t out_capacity,
                                int64_t *num_matches)
{
    OG_CHECK(handle && keys && rows && num_batches > 0, OMNI_GPU_ERR_INVALID, "null argument or no batches");
    JoinHandle &h = *static_cast<JoinHandle *>(handle);
    std::vector<omni_gpu_column> tk;
    if (h.kc.mode != kKeySingle) {  // batch b's key columns are keys[b * num_keys ..]
        OG_CHECK(h.finished, OMNI_GPU_ERR_STATE, "probe before the build side finished");
        OG_TRY(TranslateKeys<false>(h, keys, rows, num_batches, h.pkeys, h.pnulls, tk));
        keys = tk.data();
    }
    std::vector<ProbeBatch> bs(num_batches);
    int64_t blocks = 0, pblocks = 0, base = 0;
    for (int i = 0; i < num_batches; ++i) {
        OG_TRY(FillBatch(bs[i], keys[i], rows[i], base, blocks, pblocks, h.spec.key_type));
        blocks += (rows[i] + kProbeTile - 1) / kProbeTile;
        pblocks += (rows[i] + kPartTile - 1) / kPartTile;
        base += rows[i];
    }
    OG_CHECK(blocks < 0x7FFFFFFF, OMNI_GPU_ERR_INVALID, "too many probe rows for one launch");
    return RunProbe(h, bs.data(), num_batches, blocks, pblocks, reinterpret_cast<long long *>(out_probe_rows),
                    out_payloads, out_capacity, num_matches);
}

int omni_gpu_join_set_type(void *handle, int32_t join_type, int32_t build_side)
{
    OG_CHECK(handle, OMNI_GPU_ERR_INVALID, "null handle");
    JoinHandle &h = *static_cast<JoinHandle *>(handle);
    OG_CHECK(!h.finished, OMNI_GPU_ERR_STATE, "the join type is set before the build side finishes");
    OG_CHECK(join_type >= OMNI_GPU_JOIN_INNER && join_type <= OMNI_GPU_JOIN_EXISTENCE, OMNI_GPU_ERR_INVALID,
             "join type %d not supported", join_type);
    OG_CHECK(build_side >= OMNI_GPU_BUILD_UNKNOWN && build_side <= OMNI_GPU_BUILD_RIGHT, OMNI_GPU_ERR_INVALID,
             "build side %d not supported", build_side);
    // lookup_join.cpp:344-367 and hash_builder.cpp:36-40: LEFT / RIGHT keep the build side's rows when the build
    // is on that side, else the probe side's; FULL keeps both
    const bool same = (join_type == OMNI_GPU_JOIN_LEFT && build_side == OMNI_GPU_BUILD_LEFT) ||
                      (join_type == OMNI_GPU_JOIN_RIGHT && build_side == OMNI_GPU_BUILD_RIGHT);
    h.mode = kModeInner;
    h.preserve_build = false;
    h.existence = false;
    switch (join_type) {
        case OMNI_GPU_JOIN_LEFT:
        case OMNI_GPU_JOIN_RIGHT:
            if (same) h.preserve_build = true;
            else h.mode = kModeLeft;
            break;
        case OMNI_GPU_JOIN_FULL:
            h.mode = kModeLeft;
            h.preserve_build = true;
            break;
        case OMNI_GPU_JOIN_LEFT_SEMI: h.mode = kModeSemi; break;
        case OMNI_GPU_JOIN_LEFT_ANTI: h.mode = kModeAnti; break;
        case OMNI_GPU_JOIN_EXISTENCE: h.existence = true; break;
        default: break;
    }
    return OMNI_GPU_OK;
}

int omni_gpu_join_probe_exists(void *handle, const omni_gpu_column *probe_key, int64_t rows, uint8_t *out_exists)
{
    OG_CHECK(handle && probe_key, OMNI_GPU_ERR_INVALID, "null argument");
    JoinHandle &h = *static_cast<JoinHandle *>(handle);
    OG_CHECK(h.finished, OMNI_GPU_ERR_STATE, "probe before the build side finished");
    OG_CHECK(rows == 0 || out_exists, OMNI_GPU_ERR_INVALID, "null out_exists");
    OG_TRY(ResolveBuildFlags(h));
    std::vector<omni_gpu_column> tk;
    if (h.kc.mode != kKeySingle) {
        OG_TRY(TranslateKeys<false>(h, probe_key, &rows, 1, h.pkeys, h.pnulls, tk));
        probe_key = tk.data();
    }
    ProbeBatch b{};
    OG_TRY(FillBatch(b, *probe_key, rows, 0, 0, 0, h.spec.key_type));
    OG_CHECK(rows < (1LL << 31) * kBlock, OMNI_GPU_ERR_INVALID, "probe batch too large");
    hipStream_t s = h.stream;
    if (rows > 0) {
        const TableDev t = MakeTable(h);
        const int empty = h.n == 0 ? 1 : 0;
        KernelScope ks(s, "join_probe_exists");
        if (h.is_array) join_probe_exists<true><<<Blocks(rows), kBlock, 0, s>>>(b, t, empty, out_exists);
        else join_probe_exists<false><<<Blocks(rows), kBlock, 0, s>>>(b, t, empty, out_exists);
        OG_TRY(CheckLaunch("join_probe_exists"));
    }
    OG_HIP(hipStreamSynchronize(s));
    return OMNI_GPU_OK;
}

int omni_gpu_join_outer_output(void *handle, int64_t *out_build_rows, omni_gpu_out_column *out_payloads,
                               int64_t out_capacity, int64_t *num_rows)
{
    OG_CHECK(handle && num_rows, OMNI_GPU_ERR_INVALID, "null argument");
    JoinHandle &h = *static_cast<JoinHandle *>(handle);
    OG_CHECK(h.finished, OMNI_GPU_ERR_STATE, "outer output before the build side finished");
    OG_CHECK(h.preserve_build, OMNI_GPU_ERR_STATE, "outer output needs a build-preserving outer join type");
    OG_CHECK(out_capacity >= 0, OMNI_GPU_ERR_INVALID, "negative out_capacity");
    OG_TRY(ResolveBuildFlags(h));
    hipStream_t s = h.stream;
    OuterArgs g{};
    g.visited = h.visited.as<uint8_t>();
    g.n = h.n;
    g.out_rows = reinterpret_cast<long long *>(out_build_rows);
    g.npay = h.spec.num_payloads;
    g.counter = reinterpret_cast<unsigned long long *>(h.scratch.as<uint8_t>() + 32);
    g.cap = out_capacity;
    OG_TRY(SetupPayloadOutputs(h, out_payloads, out_capacity, g.pay));
    OG_HIP(hipMemsetAsync(g.counter, 0, 8, s));
    if (h.n > 0) {
        KernelScope ks(s, "join_outer_emit");
        join_outer_emit<<<Blocks(h.n), kBlock, 0, s>>>(g);
        OG_TRY(CheckLaunch("join_outer_emit"));
    }
    unsigned long long m = 0;
    OG_TRY(CopyToHost(s, {{&m, g.counter, 8}}));
    *num_rows = static_cast<int64_t>(m);
    if (static_cast<int64_t>(m) > out_capacity) {
        return SetError(OMNI_GPU_ERR_OUTPUT_FULL, "outer output has %llu rows for an output capacity of %lld", m,
                        static_cast<long long>(out_capacity));
    }
    return OMNI_GPU_OK;
}

int omni_gpu_join_gather_rows(const omni_gpu_column *cols, int32_t num_cols, const int64_t *rows, int64_t n,
                              omni_gpu_out_column *out)
{
    OG_CHECK(num_cols >= 0 && n >= 0 && (num_cols == 0 || (cols && out)) && (n == 0 || rows), OMNI_GPU_ERR_INVALID,
             "null argument");
    if (n == 0 || num_cols == 0) return OMNI_GPU_OK;
    for (int32_t c = 0; c < num_cols; ++c) {
        const int w = TypeWidth(cols[c].type_id);
        OG_CHECK(w > 0 && cols[c].values && out[c].values, OMNI_GPU_ERR_INVALID, "column %d: fixed-width values needed", c);
        OG_CHECK(TypeWidth(out[c].type_id) == w, OMNI_GPU_ERR_INVALID, "column %d: output width differs", c);
    }
    hipStream_t s;
    OG_TRY(AcquireStream(&s));
    int rc = OMNI_GPU_OK;
    const unsigned int grid = static_cast<unsigned int>((n + 255) / 256);
    for (int32_t c0 = 0; c0 < num_cols && rc == OMNI_GPU_OK; c0 += kRowGatherCols) {
        RowGatherArgs g{};
        g.ncols = std::min<int32_t>(kRowGatherCols, num_cols - c0);
        g.rows = reinterpret_cast<const long long *>(rows);
        g.n = n;
        for (int i = 0; i < g.ncols; ++i) {
            const omni_gpu_column &src = cols[c0 + i];
            g.c[i] = RowGatherCol{src.values, src.nulls, out[c0 + i].values,
                               reinterpret_cast<unsigned int *>(out[c0 + i].nulls), TypeWidth(src.type_id)};
        }
        join_gather_rows<<<grid, 256, 0, s>>>(g);
        rc = CheckLaunch("join_gather_rows");
    }
    if (rc == OMNI_GPU_OK && hipStreamSynchronize(s) != hipSuccess) rc = SetError(OMNI_GPU_ERR_HIP, "gather failed");
    ReleaseStream(s);
    return rc;
}

int omni_gpu_join_partition_ids(const omni_gpu_column *keys, int32_t num_keys, int64_t rows, uint32_t mask,
                                uint32_t *out_partitions)
{
    OG_CHECK(keys && out_partitions && num_keys >= 1 && num_keys <= OMNI_GPU_MAX_KEYS && rows >= 0,
             OMNI_GPU_ERR_INVALID, "bad argument");
    if (rows == 0) return OMNI_GPU_OK;
    PartKeyCols kc{};
    kc.n = num_keys;
    for (int32_t c = 0; c < num_keys; ++c) {
        const int t = keys[c].type_id;
        OG_CHECK(t == OMNI_TYPE_SHORT || t == OMNI_TYPE_INT || t == OMNI_TYPE_DATE32 || t == OMNI_TYPE_LONG ||
                     t == OMNI_TYPE_TIMESTAMP || t == OMNI_TYPE_DECIMAL64 || t == OMNI_TYPE_DOUBLE ||
                     t == OMNI_TYPE_BOOLEAN || t == OMNI_TYPE_DECIMAL128,
                 OMNI_GPU_ERR_INVALID, "key column %d: type %d has no partition hash here", c, t);
        OG_CHECK(keys[c].values, OMNI_GPU_ERR_INVALID, "key column %d has no values", c);
        kc.types[c] = t;
        kc.values[c] = keys[c].values;
        kc.nulls[c] = keys[c].nulls;
    }
    hipStream_t s;
    OG_TRY(AcquireStream(&s));
    join_partition_ids<<<static_cast<unsigned int>((rows + 255) / 256), 256, 0, s>>>(kc, rows, mask, out_partitions);
    int rc = CheckLaunch("join_partition_ids");
    if (rc == OMNI_GPU_OK && hipStreamSynchronize(s) != hipSuccess) rc = SetError(OMNI_GPU_ERR_HIP, "partition ids failed");
    ReleaseStream(s);
    return rc;
}

int omni_gpu_join_destroy(void *handle)
{
    if (!handle) return OMNI_GPU_OK;
    JoinHandle *h = static_cast<JoinHandle *>(handle);
    hipStream_t s = h->stream;
    if (h->flags_ev) {  // the flag copy into the cell must land before the cell is handed out again
        (void)hipEventSynchronize(h->flags_ev);
        (void)hipEventDestroy(h->flags_ev);
    }
    if (h->flag_cell) ReturnFlagCell(h->flag_cell);
    h->ForEachBuffer([s](DevBuf &b) { b.Release(s); });
    ReleaseStream(s);
    delete h;
    return OMNI_GPU_OK;
}

}  // extern "C"
