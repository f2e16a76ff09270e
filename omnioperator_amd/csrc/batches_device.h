// Device code of the batched group-by: many independent group-bys (one per VectorBatch).
//
// Phase 1 (one workgroup per batch): the batch is aggregated in an LDS table; if any row finds no LDS slot
// within the probe limit, the workgroup abandons the batch (writes nothing) and flags it (spilled = 1).
// Phase 2 (flagged batches): the batch's keys are split by hash into P partitions, P a power of two sized
// so that a partition's rows fit half an LDS table.  2a: one workgroup per batch writes each row's key and
// row number to its partition's bucket (in the batch's region of the pool).  2b: one workgroup per (batch,
// partition) item aggregates its bucket in LDS and emits its groups.  A batch's 2a workgroup and all its 2b
// items run on the same XCD (workgroup id mod 8), so the buckets are read back from that XCD's L2.  Phases 1
// and 2a take the batches by decreasing rows.  A
// partition that still overflows its table flags the batch (spilled = 2) for phase 3.
// Phase 3 (batches flagged 2; batches flagged 1 when phase 2 is not used): the batch is aggregated exactly in
// its own HBM table of next_pow2(2 * rows) slots (plus 2 side slots) carved from a region pool at a scanned
// offset, by workgroups of kHbmTile rows (device-wide atomics); then one workgroup per batch emits all its
// groups again from output row off[b], setting and clearing NULL bits, so partial phase-2 output is replaced.
// Every phase emits batch b's groups to output rows [off[b], off[b] + groups[b]).
// The configuration lives in device memory and is read through one pointer (no large kernel arguments).
#pragma once

#include "agg_device.h"

namespace omni_gpu {

struct BatchConfig {
    KeyDev kd;
    AggArgs args;
    LdsLayout L;               // LDS table: probe slots [0, L.S), side slots L.S (NULL key), L.S + 1 (sentinel key)
    LdsLayout R;               // per-slot layout of the phase-2 HBM tables (UnitLayout)
    OutArgs out;
    const int64_t *offsets;    // nb + 1
    const int32_t *modes;      // per batch: 2 = skip LDS (straight to phase 2); nullptr = LDS first everywhere
    int32_t *sketch_modes;     // per batch, written by the sketch launch (same meaning); nullptr = no sketch
    int32_t nbatches;
    int32_t max_probe;
    int32_t overflow_as_null;
    int32_t pad;
    const int64_t *tile_start; // nb + 1: first phase-3 tile of each batch (kHbmTile rows per tile, at least one
                               // tile per batch, all batches)
    int64_t total_tiles;
    LdsLayout L2;              // phase-2 LDS table (probe slots [0, L2.S), side slots L2.S, L2.S + 1)
    const int4 *items;         // phase-2 work items {batch, partition, partitions, 0}; batch -1: no work
    int64_t nitems;
    const int32_t *parts;      // per batch: phase-2 partition items launched (a power of two, <= kMaxParts)
    int32_t *parts_used;       // per batch: partitions phase 2a chose (<= parts[b], from its distinct estimate)
    unsigned int *redo;        // set when some phase-2 partition overflowed (phase 3 has work)
    int32_t hbm_flag;          // phase 3 aggregates the batches whose spilled flag equals this value
    int32_t pad2;
    unsigned int *spilled;     // per batch: 1 = phase 2 must aggregate it
    unsigned long long *groups;  // per batch
    unsigned long long *region_off;  // per batch byte offset into the region pool (phase 2)
    uint8_t *region_pool;
    unsigned int *error_flag;
    // one-workgroup-per-batch form (BatchFused): the rows of each partitioned batch in partition order, at the
    // batch's own row positions -- their keys and either the value word every aggregate reads (carry = 1) or
    // the batch-relative row number (carry = 0)
    unsigned long long *perm_keys;
    unsigned long long *perm_vals;
    unsigned int *perm_rows;
    int32_t carry;
    int32_t carry_agg;         // the aggregate whose column is carried (-1: every aggregate is COUNT(*))
    int32_t part_fit;          // distinct keys aimed at per partition (0: kFusedLoad of the table's slots)
    int32_t lite_na;           // aggregates of the lite table (0: the generic table)
    int32_t lite_smax;         // probe slots of the lite table
    int32_t lite_bytes;        // its LDS bytes
    int32_t lite_op[4];        // per aggregate: COUNT, SUM, MIN or MAX of int64 (LiteOp)
    const float *pred;         // advisor outputs (4 per batch; [0] = log2(groups + 1) / 32), or nullptr
};

// Probe slots of the phase-2 table of a batch with `rows` rows: next power of two >= 2 * rows (at least 64).
__host__ __device__ __forceinline__ int RegionSlots(int64_t rows)
{
    int s = 64;
    while (static_cast<int64_t>(s) < 2 * rows) s <<= 1;
    return s;
}

// Slots per array of that table: the probe slots, two side slots, rounded up to a multiple of 256.
__host__ __device__ __forceinline__ long long RegionStride(int64_t rows)
{
    return (static_cast<long long>(RegionSlots(rows)) + 2 + 255) & ~255LL;
}

__host__ __device__ __forceinline__ unsigned long long RegionBytes(const LdsLayout &R, int64_t rows)
{
    return static_cast<unsigned long long>(R.bytes) * static_cast<unsigned long long>(RegionStride(rows));
}

// Resets the aggregate state of one slot.
template <typename Ref> __device__ __forceinline__ void InitSlot(const AggArgs &args, const Ref &r)
{
    for (int a = 0; a < args.n; ++a) {
        int kind = args.a[a].kind;
        if (Is32(kind)) *r.template p<int>(a, 0) = static_cast<int>(InitWord(kind));
        else *r.template p<unsigned long long>(a, 0) = InitWord(kind);
        if (ExtraWords(kind) >= 1) *r.template p<unsigned long long>(a, 1) = InitWord1(kind);
        if (ExtraWords(kind) >= 2) *r.template p<unsigned long long>(a, 2) = 0;
        if (kind != K_COUNT_ALL && kind != K_COUNT_COL && kind != K_AVG) *r.flag(a) = 0;
    }
}

// Writes one group (key + aggregates) to output row r; NULL bits are OR-ed into pre-cleared bitmaps.
// CLEAR: also clears the NULL bits of non-NULL results (the row may hold an earlier partial emission).
template <bool CLEAR = false, typename Ref>
__device__ __forceinline__ void EmitGroup(const BatchConfig &c, const Ref &s, unsigned long long key, bool keyNull,
                                          int64_t r)
{
    const OutDev &ko = c.out.o[0];
    StoreWidth(ko.values, ko.width, r, keyNull ? 0ULL : key);
    unsigned int *kw = reinterpret_cast<unsigned int *>(ko.nulls) + (r >> 5);
    if (keyNull) atomicOr(kw, 1u << (r & 31));
    else if (CLEAR) atomicAnd(kw, ~(1u << (r & 31)));
    for (int i = 0; i < c.args.n; ++i) {
        const OutDev &o = c.out.o[1 + i];
        bool isNull = FinalizeAgg(c.args.a[i], i, s, o, r, c.overflow_as_null, c.error_flag);
        unsigned int *w = reinterpret_cast<unsigned int *>(o.nulls) + (r >> 5);
        if (isNull) atomicOr(w, 1u << (r & 31));
        else if (CLEAR) atomicAnd(w, ~(1u << (r & 31)));
    }
}

// Emits every used slot of a table with probe slots [0, S) and side slots S, S + 1; ref(i) addresses slot i.
template <int BLOCK, bool CLEAR = false, typename MakeRef>
__device__ __forceinline__ void EmitTable(const BatchConfig &c, const unsigned long long *keys, int S, int64_t start,
                                          unsigned int *counter, MakeRef ref)
{
    // whole waves walk the slots so that each wave reserves its groups' output rows with one LDS atomic
    for (int base = threadIdx.x - LaneId(); base < S + 2; base += BLOCK) {
        const int i = base + LaneId();
        const unsigned long long k = i < S + 2 ? keys[i] : kEmptyKey;
        const bool used = k != kEmptyKey;
        const unsigned long long mask = __ballot(used);
        unsigned int first = 0;
        if (LaneId() == 0 && mask) first = atomicAdd(counter, static_cast<unsigned int>(__popcll(mask)));
        first = __shfl(first, 0, 64);
        if (!used) continue;
        const unsigned int r = first + static_cast<unsigned int>(__popcll(mask & LaneMaskLt()));
        const bool side = i >= S;
        EmitGroup<CLEAR>(c, ref(i), side ? (i == S ? 0ULL : kEmptyKey) : k, i == S, start + r);
    }
}

// Phase 1: LDS aggregation of batch b; abandons the batch when a row finds no slot.
template <int BLOCK>
__device__ __forceinline__ void BatchLds(const BatchConfig &c, int b, uint8_t *lds, unsigned int *sh)
{
    const LdsLayout &L = c.L;
    const int S = L.S;
    const int64_t start = c.offsets[b], end = c.offsets[b + 1];
    if ((c.modes != nullptr && c.modes[b] == 2) || (c.sketch_modes != nullptr && c.sketch_modes[b] == 2)) {
        if (threadIdx.x == 0) {
            c.spilled[b] = 1u;
            c.groups[b] = 0;
        }
        return;
    }
    unsigned long long *lkeys = reinterpret_cast<unsigned long long *>(lds + L.keys_off);
    for (int i = threadIdx.x; i < S + 2; i += BLOCK) {
        lkeys[i] = kEmptyKey;
        InitSlot(c.args, LdsRef{lds, L, i});
    }
    if (threadIdx.x == 0) {
        sh[0] = 0;  // a row found no LDS slot
        sh[1] = 0;  // groups emitted
    }
    __syncthreads();
    for (int64_t row = start + threadIdx.x; row < end; row += BLOCK) {
        if (sh[0]) break;  // the batch goes to phase 2 anyway
        bool isNull = NullBit(c.kd.nulls[0], row);
        unsigned long long key = isNull ? 0ULL : LoadWord(c.kd.values[0], c.kd.width[0], row);
        int slot = -1;
        if (isNull) {
            slot = S;
            lkeys[S] = 0;  // marks the side slot used
        } else if (key == kEmptyKey) {
            slot = S + 1;
            lkeys[S + 1] = 0;
        } else {
            unsigned long long h = Mix64(key);
            unsigned int pos = static_cast<unsigned int>(((h >> 32) * static_cast<unsigned long long>(S)) >> 32);
            for (int step = 0; step < c.max_probe; ++step) {
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
        if (slot < 0) {
            sh[0] = 1u;
            break;
        }
        UpdateAggsT(c.args, LdsRef{lds, L, slot}, row);
    }
    __syncthreads();
    if (sh[0]) {
        if (threadIdx.x == 0) {
            c.spilled[b] = 1u;
            c.groups[b] = 0;  // phase 2 items add their group counts
        }
        return;
    }
    EmitTable<BLOCK>(c, lkeys, S, start, &sh[1], [&](int i) { return LdsRef{lds, L, i}; });
    __syncthreads();
    if (threadIdx.x == 0) {
        c.spilled[b] = 0u;
        c.groups[b] = sh[1];
    }
}

// Partition of a key among P = 2^k partitions of a phase-2 batch: the low bits of its hash (the LDS slot uses
// the high bits).  The NULL key and the slot sentinel belong to partition 0.
__device__ __forceinline__ int PartOfKey(unsigned long long key, bool isNull, int P)
{
    if (isNull || key == kEmptyKey) return 0;
    return static_cast<int>(Mix64(key) & static_cast<unsigned long long>(P - 1));
}

constexpr int kMaxParts = 1024;  // phase-2 partitions of one batch

// Phase-2 buckets of batch b live in its region of the pool (the region is not used by phase 3 before every
// phase-2 launch has finished): u32 bucket starts [P + 1], then the keys (u64) and the batch-relative row
// numbers (u32) of the batch's rows in partition order.
struct Buckets {
    unsigned int *start;
    unsigned long long *keys;
    unsigned int *rows;
};

__device__ __forceinline__ Buckets BucketsOf(const BatchConfig &c, int b, int P, int64_t rows)
{
    uint8_t *base = c.region_pool + c.region_off[b];
    Buckets k;
    k.start = reinterpret_cast<unsigned int *>(base);
    k.keys = reinterpret_cast<unsigned long long *>(base + ((static_cast<size_t>(P + 1) * 4 + 255) & ~size_t(255)));
    k.rows = reinterpret_cast<unsigned int *>(k.keys + rows);
    return k;
}

constexpr int kLcWords = 1024;  // linear-counting bitmap of phase 2a: 32,768 bits

// Phase 2a (one workgroup per batch; batches that phase 1 finished return): splits the batch's rows into
// partitions -- one pass builds an LDS histogram over the launched partitions (Pmax) and a linear-counting
// bitmap of the keys; the distinct estimate picks the partitions used, P <= Pmax, so that a partition holds
// at most 40% of the table's slots (P divides Pmax and partitions are the low hash bits, so P's histogram
// is Pmax's folded); a second pass writes each row's key and row number to its bucket.  The workgroup runs on
// the XCD of the batch's phase-2b items, so the buckets are read back from that XCD's L2.
template <int BLOCK>
__device__ __forceinline__ void BatchSplit(const BatchConfig &c, int b, unsigned int *hist, unsigned int *bits,
                                           unsigned int *sh)
{
    if (c.spilled[b] == 0u) return;
    const int Pmax = c.parts[b];
    const int64_t start = c.offsets[b], end = c.offsets[b + 1];
    for (int i = threadIdx.x; i < Pmax; i += BLOCK) hist[i] = 0;
    for (int i = threadIdx.x; i < kLcWords; i += BLOCK) bits[i] = 0;
    if (threadIdx.x == 0) sh[0] = 0;
    __syncthreads();
    for (int64_t row = start + threadIdx.x; row < end; row += BLOCK) {
        const bool isNull = NullBit(c.kd.nulls[0], row);
        const unsigned long long key = isNull ? 0ULL : LoadWord(c.kd.values[0], c.kd.width[0], row);
        atomicAdd(&hist[PartOfKey(key, isNull, Pmax)], 1u);
        const unsigned int bit = static_cast<unsigned int>(Mix64(key ^ (isNull ? 0x5bd1e995ULL : 0ULL)) >> 20) &
                                 (kLcWords * 32 - 1);
        atomicOr(&bits[bit >> 5], 1u << (bit & 31));
    }
    __syncthreads();
    unsigned int ones = 0;
    for (int i = threadIdx.x; i < kLcWords; i += BLOCK) ones += __popc(bits[i]);
    if (ones) atomicAdd(&sh[0], ones);
    __syncthreads();
    if (threadIdx.x == 0) {
        const double m = kLcWords * 32.0;
        const double zeros = m - static_cast<double>(sh[0]);
        const double rows = static_cast<double>(end - start);
        const double est = zeros >= 1.0 ? fmin(rows, -m * log(zeros / m)) : rows;
        int P = 1;
        while (P < Pmax && static_cast<double>(P) * 0.4 * c.L2.S < est) P <<= 1;
        c.parts_used[b] = P;
        const Buckets k = BucketsOf(c, b, P, end - start);
        unsigned int run = 0;  // fold Pmax's histogram onto P partitions, then scan (P <= kMaxParts: short)
        for (int i = 0; i < P; ++i) {
            unsigned int v = 0;
            for (int j = i; j < Pmax; j += P) v += hist[j];
            hist[i] = run;
            k.start[i] = run;
            run += v;
        }
        k.start[P] = run;
        sh[1] = static_cast<unsigned int>(P);
    }
    __syncthreads();
    const int P = static_cast<int>(sh[1]);
    const Buckets k = BucketsOf(c, b, P, end - start);
    for (int64_t row = start + threadIdx.x; row < end; row += BLOCK) {
        const bool isNull = NullBit(c.kd.nulls[0], row);
        const unsigned long long key = isNull ? 0ULL : LoadWord(c.kd.values[0], c.kd.width[0], row);
        const unsigned int at = atomicAdd(&hist[PartOfKey(key, isNull, P)], 1u);
        k.keys[at] = key;
        k.rows[at] = static_cast<unsigned int>(row - start);
    }
}

// Phase 2b: work item `it` = (batch b, partition p of P) aggregates bucket p of batch b in an LDS table of
// L2.S slots (probing the whole table), then reserves its output rows with one atomic on groups[b] and
// emits.  A partition whose keys do not fit flags the batch for phase 3 and emits nothing.
template <int BLOCK>
__device__ __forceinline__ void BatchPart(const BatchConfig &c, int64_t it, uint8_t *lds, unsigned int *sh)
{
    const int4 item = c.items[it];
    const int b = item.x;
    if (b < 0 || c.spilled[b] == 0u) return;
    const int p = item.y, P = c.parts_used[b];
    if (p >= P) return;  // phase 2a chose fewer partitions than launched
    const LdsLayout &L = c.L2;
    const int S = L.S;
    const int64_t start = c.offsets[b], end = c.offsets[b + 1];
    const Buckets k = BucketsOf(c, b, P, end - start);
    unsigned long long *lkeys = reinterpret_cast<unsigned long long *>(lds + L.keys_off);
    for (int i = threadIdx.x; i < S + 2; i += BLOCK) {
        lkeys[i] = kEmptyKey;
        InitSlot(c.args, LdsRef{lds, L, i});
    }
    if (threadIdx.x == 0) {
        sh[0] = 0;  // a key found no slot
        sh[1] = 0;  // groups emitted
        sh[2] = 0;  // occupied slots
    }
    __syncthreads();
    const unsigned int e0 = k.start[p], e1 = k.start[p + 1];
    for (unsigned int e = e0 + threadIdx.x; e < e1; e += BLOCK) {
        const unsigned long long key = k.keys[e];
        const int64_t row = start + k.rows[e];
        const bool isNull = NullBit(c.kd.nulls[0], row);
        int slot = -1;
        if (isNull) {
            slot = S;
            lkeys[S] = 0;
        } else if (key == kEmptyKey) {
            slot = S + 1;
            lkeys[S + 1] = 0;
        } else {
            unsigned long long h = Mix64(key);
            unsigned int pos = static_cast<unsigned int>(((h >> 32) * static_cast<unsigned long long>(S)) >> 32);
            for (int step = 0; step < S; ++step) {
                unsigned long long kk = lkeys[pos];
                if (kk == kEmptyKey) {
                    kk = atomicCAS(&lkeys[pos], kEmptyKey, key);
                    if (kk == kEmptyKey) kk = key;
                }
                if (kk == key) {
                    slot = static_cast<int>(pos);
                    break;
                }
                pos = pos + 1 == static_cast<unsigned int>(S) ? 0 : pos + 1;
            }
        }
        if (slot < 0) {
            sh[0] = 1u;
            break;
        }
        UpdateAggsT(c.args, LdsRef{lds, L, slot}, row);
    }
    __syncthreads();
    if (sh[0]) {
        if (threadIdx.x == 0) {
            c.spilled[b] = 2u;
            atomicOr(c.redo, 1u);
        }
        return;
    }
    unsigned int mine = 0;
    for (int i = threadIdx.x; i < S + 2; i += BLOCK) mine += lkeys[i] != kEmptyKey ? 1u : 0u;
    if (mine) atomicAdd(&sh[2], mine);
    __syncthreads();
    if (threadIdx.x == 0) sh[3] = sh[2] ? static_cast<unsigned int>(atomicAdd(&c.groups[b], static_cast<unsigned long long>(sh[2]))) : 0u;
    __syncthreads();
    EmitTable<BLOCK>(c, lkeys, S, start + sh[3], &sh[1], [&](int i) { return LdsRef{lds, L, i}; });
}

// Per-batch sketch (before phase 1): the distinct keys among the batch's first kSketchRows rows, by linear
// counting in a 4,096-bit LDS bitmap, give an estimate of the batch's distinct keys (the d solving
// d (1 - exp(-m / d)) = u for m sampled rows with u distinct); a batch estimated above the phase-1 table's
// slots starts at phase 2.  Batches with at most L.S rows always fit and are not sketched.  A wrong estimate
// only moves work between phases; results do not depend on it.
constexpr int kSketchRows = 1024;
constexpr int kSketchWords = 128;

template <int BLOCK>
__device__ __forceinline__ void BatchSketch(const BatchConfig &c, int b, unsigned int *bits, unsigned int *sh)
{
    const int64_t start = c.offsets[b], end = c.offsets[b + 1];
    const int S = c.L.S;
    if (end - start <= S) {
        if (threadIdx.x == 0) c.sketch_modes[b] = 0;
        return;
    }
    for (int i = threadIdx.x; i < kSketchWords; i += BLOCK) bits[i] = 0;
    if (threadIdx.x == 0) sh[0] = 0;
    __syncthreads();
    const int64_t m = end - start < kSketchRows ? end - start : kSketchRows;
    for (int64_t r = threadIdx.x; r < m; r += BLOCK) {
        const int64_t row = start + r;
        const bool isNull = NullBit(c.kd.nulls[0], row);
        const unsigned long long key = isNull ? 0ULL : LoadWord(c.kd.values[0], c.kd.width[0], row);
        const unsigned int bit = static_cast<unsigned int>(Mix64(key) >> 40) & (kSketchWords * 32 - 1);
        atomicOr(&bits[bit >> 5], 1u << (bit & 31));
    }
    __syncthreads();
    unsigned int ones = 0;
    for (int i = threadIdx.x; i < kSketchWords; i += BLOCK) ones += __popc(bits[i]);
    if (ones) atomicAdd(&sh[0], ones);
    __syncthreads();
    if (threadIdx.x == 0) {
        const double nb = kSketchWords * 32.0;
        const double zeros = nb - static_cast<double>(sh[0]);
        const double u = zeros >= 1.0 ? -nb * log(zeros / nb) : nb;  // distinct among the sampled rows
        const double md = static_cast<double>(m);
        double d = u;  // solve d (1 - exp(-m / d)) = u by bisection; u close to m means "far more than m"
        if (u >= 0.97 * md) {
            d = 1e18;
        } else if (u > 1.0) {
            double lo = u, hi = 1e9;
            for (int it = 0; it < 60; ++it) {
                const double mid = 0.5 * (lo + hi);
                if (mid * (1.0 - exp(-md / mid)) < u) lo = mid;
                else hi = mid;
            }
            d = lo;
        }
        c.sketch_modes[b] = d > static_cast<double>(S) ? 2 : 0;
    }
}

// Linear-probing find-or-insert of a key into a key array with probe slots [0, mask].
__device__ __forceinline__ int TableSlot(unsigned long long *keys, unsigned long long key, unsigned int mask)
{
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

constexpr int kHbmTile = 4096;  // rows (and table slots) per phase-3 tile

// Batch owning phase-3 tile `t` (binary search over the tile prefix of all batches).
__device__ __forceinline__ int TileBatch(const BatchConfig &c, int64_t t)
{
    int lo = 0, hi = c.nbatches - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (c.tile_start[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Phase 3a: tile t of a flagged batch initialises its share of the batch's table slots (and side slots).
template <int BLOCK>
__device__ __forceinline__ void BatchHbmInit(const BatchConfig &c, int64_t t)
{
    const int b = TileBatch(c, t);
    if (c.spilled[b] != static_cast<unsigned int>(c.hbm_flag)) return;
    const int64_t start = c.offsets[b], end = c.offsets[b + 1];
    const int S = RegionSlots(end - start);
    const long long stride = RegionStride(end - start);
    const int64_t ntiles = c.tile_start[b + 1] - c.tile_start[b];
    const int64_t k = t - c.tile_start[b];
    const int64_t span = (static_cast<int64_t>(S) + 2 + ntiles - 1) / ntiles;
    const int64_t s0 = k * span, s1 = s0 + span < S + 2 ? s0 + span : S + 2;
    uint8_t *base = c.region_pool + c.region_off[b];
    unsigned long long *keys = reinterpret_cast<unsigned long long *>(base);
    for (int64_t i = s0 + threadIdx.x; i < s1; i += BLOCK) {
        keys[i] = kEmptyKey;
        InitSlot(c.args, ScaledRef{base, c.R, stride, static_cast<int>(i)});
    }
}

// Phase 3b: tile t of a flagged batch inserts its kHbmTile rows (load factor <= 1/2: every probe ends).
template <int BLOCK>
__device__ __forceinline__ void BatchHbmInsert(const BatchConfig &c, int64_t t)
{
    const int b = TileBatch(c, t);
    if (c.spilled[b] != static_cast<unsigned int>(c.hbm_flag)) return;
    const int64_t start = c.offsets[b], end = c.offsets[b + 1];
    const int S = RegionSlots(end - start);
    const long long stride = RegionStride(end - start);
    const int64_t r0 = start + (t - c.tile_start[b]) * kHbmTile;
    const int64_t r1 = r0 + kHbmTile < end ? r0 + kHbmTile : end;
    uint8_t *base = c.region_pool + c.region_off[b];
    unsigned long long *keys = reinterpret_cast<unsigned long long *>(base);
    for (int64_t row = r0 + threadIdx.x; row < r1; row += BLOCK) {
        bool isNull = NullBit(c.kd.nulls[0], row);
        unsigned long long key = isNull ? 0ULL : LoadWord(c.kd.values[0], c.kd.width[0], row);
        int slot;
        if (isNull) {
            slot = S;
            keys[S] = 0;  // marks the side slot used
        } else if (key == kEmptyKey) {
            slot = S + 1;
            keys[S + 1] = 0;
        } else {
            slot = TableSlot(keys, key, static_cast<unsigned int>(S - 1));
        }
        UpdateAggsT(c.args, ScaledRef{base, c.R, stride, slot}, row);
    }
}

// Phase 3c: emits the groups of flagged batch b (a later launch, so every insert is visible).
template <int BLOCK>
__device__ __forceinline__ void BatchHbmEmit(const BatchConfig &c, int b, unsigned int *sh)
{
    if (c.spilled[b] != static_cast<unsigned int>(c.hbm_flag)) return;
    const int64_t start = c.offsets[b], end = c.offsets[b + 1];
    const int S = RegionSlots(end - start);
    const long long stride = RegionStride(end - start);
    uint8_t *base = c.region_pool + c.region_off[b];
    const unsigned long long *keys = reinterpret_cast<const unsigned long long *>(base);
    if (threadIdx.x == 0) sh[1] = 0;
    __syncthreads();
    EmitTable<BLOCK, true>(c, keys, S, start, &sh[1], [&](int i) { return ScaledRef{base, c.R, stride, i}; });
    __syncthreads();
    if (threadIdx.x == 0) c.groups[b] = sh[1];
}

// Region offsets of the flagged batches: exclusive scan of their table bytes (one workgroup of BLOCK threads).
template <int BLOCK>
__device__ __forceinline__ void ScanRegions(const BatchConfig &c, unsigned long long *wsum)
{
    const int nb = c.nbatches;
    const int per = (nb + BLOCK - 1) / BLOCK;
    const int lo = threadIdx.x * per;
    const int hi = lo + per < nb ? lo + per : nb;
    unsigned long long sum = 0;
#pragma unroll 8
    for (int b = lo; b < hi; ++b) {
        const unsigned long long r = RegionBytes(c.R, c.offsets[b + 1] - c.offsets[b]);
        sum += c.spilled[b] ? r : 0ULL;
    }
    // exclusive scan of the per-thread sums: within each wave by shuffles, then over the waves' totals
    const int lane = LaneId(), wave = WaveInBlock();
    unsigned long long incl = sum;
#pragma unroll
    for (int d = 1; d < kWaveSize; d <<= 1) {
        const unsigned long long up = __shfl_up(incl, d, kWaveSize);
        if (lane >= d) incl += up;
    }
    if (lane == kWaveSize - 1) wsum[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long run = 0;
        for (int w = 0; w < BLOCK / kWaveSize; ++w) {
            const unsigned long long v = wsum[w];
            wsum[w] = run;
            run += v;
        }
    }
    __syncthreads();
    unsigned long long off = wsum[wave] + incl - sum;
#pragma unroll 8
    for (int b = lo; b < hi; ++b) {
        c.region_off[b] = off;
        const unsigned long long r = RegionBytes(c.R, c.offsets[b + 1] - c.offsets[b]);
        off += c.spilled[b] ? r : 0ULL;
    }
}

}  // namespace omni_gpu
