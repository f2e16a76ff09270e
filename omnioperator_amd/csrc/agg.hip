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
//   per group, indexed by a dense group id handed out in insertion order:
//     gkeys[G * W], gnull[G] (single-word only), and one structure-of-arrays state block per aggregate
//
// Results are independent of the hash function and of insertion order for every integer aggregate
// (wrapping SUM, COUNT, MIN, MAX and exact 192-bit decimal accumulation); double SUM/AVG depend on the
// order of atomic additions and are held to a relative tolerance.
#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "agg_layout.h"

namespace omni_gpu {

// ---- kernels ------------------------------------------------------------------------------------------------

__global__ void __launch_bounds__(256) agg_insert_single(KeyDev kd, int64_t rows, Slot1 *slots,
                                                         unsigned long long mask, GroupCtx g, AggArgs args)
{
    int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (row >= rows) return;
    unsigned int gid;
    if (NullBit(kd.nulls[0], row)) {
        gid = SideGroup(g, 0, 0ULL, true, 1);
    } else {
        unsigned long long key = LoadWord(kd.values[0], kd.width[0], row);
        gid = key == kEmptyKey ? SideGroup(g, 1, key, false, 1) : InsertSingle(slots, mask, key, g);
    }
    if (gid == 0xFFFFFFFFu) return;
    UpdateAggs(args, gid, row);
}

__global__ void __launch_bounds__(256) agg_insert_multi(KeyDev kd, int64_t rows, unsigned int *ctrl,
                                                        unsigned long long *skeys, unsigned long long mask,
                                                        GroupCtx g, AggArgs args)
{
    int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (row >= rows) return;
    unsigned long long words[kMaxWords];
    EncodeKey(kd, row, words);
    unsigned int gid = InsertMulti(ctrl, skeys, mask, words, kd.W, HashWords(words, kd.W), g);
    if (gid == 0xFFFFFFFFu) return;
    UpdateAggs(args, gid, row);
}

constexpr int kLdsBlock = 1024;
constexpr int kLdsMaxProbe = 32;

// LDS-local partial aggregation (single-word keys).  Each workgroup aggregates a contiguous chunk of rows
// in a private LDS table of L.S slots (linear probing, 64-bit LDS CAS), then merges every occupied slot into
// the global table with one global insert and one atomic per aggregate.  Rows whose key is NULL or the
// slot sentinel, or that find no LDS slot within kLdsMaxProbe steps, go straight to the global table.
__global__ void __launch_bounds__(kLdsBlock) agg_insert_lds(KeyDev kd, int64_t rows, int64_t rows_per_block,
                                                             Slot1 *slots, unsigned long long mask, GroupCtx g,
                                                             AggArgs args, LdsLayout L,
                                                             unsigned long long *spill_counter)
{
    extern __shared__ unsigned long long lds_words[];
    uint8_t *base = reinterpret_cast<uint8_t *>(lds_words);
    unsigned long long *lkeys = reinterpret_cast<unsigned long long *>(base + L.keys_off);
    const int S = L.S;
    for (int i = threadIdx.x; i < S; i += blockDim.x) {
        lkeys[i] = kEmptyKey;
        for (int a = 0; a < args.n; ++a) {
            int kind = args.a[a].kind;
            if (Is32(kind)) {
                reinterpret_cast<int *>(base + L.off[a][0])[i] = static_cast<int>(InitWord(kind));
            } else {
                reinterpret_cast<unsigned long long *>(base + L.off[a][0])[i] = InitWord(kind);
            }
            if (L.off[a][1] >= 0) reinterpret_cast<unsigned long long *>(base + L.off[a][1])[i] = 0;
            if (L.off[a][2] >= 0) reinterpret_cast<unsigned long long *>(base + L.off[a][2])[i] = 0;
            if (L.flag_off[a] >= 0) base[L.flag_off[a] + i] = 0;
        }
    }
    __syncthreads();
    int64_t start = static_cast<int64_t>(blockIdx.x) * rows_per_block;
    int64_t end = start + rows_per_block < rows ? start + rows_per_block : rows;
    unsigned int spilled = 0;
    for (int64_t row = start + threadIdx.x; row < end; row += blockDim.x) {
        bool isNull = NullBit(kd.nulls[0], row);
        unsigned long long key = isNull ? 0ULL : LoadWord(kd.values[0], kd.width[0], row);
        int slot = -1;
        if (!isNull && key != kEmptyKey) {
            unsigned long long h = Mix64(key);
            unsigned int pos = static_cast<unsigned int>(((h >> 32) * static_cast<unsigned long long>(S)) >> 32);
            for (int step = 0; step < kLdsMaxProbe; ++step) {
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
            UpdateAggsT(args, LdsRef{base, L, slot}, row);
        } else {
            ++spilled;
            unsigned int gid = isNull ? SideGroup(g, 0, 0ULL, true, 1)
                                      : (key == kEmptyKey ? SideGroup(g, 1, key, false, 1) : InsertSingle(slots, mask, key, g));
            if (gid != 0xFFFFFFFFu) UpdateAggs(args, gid, row);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < S; i += blockDim.x) {
        unsigned long long k = lkeys[i];
        if (k == kEmptyKey) continue;
        unsigned int gid = InsertSingle(slots, mask, k, g);
        if (gid != 0xFFFFFFFFu) MergeAggs(args, gid, LdsRef{base, L, i});
    }
    if (spilled) atomicAdd(spill_counter, static_cast<unsigned long long>(spilled));
}

// Rehash after growth: re-inserts groups [0, n) with their existing ids (states do not move).
__global__ void agg_rehash_single(const unsigned long long *gkeys, const uint8_t *gnull, int64_t n, Slot1 *slots,
                                  unsigned long long mask, const unsigned int *special)
{
    int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= n) return;
    if (special[0] == gid + 2 || special[1] == gid + 2) return;  // side groups are not in the table
    unsigned long long key = gkeys[gid];
    unsigned long long pos = Mix64(key) & mask;
    while (atomicCAS(&slots[pos].key, kEmptyKey, key) != kEmptyKey) pos = (pos + 1) & mask;
    slots[pos].gid1 = static_cast<unsigned int>(gid) + 1;
}

__global__ void agg_rehash_multi(const unsigned long long *gkeys, int64_t n, int W, unsigned int *ctrl,
                                 unsigned long long *skeys, unsigned long long mask)
{
    int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= n) return;
    unsigned long long words[kMaxWords];
    for (int w = 0; w < W; ++w) words[w] = gkeys[gid * W + w];
    unsigned long long pos = HashWords(words, W) & mask;
    while (atomicCAS(&ctrl[pos], 0u, static_cast<unsigned int>(gid) + 2) != 0u) pos = (pos + 1) & mask;
    for (int w = 0; w < W; ++w) skeys[pos * W + w] = words[w];
}

__global__ void fill_u64(unsigned long long *p, unsigned long long v, int64_t n)
{
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void fill_u32(unsigned int *p, unsigned int v, int64_t n)
{
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void init_slots1(Slot1 *s, int64_t n)
{
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        s[i].key = kEmptyKey;
        s[i].gid1 = 0;
        s[i].pad = 0;
    }
}

__global__ void __launch_bounds__(256) agg_output(int64_t start, int64_t count, KeyDev kd, int single,
                                                  const unsigned long long *gkeys, const uint8_t *gnull,
                                                  AggArgs args, OutArgs out, int overflow_as_null,
                                                  unsigned int *error_flag)
{
    int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    bool live = r < count;
    int64_t gid = start + r;
    // keys
    for (int c = 0; c < kd.num; ++c) {
        bool isNull = false;
        unsigned long long lo = 0, hi = 0;
        if (live) {
            if (single) {
                isNull = gnull[gid] != 0;
                lo = gkeys[gid];
            } else {
                const unsigned long long *w = gkeys + gid * kd.W;
                isNull = (w[0] >> c) & 1ULL;
                if (kd.width[c] == 16) {
                    lo = w[kd.word[c]];
                    hi = w[kd.word[c] + 1];
                } else if (kd.in_word0[c]) {
                    lo = w[0] >> 32;
                } else {
                    lo = w[kd.word[c]];
                }
            }
            if (isNull) lo = hi = 0;
            OutDev &o = out.o[c];
            if (o.width == 16) {
                static_cast<unsigned long long *>(o.values)[2 * r] = lo;
                static_cast<unsigned long long *>(o.values)[2 * r + 1] = hi;
            } else {
                StoreWidth(o.values, o.width, r, lo);
            }
        }
        WriteNullBits(out.o[c].nulls, r, count, live && isNull);
    }
    // aggregates
    for (int i = 0; i < args.n; ++i) {
        const AggDev &d = args.a[i];
        OutDev &o = out.o[kd.num + i];
        bool isNull = false;
        if (live) {
            isNull = FinalizeAgg(d, i, GlobalRef{args, static_cast<unsigned int>(gid)}, o, r, overflow_as_null,
                                 error_flag);
        }
        WriteNullBits(o.nulls, r, count, live && isNull);
    }
}

// ---- host side ----------------------------------------------------------------------------------------------

constexpr int kBlock = 256;
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
    DevBuf slots, ctrl, skeys, gkeys, gnull;
    DevBuf state[OMNI_GPU_MAX_AGGS][4];  // s0, s1, s2, flag
    DevBuf counters;  // [0] ngroups u64, [8] overflow u32, [12] error u32, [16..24) special u32[2], [24] spills u64
    int64_t known_groups = 0;
    int64_t pending_rows = 0;
    bool use_lds = false;
    LdsLayout lds{};
    int64_t lds_rows = 0;    // rows sent through the LDS kernel since the last spill check
    int64_t spills_seen = 0;

    unsigned long long *ngroups() const { return reinterpret_cast<unsigned long long *>(counters.as<uint8_t>()); }
    unsigned int *overflow() const { return reinterpret_cast<unsigned int *>(counters.as<uint8_t>() + 8); }
    unsigned int *error_flag() const { return reinterpret_cast<unsigned int *>(counters.as<uint8_t>() + 12); }
    unsigned int *special() const { return reinterpret_cast<unsigned int *>(counters.as<uint8_t>() + 16); }
    unsigned long long *spills() const { return reinterpret_cast<unsigned long long *>(counters.as<uint8_t>() + 24); }

    GroupCtx Ctx() const
    {
        GroupCtx g;
        g.ngroups = ngroups();
        g.max_groups = static_cast<unsigned long long>(max_groups);
        g.overflow = overflow();
        g.gkeys = gkeys.as<unsigned long long>();
        g.gnull = single ? gnull.as<uint8_t>() : nullptr;
        g.special = special();
        return g;
    }
};

int Fill64(hipStream_t s, void *p, unsigned long long v, int64_t n)
{
    if (n <= 0) return OMNI_GPU_OK;
    KernelScope ks(s, "fill_u64");
    fill_u64<<<Blocks(n), kBlock, 0, s>>>(static_cast<unsigned long long *>(p), v, n);
    return CheckLaunch("fill_u64");
}

int Fill32(hipStream_t s, void *p, unsigned int v, int64_t n)
{
    if (n <= 0) return OMNI_GPU_OK;
    KernelScope ks(s, "fill_u32");
    fill_u32<<<Blocks(n), kBlock, 0, s>>>(static_cast<unsigned int *>(p), v, n);
    return CheckLaunch("fill_u32");
}

// (Re)initialises the per-group state of groups [from, to).
int InitStates(AggHandle &h, int64_t from, int64_t to)
{
    int64_t n = to - from;
    if (n <= 0) return OMNI_GPU_OK;
    for (int i = 0; i < h.args.n; ++i) {
        AggDev &d = h.args.a[i];
        unsigned long long init = InitWord(d.kind);
        if (Is32(d.kind)) OG_TRY(Fill32(h.stream, static_cast<unsigned int *>(d.s0) + from, static_cast<unsigned int>(init), n));
        else OG_TRY(Fill64(h.stream, static_cast<unsigned long long *>(d.s0) + from, init, n));
        if (d.s1) OG_TRY(Fill64(h.stream, static_cast<unsigned long long *>(d.s1) + from, 0, n));
        if (d.s2) OG_TRY(Fill64(h.stream, static_cast<unsigned long long *>(d.s2) + from, 0, n));
        if (d.flag) OG_HIP(hipMemsetAsync(d.flag + from, 0, n, h.stream));
    }
    return OMNI_GPU_OK;
}

// Allocates a table of `cap` slots and group arrays for cap * kMaxLoad groups, keeping groups [0, keep).
int Resize(AggHandle &h, uint64_t cap, int64_t keep)
{
    hipStream_t s = h.stream;
    int64_t newMax = static_cast<int64_t>(static_cast<double>(cap) * kMaxLoad);
    int W = h.single ? 1 : h.kd.W;
    // group arrays (contents kept)
    OG_TRY(h.gkeys.Grow(sizeof(unsigned long long) * W * newMax, sizeof(unsigned long long) * W * keep, s));
    if (h.single) OG_TRY(h.gnull.Grow(newMax, keep, s));
    for (int i = 0; i < h.args.n; ++i) {
        AggDev &d = h.args.a[i];
        bool is32 = d.kind == K_MIN_I32 || d.kind == K_MAX_I32;
        size_t w0 = is32 ? 4 : 8;
        OG_TRY(h.state[i][0].Grow(w0 * newMax, w0 * keep, s));
        d.s0 = h.state[i][0].p;
        if (d.kind == K_AVG || d.kind == K_SUM_DEC) {
            OG_TRY(h.state[i][1].Grow(8 * newMax, 8 * keep, s));
            d.s1 = h.state[i][1].p;
        }
        if (d.kind == K_SUM_DEC) {
            OG_TRY(h.state[i][2].Grow(8 * newMax, 8 * keep, s));
            d.s2 = h.state[i][2].p;
        }
        if (d.kind != K_COUNT_ALL && d.kind != K_COUNT_COL && d.kind != K_AVG) {
            OG_TRY(h.state[i][3].Grow(newMax, keep, s));
            d.flag = h.state[i][3].as<uint8_t>();
        }
    }
    OG_TRY(InitStates(h, keep, newMax));
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
            agg_rehash_single<<<Blocks(keep), kBlock, 0, s>>>(h.gkeys.as<unsigned long long>(), h.gnull.as<uint8_t>(),
                                                               keep, h.slots.as<Slot1>(), mask, h.special());
            OG_TRY(CheckLaunch("agg_rehash_single"));
        }
    } else {
        OG_TRY(h.ctrl.Reserve(sizeof(unsigned int) * cap, s));
        OG_TRY(h.skeys.Reserve(sizeof(unsigned long long) * W * cap, s));
        OG_TRY(Fill32(s, h.ctrl.p, 0, static_cast<int64_t>(cap)));
        if (keep > 0) {
            KernelScope ks(s, "agg_rehash");
            agg_rehash_multi<<<Blocks(keep), kBlock, 0, s>>>(h.gkeys.as<unsigned long long>(), keep, W,
                                                              h.ctrl.as<unsigned int>(), h.skeys.as<unsigned long long>(), mask);
            OG_TRY(CheckLaunch("agg_rehash_multi"));
        }
    }
    h.cap = cap;
    h.max_groups = newMax;
    return OMNI_GPU_OK;
}

int SyncGroups(AggHandle &h)
{
    unsigned long long g = 0;
    unsigned int ovf = 0;
    OG_HIP(hipMemcpyAsync(&g, h.ngroups(), 8, hipMemcpyDeviceToHost, h.stream));
    OG_HIP(hipMemcpyAsync(&ovf, h.overflow(), 4, hipMemcpyDeviceToHost, h.stream));
    OG_HIP(hipStreamSynchronize(h.stream));
    OG_CHECK(ovf == 0, OMNI_GPU_ERR_STATE, "aggregation table overflowed its group arrays (internal sizing error)");
    h.known_groups = static_cast<int64_t>(g);
    h.pending_rows = 0;
    if (h.use_lds && h.spec.strategy == OMNI_GPU_STRATEGY_AUTO && h.lds_rows > 0) {
        // Adaptive fallback: when most rows miss the LDS table (too many groups per workgroup chunk),
        // later batches go straight to the global table.
        unsigned long long sp = 0;
        OG_HIP(hipMemcpy(&sp, h.spills(), 8, hipMemcpyDeviceToHost));
        int64_t delta = static_cast<int64_t>(sp) - h.spills_seen;
        h.spills_seen = static_cast<int64_t>(sp);
        if (delta * 2 > h.lds_rows) h.use_lds = false;
        h.lds_rows = 0;
    }
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
    for (int i = 0; i < spec->num_aggs; ++i) OG_TRY(ClassifyAgg(*spec, i, h->args.a[i]));
    OG_TRY(AcquireStream(&h->stream));
    OG_TRY(h->counters.Reserve(64, h->stream));
    OG_HIP(hipMemsetAsync(h->counters.p, 0, 64, h->stream));
    uint64_t want = spec->capacity_hint > 0 ? static_cast<uint64_t>(spec->capacity_hint) : 16384;
    uint64_t cap = NextPow2(static_cast<uint64_t>(static_cast<double>(want + 2) / kMaxLoad) + 1);
    cap = std::max<uint64_t>(cap, 1024);
    OG_TRY(Resize(*h, cap, 0));
    // Strategy: the LDS-local table serves single-word keys whose group count fits it (a capacity hint
    // above half the LDS slots, or a multi-word key, selects the global table directly).
    h->lds = MakeLdsLayout(h->args, kLdsBudget);
    bool ldsFits = h->single && h->lds.S >= 256;
    if (spec->strategy == OMNI_GPU_STRATEGY_LDS) {
        OG_CHECK(ldsFits, OMNI_GPU_ERR_INVALID, "the LDS strategy needs one key column of <= 8 bytes");
        h->use_lds = true;
    } else if (spec->strategy == OMNI_GPU_STRATEGY_AUTO) {
        h->use_lds = ldsFits && (spec->capacity_hint <= 0 || spec->capacity_hint * 2 <= h->lds.S);
    }
    *handle = h.release();
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
    // Capacity: groups after this batch <= known + pending + rows + 2 side groups.
    if (h.known_groups + h.pending_rows + rows + 2 > h.max_groups) {
        OG_TRY(SyncGroups(h));
        int64_t need = h.known_groups + rows + 2;
        if (need > h.max_groups) {
            uint64_t cap = h.cap;
            while (static_cast<int64_t>(static_cast<double>(cap) * kMaxLoad) < need) cap *= 4;
            OG_TRY(Resize(h, cap, h.known_groups));
        }
    }
    GroupCtx g = h.Ctx();
    if (h.use_lds) {
        // chunks of >= 8 rows per thread so that the per-workgroup merge stays small against the rows
        int64_t blocks = (rows + kLdsBlock * 8 - 1) / (kLdsBlock * 8);
        blocks = std::max<int64_t>(1, std::min<int64_t>(blocks, 4096));
        int64_t per = (rows + blocks - 1) / blocks;
        KernelScope ks(h.stream, "agg_insert_lds");
        agg_insert_lds<<<static_cast<unsigned int>(blocks), kLdsBlock, h.lds.bytes, h.stream>>>(
            h.kd, rows, per, h.slots.as<Slot1>(), h.cap - 1, g, h.args, h.lds, h.spills());
        OG_TRY(CheckLaunch("agg_insert_lds"));
        h.lds_rows += rows;
        h.pending_rows += rows;
        return OMNI_GPU_OK;
    }
    {
        KernelScope ks(h.stream, h.single ? "agg_insert_single" : "agg_insert_multi");
        if (h.single) {
            agg_insert_single<<<Blocks(rows), kBlock, 0, h.stream>>>(h.kd, rows, h.slots.as<Slot1>(), h.cap - 1, g, h.args);
        } else {
            agg_insert_multi<<<Blocks(rows), kBlock, 0, h.stream>>>(h.kd, rows, h.ctrl.as<unsigned int>(),
                                                                   h.skeys.as<unsigned long long>(), h.cap - 1, g, h.args);
        }
        OG_TRY(CheckLaunch("agg_insert"));
    }
    h.pending_rows += rows;
    return OMNI_GPU_OK;
}

int omni_gpu_agg_wait(void *handle)
{
    OG_CHECK(handle, OMNI_GPU_ERR_INVALID, "null handle");
    OG_HIP(hipStreamSynchronize(static_cast<AggHandle *>(handle)->stream));
    return OMNI_GPU_OK;
}

int omni_gpu_agg_num_groups(void *handle, int64_t *groups)
{
    OG_CHECK(handle && groups, OMNI_GPU_ERR_INVALID, "null argument");
    AggHandle &h = *static_cast<AggHandle *>(handle);
    OG_TRY(SyncGroups(h));
    *groups = h.known_groups;
    return OMNI_GPU_OK;
}

int omni_gpu_agg_get_output(void *handle, int64_t start, int64_t count, omni_gpu_out_column *outs)
{
    OG_CHECK(handle && outs, OMNI_GPU_ERR_INVALID, "null argument");
    AggHandle &h = *static_cast<AggHandle *>(handle);
    OG_TRY(SyncGroups(h));
    OG_CHECK(start >= 0 && count >= 0 && start + count <= h.known_groups, OMNI_GPU_ERR_INVALID,
             "output range [%lld, %lld) outside [0, %lld)", static_cast<long long>(start),
             static_cast<long long>(start + count), static_cast<long long>(h.known_groups));
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
                                                          h.gkeys.as<unsigned long long>(), h.gnull.as<uint8_t>(),
                                                          h.args, oa, h.spec.overflow_as_null, h.error_flag());
        OG_TRY(CheckLaunch("agg_output"));
    }
    unsigned int err = 0;
    OG_HIP(hipMemcpyAsync(&err, h.error_flag(), 4, hipMemcpyDeviceToHost, h.stream));
    OG_HIP(hipStreamSynchronize(h.stream));
    if (err) return SetError(OMNI_GPU_ERR_OVERFLOW, "decimal overflow in SUM (overflowAsNull is false)");
    return OMNI_GPU_OK;
}

int omni_gpu_agg_destroy(void *handle)
{
    if (!handle) return OMNI_GPU_OK;
    AggHandle *h = static_cast<AggHandle *>(handle);
    hipStream_t s = h->stream;
    h->slots.Release(s);
    h->ctrl.Release(s);
    h->skeys.Release(s);
    h->gkeys.Release(s);
    h->gnull.Release(s);
    for (auto &st : h->state)
        for (auto &b : st) b.Release(s);
    h->counters.Release(s);
    ReleaseStream(s);
    delete h;
    return OMNI_GPU_OK;
}

}  // extern "C"
