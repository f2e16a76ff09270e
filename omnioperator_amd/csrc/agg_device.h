// Device-side data structures and table operations of the hash aggregation, shared by the
// per-batch kernels (agg.hip) and the batched persistent kernel (batched.hip).
#pragma once

#include "omni_gpu_internal.h"

namespace omni_gpu {


enum AggKind : int {
    K_SUM_I64 = 0,
    K_SUM_F64,
    K_SUM_DEC,
    K_COUNT_ALL,
    K_COUNT_COL,
    K_AVG,
    K_MIN_I64,
    K_MAX_I64,
    K_MIN_I32,
    K_MAX_I32,
    K_MIN_F64,
    K_MAX_F64,
    K_SUM_DEC_CHK,  // decimal SUM whose running sum can leave its limits: DECIMAL64 -> DECIMAL64, DECIMAL128 input
};

// State of a checked decimal SUM (K_SUM_DEC_CHK).  The reference adds in row order with a checked add and the
// first add that leaves the limits marks the group OVERFLOWED for good (AddDecimalUseRowIndex / SumOp,
// sum_spark_decimal_aggregator.h:25-62, sum_aggregator.h:12-54), so the result depends on the order of the rows,
// not only on their total.  The GPU keeps, per group, the exact 192-bit sums of the positive inputs (P, words
// 0-2) and of the magnitudes of the negative inputs (N, words 3-5), plus a u32 word of flags:
//   - while P <= U and N <= L (U, L: the limits' magnitudes), every partial sum in ANY order lies in [-N, P], so
//     no add of the group overflowed and its value is exactly P - N;
//   - after each insert launch a group whose P or N passed its limit is CROSSED: the launch's rows of that group
//     are replayed in row order from the value before the launch (kDecOvf when an add leaves the limits), and
//     the value is stored back as P = max(v, 0), N = max(-v, 0).
constexpr unsigned int kDecOvf = 1u;      // a row-order add left the limits (sticky)
constexpr unsigned int kDecCrossed = 2u;  // P or N passed its limit during the last launch (being replayed)

struct AggDev {
    int kind;
    int in_width;     // bytes per input value
    int in_double;    // input column holds doubles
    int partial_out;  // partial output: AVG its running sum (0 when empty) instead of the mean; decimal SUM 0 when
                      // empty and NULL (never an error) when overflowed
    int out_width;    // bytes per output value
    const void *values;
    const uint8_t *nulls;
    int off[6];       // byte offsets of the state words s0..s5 within a group's state row (-1 unused)
    int flag_off;     // byte offset of the "a non-NULL value was added" flag (-1 unused)
    int cflag_off;    // byte offset of the u32 kDecOvf / kDecCrossed word of K_SUM_DEC_CHK (-1 unused)
    int dec_scale_up; // 10^(outScale - inScale) exponent for decimal SUM
    int dec_out128;   // decimal SUM output is DECIMAL128
};

struct AggArgs {
    int n;
    int row_bytes;        // bytes of one group's state row (a multiple of 8)
    uint8_t *states;      // state rows, indexed by group id (array of structures: one row per group)
    long long row_base;   // input position of row 0 of the launch (rows added before it): orders double MIN/MAX ties
    AggDev a[OMNI_GPU_MAX_AGGS];
};

struct KeyDev {
    int num;
    int W;                      // words per key (multi-word path)
    int width[OMNI_GPU_MAX_KEYS];
    int word[OMNI_GPU_MAX_KEYS];   // first word index of each column
    int in_word0[OMNI_GPU_MAX_KEYS];  // column stored in the high half of word 0
    const void *values[OMNI_GPU_MAX_KEYS];
    const uint8_t *nulls[OMNI_GPU_MAX_KEYS];
};

struct Slot1 {
    unsigned long long key;
    unsigned int gid1;
    unsigned int pad;
};

struct GroupCtx {
    unsigned long long *ngroups;   // group ids handed out so far (in chunks)
    unsigned long long max_groups; // capacity of the per-group arrays
    unsigned int *overflow;        // set when a group id would exceed max_groups (host sizing bug guard)
    unsigned long long *gkeys;
    uint8_t *gnull;
    uint8_t *gvalid;               // 1 once a group id holds a group (ids reserved but unused stay 0)
    unsigned int *special;         // [0] NULL group, [1] sentinel-key group (0 empty, 1 busy, gid+2)
    unsigned long long *wave;      // this wave's id chunk in LDS: [0] next id, [1] end of the chunk
    unsigned long long chunk;      // ids a wave takes from ngroups at a time
};

constexpr int kMaxWords = 6;

// Table words shared by all workgroups are read and written with agent-scope atomics.  On MI355X each of the
// eight XCDs has its own L2, so an agent-scope acquire invalidates the XCD's L2 and a release writes it back;
// the hot paths therefore use relaxed accesses, which go to the coherence point past the L2.
__device__ __forceinline__ unsigned int LoadRelaxed32(unsigned int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long LoadRelaxed64(unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void StoreRelaxed32(unsigned int *p, unsigned int v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void StoreRelaxed64(unsigned long long *p, unsigned long long v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Publishes `v` at *p after every earlier store of the calling wave has completed.  The key words a reader
// compares are written with agent-scope (write-through) stores, so waiting for their completion (vmcnt 0)
// before the flag store orders them for any later agent-scope reader without the agent-scope release fence,
// which would write back this XCD's whole L2 for every new group.
__device__ __forceinline__ void PublishAfterStores(unsigned int *p, unsigned int v)
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0); expcnt and lgkmcnt left at their maximum
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned long long LdsLoad(unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}
__device__ __forceinline__ void LdsStore(unsigned long long *p, unsigned long long v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

// Claims a fresh group id for every calling lane.  Each wave owns a chunk of ids [next, end) in LDS and takes a
// new chunk of g.chunk ids from the global counter only when the calling lanes need more than it holds, so the
// counter sees one atomic per chunk instead of one per wave.  The chunk cursor is read and written through
// relaxed atomics because lanes of the wave read what its leader wrote.  Returns ~0 if the per-group arrays are
// full (never expected: the host sizes them).
__device__ __forceinline__ unsigned int NewGroup(const GroupCtx &g)
{
    const unsigned long long act = __ballot(1);
    const unsigned int k = static_cast<unsigned int>(__popcll(act));
    const unsigned int r = static_cast<unsigned int>(__popcll(act & LaneMaskLt()));
    const int leader = __ffsll(static_cast<long long>(act)) - 1;
    const bool lead = LaneId() == leader;
    const unsigned long long next = LdsLoad(&g.wave[0]), end = LdsLoad(&g.wave[1]);
    const unsigned long long have = end - next;
    unsigned long long id;
    if (have >= k) {
        id = next + r;
        if (lead) LdsStore(&g.wave[0], next + k);
    } else {
        const unsigned long long need = k - have;
        const unsigned long long take = need > g.chunk ? need : g.chunk;
        unsigned long long base = 0;
        if (lead) base = atomicAdd(g.ngroups, take);
        base = __shfl(base, leader, 64);
        id = r < have ? next + r : base + (r - have);
        if (lead) {
            LdsStore(&g.wave[0], base + need);
            LdsStore(&g.wave[1], base + take);
        }
    }
    if (id >= g.max_groups) {
        atomicOr(g.overflow, 1u);
        return 0xFFFFFFFFu;
    }
    return static_cast<unsigned int>(id);
}

// Loads (at kernel start) and stores back (at kernel end) the id chunk of wave slot `slot`.  All lanes of the
// wave call both; between them the chunk lives in `lds` (two words per wave).
__device__ __forceinline__ void LoadWaveChunk(unsigned long long *lds, const unsigned long long *cursors, int slot)
{
    if (LaneId() == 0) {
        LdsStore(&lds[0], cursors[2 * slot]);
        LdsStore(&lds[1], cursors[2 * slot + 1]);
    }
}
__device__ __forceinline__ void StoreWaveChunk(const unsigned long long *lds, unsigned long long *cursors, int slot)
{
    if (LaneId() == 0) {
        cursors[2 * slot] = LdsLoad(const_cast<unsigned long long *>(&lds[0]));
        cursors[2 * slot + 1] = LdsLoad(const_cast<unsigned long long *>(&lds[1]));
    }
}

// Group of a cell (0 empty, 1 being created, gid + 2 published), created on first use: a side cell (NULL key,
// or the key equal to the slot sentinel) or a cell of the direct-index map.
static __device__ unsigned int CellGroup(const GroupCtx &g, unsigned int *cell, unsigned long long key, bool isNull,
                                         int W)
{
    while (true) {
        unsigned int c = LoadRelaxed32(cell);
        if (c >= 2) return c - 2;
        if (c == 0 && atomicCAS(cell, 0u, 1u) == 0u) {
            unsigned int gid = NewGroup(g);
            if (gid == 0xFFFFFFFFu) {
                StoreRelaxed32(cell, 0u);
                return gid;
            }
            g.gkeys[static_cast<size_t>(gid) * W] = key;
            if (g.gnull) g.gnull[gid] = isNull ? 1 : 0;
            g.gvalid[gid] = 1;
            StoreRelaxed32(cell, gid + 2);  // readers only need the id (states were initialised before)
            return gid;
        }
    }
}

static __device__ __forceinline__ unsigned int SideGroup(const GroupCtx &g, int which, unsigned long long key,
                                                         bool isNull, int W)
{
    return CellGroup(g, g.special + which, key, isNull, W);
}

// Direct-index group map of a single integer key (the ARRAY hash-table rule): key k in [min, min + range) owns
// cell k - min.  Keys outside the range use the hash table; both hand out ids from the same counter, so every
// group has one id whichever table found it.
struct ArrayMapDev {
    unsigned int *cells;        // nullptr: no map
    unsigned long long min;     // the range's first key, as a sign-extended word
    unsigned long long range;   // cells
};

// Find-or-insert of a single-word key; returns its group id.
static __device__ unsigned int InsertSingle(Slot1 *slots, unsigned long long mask, unsigned long long key,
                                     const GroupCtx &g)
{
    unsigned long long pos = Mix64(key) & mask;
    while (true) {
        unsigned long long k = slots[pos].key;
        if (k == kEmptyKey) {
            k = atomicCAS(&slots[pos].key, kEmptyKey, key);
            if (k == kEmptyKey) {
                unsigned int gid = NewGroup(g);
                if (gid == 0xFFFFFFFFu) {
                    StoreRelaxed32(&slots[pos].gid1, 0xFFFFFFFFu);  // lanes waiting on this slot give up too
                    return gid;
                }
                g.gkeys[gid] = key;
                if (g.gnull) g.gnull[gid] = 0;
                g.gvalid[gid] = 1;
                StoreRelaxed32(&slots[pos].gid1, gid + 1);  // readers only need the id
                return gid;
            }
        }
        if (k == key) {
            unsigned int g1 = LoadRelaxed32(&slots[pos].gid1);
            if (g1 == 0xFFFFFFFFu) return g1;  // the arrays were full when this key arrived
            if (g1) return g1 - 1;
            continue;  // the winning lane publishes its group id; re-read the same slot
        }
        pos = (pos + 1) & mask;
    }
}

// Group id of a non-NULL single-word key: its cell of the direct-index map when the key is in range, else the
// side cell of the slot sentinel or the hash table.
static __device__ __forceinline__ unsigned int KeyGroup(const ArrayMapDev &am, Slot1 *slots, unsigned long long mask,
                                                        unsigned long long key, const GroupCtx &g)
{
    if (am.cells != nullptr && key - am.min < am.range) return CellGroup(g, am.cells + (key - am.min), key, false, 1);
    return key == kEmptyKey ? SideGroup(g, 1, key, false, 1) : InsertSingle(slots, mask, key, g);
}

// Find-or-insert of a W-word key (ctrl-word protocol).
static __device__ unsigned int InsertMulti(unsigned int *ctrl, unsigned long long *skeys, unsigned long long mask,
                                    const unsigned long long *words, int W, unsigned long long hash,
                                    const GroupCtx &g)
{
    unsigned long long pos = hash & mask;
    while (true) {
        unsigned int c = LoadRelaxed32(&ctrl[pos]);
        if (c == 0) {
            c = atomicCAS(&ctrl[pos], 0u, 1u);
            if (c == 0) {
                unsigned int gid = NewGroup(g);
                if (gid == 0xFFFFFFFFu) {
                    StoreRelaxed32(&ctrl[pos], 0u);
                    return gid;
                }
                for (int w = 0; w < W; ++w) {
                    StoreRelaxed64(&skeys[pos * W + w], words[w]);
                    g.gkeys[static_cast<size_t>(gid) * W + w] = words[w];  // read by later kernels only
                }
                g.gvalid[gid] = 1;
                PublishAfterStores(&ctrl[pos], gid + 2);
                return gid;
            }
        }
        if (c == 1) continue;  // another lane is writing this slot's key
        bool same = true;
        for (int w = 0; w < W; ++w) same &= LoadRelaxed64(&skeys[pos * W + w]) == words[w];
        if (same) return c - 2;
        pos = (pos + 1) & mask;
    }
}

__device__ __forceinline__ void EncodeKey(const KeyDev &kd, int64_t row, unsigned long long *words)
{
    for (int w = 0; w < kd.W; ++w) words[w] = 0;
    for (int c = 0; c < kd.num; ++c) {
        if (NullBit(kd.nulls[c], row)) {
            words[0] |= 1ULL << c;
            continue;
        }
        if (kd.width[c] == 16) {
            const unsigned long long *v = static_cast<const unsigned long long *>(kd.values[c]);
            words[kd.word[c]] = v[2 * row];
            words[kd.word[c] + 1] = v[2 * row + 1];
        } else if (kd.in_word0[c]) {
            words[0] |= (LoadWord(kd.values[c], kd.width[c], row) & 0xFFFFFFFFULL) << 32;
        } else {
            words[kd.word[c]] = LoadWord(kd.values[c], kd.width[c], row);
        }
    }
}

__device__ __forceinline__ unsigned long long HashWords(const unsigned long long *words, int W)
{
    unsigned long long h = 0x9E3779B97F4A7C15ULL;
    for (int w = 0; w < W; ++w) h = Mix64(h ^ (words[w] + 0x632BE59BD9B4E019ULL * (w + 1)));
    return h;
}

// Exact 192-bit accumulation: adds top*2^128 + hi*2^64 + lo (hi, lo unsigned words) into the
// accumulator (s0 = lo, s1 = hi, s2 = top).  Every atomic's carry-out is derived from the value it
// returned, so concurrent additions in any order give the exact sum.
__device__ __forceinline__ void AddDec3(unsigned long long *lo, unsigned long long *hi, long long *top,
                                        unsigned long long vlo, unsigned long long vhi, long long vtop)
{
    unsigned long long old = atomicAdd(lo, vlo);
    unsigned long long c = (old + vlo) < old ? 1ULL : 0ULL;
    unsigned long long a = vhi + c;
    long long t = vtop + ((c && a == 0) ? 1 : 0);  // vhi == 2^64-1 plus a carry wraps into top
    if (a) {
        unsigned long long oh = atomicAdd(hi, a);
        t += (oh + a) < oh ? 1 : 0;
    }
    if (t) atomicAdd(reinterpret_cast<unsigned long long *>(top), static_cast<unsigned long long>(t));
}

// Signed 128-bit input value (lo, hi) of a decimal SUM as (lo, hi, top) words.
__device__ __forceinline__ void DecInput(const AggDev &d, int64_t row, unsigned long long &vlo,
                                         unsigned long long &vhi, long long &vtop)
{
    long long h;
    if (d.in_width == 16) {
        const unsigned long long *v = static_cast<const unsigned long long *>(d.values);
        vlo = v[2 * row];
        h = static_cast<long long>(v[2 * row + 1]);
    } else {
        long long v = static_cast<long long>(LoadWord(d.values, d.in_width, row));
        vlo = static_cast<unsigned long long>(v);
        h = v < 0 ? -1 : 0;
    }
    vhi = static_cast<unsigned long long>(h);
    vtop = h < 0 ? -1 : 0;
}

// ---- 192-bit helpers of the checked decimal SUM ---------------------------------------------------------------
struct U192 {
    unsigned long long lo, hi;
    long long top;
};

__device__ __forceinline__ U192 Neg192(U192 v)
{
    U192 r;
    r.lo = ~v.lo + 1ULL;
    unsigned long long c = r.lo == 0 ? 1ULL : 0ULL;
    r.hi = ~v.hi + c;
    c = (c && r.hi == 0) ? 1ULL : 0ULL;
    r.top = static_cast<long long>(~static_cast<unsigned long long>(v.top) + c);
    return r;
}

__device__ __forceinline__ U192 Add192(U192 a, U192 b)
{
    U192 r;
    r.lo = a.lo + b.lo;
    unsigned long long c = r.lo < a.lo ? 1ULL : 0ULL;
    unsigned long long h = a.hi + b.hi;
    unsigned long long c2 = h < a.hi ? 1ULL : 0ULL;
    r.hi = h + c;
    c2 += (r.hi < h) ? 1ULL : 0ULL;
    r.top = static_cast<long long>(static_cast<unsigned long long>(a.top) + static_cast<unsigned long long>(b.top) + c2);
    return r;
}

// a > b as signed 192-bit values
__device__ __forceinline__ bool Gt192(U192 a, U192 b)
{
    if (a.top != b.top) return a.top > b.top;
    if (a.hi != b.hi) return a.hi > b.hi;
    return a.lo > b.lo;
}

// Magnitude limit of a checked decimal SUM: 10^38 - 1 for a DECIMAL128 sum (AddCheckedOverflow,
// core/src/type/base_operations.h:20-28); INT64_MAX above zero and 2^63 below it for a DECIMAL64 sum
// (__builtin_add_overflow on int64).
__device__ __forceinline__ U192 DecUpper(int out128)
{
    return out128 ? U192{687399551400673279ULL, 5421010862427522170ULL, 0} : U192{0x7FFFFFFFFFFFFFFFULL, 0ULL, 0};
}
__device__ __forceinline__ U192 DecLowerMag(int out128)
{
    return out128 ? U192{687399551400673279ULL, 5421010862427522170ULL, 0} : U192{0x8000000000000000ULL, 0ULL, 0};
}

// Order-preserving image of a double as an unsigned word (IEEE order; -0.0 below +0.0); NaN is never stored.
__host__ __device__ __forceinline__ unsigned long long OrderedKey(unsigned long long bits)
{
    return (bits >> 63) ? ~bits : (bits | 0x8000000000000000ULL);
}
__device__ __forceinline__ unsigned long long FromOrderedKey(unsigned long long k)
{
    return (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFULL) : ~k;
}

// Per-group state addressing: GlobalRef indexes the HBM state arrays by group id, LdsRef the
// workgroup's LDS table by slot.  The same update code runs against either.
struct GlobalRef {
    static constexpr bool kGlobal = true;
    const AggArgs &args;
    unsigned int gid;
    __device__ __forceinline__ uint8_t *row() const
    {
        return args.states + static_cast<size_t>(gid) * static_cast<size_t>(args.row_bytes);
    }
    template <typename T> __device__ __forceinline__ T *p(int i, int w) const
    {
        return reinterpret_cast<T *>(row() + args.a[i].off[w]);
    }
    __device__ __forceinline__ uint8_t *flag(int i) const { return row() + args.a[i].flag_off; }
    __device__ __forceinline__ unsigned int *cflag(int i) const
    {
        return reinterpret_cast<unsigned int *>(row() + args.a[i].cflag_off);
    }
    __device__ __forceinline__ U192 word3(int i, int w) const
    {
        return U192{*p<unsigned long long>(i, w), *p<unsigned long long>(i, w + 1), *p<long long>(i, w + 2)};
    }
};

struct LdsLayout {
    int S;                              // LDS slots
    int keys_off;                       // byte offset of the u64 key array (W words per slot)
    int ctrl_off;                       // byte offset of the u32 slot states of multi-word keys (-1: one word)
    int off[OMNI_GPU_MAX_AGGS][3];      // byte offsets of each aggregate's state arrays (-1 unused)
    int flag_off[OMNI_GPU_MAX_AGGS];    // byte offsets of the flag arrays (-1 unused)
    int bytes;                          // total LDS bytes
};

struct LdsRef {
    static constexpr bool kGlobal = false;
    uint8_t *base;
    const LdsLayout &L;
    int slot;
    template <typename T> __device__ __forceinline__ T *p(int i, int w) const
    {
        return reinterpret_cast<T *>(base + L.off[i][w]) + slot;
    }
    __device__ __forceinline__ uint8_t *flag(int i) const { return base + L.flag_off[i] + slot; }
};

template <typename Ref> __device__ __forceinline__ void UpdateAggsT(const AggArgs &args, const Ref &r, int64_t row)
{
    typedef unsigned long long u64;
    for (int i = 0; i < args.n; ++i) {
        const AggDev &d = args.a[i];
        if (d.kind == K_COUNT_ALL) {
            atomicAdd(r.template p<u64>(i, 0), 1ULL);
            continue;
        }
        if (NullBit(d.nulls, row)) continue;
        switch (d.kind) {
            case K_COUNT_COL: atomicAdd(r.template p<u64>(i, 0), 1ULL); break;
            case K_SUM_I64:
                atomicAdd(r.template p<u64>(i, 0), LoadWord(d.values, d.in_width, row));
                *r.flag(i) = 1;
                break;
            case K_SUM_F64:
                atomicAdd(r.template p<double>(i, 0), static_cast<const double *>(d.values)[row]);
                *r.flag(i) = 1;
                break;
            case K_AVG: {
                double v = d.in_double ? static_cast<const double *>(d.values)[row]
                                       : static_cast<double>(static_cast<long long>(LoadWord(d.values, d.in_width, row)));
                atomicAdd(r.template p<double>(i, 0), v);
                atomicAdd(r.template p<u64>(i, 1), 1ULL);
                break;
            }
            case K_SUM_DEC: {
                u64 vlo, vhi;
                long long vtop;
                DecInput(d, row, vlo, vhi, vtop);
                AddDec3(r.template p<u64>(i, 0), r.template p<u64>(i, 1), r.template p<long long>(i, 2), vlo, vhi, vtop);
                *r.flag(i) = 1;
                break;
            }
            case K_MIN_I64:
                atomicMin(r.template p<long long>(i, 0), static_cast<long long>(LoadWord(d.values, d.in_width, row)));
                *r.flag(i) = 1;
                break;
            case K_MAX_I64:
                atomicMax(r.template p<long long>(i, 0), static_cast<long long>(LoadWord(d.values, d.in_width, row)));
                *r.flag(i) = 1;
                break;
            case K_MIN_I32:
                atomicMin(r.template p<int>(i, 0), static_cast<int>(LoadWord(d.values, d.in_width, row)));
                *r.flag(i) = 1;
                break;
            case K_MAX_I32:
                atomicMax(r.template p<int>(i, 0), static_cast<int>(LoadWord(d.values, d.in_width, row)));
                *r.flag(i) = 1;
                break;
            case K_MIN_F64:
            case K_MAX_F64: {
                // std::min / std::max over the rows in order (MaxOp, max_aggregator.h:39-44): a NaN input never
                // replaces the state, and among inputs equal to the result the first one stays -- which only shows
                // for zeros, so the first zero's sign is kept apart (its input position, sign in bit 0)
                const unsigned long long bits = static_cast<const unsigned long long *>(d.values)[row];
                *r.flag(i) = 1;
                const bool nan = (bits & 0x7FFFFFFFFFFFFFFFULL) > 0x7FF0000000000000ULL;
                if (nan) break;
                const bool zero = (bits & 0x7FFFFFFFFFFFFFFFULL) == 0;
                const unsigned long long key = OrderedKey(zero ? 0ULL : bits);
                if (d.kind == K_MIN_F64) atomicMin(r.template p<u64>(i, 0), key);
                else atomicMax(r.template p<u64>(i, 0), key);
                if (zero) {
                    const u64 pos = static_cast<u64>(args.row_base + row);
                    atomicMin(r.template p<u64>(i, 1), (pos << 1) | (bits >> 63));
                }
                break;
            }
            case K_SUM_DEC_CHK:
                if constexpr (Ref::kGlobal) {
                    u64 vlo, vhi;
                    long long vtop;
                    DecInput(d, row, vlo, vhi, vtop);
                    U192 v{vlo, vhi, vtop};
                    if (vtop < 0) {  // the magnitude of a negative input into N
                        v = Neg192(v);
                        AddDec3(r.template p<u64>(i, 3), r.template p<u64>(i, 4), r.template p<long long>(i, 5), v.lo,
                                v.hi, v.top);
                    } else {
                        AddDec3(r.template p<u64>(i, 0), r.template p<u64>(i, 1), r.template p<long long>(i, 2), v.lo,
                                v.hi, v.top);
                    }
                    *r.flag(i) = 1;
                }
                break;
            default: break;
        }
    }
}

__device__ __forceinline__ void UpdateAggs(const AggArgs &args, unsigned int gid, int64_t row)
{
    UpdateAggsT(args, GlobalRef{args, gid}, row);
}

// Merges one LDS partial state into the global group `gid`.
__device__ __forceinline__ void MergeAggs(const AggArgs &args, unsigned int gid, const LdsRef &l)
{
    typedef unsigned long long u64;
    GlobalRef g{args, gid};
    for (int i = 0; i < args.n; ++i) {
        const AggDev &d = args.a[i];
        switch (d.kind) {
            case K_COUNT_ALL:
            case K_COUNT_COL: {
                u64 v = *l.p<u64>(i, 0);
                if (v) atomicAdd(g.p<u64>(i, 0), v);
                break;
            }
            case K_AVG: {
                u64 c = *l.p<u64>(i, 1);
                if (c) {
                    atomicAdd(g.p<double>(i, 0), *l.p<double>(i, 0));
                    atomicAdd(g.p<u64>(i, 1), c);
                }
                break;
            }
            default: {
                if (!*l.flag(i)) break;
                *g.flag(i) = 1;
                switch (d.kind) {
                    case K_SUM_I64: atomicAdd(g.p<u64>(i, 0), *l.p<u64>(i, 0)); break;
                    case K_SUM_F64: atomicAdd(g.p<double>(i, 0), *l.p<double>(i, 0)); break;
                    case K_SUM_DEC:
                        AddDec3(g.p<u64>(i, 0), g.p<u64>(i, 1), g.p<long long>(i, 2), *l.p<u64>(i, 0), *l.p<u64>(i, 1),
                                *l.p<long long>(i, 2));
                        break;
                    case K_MIN_I64: atomicMin(g.p<long long>(i, 0), *l.p<long long>(i, 0)); break;
                    case K_MAX_I64: atomicMax(g.p<long long>(i, 0), *l.p<long long>(i, 0)); break;
                    case K_MIN_I32: atomicMin(g.p<int>(i, 0), *l.p<int>(i, 0)); break;
                    case K_MAX_I32: atomicMax(g.p<int>(i, 0), *l.p<int>(i, 0)); break;
                    case K_MIN_F64:
                    case K_MAX_F64: {
                        const u64 key = *l.p<u64>(i, 0), zpos = *l.p<u64>(i, 1);
                        if (d.kind == K_MIN_F64) atomicMin(g.p<u64>(i, 0), key);
                        else atomicMax(g.p<u64>(i, 0), key);
                        if (zpos != ~0ULL) atomicMin(g.p<u64>(i, 1), zpos);
                        break;
                    }
                    default: break;
                }
            }
        }
    }
}

// Initial value of an aggregate's first state word (MIN/MAX start at the type's extreme).
__host__ __device__ __forceinline__ unsigned long long InitWord(int kind)
{
    switch (kind) {
        case K_MIN_I64: return 0x7FFFFFFFFFFFFFFFULL;
        case K_MAX_I64: return 0x8000000000000000ULL;
        case K_MIN_I32: return 0x7FFFFFFFULL;
        case K_MAX_I32: return 0x80000000ULL;
        case K_MIN_F64: return OrderedKey(0x7FEFFFFFFFFFFFFFULL);  // DBL_MAX (MinAggregator's initial state)
        case K_MAX_F64: return OrderedKey(0xFFEFFFFFFFFFFFFFULL);  // lowest() = -DBL_MAX
        default: return 0;
    }
}

// Initial value of an aggregate's second state word (the first zero's position of double MIN/MAX: none yet).
__host__ __device__ __forceinline__ unsigned long long InitWord1(int kind)
{
    return (kind == K_MIN_F64 || kind == K_MAX_F64) ? ~0ULL : 0ULL;
}

// State words (8 bytes each) an aggregate keeps besides its first word in an LDS or per-slot table.
__host__ __device__ __forceinline__ int ExtraWords(int kind)
{
    switch (kind) {
        case K_AVG:
        case K_MIN_F64:
        case K_MAX_F64: return 1;
        case K_SUM_DEC: return 2;
        default: return 0;
    }
}

__host__ __device__ __forceinline__ bool Is32(int kind) { return kind == K_MIN_I32 || kind == K_MAX_I32; }

// Per-slot layout of an HBM table with the LDS-table structure (one u64 key array, then each aggregate's
// state arrays and flag bytes): offsets and `bytes` count bytes per slot.  In a table whose arrays each hold
// `stride` slots, array j starts at offset × stride, 256-byte aligned when stride is a multiple of 256.
inline LdsLayout UnitLayout(const AggArgs &args)
{
    LdsLayout U{};
    U.S = 1;
    int off = 8;  // keys at 0
    U.keys_off = 0;
    for (int i = 0; i < args.n; ++i) {
        int k = args.a[i].kind;
        U.off[i][0] = off;
        off += Is32(k) ? 4 : 8;
        U.off[i][1] = -1;
        U.off[i][2] = -1;
        U.flag_off[i] = -1;
        if (ExtraWords(k) >= 1) {
            U.off[i][1] = off;
            off += 8;
        }
        if (ExtraWords(k) >= 2) {
            U.off[i][2] = off;
            off += 8;
        }
        if (k != K_COUNT_ALL && k != K_COUNT_COL && k != K_AVG) {
            U.flag_off[i] = off;
            off += 1;
        }
    }
    U.bytes = off;
    return U;
}

// State addressing of one slot of a table laid out by UnitLayout with `stride` slots per array.
struct ScaledRef {
    static constexpr bool kGlobal = false;
    uint8_t *base;
    const LdsLayout &U;
    long long stride;
    int slot;
    template <typename T> __device__ __forceinline__ T *p(int i, int w) const
    {
        return reinterpret_cast<T *>(base + static_cast<long long>(U.off[i][w]) * stride) + slot;
    }
    __device__ __forceinline__ uint8_t *flag(int i) const
    {
        return base + static_cast<long long>(U.flag_off[i]) * stride + slot;
    }
};

struct OutDev {
    int width;
    void *values;
    uint8_t *nulls;
};

struct OutArgs {
    int nkeys;
    int naggs;
    OutDev o[OMNI_GPU_MAX_KEYS + OMNI_GPU_MAX_AGGS];
};

// Writes the null bitmap of 64 consecutive output rows handled by one wave.
__device__ __forceinline__ void WriteNullBits(uint8_t *nulls, int64_t r, int64_t count, bool isNull)
{
    unsigned long long m = __ballot(isNull);
    int lane = LaneId();
    int64_t base = r - lane;  // first row of this wave (waves cover aligned groups of 64 rows)
    if (lane < 8) {
        int64_t byte = (base >> 3) + lane;
        if (byte * 8 < count) nulls[byte] = static_cast<uint8_t>(m >> (8 * lane));
    }
}

__device__ __forceinline__ void StoreWidth(void *dst, int width, int64_t r, unsigned long long v)
{
    switch (width) {
        case 8: static_cast<unsigned long long *>(dst)[r] = v; break;
        case 4: static_cast<unsigned int *>(dst)[r] = static_cast<unsigned int>(v); break;
        case 2: static_cast<unsigned short *>(dst)[r] = static_cast<unsigned short>(v); break;
        default: static_cast<uint8_t *>(dst)[r] = static_cast<uint8_t>(v); break;
    }
}

// Decimal finalisation: the 192-bit total must lie within +-(10^38 - 1), then x 10^up checked against the same
// DECIMAL128 limits whatever the output type (MulCheckedOverflow in ExtractValuesBatchInternal,
// sum_spark_decimal_aggregator.h:246-301; a DECIMAL64 output keeps the low 64 bits of `res`).  `sum128` is the
// unchecked DECIMAL64 -> DECIMAL128 sum, whose total cannot leave those limits (|sum| <= rows * 2^63); the checked
// sums arrive here already inside their own limits.
static __device__ bool FinishDecimal(unsigned long long lo, unsigned long long hi, long long top, int up, bool sum128,
                                     __int128 &res)
{
    (void)sum128;
    bool fits128 = (top == 0 && static_cast<long long>(hi) >= 0) || (top == -1 && static_cast<long long>(hi) < 0);
    if (!fits128) return false;
    __int128 v = static_cast<__int128>((static_cast<unsigned __int128>(hi) << 64) | lo);
    const __int128 max38 = (static_cast<__int128>(5421010862427522170LL) << 64) + 687399551400673279LL;
    if (v > max38 || v < -max38) return false;
    for (int i = 0; i < up; ++i) {
        // v * 10 must stay within [-max38, max38]
        if (v > max38 / 10 || v < -(max38 / 10)) return false;
        v = (v << 3) + (v << 1);
    }
    res = v;
    return true;
}


// Writes aggregate i of one group (state addressed through `s`) into output row r of `o` and returns whether
// the value is NULL (ExtractValuesBatch of each aggregator: sum_flatim_aggregator.h:243-257,
// sum_spark_decimal_aggregator.h:246-301, max_aggregator.cpp:34-55, average_flatim_aggregator.h:275-290).
template <typename Ref>
__device__ __forceinline__ bool FinalizeAgg(const AggDev &d, int i, const Ref &s, const OutDev &o, int64_t r,
                                            int overflow_as_null, unsigned int *error_flag)
{
    typedef unsigned long long u64;
    bool isNull = false;
    switch (d.kind) {
        case K_COUNT_ALL:
        case K_COUNT_COL: static_cast<u64 *>(o.values)[r] = *s.template p<u64>(i, 0); break;
        case K_AVG: {
            u64 cnt = *s.template p<u64>(i, 1);
            if (d.partial_out) {  // partial state: the sum, 0 for an empty group (never NULL)
                static_cast<double *>(o.values)[r] = cnt == 0 ? 0.0 : *s.template p<double>(i, 0);
                break;
            }
            isNull = cnt == 0;
            static_cast<double *>(o.values)[r] = isNull ? 0.0 : *s.template p<double>(i, 0) / static_cast<double>(cnt);
            break;
        }
        case K_SUM_DEC: {
            // partial output (sum_spark_decimal_aggregator.h:257-296 with OUTPUT_PARTIAL): 0 for an empty group,
            // NULL for an overflowed one, never an error
            isNull = *s.flag(i) == 0 && !d.partial_out;
            __int128 v = 0;
            if (*s.flag(i) != 0) {
                bool ok = FinishDecimal(*s.template p<u64>(i, 0), *s.template p<u64>(i, 1),
                                        *s.template p<long long>(i, 2), d.dec_scale_up, d.dec_out128, v);
                if (!ok) {
                    isNull = true;
                    v = 0;
                    if (!overflow_as_null && !d.partial_out) atomicOr(error_flag, 1u);
                }
            }
            if (d.dec_out128) {
                static_cast<u64 *>(o.values)[2 * r] = static_cast<u64>(v);
                static_cast<u64 *>(o.values)[2 * r + 1] = static_cast<u64>(static_cast<unsigned __int128>(v) >> 64);
            } else {
                static_cast<long long *>(o.values)[r] = static_cast<long long>(v);
            }
            break;
        }
        case K_SUM_DEC_CHK: {
            isNull = *s.flag(i) == 0 && !d.partial_out;
            __int128 v = 0;
            if constexpr (Ref::kGlobal) {
                if (*s.flag(i) != 0) {
                    bool ok = (*s.cflag(i) & kDecOvf) == 0;
                    if (ok) {
                        // not overflowed: P <= U and N <= L, so P - N fits the output's int64 / int128 range
                        const U192 t = Add192(s.word3(i, 0), Neg192(s.word3(i, 3)));
                        ok = FinishDecimal(t.lo, t.hi, t.top, d.dec_scale_up, true, v);
                    }
                    if (!ok) {
                        isNull = true;
                        v = 0;
                        if (!overflow_as_null && !d.partial_out) atomicOr(error_flag, 1u);
                    }
                }
            }
            if (d.dec_out128) {
                static_cast<u64 *>(o.values)[2 * r] = static_cast<u64>(v);
                static_cast<u64 *>(o.values)[2 * r + 1] = static_cast<u64>(static_cast<unsigned __int128>(v) >> 64);
            } else {
                // SetValToVector: static_cast<int64_t> of the rescaled value (its low 64 bits)
                static_cast<long long *>(o.values)[r] = static_cast<long long>(v);
            }
            break;
        }
        case K_MIN_F64:
        case K_MAX_F64: {
            isNull = *s.flag(i) == 0;
            u64 bits = 0;
            if (!isNull) {
                bits = FromOrderedKey(*s.template p<u64>(i, 0));
                if ((bits & 0x7FFFFFFFFFFFFFFFULL) == 0) bits = (*s.template p<u64>(i, 1) & 1ULL) << 63;
            }
            static_cast<u64 *>(o.values)[r] = bits;
            break;
        }
        case K_MIN_I32:
        case K_MAX_I32: {
            isNull = *s.flag(i) == 0;
            int v = isNull ? 0 : *s.template p<int>(i, 0);
            StoreWidth(o.values, o.width, r, static_cast<u64>(static_cast<long long>(v)));
            break;
        }
        default:  // SUM_I64, SUM_F64, integer MIN/MAX 64-bit: the 8-byte state is the value
            isNull = *s.flag(i) == 0;
            static_cast<u64 *>(o.values)[r] = isNull ? 0ULL : *s.template p<u64>(i, 0);
            break;
    }
    return isNull;
}

}  // namespace omni_gpu
