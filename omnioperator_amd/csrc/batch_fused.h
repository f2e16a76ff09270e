// One workgroup per batch: the whole group-by of a VectorBatch in one workgroup, with no device-wide atomics.
//
// A batch whose distinct keys fit one LDS table (at most kFusedLoad of its slots) is aggregated straight from its
// columns ("direct").  Otherwise the workgroup
//   1. counts its rows per hash partition over kFusedParts partitions (and, without an advisor prediction,
//      estimates the batch's distinct keys by linear counting), picks P partitions so that a partition holds
//      about kFusedLoad of a table's slots, and scans the folded counts into bucket starts;
//   2. writes every row, in partition order, to the batch's own row positions of perm_keys and perm_vals (the
//      value word every aggregate reads) or perm_rows (the row number, when the aggregates read several columns
//      or NULLs), staging each chunk of rows in LDS so that the writes are coalesced;
//   3. aggregates partition after partition in its LDS table (sized to the partition's rows), emitting each
//      partition's groups to the batch's output rows through a counter in LDS; the records of the next partition
//      are loaded into registers while the current one is aggregated and emitted.
// With an advisor, a batch predicted to fill at most kDirectLoad of the table skips step 1 and goes direct (a
// prediction that is too low shows as a full table, and the batch then takes steps 1-3 with its own estimate);
// otherwise P follows the larger of the estimate and the prediction times kFusedMargin.  A partition whose keys overflow the largest table (adversarial
// hash skew or a far too low prediction) flags the batch for the exact HBM-table pass (spilled = 2), which
// re-aggregates it from its columns and overwrites its output rows.
#pragma once

#include "batches_device.h"

namespace omni_gpu {

constexpr int kFusedParts = 256;      // partitions counted in step 1 (P divides it)
constexpr float kFusedLoad = 0.45f;   // distinct keys per table slot aimed at
constexpr int kFusedAhead = 4;        // records per thread loaded ahead
constexpr int kFusedRows = 4;         // rows per thread whose loads are issued together in the row loops
constexpr float kFusedMargin = 1.25f; // an advisor prediction is scaled by this before it sizes partitions
constexpr float kDirectLoad = 0.6f;   // a batch with at most this share of the largest table's slots goes direct
constexpr int kLiteAggs = 4;          // aggregates of the lite table at most
enum LiteOp : int { kLiteCount = 0, kLiteSum = 1, kLiteMin = 2, kLiteMax = 3 };


// Probe slots of a table for at most `d` distinct keys at the target load (a multiple of 64, in [64, smax]).
__device__ __forceinline__ int FusedSlots(float d, int smax)
{
    int s = static_cast<int>(d / kFusedLoad) + 1;
    s = (s + 63) & ~63;
    return s < 64 ? 64 : (s > smax ? smax : s);
}

// Partitions for `est` distinct keys when one partition may hold `fit` of them (a power of two <= kFusedParts).
__device__ __forceinline__ int FusedParts(float est, int fit)
{
    int P = 1;
    while (P < kFusedParts && static_cast<float>(P) * static_cast<float>(fit) < est) P <<= 1;
    return P;
}

// Clears slots [0, S + 2) of an LDS table laid out by L (probe slots [0, S), side slots S and S + 1).
template <int BLOCK>
__device__ __forceinline__ void FusedClear(const AggArgs &args, uint8_t *lds, const LdsLayout &L, int S)
{
    unsigned long long *lkeys = reinterpret_cast<unsigned long long *>(lds + L.keys_off);
    for (int i = threadIdx.x; i < S + 2; i += BLOCK) {
        lkeys[i] = kEmptyKey;
        InitSlot(args, LdsRef{lds, L, i});
    }
}

// Find-or-insert of a key among probe slots [0, S) (linear probing over the whole range); the NULL key takes side
// slot S, the slot sentinel's value S + 1.  Returns -1 when every probe slot holds another key.
__device__ __forceinline__ int FusedSlot(unsigned long long *lkeys, int S, unsigned long long key, bool isNull)
{
    if (isNull) {
        lkeys[S] = 0;  // marks the side slot used
        return S;
    }
    if (key == kEmptyKey) {
        lkeys[S + 1] = 0;
        return S + 1;
    }
    unsigned int pos = static_cast<unsigned int>(((Mix64(key) >> 32) * static_cast<unsigned long long>(S)) >> 32);
    for (int step = 0; step < S; ++step) {
        unsigned long long k = lkeys[pos];
        if (k == kEmptyKey) {
            k = atomicCAS(&lkeys[pos], kEmptyKey, key);
            if (k == kEmptyKey) k = key;
        }
        if (k == key) return static_cast<int>(pos);
        pos = pos + 1 == static_cast<unsigned int>(S) ? 0 : pos + 1;
    }
    return -1;
}

// The aggregates' update from a carried value word v (the host selects the carried form only when every aggregate
// is COUNT(*) or one of these kinds reading the same column, with no NULLs, of at most 8 bytes; v is that column's
// value as LoadWord returns it: sign-extended integers, raw double bits).
template <typename Ref> __device__ __forceinline__ void UpdateAggsCarried(const AggArgs &args, const Ref &r, unsigned long long v)
{
    typedef unsigned long long u64;
    for (int i = 0; i < args.n; ++i) {
        const AggDev &d = args.a[i];
        switch (d.kind) {
            case K_COUNT_ALL:
            case K_COUNT_COL: atomicAdd(r.template p<u64>(i, 0), 1ULL); break;
            case K_SUM_I64:
                atomicAdd(r.template p<u64>(i, 0), v);
                *r.flag(i) = 1;
                break;
            case K_SUM_F64:
                atomicAdd(r.template p<double>(i, 0), __longlong_as_double(static_cast<long long>(v)));
                *r.flag(i) = 1;
                break;
            case K_AVG:
                atomicAdd(r.template p<double>(i, 0), d.in_double ? __longlong_as_double(static_cast<long long>(v))
                                                                  : static_cast<double>(static_cast<long long>(v)));
                atomicAdd(r.template p<u64>(i, 1), 1ULL);
                break;
            case K_SUM_DEC: {
                const long long s = static_cast<long long>(v);
                AddDec3(r.template p<u64>(i, 0), r.template p<u64>(i, 1), r.template p<long long>(i, 2), v,
                        s < 0 ? ~0ULL : 0ULL, s < 0 ? -1LL : 0LL);
                *r.flag(i) = 1;
                break;
            }
            case K_MIN_I64:
                atomicMin(r.template p<long long>(i, 0), static_cast<long long>(v));
                *r.flag(i) = 1;
                break;
            case K_MAX_I64:
                atomicMax(r.template p<long long>(i, 0), static_cast<long long>(v));
                *r.flag(i) = 1;
                break;
            case K_MIN_I32:
                atomicMin(r.template p<int>(i, 0), static_cast<int>(static_cast<long long>(v)));
                *r.flag(i) = 1;
                break;
            case K_MAX_I32:
                atomicMax(r.template p<int>(i, 0), static_cast<int>(static_cast<long long>(v)));
                *r.flag(i) = 1;
                break;
            default: break;
        }
    }
}

// In-place exclusive scan of v[0, n), n <= kFusedParts, by the whole workgroup; v[n] receives the total.
template <int BLOCK>
__device__ __forceinline__ void FusedScan(unsigned int *v, int n, unsigned int *wsum)
{
    constexpr int K = (kFusedParts + BLOCK - 1) / BLOCK;
    const int i0 = static_cast<int>(threadIdx.x) * K;
    unsigned int x[K];
    unsigned int mine = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        x[k] = i0 + k < n ? v[i0 + k] : 0u;
        mine += x[k];
    }
    const int lane = LaneId(), wave = WaveInBlock();
    unsigned int incl = mine;
#pragma unroll
    for (int d = 1; d < kWaveSize; d <<= 1) {
        const unsigned int up = __shfl_up(incl, d, kWaveSize);
        if (lane >= d) incl += up;
    }
    if (lane == kWaveSize - 1) wsum[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int run = 0;
        for (int w = 0; w < BLOCK / kWaveSize; ++w) {
            const unsigned int t = wsum[w];
            wsum[w] = run;
            run += t;
        }
        v[n] = run;
    }
    __syncthreads();
    unsigned int excl = wsum[wave] + incl - mine;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (i0 + k < n) v[i0 + k] = excl;
        excl += x[k];
    }
    __syncthreads();
}

// Loads and stores through global-address-space pointers: the pointers come from the configuration in memory, so
// the compiler would otherwise emit flat accesses, which also count against the LDS wait counter and make every
// wait for a load a wait for the LDS operations in flight as well.
template <typename T> __device__ __forceinline__ T LdG(const T *p, int64_t i)
{
    return ((const __attribute__((address_space(1))) T *)(p))[i];
}
template <typename T> __device__ __forceinline__ void StG(T *p, int64_t i, T v)
{
    ((__attribute__((address_space(1))) T *)(p))[i] = v;
}

// Everything the workgroup reads of the configuration's columns, loaded into registers once (the configuration
// lives in global memory, and the workgroup's own global stores would otherwise have it read again).
struct FusedIn {
    const void *kvals;
    const uint8_t *knulls;
    int kwidth;
    int carry;
    const void *cvals;   // the carried column (nullptr: every aggregate is COUNT(*))
    int cwidth;
    unsigned long long *pkeys;
    unsigned long long *pvals;
    unsigned int *prows;
};

__device__ __forceinline__ FusedIn LoadFusedIn(const BatchConfig &c)
{
    FusedIn in;
    in.kvals = c.kd.values[0];
    in.knulls = c.kd.nulls[0];
    in.kwidth = c.kd.width[0];
    in.carry = c.carry;
    in.cvals = c.carry && c.carry_agg >= 0 ? c.args.a[c.carry_agg].values : nullptr;
    in.cwidth = c.carry_agg >= 0 ? c.args.a[c.carry_agg].in_width : 8;
    in.pkeys = c.perm_keys;
    in.pvals = c.perm_vals;
    in.prows = c.perm_rows;
    return in;
}

// The table of the generic form: any aggregate list, state laid out by c.L2, updated by row (UpdateAggsT) or by
// carried value (UpdateAggsCarried), emitted through EmitGroup.
struct GenericTable {
    const BatchConfig &c;
    uint8_t *lds;
    FusedIn in;
    __device__ __forceinline__ unsigned long long Key(int64_t row, bool &nul) const
    {
        nul = NullBit(in.knulls, row);
        return nul ? 0ULL : LoadWord(in.kvals, in.kwidth, row);
    }
    __device__ __forceinline__ unsigned long long Val(int64_t row) const
    {
        return in.cvals ? LoadWord(in.cvals, in.cwidth, row) : 0ULL;
    }
    __device__ __forceinline__ unsigned long long *Keys() const
    {
        return reinterpret_cast<unsigned long long *>(lds + c.L2.keys_off);
    }
    template <int BLOCK> __device__ __forceinline__ void Clear(int S) const { FusedClear<BLOCK>(c.args, lds, c.L2, S); }
    __device__ __forceinline__ void Update(int slot, int64_t row, unsigned long long v, int carry) const
    {
        if (carry) UpdateAggsCarried(c.args, LdsRef{lds, c.L2, slot}, v);
        else UpdateAggsT(c.args, LdsRef{lds, c.L2, slot}, row);
    }
    // Emits the used slots of [0, S + 2); with RESET, also returns each of them to its initial state (the
    // other slots are untouched since the last reset, so the whole table is clean afterwards).
    template <int BLOCK, bool RESET = false>
    __device__ __forceinline__ void Emit(int S, int64_t start, unsigned int *counter) const
    {
        unsigned long long *keys = Keys();
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
            EmitGroup(c, LdsRef{lds, c.L2, i}, side ? (i == S ? 0ULL : kEmptyKey) : k, i == S, start + r);
            if (RESET) {
                keys[i] = kEmptyKey;
                InitSlot(c.args, LdsRef{lds, c.L2, i});
            }
        }
    }
};

// The table of the lite form: NA <= kLiteAggs aggregates, each COUNT / SUM / MIN / MAX into one int64 state word,
// over a carried NULL-free column, with 8-byte outputs.  LDS: the u64 keys, then one u64 state array per
// aggregate, `stride` slots each; the aggregate codes and the output columns stay in registers.
template <int NA, int KW> struct LiteTable {
    unsigned long long *lkeys;
    int stride;
    int op[NA];
    const void *kvals;
    const unsigned long long *cvals;  // the carried int64 column (nullptr: every aggregate is COUNT(*))
    void *okey;
    unsigned long long *oval[NA];
    __device__ __forceinline__ unsigned long long Key(int64_t row, bool &nul) const
    {
        nul = false;
        if constexpr (KW == 8) return LdG(static_cast<const unsigned long long *>(kvals), row);
        else return static_cast<unsigned long long>(static_cast<long long>(LdG(static_cast<const int *>(kvals), row)));
    }
    __device__ __forceinline__ unsigned long long Val(int64_t row) const { return cvals ? LdG(cvals, row) : 0ULL; }
    __device__ __forceinline__ unsigned long long *Keys() const { return lkeys; }
    __device__ __forceinline__ unsigned long long Init(int a) const
    {
        return op[a] == kLiteMin ? 0x7FFFFFFFFFFFFFFFULL : op[a] == kLiteMax ? 0x8000000000000000ULL : 0ULL;
    }
    template <int BLOCK> __device__ __forceinline__ void Clear(int S) const
    {
        for (int i = threadIdx.x; i < S + 2; i += BLOCK) {
            lkeys[i] = kEmptyKey;
#pragma unroll
            for (int a = 0; a < NA; ++a) lkeys[(a + 1) * stride + i] = Init(a);
        }
    }
    __device__ __forceinline__ void Update(int slot, int64_t, unsigned long long v, int) const
    {
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            unsigned long long *w = lkeys + (a + 1) * stride + slot;
            switch (op[a]) {
                case kLiteCount: atomicAdd(w, 1ULL); break;
                case kLiteSum: atomicAdd(w, v); break;
                case kLiteMin: atomicMin(reinterpret_cast<long long *>(w), static_cast<long long>(v)); break;
                default: atomicMax(reinterpret_cast<long long *>(w), static_cast<long long>(v)); break;
            }
        }
    }
    template <int BLOCK, bool RESET = false>
    __device__ __forceinline__ void Emit(int S, int64_t start, unsigned int *counter) const
    {
        for (int base = threadIdx.x - LaneId(); base < S + 2; base += BLOCK) {
            const int i = base + LaneId();
            const unsigned long long k = i < S + 2 ? lkeys[i] : kEmptyKey;
            const bool used = k != kEmptyKey;
            const unsigned long long mask = __ballot(used);
            unsigned int first = 0;
            if (LaneId() == 0 && mask) first = atomicAdd(counter, static_cast<unsigned int>(__popcll(mask)));
            first = __shfl(first, 0, 64);
            if (!used) continue;
            const int64_t r = start + first + static_cast<unsigned int>(__popcll(mask & LaneMaskLt()));
            const unsigned long long kout = i == S + 1 ? kEmptyKey : k;  // slot S + 1: the sentinel's own key
            if constexpr (KW == 8) StG(static_cast<unsigned long long *>(okey), r, kout);
            else StG(static_cast<unsigned int *>(okey), r, static_cast<unsigned int>(kout));
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                unsigned long long *w = lkeys + (a + 1) * stride + i;
                StG(oval[a], r, *w);
                if (RESET) *w = Init(a);
            }
            if (RESET) lkeys[i] = kEmptyKey;
        }
    }
};

template <int NA, int KW>
__device__ __forceinline__ LiteTable<NA, KW> MakeLiteTable(const BatchConfig &c, uint8_t *lds)
{
    LiteTable<NA, KW> t;
    t.lkeys = reinterpret_cast<unsigned long long *>(lds);
    t.stride = c.lite_smax + 2;
    t.kvals = c.kd.values[0];
    t.cvals = c.carry_agg >= 0 ? static_cast<const unsigned long long *>(c.args.a[c.carry_agg].values) : nullptr;
    t.okey = c.out.o[0].values;
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        t.op[a] = c.lite_op[a];
        t.oval[a] = static_cast<unsigned long long *>(c.out.o[1 + a].values);
    }
    return t;
}

// Aggregates rows [start, end) straight from the batch's columns into a table of S probe slots and, unless a key
// found no slot, emits its groups.  Returns false when the table overflowed (nothing emitted).
template <int BLOCK, typename Table>
__device__ __forceinline__ bool FusedDirect(const BatchConfig &c, const FusedIn &in, const Table &t, int b,
                                            int64_t start, int64_t end, int S, unsigned int *sh)
{
    unsigned long long *lkeys = t.Keys();
    t.template Clear<BLOCK>(S);
    if (threadIdx.x == 0) {
        sh[0] = 0;
        sh[1] = 0;
    }
    __syncthreads();
    for (int64_t base = start + threadIdx.x; base < end; base += kFusedRows * BLOCK) {
        if (sh[0]) break;
        unsigned long long key[kFusedRows], val[kFusedRows];
        bool nul[kFusedRows];
#pragma unroll
        for (int u = 0; u < kFusedRows; ++u) {  // every load of the chunk in flight before the first insert
            const int64_t row = base + u * BLOCK;
            nul[u] = false;
            key[u] = 0;
            val[u] = 0;
            if (row < end) {
                key[u] = t.Key(row, nul[u]);
                if (in.carry) val[u] = t.Val(row);
            }
        }
#pragma unroll
        for (int u = 0; u < kFusedRows; ++u) {
            const int64_t row = base + u * BLOCK;
            if (row >= end) break;
            const int slot = FusedSlot(lkeys, S, key[u], nul[u]);
            if (slot < 0) {
                sh[0] = 1u;
                break;
            }
            t.Update(slot, row, val[u], in.carry);
        }
    }
    __syncthreads();
    const bool full = sh[0] != 0u;
    if (!full) {
        t.template Emit<BLOCK>(S, start, &sh[1]);
        __syncthreads();
        if (threadIdx.x == 0) {
            c.groups[b] = sh[1];
            c.spilled[b] = 0u;
        }
    }
    __syncthreads();  // every thread has read sh[0] before it is reused
    return !full;
}

// Rows per staging chunk of step 3: kFusedRows per thread, at most what the table's LDS holds as 16-byte records
// (a multiple of the wave size).
template <int BLOCK>
__host__ __device__ __forceinline__ int FusedChunk(int lds_bytes)
{
    const int fit = (lds_bytes / 16) & ~63;
    return fit < kFusedRows * BLOCK ? fit : kFusedRows * BLOCK;
}

// Records of one partition held in registers: element k of this thread is record e0 + threadIdx.x + k * BLOCK.
struct FusedRecs {
    unsigned long long key[kFusedAhead];
    unsigned long long val[kFusedAhead];  // carried value word, or the batch-relative row number
};

template <int BLOCK>
__device__ __forceinline__ void FusedLoad(const FusedIn &in, int64_t start, unsigned int e0, unsigned int e1,
                                          FusedRecs &r)
{
#pragma unroll
    for (int k = 0; k < kFusedAhead; ++k) {
        const unsigned int e = e0 + threadIdx.x + k * BLOCK;
        if (e < e1) {
            r.key[k] = LdG(in.pkeys, start + e);
            r.val[k] = in.carry ? LdG(in.pvals, start + e) : static_cast<unsigned long long>(LdG(in.prows, start + e));
        }
    }
}

// Inserts the records e0 + threadIdx.x + k * BLOCK (< e1) of r; sets sh[0] when a key finds no slot.
template <int BLOCK, typename Table>
__device__ __forceinline__ void FusedInsert(const FusedIn &in, const Table &t, int64_t start, unsigned int e0,
                                            unsigned int e1, const FusedRecs &r, int S, unsigned int *sh)
{
    unsigned long long *lkeys = t.Keys();
#pragma unroll
    for (int k = 0; k < kFusedAhead; ++k) {
        const unsigned int e = e0 + threadIdx.x + k * BLOCK;
        if (e >= e1) break;
        const int64_t row = in.carry ? 0 : start + static_cast<int64_t>(r.val[k]);
        const bool isNull = !in.carry && NullBit(in.knulls, row);
        const int slot = FusedSlot(lkeys, S, r.key[k], isNull);
        if (slot < 0) {
            sh[0] = 1u;
            break;
        }
        t.Update(slot, row, r.val[k], in.carry);
    }
}

// The rows of one staging chunk held in registers (row base + threadIdx.x + u * BLOCK): key, carried value or
// (partition << 32) | batch-relative row, partition.
struct FusedChunkRegs {
    unsigned long long key[kFusedRows];
    unsigned long long val[kFusedRows];
    unsigned int part[kFusedRows];
};

template <int BLOCK, typename Table>
__device__ __forceinline__ void FusedLoadChunk(const FusedIn &in, const Table &t, int64_t start, int64_t end,
                                               int64_t base, int P, FusedChunkRegs &r)
{
#pragma unroll
    for (int u = 0; u < kFusedRows; ++u) {
        const int64_t row = base + threadIdx.x + u * BLOCK;
        bool nul = false;
        r.key[u] = 0;
        r.val[u] = 0;
        if (row < end) {
            r.key[u] = t.Key(row, nul);
            if (in.carry) r.val[u] = t.Val(row);
        }
        r.part[u] = static_cast<unsigned int>(PartOfKey(r.key[u], nul, P));
        if (!in.carry && row < end) {
            r.val[u] = (static_cast<unsigned long long>(r.part[u]) << 32) | static_cast<unsigned long long>(row - start);
        }
    }
}

// The whole group-by of batch b in one workgroup.  smax: the table's probe slots at most; lds_bytes: its LDS.
template <int BLOCK, typename Table>
__device__ __forceinline__ void BatchFused(const BatchConfig &c, const Table &t, int b, uint8_t *lds, int smax,
                                           int lds_bytes, unsigned int *cnt, unsigned int *aux, unsigned int *ccnt,
                                           unsigned int *wsum, unsigned int *sh)
{
    const FusedIn in = LoadFusedIn(c);
    const int64_t start = c.offsets[b], end = c.offsets[b + 1];
    const int n = static_cast<int>(end - start);
    const bool part_fit_set = c.part_fit > 0;
    const int fit = part_fit_set ? c.part_fit : static_cast<int>(kFusedLoad * static_cast<float>(smax));
    const bool forced = c.modes != nullptr && c.modes[b] == 2;
    const float *pred = c.pred;
    // 1. plan: direct when the rows (or the predicted groups) fit one table; a batch whose keys fill at most
    //    kDirectLoad of the largest table goes direct (a probe run is still short), and one of at most that many
    //    rows has at most that many keys, so it needs no estimate
    const float directMax = part_fit_set ? static_cast<float>(fit) : kDirectLoad * static_cast<float>(smax);
    bool triedDirect = false;
    if (!forced && static_cast<float>(n) <= fmaxf(static_cast<float>(fit), directMax)) {
        // at most n keys: the table cannot overflow unless part_fit sized partitions above the table
        if (FusedDirect<BLOCK>(c, in, t, b, start, end, FusedSlots(static_cast<float>(n), smax), sh)) return;
        triedDirect = true;
    }
    float predicted = 0.f;  // the advisor's group count (0: none)
    if (!forced && !triedDirect && pred != nullptr) {
        const float y = pred[static_cast<int64_t>(b) * 4];
        predicted = fminf(static_cast<float>(n), fmaxf(1.f, exp2f(fminf(fmaxf(y, 0.f), 1.f) * 32.f) - 1.f));
        if (predicted <= directMax) {
            if (FusedDirect<BLOCK>(c, in, t, b, start, end, smax, sh)) return;
            triedDirect = true;  // the prediction was too low: count and estimate below
            predicted = 0.f;
        }
    }
    // 2. counts per partition and the distinct estimate (its bitmap in the table's LDS, unused until step 4); a
    //    prediction raises the estimate, never lowers it, so a low one cannot make partitions overflow
    const bool estimate = true;
    float est = -1.f;
    unsigned int *lcbits = reinterpret_cast<unsigned int *>(lds);
    for (int i = threadIdx.x; i < kFusedParts; i += BLOCK) cnt[i] = 0;
    if (estimate) {
        for (int i = threadIdx.x; i < kLcWords; i += BLOCK) lcbits[i] = 0;
    }
    if (threadIdx.x == 0) sh[2] = 0;
    __syncthreads();
    for (int64_t base = start + threadIdx.x; base < end; base += kFusedRows * BLOCK) {
        unsigned long long key[kFusedRows];
        bool nul[kFusedRows];
#pragma unroll
        for (int u = 0; u < kFusedRows; ++u) {
            const int64_t row = base + u * BLOCK;
            nul[u] = false;
            key[u] = row < end ? t.Key(row, nul[u]) : 0ULL;
        }
#pragma unroll
        for (int u = 0; u < kFusedRows; ++u) {
            if (base + u * BLOCK >= end) break;
            // one hash for both: the partition (PartOfKey's low bits) and the counting bit (bits 20..34)
            const unsigned long long h = Mix64(key[u]);
            const bool side = nul[u] || key[u] == kEmptyKey;
            atomicAdd(&cnt[side ? 0 : static_cast<int>(h & (kFusedParts - 1))], 1u);
            if (estimate) {
                const unsigned int bit = nul[u] ? 0u : static_cast<unsigned int>(h >> 20) & (kLcWords * 32 - 1);
                atomicOr(&lcbits[bit >> 5], 1u << (bit & 31));
            }
        }
    }
    __syncthreads();
    if (estimate) {
        unsigned int ones = 0;
        for (int i = threadIdx.x; i < kLcWords; i += BLOCK) ones += __popc(lcbits[i]);
        if (ones) atomicAdd(&sh[2], ones);
        __syncthreads();
        const float m = kLcWords * 32.f;
        const float zeros = m - static_cast<float>(sh[2]);
        est = zeros >= 1.f ? fminf(static_cast<float>(n), -m * logf(zeros / m)) : static_cast<float>(n);
        est = fmaxf(est, fminf(static_cast<float>(n), kFusedMargin * predicted));
    }
    int P = est <= directMax ? 1 : FusedParts(est, fit);
    if (P == 1) {
        if (!triedDirect && FusedDirect<BLOCK>(c, in, t, b, start, end, smax, sh)) return;
        P = 2;  // the keys did not fit one table after all
    }
    // fold the counts onto P partitions (partitions are the low hash bits) and scan them into bucket starts
    for (int p = threadIdx.x; p < P; p += BLOCK) {
        unsigned int v = 0;
        for (int j = p; j < kFusedParts; j += P) v += cnt[j];
        aux[p] = v;
    }
    __syncthreads();
    FusedScan<BLOCK>(aux, P, wsum);
    for (int p = threadIdx.x; p < P; p += BLOCK) cnt[p] = aux[p];
    __syncthreads();
    // 3. rows to their partitions' buckets, staged in LDS chunk by chunk: a chunk's rows are ranked within their
    //    partitions (one LDS atomic each), placed in partition order in the table's LDS, and written out so that
    //    consecutive lanes write consecutive records of a partition (cnt[p] is bucket p's write cursor)
    const int cap = FusedChunk<BLOCK>(lds_bytes);
    unsigned long long *skeys = reinterpret_cast<unsigned long long *>(lds);
    unsigned long long *svals = skeys + cap;  // carried value, or (partition << 32) | batch-relative row
    for (int64_t base = start; base < end; base += cap) {
        const int m = end - base < cap ? static_cast<int>(end - base) : cap;
        for (int p = threadIdx.x; p <= P; p += BLOCK) ccnt[p] = 0;
        FusedChunkRegs chunk;
        FusedLoadChunk<BLOCK>(in, t, start, end, base, P, chunk);
        __syncthreads();
        unsigned int rank[kFusedRows];
#pragma unroll
        for (int u = 0; u < kFusedRows; ++u) {
            if (static_cast<int>(threadIdx.x) + u * BLOCK < m) rank[u] = atomicAdd(&ccnt[chunk.part[u]], 1u);
        }
        __syncthreads();
        FusedScan<BLOCK>(ccnt, P, wsum);  // chunk offsets per partition; ccnt[P] = m
#pragma unroll
        for (int u = 0; u < kFusedRows; ++u) {
            if (static_cast<int>(threadIdx.x) + u * BLOCK < m) {
                const unsigned int pos = ccnt[chunk.part[u]] + rank[u];
                skeys[pos] = chunk.key[u];
                svals[pos] = chunk.val[u];
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < m; i += BLOCK) {
            const unsigned long long k = skeys[i], v = svals[i];
            // carried records have no NULL keys (the host selects carrying only then), so the key gives the partition
            const unsigned int p =
                in.carry ? static_cast<unsigned int>(PartOfKey(k, false, P)) : static_cast<unsigned int>(v >> 32);
            const unsigned int at = cnt[p] + static_cast<unsigned int>(i) - ccnt[p];
            StG(in.pkeys, start + at, k);
            if (in.carry) StG(in.pvals, start + at, v);
            else StG(in.prows, start + at, static_cast<unsigned int>(v));
        }
        __syncthreads();
        for (int p = threadIdx.x; p < P; p += BLOCK) cnt[p] += ccnt[p + 1] - ccnt[p];
        __syncthreads();
    }
    // 4. partition by partition (the barriers above made the workgroup's bucket writes visible to its loads)
    unsigned long long *lkeys = t.Keys();
    if (threadIdx.x == 0) {
        sh[0] = 0;
        sh[1] = 0;
    }
    FusedRecs cur;
    FusedLoad<BLOCK>(in, start, aux[0], aux[1], cur);
    t.template Clear<BLOCK>(smax);  // once: each partition's emission resets the slots it used
    __syncthreads();
    bool full = false;
    for (int p = 0; p < P; ++p) {
        const unsigned int e0 = aux[p], e1 = aux[p + 1];
        const int S = FusedSlots(static_cast<float>(e1 - e0), smax);
        FusedInsert<BLOCK>(in, t, start, e0, e1, cur, S, sh);
        // records past the ones held in registers (a partition of many duplicate keys), kFusedAhead per thread at a time
        for (unsigned int base = e0 + kFusedAhead * BLOCK; base < e1; base += kFusedAhead * BLOCK) {
            FusedRecs more;
            FusedLoad<BLOCK>(in, start, base, e1, more);
            FusedInsert<BLOCK>(in, t, start, base, e1, more, S, sh);
        }
        if (p + 1 < P) FusedLoad<BLOCK>(in, start, aux[p + 1], aux[p + 2], cur);
        __syncthreads();
        full = sh[0] != 0u;
        if (full) break;
        t.template Emit<BLOCK, true>(S, start, &sh[1]);
        __syncthreads();
    }
    (void)lkeys;
    if (threadIdx.x == 0) {
        if (full) {
            c.spilled[b] = 2u;  // the exact HBM pass re-aggregates the batch and overwrites its output rows
            atomicOr(c.redo, 1u);
        } else {
            c.spilled[b] = 1u;
            c.groups[b] = sh[1];
        }
    }
}

}  // namespace omni_gpu
