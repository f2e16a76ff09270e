// Device code of the inner hash join: table layouts, build and probe kernels (included by join.hip).
#pragma once

#include <limits>

#include "scan_device.h"

namespace omni_gpu {

struct JoinSlot {
    long long key;
    int head;
    unsigned int pay;
};

struct PayloadDev {
    int width;
    const void *values;     // build store, width bytes per build row
    const uint8_t *nulls;   // build store NULL bytes, or nullptr if the column has no NULLs
    void *out;              // output values
    uint8_t *out_nulls;     // output bitmap (pre-cleared) or nullptr
};

struct ProbeBatch {
    const void *keys;
    const uint8_t *nulls;
    int64_t rows;
    int64_t row_base;
    int64_t block_start;   // first workgroup of this batch in a launch tiled by kProbeTile rows
    int64_t pblock_start;  // first workgroup of this batch in a launch tiled by kPartTile rows
    int width;
    int pad;
};

struct ProbeArgs {
    int nb;
    const ProbeBatch *batches;  // device array when nb > 1
    ProbeBatch single;
    int npay;
    PayloadDev pay[OMNI_GPU_MAX_AGGS];
    long long *out_rows;
    unsigned long long *counter;
    long long cap;
};

struct TableDev {
    const JoinSlot *slots;
    unsigned long long mask;
    const int *heads;      // array table
    long long min_key;
    unsigned long long range;
    const int *next;
    const int *sentinel_head;  // chain of build rows whose key equals the slot sentinel
    const unsigned int *sentinel_pay;  // that key's inline payload (inline-payload tables only)
};

constexpr int kBlock = 256;
inline unsigned int Blocks(int64_t n) { return static_cast<unsigned int>((n + kBlock - 1) / kBlock); }

// Home slot of a join key in a table of mask + 1 = 2^b slots: the top b bits of a multiplicative (Fibonacci)
// hash of the key, after folding its high half onto its low half so that keys differing only in high bits
// still spread.  One 64-bit multiply per key; the mask is uniform, so the shift is a scalar value.
__device__ __forceinline__ unsigned long long JoinHome(long long key, unsigned long long mask)
{
    unsigned long long k = static_cast<unsigned long long>(key);
    k ^= k >> 29;
    return (k * 0x9E3779B97F4A7C15ULL) >> (64 - __popcll(mask));
}

// One retained build batch (HashBuilderOperator::AddInput keeps the VectorBatch; here its device columns).
struct BuildBatch {
    const void *key;
    const uint8_t *knull;       // bitmap or nullptr
    int64_t rows;
    int64_t row_base;           // first build row of the batch
    int64_t block_start;        // first workgroup of the gather launch
    const void *pay[OMNI_GPU_MAX_AGGS];
    const uint8_t *pnull[OMNI_GPU_MAX_AGGS];
};

struct GatherArgs {
    const BuildBatch *batches;
    int nb;
    int key_width;
    int npay;
    int pay_width[OMNI_GPU_MAX_AGGS];
    long long *keys;            // build store
    uint8_t *knull;
    void *pay[OMNI_GPU_MAX_AGGS];
    uint8_t *pnull[OMNI_GPU_MAX_AGGS];  // nullptr: the column has no NULLs in any batch
    long long *mm;              // per-workgroup min/max pairs of the non-NULL keys, or nullptr
};

constexpr int kGatherRows = 8;                          // rows per thread
constexpr int kGatherTile = 256 * kGatherRows;          // rows per workgroup

__device__ __forceinline__ void CopyValue(const void *src, void *dst, int width, int64_t i, int64_t o)
{
    switch (width) {
        case 8: static_cast<unsigned long long *>(dst)[o] = static_cast<const unsigned long long *>(src)[i]; break;
        case 4: static_cast<unsigned int *>(dst)[o] = static_cast<const unsigned int *>(src)[i]; break;
        case 2: static_cast<unsigned short *>(dst)[o] = static_cast<const unsigned short *>(src)[i]; break;
        case 16: {
            const unsigned long long *s = static_cast<const unsigned long long *>(src);
            unsigned long long *d = static_cast<unsigned long long *>(dst);
            d[2 * o] = s[2 * i];
            d[2 * o + 1] = s[2 * i + 1];
            break;
        }
        default: static_cast<uint8_t *>(dst)[o] = static_cast<const uint8_t *>(src)[i]; break;
    }
}

// Stores the low `width` bytes of v as element o of a column of width 1/2/4/8 (the inverse of LoadWord).
__device__ __forceinline__ void StoreWordRaw(void *dst, int width, int64_t o, unsigned long long v)
{
    switch (width) {
        case 8: static_cast<unsigned long long *>(dst)[o] = v; break;
        case 4: static_cast<unsigned int *>(dst)[o] = static_cast<unsigned int>(v); break;
        case 2: static_cast<unsigned short *>(dst)[o] = static_cast<unsigned short>(v); break;
        default: static_cast<uint8_t *>(dst)[o] = static_cast<uint8_t>(v); break;
    }
}

// Concatenates every retained build batch into the build store (keys sign-extended to int64, NULL bytes,
// payload columns) in one launch. Each thread loads all of its rows of a column before storing any of them
// (the stores could alias the loads as far as the compiler knows, so a load-store loop would wait for one
// round trip per row). With g.mm set, workgroup b stores the min / max of its non-NULL keys to
// g.mm[2b], g.mm[2b + 1] ({INT64_MAX, INT64_MIN} when it has none), reduced by join_minmax_reduce: one
// contended atomic pair per workgroup cost about half of this kernel.
__global__ void __launch_bounds__(256) join_gather_build(GatherArgs g)
{
    __shared__ long long slo[4], shi[4];
    int lo_i = 0, hi_i = g.nb - 1;
    while (lo_i < hi_i) {
        int mid = (lo_i + hi_i + 1) >> 1;
        if (g.batches[mid].block_start <= static_cast<int64_t>(blockIdx.x)) lo_i = mid;
        else hi_i = mid - 1;
    }
    const BuildBatch &b = g.batches[lo_i];
    const int64_t tile0 = (static_cast<int64_t>(blockIdx.x) - b.block_start) * kGatherTile + threadIdx.x;
    long long k[kGatherRows];
    bool nul[kGatherRows];
#pragma unroll
    for (int j = 0; j < kGatherRows; ++j) {
        const int64_t i = tile0 + j * 256;
        const bool ok = i < b.rows;
        nul[j] = ok && NullBit(b.knull, i);
        k[j] = ok && !nul[j] ? static_cast<long long>(LoadWord(b.key, g.key_width, i)) : 0;
    }
    long long lo = std::numeric_limits<long long>::max(), hi = std::numeric_limits<long long>::min();
#pragma unroll
    for (int j = 0; j < kGatherRows; ++j) {
        const int64_t i = tile0 + j * 256;
        if (i < b.rows) {
            const int64_t o = b.row_base + i;
            g.keys[o] = k[j];
            g.knull[o] = nul[j] ? 1 : 0;
            if (!nul[j]) {
                lo = min(lo, k[j]);
                hi = max(hi, k[j]);
            }
        }
    }
    for (int p = 0; p < g.npay; ++p) {
        const int w = g.pay_width[p];
        if (w <= 8) {
            unsigned long long v[kGatherRows];
#pragma unroll
            for (int j = 0; j < kGatherRows; ++j) {
                const int64_t i = tile0 + j * 256;
                v[j] = i < b.rows ? LoadWord(b.pay[p], w, i) : 0ULL;
            }
#pragma unroll
            for (int j = 0; j < kGatherRows; ++j) {
                const int64_t i = tile0 + j * 256;
                if (i < b.rows) StoreWordRaw(g.pay[p], w, b.row_base + i, v[j]);
            }
        } else {
            for (int j = 0; j < kGatherRows; ++j) {
                const int64_t i = tile0 + j * 256;
                if (i < b.rows) CopyValue(b.pay[p], g.pay[p], w, i, b.row_base + i);
            }
        }
        if (g.pnull[p]) {
#pragma unroll
            for (int j = 0; j < kGatherRows; ++j) {
                const int64_t i = tile0 + j * 256;
                if (i < b.rows) g.pnull[p][b.row_base + i] = NullBit(b.pnull[p], i) ? 1 : 0;
            }
        }
    }
    if (g.mm == nullptr) return;  // uniform across the launch
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        lo = min(lo, __shfl_xor(lo, d, 64));
        hi = max(hi, __shfl_xor(hi, d, 64));
    }
    if (LaneId() == 0) {
        slo[threadIdx.x >> 6] = lo;
        shi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) {
            lo = min(lo, slo[w]);
            hi = max(hi, shi[w]);
        }
        g.mm[2 * blockIdx.x] = lo;
        g.mm[2 * blockIdx.x + 1] = hi;
    }
}

// One workgroup of 1024 lanes: mm = {min, max} over the n pairs of parts.
__global__ void __launch_bounds__(1024) join_minmax_reduce(const long long *parts, int64_t n, long long *mm)
{
    __shared__ long long slo[16], shi[16];
    long long lo = std::numeric_limits<long long>::max(), hi = std::numeric_limits<long long>::min();
    for (int64_t i = threadIdx.x; i < n; i += 1024) {
        lo = min(lo, parts[2 * i]);
        hi = max(hi, parts[2 * i + 1]);
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        lo = min(lo, __shfl_xor(lo, d, 64));
        hi = max(hi, __shfl_xor(hi, d, 64));
    }
    if (LaneId() == 0) {
        slo[threadIdx.x >> 6] = lo;
        shi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 16; ++w) {
            lo = min(lo, slo[w]);
            hi = max(hi, shi[w]);
        }
        mm[0] = lo;
        mm[1] = hi;
    }
}

__global__ void join_init_slots(JoinSlot *s, int64_t n)
{
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        s[i].key = static_cast<long long>(kEmptyKey);
        s[i].head = -1;
        s[i].pay = 0;
    }
}

__global__ void join_fill_i32(int *p, int v, int64_t n)
{
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__device__ __forceinline__ unsigned int PayloadWord(const void *v, int width, int64_t row)
{
    switch (width) {
        case 4: return static_cast<const unsigned int *>(v)[row];
        case 2: return static_cast<const unsigned short *>(v)[row];
        default: return static_cast<const uint8_t *>(v)[row];
    }
}

__global__ void join_build_hash(const long long *keys, const uint8_t *knull, int64_t n, JoinSlot *slots,
                                unsigned long long mask, int *next, int *sentinel_head, unsigned int *sentinel_pay,
                                unsigned int *has_dup, const void *inline_pay, int inline_width)
{
    int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (row >= n || knull[row]) return;
    long long key = keys[row];
    int old;
    if (static_cast<unsigned long long>(key) == kEmptyKey) {
        old = atomicExch(sentinel_head, static_cast<int>(row));
        // the sentinel key's chain lives outside the table, so its inline payload does too (inline payloads
        // are only used when every key is unique)
        if (inline_pay) *sentinel_pay = PayloadWord(inline_pay, inline_width, row);
    } else {
        unsigned long long pos = JoinHome(key, mask);
        while (true) {
            long long k = slots[pos].key;
            if (static_cast<unsigned long long>(k) == kEmptyKey) {
                k = static_cast<long long>(atomicCAS(reinterpret_cast<unsigned long long *>(&slots[pos].key), kEmptyKey,
                                                     static_cast<unsigned long long>(key)));
                if (static_cast<unsigned long long>(k) == kEmptyKey) k = key;
            }
            if (k == key) break;
            pos = (pos + 1) & mask;
        }
        old = atomicExch(&slots[pos].head, static_cast<int>(row));
        if (inline_pay) slots[pos].pay = PayloadWord(inline_pay, inline_width, row);
    }
    next[row] = old;
    if (old >= 0) *has_dup = 1u;
}

__global__ void join_build_array(const long long *keys, const uint8_t *knull, int64_t n, int *heads,
                                 long long min_key, int *next, unsigned int *has_dup)
{
    int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (row >= n || knull[row]) return;
    unsigned long long idx = static_cast<unsigned long long>(keys[row]) - static_cast<unsigned long long>(min_key);
    int old = atomicExch(&heads[idx], static_cast<int>(row));
    next[row] = old;
    if (old >= 0) *has_dup = 1u;
}

// Longest duplicate chain (only run when some key repeats).
__global__ void join_max_chain(const int *heads, int64_t nheads, int stride, const int *next, unsigned int *max_len)
{
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nheads) return;
    int h = heads[i * stride];
    unsigned int len = 0;
    for (; h >= 0; h = next[h]) ++len;
    if (len) atomicMax(max_len, len);
}

__device__ __forceinline__ int FindBatch(const ProbeArgs &a, int64_t block)
{
    int lo = 0, hi = a.nb - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (a.batches[mid].block_start <= block) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

constexpr int kProbeRows = 16;                     // probe rows per thread
constexpr int kProbeTile = kBlock * kProbeRows;    // probe rows per workgroup

// G (the probe width, a tuning parameter in {1, 2, 4, 8, 16}): lookups issued together before any is resolved.

// Keys of probe rows first, first + kBlock, ... (G rows, coalesced across the wave); the key width
// and the presence of a NULL bitmap are uniform, so each is one branch around the unrolled loads.
template <int G, typename T, int STRIDE = kBlock>
__device__ __forceinline__ void LoadKeysT(const void *base, int64_t rows, int64_t first, long long *key, bool *live)
{
    // rows >= 1 for every batch that owns a workgroup: out-of-range lanes re-read the last row (no branch
    // around the loads, so all of them are in flight together) and are masked off
    const T OG_GLOBAL *v = reinterpret_cast<const T OG_GLOBAL *>(
        reinterpret_cast<uintptr_t>(base));
#pragma unroll
    for (int j = 0; j < G; ++j) {
        int64_t row = first + j * STRIDE;
        live[j] = row < rows;
        key[j] = static_cast<long long>(v[live[j] ? row : rows - 1]);
    }
}

template <int G, int STRIDE = kBlock>
__device__ __forceinline__ void LoadProbeKeys(const ProbeBatch &b, int64_t first, long long *key, bool *live)
{
    switch (b.width) {
        case 8: LoadKeysT<G, long long, STRIDE>(b.keys, b.rows, first, key, live); break;
        case 4: LoadKeysT<G, int, STRIDE>(b.keys, b.rows, first, key, live); break;
        case 2: LoadKeysT<G, short, STRIDE>(b.keys, b.rows, first, key, live); break;
        default:  // BYTE, or BOOLEAN (0 or 1): int8, as LoadWord reads it on the build side
            LoadKeysT<G, int8_t, STRIDE>(b.keys, b.rows, first, key, live);
            break;
    }
    if (b.nulls != nullptr) {
#pragma unroll
        for (int j = 0; j < G; ++j) {
            int64_t row = first + j * STRIDE;
            if (live[j] && NullBit(b.nulls, row)) live[j] = false;
        }
    }
}

// Looks up G probe rows at once: every first table read is issued before any is examined, so a lane keeps
// G independent loads in flight; the rare longer probe sequences are then walked.
// head[j] = first build row of the key (-1: no match); pay[j] = inline payload of a hash-table hit.
template <bool ARRAY, int G>
__device__ __forceinline__ void LookupGroup(const TableDev &t, const long long *key, const bool *live, int *head,
                                            unsigned int *pay)
{
    if (ARRAY) {
#pragma unroll
        for (int j = 0; j < G; ++j) {
            unsigned long long idx = static_cast<unsigned long long>(key[j]) - static_cast<unsigned long long>(t.min_key);
            head[j] = (live[j] && idx < t.range) ? t.heads[idx] : -1;
            pay[j] = 0;
        }
        return;
    }
    unsigned long long pos[G];
    long long fk[G];
    int fh[G];
    unsigned int fp[G];
#pragma unroll
    for (int j = 0; j < G; ++j) {
        pos[j] = JoinHome(key[j], t.mask);
        JoinSlot sl = t.slots[pos[j]];  // read for every lane (pos is always in the table): no branch
        fk[j] = sl.key;
        fh[j] = sl.head;
        fp[j] = sl.pay;
    }
#pragma unroll
    for (int j = 0; j < G; ++j) {
        int hd = -1;
        unsigned int py = 0;
        if (live[j] && static_cast<unsigned long long>(key[j]) == kEmptyKey) {
            hd = *t.sentinel_head;
            py = *t.sentinel_pay;
        } else if (live[j]) {
            long long k = fk[j];
            int h0 = fh[j];
            unsigned int p0 = fp[j];
            unsigned long long p = pos[j];
            while (k != key[j] && static_cast<unsigned long long>(k) != kEmptyKey) {
                p = (p + 1) & t.mask;
                JoinSlot sl = t.slots[p];
                k = sl.key;
                h0 = sl.head;
                p0 = sl.pay;
            }
            if (k == key[j]) {
                hd = h0;
                py = p0;
            }
        }
        head[j] = hd;
        pay[j] = py;
    }
}

template <bool INLINE>
__device__ __forceinline__ void WriteMatch(const ProbeArgs &a, long long o, long long probeRow, int h, unsigned int pay)
{
    if (o >= a.cap) return;
    a.out_rows[o] = probeRow;
    if (INLINE) {  // inline payloads: exactly one <= 4-byte payload column without NULLs
        const PayloadDev &pd = a.pay[0];
        if (pd.out == nullptr) return;
        if (pd.width == 4) static_cast<unsigned int *>(pd.out)[o] = pay;
        else if (pd.width == 2) static_cast<unsigned short *>(pd.out)[o] = static_cast<unsigned short>(pay);
        else static_cast<uint8_t *>(pd.out)[o] = static_cast<uint8_t>(pay);
        return;
    }
    for (int p = 0; p < a.npay; ++p) {
        const PayloadDev &pd = a.pay[p];
        if (pd.out == nullptr) continue;
        if (INLINE) {
            if (pd.width == 4) static_cast<unsigned int *>(pd.out)[o] = pay;
            else if (pd.width == 2) static_cast<unsigned short *>(pd.out)[o] = static_cast<unsigned short>(pay);
            else static_cast<uint8_t *>(pd.out)[o] = static_cast<uint8_t>(pay);
            continue;
        }
        switch (pd.width) {
            case 8: static_cast<unsigned long long *>(pd.out)[o] = static_cast<const unsigned long long *>(pd.values)[h]; break;
            case 4: static_cast<unsigned int *>(pd.out)[o] = static_cast<const unsigned int *>(pd.values)[h]; break;
            case 2: static_cast<unsigned short *>(pd.out)[o] = static_cast<const unsigned short *>(pd.values)[h]; break;
            case 16: {
                const unsigned long long *src = static_cast<const unsigned long long *>(pd.values);
                unsigned long long *dst = static_cast<unsigned long long *>(pd.out);
                dst[2 * o] = src[2 * h];
                dst[2 * o + 1] = src[2 * h + 1];
                break;
            }
            default: static_cast<uint8_t *>(pd.out)[o] = static_cast<const uint8_t *>(pd.values)[h]; break;
        }
        if (pd.nulls && pd.out_nulls && pd.nulls[h]) {
            atomicOr(reinterpret_cast<unsigned int *>(pd.out_nulls) + (o >> 5), 1u << (o & 31));
        }
    }
}

// One global atomic per workgroup: every wave passes its match total; returns the wave's first output row.
// Must be reached by all threads of the workgroup.
template <int BLOCK = kBlock>
__device__ __forceinline__ unsigned long long BlockReserve(unsigned long long *counter, unsigned int wave_total,
                                                           unsigned long long *sh)
{
    constexpr int kWaves = BLOCK / 64;
    const int wave = threadIdx.x >> 6;
    if (LaneId() == 0) sh[wave] = wave_total;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long run = 0;
        for (int w = 0; w < kWaves; ++w) {
            unsigned long long v = sh[w];
            sh[w] = run;
            run += v;
        }
        sh[kWaves] = run ? atomicAdd(counter, run) : 0ULL;
    }
    __syncthreads();
    return sh[kWaves] + sh[wave];
}

// Probes kProbeTile rows per workgroup (kProbeRows per thread, coalesced per row slot), keeps the hits in
// registers, then reserves the workgroup's output range with one atomic and writes the matches contiguously.
// ARRAY: direct-index table; INLINE: the single <= 4-byte payload lives in the slot (unique keys only);
// DUPS: some key has several build rows (walk next[]).
// Writes the matches of one tile: lane row j has first build row head[j] (-1: none), inline payload pay[j] and
// output probe row id rowid(j).  Reserves the workgroup's output range with one atomic.  All threads of the
// workgroup must call it.
template <bool INLINE, bool DUPS, int BLOCK = kBlock, int ROWS = kProbeRows, typename RowId>
__device__ __forceinline__ void EmitTile(const ProbeArgs &a, const TableDev &t, const int *head, const unsigned int *pay,
                                         RowId rowid, unsigned long long *sh)
{
    if (!DUPS) {
        unsigned long long m[ROWS];
        unsigned int total = 0;
#pragma unroll
        for (int j = 0; j < ROWS; ++j) {
            m[j] = __ballot(head[j] >= 0);
            total += __popcll(m[j]);
        }
        unsigned long long base = BlockReserve<BLOCK>(a.counter, total, sh);
        const unsigned long long lt = LaneMaskLt();
#pragma unroll
        for (int j = 0; j < ROWS; ++j) {
            if (head[j] >= 0) {
                WriteMatch<INLINE>(a, static_cast<long long>(base + __popcll(m[j] & lt)), rowid(j), head[j], pay[j]);
            }
            base += __popcll(m[j]);
        }
        return;
    }
    // duplicate keys: per-row match counts, wave-level exclusive scan of each lane's total
    unsigned int mine = 0;
#pragma unroll
    for (int j = 0; j < ROWS; ++j) {
        for (int h = head[j]; h >= 0; h = t.next[h]) ++mine;
    }
    const int lane = LaneId();
    unsigned int incl = mine;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        unsigned int up = __shfl_up(incl, d, 64);
        if (lane >= d) incl += up;
    }
    unsigned int waveTotal = __builtin_amdgcn_readlane(incl, 63);
    unsigned long long o = BlockReserve<BLOCK>(a.counter, waveTotal, sh) + (incl - mine);
#pragma unroll
    for (int j = 0; j < ROWS; ++j) {
        for (int h = head[j]; h >= 0; h = t.next[h]) WriteMatch<false>(a, static_cast<long long>(o++), rowid(j), h, 0);
    }
}

// The batch owning workgroup `block` of a tiled launch over the probe batches (uniform per workgroup).
__device__ __forceinline__ ProbeBatch BatchOf(const ProbeArgs &a, unsigned int block)
{
    ProbeBatch b;
    if (a.nb == 1) b = a.single;
    else b = a.batches[FindBatch(a, block)];
    return b;
}

template <bool ARRAY, bool INLINE, bool DUPS, int G>
__global__ void __launch_bounds__(kBlock) join_probe(ProbeArgs a, TableDev t)
{
    __shared__ unsigned long long sh[kBlock / 64 + 1];
    const ProbeBatch b = BatchOf(a, blockIdx.x);  // uniform: held in scalar registers
    const int64_t tile0 = (static_cast<int64_t>(blockIdx.x) - b.block_start) * kProbeTile;
    int head[kProbeRows];
    unsigned int pay[kProbeRows];
#pragma unroll
    for (int g = 0; g < kProbeRows; g += G) {
        long long key[G];
        bool live[G];
        LoadProbeKeys<G>(b, tile0 + g * kBlock + threadIdx.x, key, live);
        LookupGroup<ARRAY, G>(t, key, live, head + g, pay + g);
    }
    const long long rowBase = b.row_base + tile0 + threadIdx.x;
    EmitTile<INLINE, DUPS>(a, t, head, pay, [&](int j) { return rowBase + j * kBlock; }, sh);
}

// ---- join types other than inner (lookup_join.cpp:1037-1438, lookup_outer_join.cpp:155-320) ------------------
// Probe-side emission rules per probe row (a NULL key matches nothing):
//   kModeInner  one output row per matching build row;
//   kModeLeft   the same, or one row with NULL payloads when nothing matches (probe-preserving outer joins:
//               LEFT / RIGHT with the build on the other side, and FULL; ProbeBatchForOppositeSideOuterJoin,
//               ProbeBatchForFullJoin);
//   kModeSemi   one row (the probe row alone) when any build row matches (ProbeBatchForLeftSemiJoin);
//   kModeAnti   one row (the probe row alone) when none matches (ProbeBatchForLeftAntiJoin).
// With `visited` set, every matching build row is marked (build-preserving outer joins: LEFT / RIGHT with the
// build on the same side, and FULL; RowRefListWithFlags::visited, join_hash_table_variants.h:129-137), and
// join_outer_emit later writes the build rows never marked.
constexpr int kModeInner = 0, kModeLeft = 1, kModeSemi = 2, kModeAnti = 3;

struct TypedArgs {
    int mode;
    int empty;           // the build side has no rows: nothing matches (no table to read)
    uint8_t *visited;    // one byte per build row, or nullptr
};

// Payload values of build row h into output row o (NULL bits from the build store's NULL bytes).
__device__ __forceinline__ void CopyPayloads(const PayloadDev *pay, int npay, long long o, long long h)
{
    for (int p = 0; p < npay; ++p) {
        const PayloadDev &pd = pay[p];
        if (pd.out == nullptr) continue;
        CopyValue(pd.values, pd.out, pd.width, h, o);
        if (pd.nulls && pd.out_nulls && pd.nulls[h]) {
            atomicOr(reinterpret_cast<unsigned int *>(pd.out_nulls) + (o >> 5), 1u << (o & 31));
        }
    }
}

// A probe row without a build row: payload values zero, NULL bits set (the output bitmaps are required).
__device__ __forceinline__ void WriteNullMatch(const ProbeArgs &a, long long o, long long probeRow)
{
    if (o >= a.cap) return;
    a.out_rows[o] = probeRow;
    for (int p = 0; p < a.npay; ++p) {
        const PayloadDev &pd = a.pay[p];
        if (pd.out == nullptr) continue;
        switch (pd.width) {
            case 8: static_cast<unsigned long long *>(pd.out)[o] = 0; break;
            case 4: static_cast<unsigned int *>(pd.out)[o] = 0; break;
            case 2: static_cast<unsigned short *>(pd.out)[o] = 0; break;
            case 16:
                static_cast<unsigned long long *>(pd.out)[2 * o] = 0;
                static_cast<unsigned long long *>(pd.out)[2 * o + 1] = 0;
                break;
            default: static_cast<uint8_t *>(pd.out)[o] = 0; break;
        }
        atomicOr(reinterpret_cast<unsigned int *>(pd.out_nulls) + (o >> 5), 1u << (o & 31));
    }
}

// Marks build row h matched; the load skips the store once the row is marked (every writer stores 1).
__device__ __forceinline__ void MarkVisited(uint8_t *visited, int h)
{
    if (visited && visited[h] == 0) visited[h] = 1;
}

// Writes the output rows of one tile under the rules above: lane row j has first build row head[j] (-1: none),
// inline payload pay[j] (INLINE tables), counts as a probe row when inr[j] (a NULL key is a probe row without a
// match), and has output probe row id rowid(j).  Reserves the workgroup's output range with one atomic.  Unique
// keys give every row at most one output row, so row j's output rows are compacted across the wave with a
// ballot (coalesced stores); with duplicate keys each lane writes its rows in order after a wave-level scan of
// per-lane counts.  All threads of the workgroup must call it.
template <bool DUPS, bool INLINE, int BLOCK, int ROWS, typename RowId>
__device__ __forceinline__ void EmitTileTyped(const ProbeArgs &a, const TableDev &t, const TypedArgs &ty,
                                              const int *head, const unsigned int *pay, const bool *inr, RowId rowid,
                                              unsigned long long *sh)
{
    const bool probeOnly = ty.mode == kModeSemi || ty.mode == kModeAnti;
    if (!DUPS) {
        unsigned long long m[ROWS];
        unsigned int total = 0;
#pragma unroll
        for (int j = 0; j < ROWS; ++j) {
            const bool out = ty.mode == kModeLeft ? inr[j]
                           : ty.mode == kModeAnti ? (inr[j] && head[j] < 0)
                                                  : head[j] >= 0;  // inner, semi
            m[j] = __ballot(out);
            total += static_cast<unsigned int>(__popcll(m[j]));
        }
        unsigned long long base = BlockReserve<BLOCK>(a.counter, total, sh);
        const unsigned long long lt = LaneMaskLt();
        const int lane = LaneId();
#pragma unroll
        for (int j = 0; j < ROWS; ++j) {
            if ((m[j] >> lane) & 1ULL) {
                const long long o = static_cast<long long>(base + __popcll(m[j] & lt));
                const long long pr = rowid(j);
                if (probeOnly) {
                    if (o < a.cap) a.out_rows[o] = pr;
                } else if (head[j] >= 0) {
                    if (INLINE) {
                        WriteMatch<true>(a, o, pr, head[j], pay[j]);
                    } else if (o < a.cap) {
                        a.out_rows[o] = pr;
                        CopyPayloads(a.pay, a.npay, o, head[j]);
                    }
                    MarkVisited(ty.visited, head[j]);
                } else {
                    WriteNullMatch(a, o, pr);
                }
            }
            base += __popcll(m[j]);
        }
        return;
    }
    unsigned int mine = 0;
#pragma unroll
    for (int j = 0; j < ROWS; ++j) {
        unsigned int len = 0;
        for (int h = head[j]; h >= 0; h = t.next[h]) ++len;
        if (ty.mode == kModeInner) mine += len;
        else if (ty.mode == kModeLeft) mine += inr[j] ? (len ? len : 1u) : 0u;
        else if (ty.mode == kModeSemi) mine += len ? 1u : 0u;
        else mine += (inr[j] && !len) ? 1u : 0u;
    }
    const int lane = LaneId();
    unsigned int incl = mine;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        unsigned int up = __shfl_up(incl, d, 64);
        if (lane >= d) incl += up;
    }
    const unsigned int waveTotal = __builtin_amdgcn_readlane(incl, 63);
    long long o = static_cast<long long>(BlockReserve<BLOCK>(a.counter, waveTotal, sh) + (incl - mine));
#pragma unroll
    for (int j = 0; j < ROWS; ++j) {
        const long long pr = rowid(j);
        if (probeOnly) {
            const bool out = ty.mode == kModeSemi ? head[j] >= 0 : (inr[j] && head[j] < 0);
            if (out) {
                if (o < a.cap) a.out_rows[o] = pr;
                ++o;
            }
            continue;
        }
        for (int h = head[j]; h >= 0; h = t.next[h]) {
            if (o < a.cap) {
                a.out_rows[o] = pr;
                CopyPayloads(a.pay, a.npay, o, h);
            }
            ++o;
            MarkVisited(ty.visited, h);
        }
        if (ty.mode == kModeLeft && inr[j] && head[j] < 0) WriteNullMatch(a, o++, pr);
    }
}

// Direct probe of one tile (kProbeTile rows per workgroup, the launch shape of join_probe) under the rules
// above.
template <bool ARRAY, bool DUPS, bool INLINE>
__global__ void __launch_bounds__(kBlock) join_probe_typed(ProbeArgs a, TableDev t, TypedArgs ty)
{
    __shared__ unsigned long long sh[kBlock / 64 + 1];
    const ProbeBatch b = BatchOf(a, blockIdx.x);
    const int64_t tile0 = (static_cast<int64_t>(blockIdx.x) - b.block_start) * kProbeTile;
    int head[kProbeRows];
    unsigned int pay[kProbeRows];
    bool inr[kProbeRows];
#pragma unroll
    for (int j = 0; j < kProbeRows; ++j) {
        const int64_t row = tile0 + j * kBlock + threadIdx.x;
        long long key[1];
        bool live[1];
        LoadProbeKeys<1>(b, row, key, live);
        inr[j] = row < b.rows;
        head[j] = -1;
        pay[j] = 0;
        if (!ty.empty) LookupGroup<ARRAY, 1>(t, key, live, head + j, pay + j);
    }
    const long long rowBase = b.row_base + tile0 + threadIdx.x;
    EmitTileTyped<DUPS, INLINE, kBlock, kProbeRows>(a, t, ty, head, pay, inr,
                                                    [&](int j) { return rowBase + j * kBlock; }, sh);
}

// The NULL-key probe rows of the outer (probe-preserving) and anti joins when the probe is radix-partitioned:
// the partitioning skips NULL keys, so this launch (the direct probe's tiling) writes each of them as a row
// without a build row.
__global__ void __launch_bounds__(kBlock) join_emit_null_keys(ProbeArgs a)
{
    __shared__ unsigned long long sh[kBlock / 64 + 1];
    const ProbeBatch b = BatchOf(a, blockIdx.x);
    const int64_t tile0 = (static_cast<int64_t>(blockIdx.x) - b.block_start) * kProbeTile;
    unsigned long long m[kProbeRows];
    unsigned int total = 0;
#pragma unroll
    for (int j = 0; j < kProbeRows; ++j) {
        const int64_t row = tile0 + j * kBlock + threadIdx.x;
        m[j] = __ballot(b.nulls != nullptr && row < b.rows && NullBit(b.nulls, row));
        total += static_cast<unsigned int>(__popcll(m[j]));
    }
    unsigned long long base = BlockReserve<kBlock>(a.counter, total, sh);
    const unsigned long long lt = LaneMaskLt();
    const long long rowBase = b.row_base + tile0 + threadIdx.x;
#pragma unroll
    for (int j = 0; j < kProbeRows; ++j) {
        if ((m[j] >> LaneId()) & 1ULL) {
            const long long o = static_cast<long long>(base + __popcll(m[j] & lt));
            const long long pr = rowBase + j * kBlock;
            if (a.npay > 0) WriteNullMatch(a, o, pr);
            else if (o < a.cap) a.out_rows[o] = pr;
        }
        base += __popcll(m[j]);
    }
}

// Existence join (ProbeBatchForExistenceJoin, lookup_join.cpp:1393-1438): out[i] = 1 when probe row i has a
// matching build row, else 0 (also for a NULL key).  One row per thread.
template <bool ARRAY>
__global__ void __launch_bounds__(kBlock) join_probe_exists(ProbeBatch b, TableDev t, int empty, uint8_t *out)
{
    const int64_t row = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (row >= b.rows) return;
    long long key[1];
    bool live[1];
    int head[1] = {-1};
    unsigned int pay[1];
    LoadProbeKeys<1>(b, row, key, live);
    if (!empty) LookupGroup<ARRAY, 1>(t, key, live, head, pay);
    out[row] = head[0] >= 0 ? 1 : 0;
}

struct OuterArgs {
    const uint8_t *visited;
    int64_t n;                  // build rows
    long long *out_rows;        // build row ids, or nullptr
    int npay;
    PayloadDev pay[OMNI_GPU_MAX_AGGS];
    unsigned long long *counter;
    long long cap;
};

// LookupOuterJoinOperator::GetOutput (lookup_outer_join.cpp:155-320): every build row never marked by a probe,
// NULL-key rows included (they never match), with its payloads; output order unspecified.  One row per thread.
__global__ void __launch_bounds__(kBlock) join_outer_emit(OuterArgs g)
{
    __shared__ unsigned long long sh[kBlock / 64 + 1];
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    const bool emit = i < g.n && g.visited[i] == 0;
    const unsigned long long m = __ballot(emit);
    const unsigned long long base = BlockReserve<kBlock>(g.counter, static_cast<unsigned int>(__popcll(m)), sh);
    if (!emit) return;
    const long long o = static_cast<long long>(base + __popcll(m & LaneMaskLt()));
    if (o >= g.cap) return;
    if (g.out_rows) g.out_rows[o] = i;
    CopyPayloads(g.pay, g.npay, o, i);
}

// ---- radix-partitioned probe (hash tables larger than the L2 caches) -------------------------------------
// Partition p of a probe key is the top bits of its home slot (slot >> shift), so partition p only reads
// table region [p << shift, (p + 1) << shift) (plus the rare probe run past its end).  The probe rows are
// histogrammed and scattered into partition order (key + row id), then one of two probes runs:
//  - window probe: the region is staged in LDS by each workgroup serving the partition, and lookups read
//    LDS (the table is read once per workgroup instead of one 128-byte line per lookup);
//  - L2 probe (regions too big for LDS): a persistent probe in which XCD x (workgroup id mod 8) walks
//    partitions x, x + 8, ... in order, so the region it reads stays in its own L2.

constexpr int kPartMax = 4096;
constexpr int kPartTile = 8192;                     // probe rows per workgroup of the histogram / scatter launches
constexpr int kHistRows = kPartTile / kBlock;       // rows per thread of the histogram
constexpr int kScatterBlock = 1024;                 // threads of the scatter
constexpr int kScatterRows = kPartTile / kScatterBlock;
constexpr int kWinSlots = 8192;                     // largest region the window probe stages in LDS
constexpr int kWinExtra = 256;                      // slots past the region end staged with it
constexpr int kWinBlock = 1024;                     // threads of the window probe
constexpr int kWinTile = kWinBlock * kProbeRows;    // entries per window-probe iteration
constexpr int kWinStage = (kWinSlots + kWinExtra + kWinBlock - 1) / kWinBlock;  // staged slots per thread
constexpr size_t kWinLdsBytes = sizeof(JoinSlot) * kWinStage * kWinBlock;        // window probe dynamic LDS

// Dynamic LDS bytes of the scatter for P partitions.
inline size_t ScatterLdsBytes(int P)
{
    return static_cast<size_t>(kPartTile) * (sizeof(long long) + sizeof(unsigned int)) + 2 * sizeof(unsigned int) * P;
}

struct PartArgs {
    int nparts;                  // P, a power of two in [8, kPartMax]
    int shift;                   // partition = home slot >> shift
    long long ntiles;            // tiles of kPartTile probe rows (histogram workgroups with rows)
    long long ncol;              // columns of the count matrix: ntiles rounded up to 8
    long long xcd_cols;          // ncol / 8
    unsigned int *counts;        // [P][ncol] counts, scanned in place into output offsets
    long long *pkeys;            // probe keys in partition order
    unsigned int *prows;         // their probe row ids
    unsigned int *total;         // rows partitioned (non-NULL probe rows)
    unsigned int *part_start;    // [P + 1] first entry of each partition
    unsigned int *xcd_prefix;    // [kXcds][P / kXcds + 1] tile prefix over each XCD's partitions (L2 probe)
    unsigned int *xcd_next;      // [kXcds] next tile to take (work counters, L2 probe)
};

__device__ __forceinline__ unsigned int PartOf(const TableDev &t, const PartArgs &pa, long long key)
{
    return static_cast<unsigned int>(JoinHome(key, t.mask) >> pa.shift);
}

// Count-matrix column of a histogram / scatter workgroup.  Workgroup w runs on XCD w mod 8, and the columns
// of one XCD are adjacent, so within a partition the rows written by one XCD's workgroups are contiguous and
// their partial lines meet in that XCD's L2 before being written back.
__device__ __forceinline__ long long PartCol(const PartArgs &pa, unsigned int block)
{
    return static_cast<long long>(block % kXcds) * pa.xcd_cols + block / kXcds;
}

// The batch owning workgroup `block` of a launch tiled by kPartTile rows (uniform per workgroup).
__device__ __forceinline__ ProbeBatch PartBatchOf(const ProbeArgs &a, unsigned int block)
{
    if (a.nb == 1) return a.single;
    int lo = 0, hi = a.nb - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (a.batches[mid].pblock_start <= block) lo = mid;
        else hi = mid - 1;
    }
    return a.batches[lo];
}

// Launched with ncol workgroups: those past ntiles only write the zero counts of their padding column.
__global__ void __launch_bounds__(kBlock) join_part_hist(ProbeArgs a, TableDev t, PartArgs pa)
{
    __shared__ unsigned int hist[kPartMax];
    if (blockIdx.x >= pa.ntiles) {
        const long long col = PartCol(pa, blockIdx.x);
        for (int i = threadIdx.x; i < pa.nparts; i += kBlock) pa.counts[static_cast<long long>(i) * pa.ncol + col] = 0;
        return;
    }
    for (int i = threadIdx.x; i < pa.nparts; i += kBlock) hist[i] = 0;
    __syncthreads();
    const ProbeBatch b = PartBatchOf(a, blockIdx.x);
    const int64_t tile0 = (static_cast<int64_t>(blockIdx.x) - b.pblock_start) * kPartTile;
    for (int g = 0; g < kHistRows; g += 8) {
        long long key[8];
        bool live[8];
        LoadProbeKeys<8>(b, tile0 + g * kBlock + threadIdx.x, key, live);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if (live[j]) atomicAdd(&hist[PartOf(t, pa, key[j])], 1u);
        }
    }
    __syncthreads();
    const long long col = PartCol(pa, blockIdx.x);
    for (int i = threadIdx.x; i < pa.nparts; i += kBlock) pa.counts[static_cast<long long>(i) * pa.ncol + col] = hist[i];
}

// Partition starts and, per XCD, the tile prefix over its partitions; resets the XCD work counters.
// One workgroup of kBlock threads: the starts are gathered in parallel into LDS, then one thread per XCD
// walks that XCD's partitions.
__global__ void __launch_bounds__(kBlock) join_part_plan(PartArgs pa)
{
    __shared__ unsigned int st[kPartMax + 1];
    const int P = pa.nparts, per = P / kXcds;
    for (int p = threadIdx.x; p < P; p += kBlock) {
        st[p] = pa.counts[static_cast<long long>(p) * pa.ncol];
        pa.part_start[p] = st[p];
    }
    if (threadIdx.x == 0) {
        st[P] = *pa.total;
        pa.part_start[P] = st[P];
    }
    __syncthreads();
    if (threadIdx.x < kXcds) {
        const int x = threadIdx.x;
        unsigned int run = 0;
        for (int i = 0; i < per; ++i) {
            int p = x + i * kXcds;
            pa.xcd_prefix[x * (per + 1) + i] = run;
            run += (st[p + 1] - st[p] + kProbeTile - 1) / kProbeTile;
        }
        pa.xcd_prefix[x * (per + 1) + per] = run;
        pa.xcd_next[x] = 0;
    }
}

// LDS-staged scatter of one tile of kPartTile probe rows: the rows are ranked within their partition (LDS
// histogram), the partition counts are scanned, the rows are placed in partition order in LDS, and then
// written out in that order, so consecutive lanes store consecutive entries of one partition's run instead of
// every store of a wave touching a different partition.  Dynamic LDS: ScatterLdsBytes(P).
__global__ void __launch_bounds__(kScatterBlock) join_part_scatter(ProbeArgs a, TableDev t, PartArgs pa)
{
    OG_DYNAMIC_LDS(unsigned long long, lds);
    __shared__ unsigned int wsum[kScatterBlock / 64];
    __shared__ unsigned int ntot;
    long long *skey = reinterpret_cast<long long *>(lds);
    unsigned int *srid = reinterpret_cast<unsigned int *>(skey + kPartTile);
    unsigned int *loff = srid + kPartTile;  // partition counts of the tile, then their exclusive offsets
    unsigned int *gbase = loff + pa.nparts;  // the tile's first output entry in each partition
    const int P = pa.nparts;
    const long long col = PartCol(pa, blockIdx.x);
    for (int i = threadIdx.x; i < P; i += kScatterBlock) {
        loff[i] = 0;
        gbase[i] = pa.counts[static_cast<long long>(i) * pa.ncol + col];
    }
    __syncthreads();
    const ProbeBatch b = PartBatchOf(a, blockIdx.x);
    const int64_t first = (static_cast<int64_t>(blockIdx.x) - b.pblock_start) * kPartTile + threadIdx.x;
    long long key[kScatterRows];
    bool live[kScatterRows];
    LoadProbeKeys<kScatterRows, kScatterBlock>(b, first, key, live);
    unsigned int part[kScatterRows], rank[kScatterRows];
#pragma unroll
    for (int j = 0; j < kScatterRows; ++j) {
        part[j] = live[j] ? PartOf(t, pa, key[j]) : 0u;
        rank[j] = live[j] ? atomicAdd(&loff[part[j]], 1u) : 0u;
    }
    __syncthreads();
    // exclusive scan of the P counts: thread i owns counts [i * per, (i + 1) * per)
    const int per = (P + kScatterBlock - 1) / kScatterBlock;
    const int c0 = threadIdx.x * per;
    unsigned int mine = 0;
    for (int i = 0; i < per; ++i) mine += (c0 + i < P) ? loff[c0 + i] : 0u;
    const int lane = LaneId(), wave = threadIdx.x / 64;
    unsigned int incl = mine;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        unsigned int up = __shfl_up(incl, d, 64);
        if (lane >= d) incl += up;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int run = 0;
        for (int w = 0; w < kScatterBlock / 64; ++w) {
            unsigned int v = wsum[w];
            wsum[w] = run;
            run += v;
        }
        ntot = run;
    }
    __syncthreads();
    unsigned int run = wsum[wave] + incl - mine;
    for (int i = 0; i < per; ++i) {
        if (c0 + i < P) {
            unsigned int c = loff[c0 + i];
            loff[c0 + i] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kScatterRows; ++j) {
        if (live[j]) {
            const unsigned int pos = loff[part[j]] + rank[j];
            skey[pos] = key[j];
            srid[pos] = static_cast<unsigned int>(b.row_base + first + j * kScatterBlock);
        }
    }
    __syncthreads();
    const unsigned int n = ntot;
    for (unsigned int i = threadIdx.x; i < n; i += kScatterBlock) {
        const long long k = skey[i];
        const unsigned int p = PartOf(t, pa, k);
        const unsigned int dst = gbase[p] + (i - loff[p]);
        pa.pkeys[dst] = k;
        pa.prows[dst] = srid[i];
    }
}

// Persistent L2 probe over the partitioned rows: workgroup w serves XCD w % 8 and takes that XCD's tiles in
// partition order from its work counter until none is left.
template <bool INLINE, bool DUPS>
__global__ void __launch_bounds__(kBlock) join_part_probe(ProbeArgs a, TableDev t, PartArgs pa)
{
    __shared__ unsigned long long sh[kBlock / 64 + 1];
    __shared__ unsigned int task;
    const int x = blockIdx.x % kXcds;
    const int per = pa.nparts / kXcds;
    const unsigned int *prefix = pa.xcd_prefix + x * (per + 1);
    const unsigned int ntiles = prefix[per];
    while (true) {
        if (threadIdx.x == 0) task = atomicAdd(&pa.xcd_next[x], 1u);
        __syncthreads();
        const unsigned int tk = task;
        __syncthreads();
        if (tk >= ntiles) break;  // uniform: every workgroup reaches it once the counter passes ntiles
        int lo = 0, hi = per - 1;  // last partition i of this XCD with prefix[i] <= tk
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (prefix[mid] <= tk) lo = mid;
            else hi = mid - 1;
        }
        const int p = x + lo * kXcds;
        const unsigned int e0 = pa.part_start[p] + (tk - prefix[lo]) * kProbeTile;
        const unsigned int e1 = pa.part_start[p + 1];
        int head[kProbeRows];
        unsigned int pay[kProbeRows];
        long long rid[kProbeRows];
#pragma unroll
        for (int g = 0; g < kProbeRows; g += 8) {
            long long key[8];
            bool live[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                unsigned int e = e0 + (g + j) * kBlock + threadIdx.x;
                live[j] = e < e1;
                unsigned int ec = live[j] ? e : e1 - 1;  // e1 > e0: the tile holds at least one entry
                key[j] = pa.pkeys[ec];
                rid[g + j] = pa.prows[ec];
            }
            LookupGroup<false, 8>(t, key, live, head + g, pay + g);
        }
        EmitTile<INLINE, DUPS>(a, t, head, pay, [&](int j) { return rid[j]; }, sh);
    }
}

// ---- partitioned build ----------------------------------------------------------------------------------------
// The build rows are histogrammed and scattered into partitions like probe rows (same kernels, the build store
// as one batch), then one workgroup per partition builds its whole table region in LDS and writes it out with
// coalesced stores, instead of one device-wide CAS per row into HBM.  A key whose probe run reaches the end of
// its region is appended to an overflow list and inserted into the global table by a second launch (the run
// continues into the next region, exactly where the global build would have put it).  Every slot of the table
// is written by its region's workgroup, so the table needs no separate initialisation.

struct BuildArgs {
    const uint8_t *knull;           // build NULL flags (NULL keys are not inserted), nullptr when none is NULL
    int *next;                      // build row chains
    int *sentinel_head;             // chain of rows whose key is the slot sentinel
    unsigned int *sentinel_pay;
    unsigned int *has_dup;
    const void *inline_pay;         // the payload kept in the slot, or nullptr
    int inline_width;
    long long *ovf_keys;            // overflow list: keys and rows, count in *ovf_count
    unsigned int *ovf_rows;
    unsigned int *ovf_count;
};

// Window build: workgroup p builds region p (1 << shift slots) of the table in LDS from partition p's
// entries.  Dynamic LDS: (1 << shift) slots.
__global__ void __launch_bounds__(kWinBlock) join_window_build(TableDev t, PartArgs pa, BuildArgs ba, JoinSlot *slots)
{
    OG_DYNAMIC_LDS(JoinSlot, win);
    const unsigned int p = blockIdx.x;
    const unsigned int R = 1u << pa.shift;
    const unsigned long long wbase = static_cast<unsigned long long>(p) << pa.shift;
    for (unsigned int i = threadIdx.x; i < R; i += kWinBlock) {
        win[i].key = static_cast<long long>(kEmptyKey);
        win[i].head = -1;
        win[i].pay = 0;
    }
    __syncthreads();
    const unsigned int lo = pa.part_start[p], hi = pa.part_start[p + 1];
    for (unsigned int e = lo + threadIdx.x; e < hi; e += kWinBlock) {
        const long long key = pa.pkeys[e];
        const unsigned int row = pa.prows[e];
        // both gathers by row id are issued before either is used, so a row waits for one round trip
        const bool isNull = ba.knull != nullptr && ba.knull[row];
        const unsigned int pv = ba.inline_pay ? PayloadWord(ba.inline_pay, ba.inline_width, row) : 0u;
        if (isNull) continue;
        int old;
        if (static_cast<unsigned long long>(key) == kEmptyKey) {
            old = atomicExch(ba.sentinel_head, static_cast<int>(row));
            if (ba.inline_pay) *ba.sentinel_pay = pv;
        } else {
            unsigned int pos = static_cast<unsigned int>((JoinHome(key, t.mask) - wbase) & t.mask);
            while (pos < R) {
                unsigned long long k = static_cast<unsigned long long>(win[pos].key);
                if (k == kEmptyKey) {
                    k = atomicCAS(reinterpret_cast<unsigned long long *>(&win[pos].key), kEmptyKey,
                                  static_cast<unsigned long long>(key));
                    if (k == kEmptyKey) k = static_cast<unsigned long long>(key);
                }
                if (k == static_cast<unsigned long long>(key)) break;
                ++pos;
            }
            if (pos >= R) {  // the run leaves the region: inserted by join_build_overflow
                const unsigned int at = atomicAdd(ba.ovf_count, 1u);
                ba.ovf_keys[at] = key;
                ba.ovf_rows[at] = row;
                continue;
            }
            old = atomicExch(&win[pos].head, static_cast<int>(row));
            if (ba.inline_pay) win[pos].pay = pv;
        }
        ba.next[row] = old;
        if (old >= 0) *ba.has_dup = 1u;
    }
    __syncthreads();
    for (unsigned int i = threadIdx.x; i < R; i += kWinBlock) slots[wbase + i] = win[i];
}

// Inserts the overflow list into the global table (after every region is written).  Grid-stride over the
// device count.
__global__ void __launch_bounds__(kBlock) join_build_overflow(TableDev t, BuildArgs ba, JoinSlot *slots)
{
    const unsigned int n = *ba.ovf_count;
    for (unsigned int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const long long key = ba.ovf_keys[i];
        const unsigned int row = ba.ovf_rows[i];
        unsigned long long pos = JoinHome(key, t.mask);
        while (true) {
            long long k = slots[pos].key;
            if (static_cast<unsigned long long>(k) == kEmptyKey) {
                k = static_cast<long long>(atomicCAS(reinterpret_cast<unsigned long long *>(&slots[pos].key), kEmptyKey,
                                                     static_cast<unsigned long long>(key)));
                if (static_cast<unsigned long long>(k) == kEmptyKey) k = key;
            }
            if (k == key) break;
            pos = (pos + 1) & t.mask;
        }
        const int old = atomicExch(&slots[pos].head, static_cast<int>(row));
        if (ba.inline_pay) slots[pos].pay = PayloadWord(ba.inline_pay, ba.inline_width, row);
        ba.next[row] = old;
        if (old >= 0) *ba.has_dup = 1u;
    }
}

// Continues a probe run that left the LDS window, in the global table from slot `pos`.
__device__ __forceinline__ int WalkGlobal(const TableDev &t, long long key, unsigned long long pos, unsigned int *pay)
{
    while (true) {
        const JoinSlot sl = t.slots[pos & t.mask];
        if (sl.key == key) {
            *pay = sl.pay;
            return sl.head;
        }
        if (static_cast<unsigned long long>(sl.key) == kEmptyKey) return -1;
        ++pos;
    }
}

// Window probe: `splits` workgroups serve each partition (workgroup w: partition w / splits, its share
// w % splits of the partition's entries).  TYPED: the join-type emission rules (EmitTileTyped) instead of the
// inner join's.  Each stages the partition's region and the kWinExtra slots after it
// (wrapping at the table end) in LDS, then resolves its entries against LDS; a probe run that leaves the
// window continues in the global table.  Dynamic LDS: kWinLdsBytes (shift <= log2 kWinSlots).
template <bool INLINE, bool DUPS, bool TYPED>
__global__ void __launch_bounds__(kWinBlock) join_window_probe(ProbeArgs a, TableDev t, PartArgs pa, int splits,
                                                               TypedArgs ty)
{
    OG_DYNAMIC_LDS(JoinSlot, win);
    __shared__ unsigned long long sh[kWinBlock / 64 + 1];
    const unsigned int p = blockIdx.x / static_cast<unsigned int>(splits);
    const unsigned int sp = blockIdx.x % static_cast<unsigned int>(splits);
    const unsigned int W = (1u << pa.shift) + kWinExtra;
    const unsigned long long wbase = static_cast<unsigned long long>(p) << pa.shift;
    {
        // every staging load is issued before the first LDS store; the LDS window holds kWinStage * kWinBlock
        // slots, so the stores need no bound (slots past W repeat slot W - 1 and are never read)
        JoinSlot tmp[kWinStage];
#pragma unroll
        for (int u = 0; u < kWinStage; ++u) {
            const unsigned int i = u * kWinBlock + threadIdx.x;
            tmp[u] = t.slots[(wbase + (i < W ? i : W - 1)) & t.mask];
        }
#pragma unroll
        for (int u = 0; u < kWinStage; ++u) win[u * kWinBlock + threadIdx.x] = tmp[u];
    }
    const int sentHead = *t.sentinel_head;
    const unsigned int sentPay = *t.sentinel_pay;
    __syncthreads();
    const unsigned int lo = pa.part_start[p], n = pa.part_start[p + 1] - lo;
    const unsigned int s0 = lo + static_cast<unsigned int>(static_cast<unsigned long long>(n) * sp / splits);
    const unsigned int s1 = lo + static_cast<unsigned int>(static_cast<unsigned long long>(n) * (sp + 1) / splits);
    for (unsigned int e0 = s0; e0 < s1; e0 += kWinTile) {  // bounds uniform across the workgroup
        long long key[kProbeRows];
        unsigned int rid[kProbeRows];
        bool live[kProbeRows];
#pragma unroll
        for (int j = 0; j < kProbeRows; ++j) {
            const unsigned int e = e0 + j * kWinBlock + threadIdx.x;
            live[j] = e < s1;
            const unsigned int ec = live[j] ? e : s1 - 1;  // s1 > e0: the share holds this entry or an earlier one
            key[j] = pa.pkeys[ec];
            rid[j] = pa.prows[ec];
        }
        int head[kProbeRows];
        unsigned int pay[kProbeRows];
#pragma unroll
        for (int j = 0; j < kProbeRows; ++j) {
            int hd = -1;
            unsigned int py = 0;
            if (live[j] && static_cast<unsigned long long>(key[j]) == kEmptyKey) {
                hd = sentHead;
                py = sentPay;
            } else if (live[j]) {
                // home - wbase lies in [0, 1 << shift) for every entry of this partition
                unsigned int pos = static_cast<unsigned int>(
                    (JoinHome(key[j], t.mask) - wbase) & t.mask);
                bool open = true;
                while (open && pos < W) {
                    const JoinSlot sl = win[pos];
                    if (sl.key == key[j]) {
                        hd = sl.head;
                        py = sl.pay;
                        open = false;
                    } else if (static_cast<unsigned long long>(sl.key) == kEmptyKey) {
                        open = false;
                    } else {
                        ++pos;
                    }
                }
                if (open) hd = WalkGlobal(t, key[j], wbase + pos, &py);
            }
            head[j] = hd;
            pay[j] = py;
        }
        if (TYPED) {  // join types other than inner: every partitioned entry is a probe row with a non-NULL key
            EmitTileTyped<DUPS, INLINE, kWinBlock, kProbeRows>(
                a, t, ty, head, pay, live, [&](int j) { return static_cast<long long>(rid[j]); }, sh);
        } else {
            EmitTile<INLINE, DUPS, kWinBlock>(a, t, head, pay, [&](int j) { return static_cast<long long>(rid[j]); },
                                              sh);
        }
    }
}

}  // namespace omni_gpu
