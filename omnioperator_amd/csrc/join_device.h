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
    int64_t block_start;
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
    long long *mm;              // min/max of the non-NULL keys, or nullptr
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

// Concatenates every retained build batch into the build store (keys sign-extended to int64, NULL bytes,
// payload columns) in one launch, and reduces the min/max of the non-NULL keys with one atomic pair per
// workgroup.
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
    const int64_t tile0 = (static_cast<int64_t>(blockIdx.x) - b.block_start) * kGatherTile;
    long long lo = std::numeric_limits<long long>::max(), hi = std::numeric_limits<long long>::min();
    for (int j = 0; j < kGatherRows; ++j) {
        int64_t i = tile0 + j * 256 + threadIdx.x;
        if (i >= b.rows) break;
        int64_t o = b.row_base + i;
        bool isNull = NullBit(b.knull, i);
        long long k = isNull ? 0 : static_cast<long long>(LoadWord(b.key, g.key_width, i));
        g.keys[o] = k;
        g.knull[o] = isNull ? 1 : 0;
        if (!isNull) {
            lo = min(lo, k);
            hi = max(hi, k);
        }
        for (int p = 0; p < g.npay; ++p) {
            CopyValue(b.pay[p], g.pay[p], g.pay_width[p], i, o);
            if (g.pnull[p]) g.pnull[p][o] = NullBit(b.pnull[p], i) ? 1 : 0;
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
        if (lo <= hi) {
            atomicMin(&g.mm[0], lo);
            atomicMax(&g.mm[1], hi);
        }
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
        unsigned long long pos = Mix64(static_cast<unsigned long long>(key)) & mask;
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
template <int G, typename T>
__device__ __forceinline__ void LoadKeysT(const void *base, int64_t rows, int64_t first, long long *key, bool *live)
{
    // rows >= 1 for every batch that owns a workgroup: out-of-range lanes re-read the last row (no branch
    // around the loads, so all of them are in flight together) and are masked off
    const T OG_GLOBAL *v = reinterpret_cast<const T OG_GLOBAL *>(
        reinterpret_cast<uintptr_t>(base));
#pragma unroll
    for (int j = 0; j < G; ++j) {
        int64_t row = first + j * kBlock;
        live[j] = row < rows;
        key[j] = static_cast<long long>(v[live[j] ? row : rows - 1]);
    }
}

template <int G>
__device__ __forceinline__ void LoadProbeKeys(const ProbeBatch &b, int64_t first, long long *key, bool *live)
{
    switch (b.width) {
        case 8: LoadKeysT<G, long long>(b.keys, b.rows, first, key, live); break;
        case 4: LoadKeysT<G, int>(b.keys, b.rows, first, key, live); break;
        case 2: LoadKeysT<G, short>(b.keys, b.rows, first, key, live); break;
        default:  // BOOLEAN: any non-zero byte is true (1), as LoadWord reads it on the build side
            LoadKeysT<G, uint8_t>(b.keys, b.rows, first, key, live);
#pragma unroll
            for (int j = 0; j < G; ++j) key[j] = key[j] != 0 ? 1 : 0;
            break;
    }
    if (b.nulls != nullptr) {
#pragma unroll
        for (int j = 0; j < G; ++j) {
            int64_t row = first + j * kBlock;
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
        pos[j] = Mix64(static_cast<unsigned long long>(key[j])) & t.mask;
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
__device__ __forceinline__ unsigned long long BlockReserve(unsigned long long *counter, unsigned int wave_total,
                                                           unsigned long long *sh)
{
    constexpr int kWaves = kBlock / 64;
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
template <bool INLINE, bool DUPS, typename RowId>
__device__ __forceinline__ void EmitTile(const ProbeArgs &a, const TableDev &t, const int *head, const unsigned int *pay,
                                         RowId rowid, unsigned long long *sh)
{
    if (!DUPS) {
        unsigned long long m[kProbeRows];
        unsigned int total = 0;
#pragma unroll
        for (int j = 0; j < kProbeRows; ++j) {
            m[j] = __ballot(head[j] >= 0);
            total += __popcll(m[j]);
        }
        unsigned long long base = BlockReserve(a.counter, total, sh);
        const unsigned long long lt = LaneMaskLt();
#pragma unroll
        for (int j = 0; j < kProbeRows; ++j) {
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
    for (int j = 0; j < kProbeRows; ++j) {
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
    unsigned long long o = BlockReserve(a.counter, waveTotal, sh) + (incl - mine);
#pragma unroll
    for (int j = 0; j < kProbeRows; ++j) {
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

// ---- radix-partitioned probe (hash tables larger than the L2 caches) -------------------------------------
// Partition p of a probe key is the top bits of its home slot (slot >> shift), so partition p only reads
// table region [p << shift, (p + 1) << shift) (plus the rare probe run past its end).  The probe rows are
// histogrammed and scattered into partition order, then a persistent probe walks the partitions with each
// XCD (workgroup id mod 8) owning partitions p = x, x + 8, ... one after another, so the region it reads stays
// in its own L2 instead of every lookup fetching a 128-byte line from the Infinity Cache or HBM.

constexpr int kPartMax = 1024;
constexpr int kXcds = 8;

struct PartArgs {
    int nparts;                  // P, a power of two in [8, kPartMax]
    int shift;                   // partition = home slot >> shift
    long long nwg;               // workgroups of the histogram / scatter launches (same tiling as join_probe)
    unsigned int *counts;        // [P][nwg] counts, scanned in place into output offsets
    long long *pkeys;            // probe keys in partition order
    unsigned int *prows;         // their probe row ids
    unsigned int *total;         // rows partitioned (non-NULL probe rows)
    unsigned int *part_start;    // [P + 1] first entry of each partition
    unsigned int *xcd_prefix;    // [kXcds][P / kXcds + 1] tile prefix over each XCD's partitions
    unsigned int *xcd_next;      // [kXcds] next tile to take (work counters)
};

__device__ __forceinline__ unsigned int PartOf(const TableDev &t, const PartArgs &pa, long long key)
{
    return static_cast<unsigned int>((Mix64(static_cast<unsigned long long>(key)) & t.mask) >> pa.shift);
}

__global__ void __launch_bounds__(kBlock) join_part_hist(ProbeArgs a, TableDev t, PartArgs pa)
{
    __shared__ unsigned int hist[kPartMax];
    for (int i = threadIdx.x; i < pa.nparts; i += kBlock) hist[i] = 0;
    __syncthreads();
    const ProbeBatch b = BatchOf(a, blockIdx.x);
    const int64_t tile0 = (static_cast<int64_t>(blockIdx.x) - b.block_start) * kProbeTile;
#pragma unroll
    for (int g = 0; g < kProbeRows; g += 8) {
        long long key[8];
        bool live[8];
        LoadProbeKeys<8>(b, tile0 + g * kBlock + threadIdx.x, key, live);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if (live[j]) atomicAdd(&hist[PartOf(t, pa, key[j])], 1u);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < pa.nparts; i += kBlock) pa.counts[static_cast<long long>(i) * pa.nwg + blockIdx.x] = hist[i];
}

// Partition starts and, per XCD, the tile prefix over its partitions; resets the XCD work counters.
__global__ void join_part_plan(PartArgs pa)
{
    if (threadIdx.x != 0) return;
    const int P = pa.nparts, per = P / kXcds;
    for (int p = 0; p < P; ++p) pa.part_start[p] = pa.counts[static_cast<long long>(p) * pa.nwg];
    pa.part_start[P] = *pa.total;
    for (int x = 0; x < kXcds; ++x) {
        unsigned int run = 0;
        for (int i = 0; i < per; ++i) {
            int p = x + i * kXcds;
            pa.xcd_prefix[x * (per + 1) + i] = run;
            run += (pa.part_start[p + 1] - pa.part_start[p] + kProbeTile - 1) / kProbeTile;
        }
        pa.xcd_prefix[x * (per + 1) + per] = run;
        pa.xcd_next[x] = 0;
    }
}

__global__ void __launch_bounds__(kBlock) join_part_scatter(ProbeArgs a, TableDev t, PartArgs pa)
{
    __shared__ unsigned int base[kPartMax];
    for (int i = threadIdx.x; i < pa.nparts; i += kBlock) base[i] = pa.counts[static_cast<long long>(i) * pa.nwg + blockIdx.x];
    __syncthreads();
    const ProbeBatch b = BatchOf(a, blockIdx.x);
    const int64_t tile0 = (static_cast<int64_t>(blockIdx.x) - b.block_start) * kProbeTile;
#pragma unroll
    for (int g = 0; g < kProbeRows; g += 8) {
        long long key[8];
        bool live[8];
        const int64_t first = tile0 + g * kBlock + threadIdx.x;
        LoadProbeKeys<8>(b, first, key, live);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if (live[j]) {
                unsigned int dst = atomicAdd(&base[PartOf(t, pa, key[j])], 1u);
                pa.pkeys[dst] = key[j];
                pa.prows[dst] = static_cast<unsigned int>(b.row_base + first + j * kBlock);
            }
        }
    }
}

// Persistent probe over the partitioned rows: workgroup w serves XCD w % 8 and takes that XCD's tiles in
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

}  // namespace omni_gpu
