// Radix-partitioned group-by (OMNI_GPU_STRATEGY_RADIX): the rows are partitioned by bits of their key hash into
// pieces small enough that each one's groups fit a workgroup's LDS table, and every piece is aggregated there by
// one workgroup, which then writes its groups out under ids of its own.  No group is ever touched by two
// workgroups, so the aggregation does no device-wide atomics on group state: the global-table kernels spend
// their time on such atomics once the groups no longer fit any cache (many millions of groups).
//
//   level 1   all input batches -> store 1, in 2^8 partitions (hash bits 63..56)
//   level 2   each level-1 partition -> store 2, in P2 sub-partitions (the next bits)
//   final     one workgroup per sub-partition: LDS table, then its groups written to the group arrays
//   overflow  rows whose key found no LDS slot (a sub-partition with more groups than its table holds) are
//             gathered and go through one more level on lower hash bits, then a final step again
//
// Partitioning a level: a histogram launch counts each tile's rows per child partition into a count matrix
// (parent-major, then child, then the tile), a device-wide exclusive scan turns the counts into output
// offsets, and a scatter launch ranks each tile's rows per child in LDS and copies every column of the row
// into place.  A key's rows always land in one final piece, and a key that finds no slot there finds none for
// any of its rows (slots are never released), so each group is created exactly once.
#pragma once

#include <type_traits>

#include "agg_kernels.h"

namespace omni_gpu {

constexpr int kRadixCols = OMNI_GPU_MAX_KEYS + OMNI_GPU_MAX_AGGS;  // stored columns: keys, then aggregate inputs
constexpr int kRadixTile = 4096;                                   // rows per partitioning workgroup
constexpr int kRadixThreads = 512;
constexpr int kRadixRows = kRadixTile / kRadixThreads;             // rows per thread of a tile
constexpr int kRadixMaxFanout = 2048;                              // child partitions of one level, at most
constexpr int kRadixFirstFanout = 256;                             // ... of the first level (runs of 16 rows a tile)
constexpr int kRadixFinalThreads = 256;
constexpr int kHllBits = 10;                                       // HyperLogLog registers of the group estimate
constexpr int kHllRegs = 1 << kHllBits;

// Columns of one input batch or of a store: values (and LSB-first NULL bitmaps, nullptr when none) from the
// segment's row 0.
struct RadixSeg {
    const void *values[kRadixCols];
    const uint8_t *nulls[kRadixCols];
};

// A tile of at most kRadixTile rows of one segment and one parent partition.  Its count of child c is the
// count-matrix entry cbase + c * tpar + tcol.
struct RadixTile {
    long long row0;   // first row within the segment
    long long cbase;  // the parent's first count-matrix entry
    int seg;
    int rows;
    int tpar;         // tiles of the parent
    int tcol;         // this tile among them
};

// A stored column: width and destination arrays of one partitioning level.
struct RadixCol {
    int width;
    void *values;
    unsigned int *nulls;  // 32-bit words, cleared before the scatter; nullptr: the column has no NULLs
};

struct RadixPass {
    const RadixSeg *segs;
    const RadixTile *tiles;
    const RadixCol *cols;  // nc stored columns (the destination of the scatter)
    int nc;
    int shift;             // child = (hash >> shift) & (P - 1)
    int P;
    unsigned int *counts;
    KeyDev kd;             // key layout of the encoding (pointers come from each tile's segment)
};

// EncodeKey with the column pointers of a segment (read from device memory; the layout comes from `kd`).
__device__ __forceinline__ void EncodeSegKey(const KeyDev &kd, const RadixSeg &s, int64_t row, unsigned long long *words)
{
    for (int w = 0; w < kd.W; ++w) words[w] = 0;
    for (int c = 0; c < kd.num; ++c) {
        if (NullBit(s.nulls[c], row)) {
            words[0] |= 1ULL << c;
            continue;
        }
        if (kd.width[c] == 16) {
            const unsigned long long *v = static_cast<const unsigned long long *>(s.values[c]);
            words[kd.word[c]] = v[2 * row];
            words[kd.word[c] + 1] = v[2 * row + 1];
        } else if (kd.in_word0[c]) {
            words[0] |= (LoadWord(s.values[c], kd.width[c], row) & 0xFFFFFFFFULL) << 32;
        } else {
            words[kd.word[c]] = LoadWord(s.values[c], kd.width[c], row);
        }
    }
}

__device__ __forceinline__ unsigned int RadixChild(const KeyDev &kd, const RadixSeg &s, int64_t row, int shift, int P)
{
    unsigned long long words[kMaxWords];
    EncodeSegKey(kd, s, row, words);
    return static_cast<unsigned int>((HashWords(words, kd.W) >> shift) & static_cast<unsigned long long>(P - 1));
}

// Counts of each tile's rows per child partition into the count matrix.  With `hll` (the first level), each tile
// also writes its HyperLogLog registers of the key hashes (register = low kHllBits bits, rank = position of the
// lowest set bit above them) to hll[tile * kHllRegs ..], from which the host estimates the number of groups.
__global__ void __launch_bounds__(kRadixThreads) agg_radix_hist(RadixPass a, uint8_t *hll)
{
    __shared__ unsigned int hist[kRadixMaxFanout];
    __shared__ unsigned int regs[kHllRegs];
    const RadixTile t = a.tiles[blockIdx.x];
    for (int i = threadIdx.x; i < a.P; i += kRadixThreads) hist[i] = 0;
    if (hll) {
        for (int i = threadIdx.x; i < kHllRegs; i += kRadixThreads) regs[i] = 0;
    }
    __syncthreads();
    const RadixSeg &seg = a.segs[t.seg];
    for (int j = 0; j < kRadixRows; ++j) {
        const int r = j * kRadixThreads + threadIdx.x;
        if (r < t.rows) {
            unsigned long long words[kMaxWords];
            EncodeSegKey(a.kd, seg, t.row0 + r, words);
            const unsigned long long h = HashWords(words, a.kd.W);
            atomicAdd(&hist[(h >> a.shift) & static_cast<unsigned long long>(a.P - 1)], 1u);
            if (hll) {
                const unsigned long long rest = (h >> kHllBits) | (1ULL << (63 - kHllBits));
                atomicMax(&regs[h & (kHllRegs - 1)], static_cast<unsigned int>(__builtin_ctzll(rest)) + 1u);
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < a.P; i += kRadixThreads) {
        a.counts[t.cbase + static_cast<long long>(i) * t.tpar + t.tcol] = hist[i];
    }
    if (hll) {
        for (int i = threadIdx.x; i < kHllRegs; i += kRadixThreads) {
            hll[static_cast<long long>(blockIdx.x) * kHllRegs + i] = static_cast<uint8_t>(regs[i]);
        }
    }
}

// Maximum of every tile's HyperLogLog registers into out[kHllRegs] (u32, cleared before): each workgroup takes
// kHllTilesPerBlock tiles and merges its maxima with one atomicMax per register.
constexpr int kHllTilesPerBlock = 64;
__global__ void __launch_bounds__(256) agg_radix_hll_reduce(const uint8_t *hll, int tiles, unsigned int *out)
{
    const int t0 = blockIdx.x * kHllTilesPerBlock;
    const int t1 = t0 + kHllTilesPerBlock < tiles ? t0 + kHllTilesPerBlock : tiles;
    for (int i = threadIdx.x; i < kHllRegs; i += 256) {
        unsigned int m = 0;
        for (int t = t0; t < t1; ++t) {
            const unsigned int v = hll[static_cast<long long>(t) * kHllRegs + i];
            m = v > m ? v : m;
        }
        if (m) atomicMax(&out[i], m);
    }
}

// Copies each tile's rows into the partitioned store.  The rows are ranked within their child partition in LDS,
// the tile's child counts are scanned, and the tile's row order by child is kept in LDS; then each column (8 bytes
// at a time) is read in row order into LDS and written out in child order, so both the loads and the stores of a
// wave are consecutive.  The count matrix holds the scanned output offsets.
template <int MAXP>
__global__ void __launch_bounds__(kRadixThreads) agg_radix_scatter(RadixPass a)
{
    typedef typename std::conditional<(MAXP <= 256), unsigned char, unsigned short>::type ChildT;
    __shared__ unsigned int off[MAXP];               // the tile's child counts, then their exclusive offsets
    __shared__ unsigned short perm[kRadixTile];      // tile row at each child-ordered position
    __shared__ ChildT pchild[kRadixTile];            // child of each position
    __shared__ unsigned long long stage[kRadixTile]; // one 8-byte word of a column, in row order
    __shared__ unsigned int wsum[kRadixThreads / kWaveSize];
    const RadixTile t = a.tiles[blockIdx.x];
    const int P = a.P;
    for (int i = threadIdx.x; i < P; i += kRadixThreads) off[i] = 0;
    __syncthreads();
    const RadixSeg &seg = a.segs[t.seg];
    unsigned int child[kRadixRows], rank[kRadixRows];
#pragma unroll
    for (int j = 0; j < kRadixRows; ++j) {
        const int r = j * kRadixThreads + threadIdx.x;
        child[j] = 0;
        rank[j] = 0;
        if (r < t.rows) {
            child[j] = RadixChild(a.kd, seg, t.row0 + r, a.shift, P);
            rank[j] = atomicAdd(&off[child[j]], 1u);
        }
    }
    __syncthreads();
    // exclusive scan of the P child counts: thread i owns counts [i * per, (i + 1) * per)
    const int per = (P + kRadixThreads - 1) / kRadixThreads;
    const int c0 = threadIdx.x * per;
    unsigned int mine = 0;
    for (int i = 0; i < per; ++i) mine += c0 + i < P ? off[c0 + i] : 0u;
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
        for (int w = 0; w < kRadixThreads / kWaveSize; ++w) {
            const unsigned int v = wsum[w];
            wsum[w] = run;
            run += v;
        }
    }
    __syncthreads();
    unsigned int run = wsum[wave] + incl - mine;
    for (int i = 0; i < per; ++i) {
        if (c0 + i < P) {
            const unsigned int c = off[c0 + i];
            off[c0 + i] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kRadixRows; ++j) {
        const int r = j * kRadixThreads + threadIdx.x;
        if (r < t.rows) {
            const unsigned int pos = off[child[j]] + rank[j];
            perm[pos] = static_cast<unsigned short>(r);
            pchild[pos] = static_cast<ChildT>(child[j]);
        }
    }
    __syncthreads();
    // a position's destination: the tile's first entry in its child (from the count matrix, the same for the
    // consecutive positions of one child) plus its place in the child's run
    const unsigned int *gbase = a.counts + t.cbase + t.tcol;
    for (int c = 0; c < a.nc; ++c) {
        const RadixCol col = a.cols[c];
        const int words = col.width == 16 ? 2 : 1;
        for (int w = 0; w < words; ++w) {
            for (int r = threadIdx.x; r < t.rows; r += kRadixThreads) {
                const int64_t row = t.row0 + r;
                const void *src = seg.values[c];
                unsigned long long v;
                switch (col.width) {
                    case 16: v = static_cast<const unsigned long long *>(src)[2 * row + w]; break;
                    case 8: v = static_cast<const unsigned long long *>(src)[row]; break;
                    case 4: v = static_cast<const unsigned int *>(src)[row]; break;
                    case 2: v = static_cast<const unsigned short *>(src)[row]; break;
                    default: v = static_cast<const uint8_t *>(src)[row]; break;
                }
                stage[r] = v;
            }
            __syncthreads();
            for (int i = threadIdx.x; i < t.rows; i += kRadixThreads) {
                const unsigned int ch = pchild[i];
                const int64_t dst = static_cast<int64_t>(gbase[static_cast<long long>(ch) * t.tpar]) + (i - off[ch]);
                const unsigned long long v = stage[perm[i]];
                switch (col.width) {
                    case 16: static_cast<unsigned long long *>(col.values)[2 * dst + w] = v; break;
                    case 8: static_cast<unsigned long long *>(col.values)[dst] = v; break;
                    case 4: static_cast<unsigned int *>(col.values)[dst] = static_cast<unsigned int>(v); break;
                    case 2: static_cast<unsigned short *>(col.values)[dst] = static_cast<unsigned short>(v); break;
                    default: static_cast<uint8_t *>(col.values)[dst] = static_cast<uint8_t>(v); break;
                }
                if (w == 0 && col.nulls && NullBit(seg.nulls[c], t.row0 + perm[i])) {
                    atomicOr(col.nulls + (dst >> 5), 1u << (dst & 31));
                }
            }
            __syncthreads();
        }
    }
}

// Count-matrix positions of the first tile of every child, as the output offsets of the final pieces:
// piece g = parent * P + child starts at the scanned entry cbase[parent] + child * tpar[parent] (an empty parent's
// pieces start at the parent's start).  fstart[npieces] = total.
__global__ void agg_radix_piece_starts(const unsigned int *counts, const long long *cbase, const int *tpar,
                                       const unsigned int *pstart, int parents, int P, unsigned int total,
                                       unsigned int *fstart)
{
    const long long g = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long n = static_cast<long long>(parents) * P;
    if (g > n) return;
    if (g == n) {
        fstart[n] = total;
        return;
    }
    const int p = static_cast<int>(g / P), c = static_cast<int>(g % P);
    fstart[g] = tpar[p] > 0 ? counts[cbase[p] + static_cast<long long>(c) * tpar[p]] : pstart[p];
}

// Gathers rows[0, n) of a store into a contiguous store (the rows that found no LDS slot).
__global__ void agg_radix_gather(const RadixSeg *srcp, const RadixCol *cols, int nc, const unsigned int *rows,
                                 unsigned int n)
{
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const RadixSeg &src = *srcp;
    const int64_t row = rows[i];
    for (int c = 0; c < nc; ++c) {
        const RadixCol col = cols[c];
        switch (col.width) {
            case 16: {
                const unsigned long long *s2 = static_cast<const unsigned long long *>(src.values[c]) + 2 * row;
                unsigned long long *d2 = static_cast<unsigned long long *>(col.values) + 2 * static_cast<int64_t>(i);
                d2[0] = s2[0];
                d2[1] = s2[1];
                break;
            }
            case 8:
                static_cast<unsigned long long *>(col.values)[i] = static_cast<const unsigned long long *>(src.values[c])[row];
                break;
            case 4: static_cast<unsigned int *>(col.values)[i] = static_cast<const unsigned int *>(src.values[c])[row]; break;
            case 2:
                static_cast<unsigned short *>(col.values)[i] = static_cast<const unsigned short *>(src.values[c])[row];
                break;
            default: static_cast<uint8_t *>(col.values)[i] = static_cast<const uint8_t *>(src.values[c])[row]; break;
        }
        if (col.nulls && NullBit(src.nulls[c], row)) atomicOr(col.nulls + (i >> 5), 1u << (i & 31));
    }
}

struct RadixFinal {
    KeyDev kd;                 // the store's key columns (multi-word encoding)
    const unsigned int *fstart;
    LdsLayout L;               // S slots at most; a piece uses the first S_eff
    unsigned long long *ngroups;
    unsigned long long max_groups;
    unsigned int *overflow;    // set when an id would pass max_groups (host sizing guard)
    unsigned long long *gkeys;
    uint8_t *gnull;            // single-word handles: NULL flag per group (and gkeys holds the one key word)
    uint8_t *gvalid;
    unsigned int *special;     // single-word handles: side cells of the NULL group and of the slot-sentinel key
    int single;                // the handle keeps single-word keys (gkeys[gid], gnull, side cells)
    int single_hi;             // single-word handles: the key sits in the high half of word 0 (width <= 4)
    unsigned int *ovf_rows;    // rows that found no LDS slot
    unsigned int *ovf_count;
    int sub_bits;              // the piece is aggregated in 2^sub_bits passes, by hash bits [sub_shift, + sub_bits)
    int sub_shift;
};

// Key word a single-word handle stores for a group (the sign-extended value), from the radix encoding.
__device__ __forceinline__ unsigned long long SingleKeyWord(const RadixFinal &f, const unsigned long long *w)
{
    return f.single_hi ? static_cast<unsigned long long>(static_cast<long long>(static_cast<int>(w[0] >> 32))) : w[1];
}

// Copies a group's LDS partial state into its (fresh) state row: both start from the same initial words, so the
// copy is the merge.
__device__ __forceinline__ void StoreLdsState(const AggArgs &args, unsigned int gid, const LdsRef &l)
{
    GlobalRef g{args, gid};
    for (int i = 0; i < args.n; ++i) {
        const int kind = args.a[i].kind;
        if (Is32(kind)) *g.p<int>(i, 0) = *l.p<int>(i, 0);
        else *g.p<unsigned long long>(i, 0) = *l.p<unsigned long long>(i, 0);
        if (l.L.off[i][1] >= 0) *g.p<unsigned long long>(i, 1) = *l.p<unsigned long long>(i, 1);
        if (l.L.off[i][2] >= 0) *g.p<unsigned long long>(i, 2) = *l.p<unsigned long long>(i, 2);
        if (l.L.flag_off[i] >= 0) *g.flag(i) = *l.flag(i);
    }
}

constexpr int kRadixListRows = 8192;  // pieces of at most this many rows are bucketed by pass in LDS

// Dynamic LDS of agg_radix_final: the table, the list of a pass's groups, the piece's row list by pass (only
// with several passes).
inline size_t RadixFinalLds(const LdsLayout &L, int passes)
{
    return static_cast<size_t>(L.bytes) + sizeof(unsigned short) * static_cast<size_t>(L.S) +
           (passes > 1 ? sizeof(unsigned short) * kRadixListRows : 0);
}

// Initial words of slot i of an LDS table.
__device__ __forceinline__ void InitLdsSlot(uint8_t *base, const LdsLayout &L, const AggArgs &args, int i)
{
    for (int a = 0; a < args.n; ++a) {
        const int kind = args.a[a].kind;
        if (Is32(kind)) reinterpret_cast<int *>(base + L.off[a][0])[i] = static_cast<int>(InitWord(kind));
        else reinterpret_cast<unsigned long long *>(base + L.off[a][0])[i] = InitWord(kind);
        if (L.off[a][1] >= 0) reinterpret_cast<unsigned long long *>(base + L.off[a][1])[i] = InitWord1(kind);
        if (L.off[a][2] >= 0) reinterpret_cast<unsigned long long *>(base + L.off[a][2])[i] = 0;
        if (L.flag_off[a] >= 0) base[L.flag_off[a] + i] = 0;
    }
}

// One workgroup per final piece: the piece's rows [fstart[g], fstart[g + 1]) of the store are aggregated in
// 2^sub_bits passes, pass j taking the keys whose hash bits [sub_shift, sub_shift + sub_bits) equal j.  A piece of
// at most kRadixListRows rows is first bucketed by pass in LDS (each row hashed twice, to count and to place), so
// each pass visits only its own rows; a larger piece (a heavy key's rows) has every pass scan all its rows.  A pass
// aggregates in an LDS table of S_eff = min(S, 2 * its rows rounded up to 64) slots (linear probing from the low 32
// hash bits -- the high ones select the piece and the pass); the lane that claims a slot appends it to the pass's
// group list.  A row whose key finds no slot within kLdsMaxProbe steps is listed in
// ovf_rows.  Each pass's groups then take consecutive ids from one atomic add and lanes write consecutive groups.
__global__ void __launch_bounds__(kRadixFinalThreads) agg_radix_final(RadixFinal f, AggArgs args)
{
    OG_DYNAMIC_LDS(unsigned long long, lds_words);
    __shared__ unsigned int pcount[64], pstart[64];
    __shared__ unsigned int ngroup;
    __shared__ unsigned long long idbase;
    const unsigned int r0 = f.fstart[blockIdx.x], r1 = f.fstart[blockIdx.x + 1];
    if (r1 <= r0) return;
    const unsigned int rows = r1 - r0;
    const LdsLayout &L = f.L;
    const int W = f.kd.W;
    const unsigned int passes = 1u << f.sub_bits;
    const unsigned long long passMask = passes - 1;
    uint8_t *base = reinterpret_cast<uint8_t *>(lds_words);
    unsigned long long *lkeys = reinterpret_cast<unsigned long long *>(base + L.keys_off);
    unsigned int *lctrl = reinterpret_cast<unsigned int *>(base + L.ctrl_off);
    unsigned short *glist = reinterpret_cast<unsigned short *>(base + L.bytes);
    unsigned short *rlist = glist + L.S;
    const bool listed = passes > 1 && rows <= static_cast<unsigned int>(kRadixListRows);
    if (threadIdx.x == 0) ngroup = 0;
    unsigned int used = 0;
    if (listed) {
        for (unsigned int p = threadIdx.x; p < passes; p += kRadixFinalThreads) pcount[p] = 0;
        __syncthreads();
        for (unsigned int r = threadIdx.x; r < rows; r += kRadixFinalThreads) {
            unsigned long long words[kMaxWords];
            EncodeKey(f.kd, r0 + r, words);
            atomicAdd(&pcount[(HashWords(words, W) >> f.sub_shift) & passMask], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned int run = 0;
            for (unsigned int p = 0; p < passes; ++p) {
                pstart[p] = run;
                run += pcount[p];
                pcount[p] = pstart[p];  // the placing cursor
            }
        }
        __syncthreads();
        for (unsigned int r = threadIdx.x; r < rows; r += kRadixFinalThreads) {
            unsigned long long words[kMaxWords];
            EncodeKey(f.kd, r0 + r, words);
            rlist[atomicAdd(&pcount[(HashWords(words, W) >> f.sub_shift) & passMask], 1u)] =
                static_cast<unsigned short>(r);
        }
        __syncthreads();
    }
    for (unsigned int pass = 0; pass < passes; ++pass) {
        // rows of the pass: rlist[lo, hi) when listed, else every row of the piece (filtered by hash)
        const unsigned int lo = listed ? pstart[pass] : 0u;
        const unsigned int hi = listed ? pcount[pass] : rows;  // pcount ends at the pass's end after placing
        const unsigned int n = hi - lo;
        if (n == 0) continue;
        const unsigned long long want = listed || passes == 1 ? n : (n + passes - 1) / passes;
        int S = static_cast<int>(((want * 2 + 63) / 64) * 64);
        S = S < L.S ? S : L.S;
        for (int i = threadIdx.x; i < S; i += kRadixFinalThreads) {
            lctrl[i] = 0;
            InitLdsSlot(base, L, args, i);
        }
        __syncthreads();
        for (unsigned int k = lo + threadIdx.x; k < hi; k += kRadixFinalThreads) {
            const unsigned int row = r0 + (listed ? rlist[k] : k);
            unsigned long long words[kMaxWords];
            EncodeKey(f.kd, row, words);
            const unsigned long long h = HashWords(words, W);
            if (!listed && ((h >> f.sub_shift) & passMask) != pass) continue;
            unsigned int pos =
                static_cast<unsigned int>(((h & 0xFFFFFFFFULL) * static_cast<unsigned long long>(S)) >> 32);
            int slot = -1;
            int step = 0;
            while (step < kLdsMaxProbe) {
                unsigned int c = __hip_atomic_load(&lctrl[pos], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (c == 0) {
                    unsigned int exp = 0;
                    if (__hip_atomic_compare_exchange_strong(&lctrl[pos], &exp, 1u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_WORKGROUP)) {
                        for (int w = 0; w < W; ++w) lkeys[pos * W + w] = words[w];
                        glist[atomicAdd(&ngroup, 1u) - used] = static_cast<unsigned short>(pos);
                        __hip_atomic_store(&lctrl[pos], 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                        slot = static_cast<int>(pos);
                        break;
                    }
                    c = exp;
                }
                if (c == 1) continue;  // another lane is writing this slot: read the state again
                bool same = true;
                for (int w = 0; w < W; ++w) same &= lkeys[pos * W + w] == words[w];
                if (same) {
                    slot = static_cast<int>(pos);
                    break;
                }
                pos = pos + 1 == static_cast<unsigned int>(S) ? 0 : pos + 1;
                ++step;
            }
            if (slot >= 0) {
                UpdateAggsT(args, LdsRef{base, L, slot}, row);
            } else {
                f.ovf_rows[atomicAdd(f.ovf_count, 1u)] = row;
            }
        }
        __syncthreads();
        const unsigned int ng = ngroup - used;
        if (threadIdx.x == 0) idbase = ng ? atomicAdd(f.ngroups, static_cast<unsigned long long>(ng)) : 0ULL;
        __syncthreads();
        const unsigned long long b0 = idbase;
        if (b0 + ng > f.max_groups) {
            if (threadIdx.x == 0) atomicOr(f.overflow, 1u);
            return;
        }
        for (unsigned int k = threadIdx.x; k < ng; k += kRadixFinalThreads) {
            const int i = glist[k];
            const unsigned int gid = static_cast<unsigned int>(b0 + k);
            const unsigned long long *w = lkeys + static_cast<size_t>(i) * W;
            if (f.single) {
                const bool isNull = (w[0] & 1ULL) != 0;
                const unsigned long long key = isNull ? 0ULL : SingleKeyWord(f, w);
                f.gkeys[gid] = key;
                f.gnull[gid] = isNull ? 1 : 0;
                if (isNull) f.special[0] = gid + 2;
                else if (key == kEmptyKey) f.special[1] = gid + 2;
            } else {
                for (int x = 0; x < W; ++x) f.gkeys[static_cast<size_t>(gid) * W + x] = w[x];
            }
            f.gvalid[gid] = 1;
            StoreLdsState(args, gid, LdsRef{base, L, i});
        }
        used += ng;
        __syncthreads();
    }
}

}  // namespace omni_gpu
