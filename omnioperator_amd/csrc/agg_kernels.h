// Kernels of the group-by hash aggregation (included by agg.hip).
#pragma once

#include "agg_device.h"

namespace omni_gpu {

// ---- kernels ------------------------------------------------------------------------------------------------

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWaveSize;
constexpr int kCursorSlots = 16384;       // wave slots of the persistent insert launches (id chunks live there)
constexpr int kMaxInsertBlocks = kCursorSlots / kWavesPerBlock;

// Min / max of the non-NULL values of a single key column (the first batch the direct-index map is planned
// from). Each workgroup stores its pair {min, max} to mm[blockIdx.x] ({INT64_MAX, INT64_MIN} when it saw no
// key); the caller reads the gridDim.x pairs back and reduces them. No atomics: one contended pair per wave
// costs far more than the read of the keys. Each thread keeps four loads in flight per iteration.
constexpr int kMinMaxUnroll = 4;
constexpr int kMinMaxMaxBlocks = 1024;
__global__ void __launch_bounds__(kBlock) agg_key_minmax(KeyDev kd, int64_t rows, long long *mm)
{
    __shared__ long long red[kWavesPerBlock][2];
    long long lo = 0x7FFFFFFFFFFFFFFFLL, hi = -0x7FFFFFFFFFFFFFFFLL - 1;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock * kMinMaxUnroll;
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * kBlock * kMinMaxUnroll + threadIdx.x; base < rows;
         base += stride) {
        long long k[kMinMaxUnroll];
        bool ok[kMinMaxUnroll];
#pragma unroll
        for (int j = 0; j < kMinMaxUnroll; ++j) {
            const int64_t row = base + static_cast<int64_t>(j) * kBlock;
            ok[j] = row < rows && !NullBit(kd.nulls[0], row);
            k[j] = ok[j] ? static_cast<long long>(LoadWord(kd.values[0], kd.width[0], row)) : 0;
        }
#pragma unroll
        for (int j = 0; j < kMinMaxUnroll; ++j) {
            if (ok[j]) {
                lo = k[j] < lo ? k[j] : lo;
                hi = k[j] > hi ? k[j] : hi;
            }
        }
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const long long l2 = __shfl_xor(lo, d, 64), h2 = __shfl_xor(hi, d, 64);
        lo = l2 < lo ? l2 : lo;
        hi = h2 > hi ? h2 : hi;
    }
    if (LaneId() == 0) {
        red[WaveInBlock()][0] = lo;
        red[WaveInBlock()][1] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kWavesPerBlock; ++w) {
            lo = red[w][0] < lo ? red[w][0] : lo;
            hi = red[w][1] > hi ? red[w][1] : hi;
        }
        mm[2 * blockIdx.x] = lo;
        mm[2 * blockIdx.x + 1] = hi;
    }
}

// Rows are visited grid-stride by a fixed grid, so wave slot w keeps its id chunk across launches.
// row_gid (optional): receives each row's group id (the replay of checked decimal sums reads it).
__global__ void __launch_bounds__(kBlock) agg_insert_single(KeyDev kd, int64_t rows, Slot1 *slots,
                                                            unsigned long long mask, GroupCtx g, AggArgs args,
                                                            unsigned long long *cursors, ArrayMapDev am,
                                                            unsigned int *row_gid)
{
    __shared__ unsigned long long wc[kWavesPerBlock][2];
    const int slot = blockIdx.x * kWavesPerBlock + (WaveInBlock());
    LoadWaveChunk(wc[WaveInBlock()], cursors, slot);
    g.wave = wc[WaveInBlock()];
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t row = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; row < rows; row += stride) {
        unsigned int gid;
        if (NullBit(kd.nulls[0], row)) {
            gid = SideGroup(g, 0, 0ULL, true, 1);
        } else {
            gid = KeyGroup(am, slots, mask, LoadWord(kd.values[0], kd.width[0], row), g);
        }
        if (gid != 0xFFFFFFFFu) UpdateAggs(args, gid, row);
        if (row_gid) row_gid[row] = gid;
    }
    StoreWaveChunk(wc[WaveInBlock()], cursors, slot);
}

__global__ void __launch_bounds__(kBlock) agg_insert_multi(KeyDev kd, int64_t rows, unsigned int *ctrl,
                                                           unsigned long long *skeys, unsigned long long mask,
                                                           GroupCtx g, AggArgs args, unsigned long long *cursors,
                                                           unsigned int *row_gid)
{
    __shared__ unsigned long long wc[kWavesPerBlock][2];
    const int slot = blockIdx.x * kWavesPerBlock + (WaveInBlock());
    LoadWaveChunk(wc[WaveInBlock()], cursors, slot);
    g.wave = wc[WaveInBlock()];
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t row = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; row < rows; row += stride) {
        unsigned long long words[kMaxWords];
        EncodeKey(kd, row, words);
        unsigned int gid = InsertMulti(ctrl, skeys, mask, words, kd.W, HashWords(words, kd.W), g);
        if (gid != 0xFFFFFFFFu) UpdateAggs(args, gid, row);
        if (row_gid) row_gid[row] = gid;
    }
    StoreWaveChunk(wc[WaveInBlock()], cursors, slot);
}

// ---- checked decimal SUM: settle and row-order replay after each insert launch ------------------------------------
// (the state and its invariant are described at K_SUM_DEC_CHK)

// Marks CROSSED every group of the launch whose P or N passed its limit (and that has not overflowed); sets
// *crossed when any did.  One thread per row of the launch (a group is checked once per row it has).
__global__ void __launch_bounds__(kBlock) agg_dec_settle(int64_t rows, const unsigned int *row_gid, AggArgs args,
                                                         unsigned int *crossed)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t row = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; row < rows; row += stride) {
        const unsigned int gid = row_gid[row];
        if (gid == 0xFFFFFFFFu) continue;
        GlobalRef g{args, gid};
        for (int i = 0; i < args.n; ++i) {
            const AggDev &d = args.a[i];
            if (d.kind != K_SUM_DEC_CHK || NullBit(d.nulls, row)) continue;
            const unsigned int cf = *g.cflag(i);
            if (cf & (kDecOvf | kDecCrossed)) continue;
            if (Gt192(g.word3(i, 0), DecUpper(d.dec_out128)) || Gt192(g.word3(i, 3), DecLowerMag(d.dec_out128))) {
                atomicOr(g.cflag(i), kDecCrossed);
                atomicOr(crossed, 1u);
            }
        }
    }
}

constexpr int kReplayBlock = 1024;

// One workgroup.  Nothing to do unless agg_dec_settle marked a group.  Then, for the CROSSED groups:
//  1. subtract the launch's inputs from P / N again (any order: exact sums), which leaves the value before the launch;
//  2. replay the launch's rows in row order, kReplayBlock rows at a time: the chunk's rows of crossed groups are
//     sorted by (group, row) in LDS and the first row of each group's run adds the run's inputs in order to the
//     group's value with the reference's checked add, marking kDecOvf at the first add that leaves the limits;
//  3. clear CROSSED.
__global__ void __launch_bounds__(kReplayBlock) agg_dec_replay(int64_t rows, const unsigned int *row_gid, AggArgs args,
                                                              unsigned int *crossed)
{
    __shared__ unsigned long long sk[kReplayBlock];
    if (*crossed == 0u) return;
    const int t = threadIdx.x;
    // State words are read with L1-bypassing loads: the atomics of step 1 and the stores of other threads in step 2
    // are made at the L2, where a plain load could miss them behind a line this CU's L1 still holds.
    auto isCrossed = [&](unsigned int gid, int i) {
        return args.a[i].kind == K_SUM_DEC_CHK &&
               (LoadRelaxed32(GlobalRef{args, gid}.cflag(i)) & (kDecOvf | kDecCrossed)) == kDecCrossed;
    };
    auto load3 = [&](const GlobalRef &g, int i, int w) {
        return U192{LoadRelaxed64(g.p<unsigned long long>(i, w)), LoadRelaxed64(g.p<unsigned long long>(i, w + 1)),
                    static_cast<long long>(LoadRelaxed64(g.p<unsigned long long>(i, w + 2)))};
    };
    auto store3 = [&](const GlobalRef &g, int i, int w, U192 v) {
        StoreRelaxed64(g.p<unsigned long long>(i, w), v.lo);
        StoreRelaxed64(g.p<unsigned long long>(i, w + 1), v.hi);
        StoreRelaxed64(g.p<unsigned long long>(i, w + 2), static_cast<unsigned long long>(v.top));
    };
    // 1. take the launch's inputs out of the crossed groups' sums
    for (int64_t row = t; row < rows; row += kReplayBlock) {
        const unsigned int gid = row_gid[row];
        if (gid == 0xFFFFFFFFu) continue;
        GlobalRef g{args, gid};
        for (int i = 0; i < args.n; ++i) {
            const AggDev &d = args.a[i];
            if (!isCrossed(gid, i) || NullBit(d.nulls, row)) continue;
            unsigned long long vlo, vhi;
            long long vtop;
            DecInput(d, row, vlo, vhi, vtop);
            U192 v{vlo, vhi, vtop};
            const int w = vtop < 0 ? 3 : 0;
            if (vtop < 0) v = Neg192(v);
            const U192 m = Neg192(v);
            AddDec3(g.p<unsigned long long>(i, w), g.p<unsigned long long>(i, w + 1), g.p<long long>(i, w + 2), m.lo,
                    m.hi, m.top);
        }
    }
    __syncthreads();
    // 2. row-order replay
    for (int64_t base = 0; base < rows; base += kReplayBlock) {
        const int64_t row = base + t;
        unsigned long long key = ~0ULL;
        if (row < rows) {
            const unsigned int gid = row_gid[row];
            bool any = false;
            if (gid != 0xFFFFFFFFu) {
                for (int i = 0; i < args.n; ++i) any |= isCrossed(gid, i);
            }
            if (any) key = (static_cast<unsigned long long>(gid) << 10) | static_cast<unsigned long long>(t);
        }
        sk[t] = key;
        __syncthreads();
        for (int k = 2; k <= kReplayBlock; k <<= 1) {  // bitonic sort, ascending
            for (int j = k >> 1; j > 0; j >>= 1) {
                const int p = t ^ j;
                if (p > t) {
                    const unsigned long long a = sk[t], b = sk[p];
                    const bool up = (t & k) == 0;
                    if ((a > b) == up) {
                        sk[t] = b;
                        sk[p] = a;
                    }
                }
                __syncthreads();
            }
        }
        const unsigned long long mine = sk[t];
        if (mine != ~0ULL && (t == 0 || (sk[t - 1] >> 10) != (mine >> 10))) {
            const unsigned int gid = static_cast<unsigned int>(mine >> 10);
            GlobalRef g{args, gid};
            for (int i = 0; i < args.n; ++i) {
                const AggDev &d = args.a[i];
                if (!isCrossed(gid, i)) continue;
                const U192 up = DecUpper(d.dec_out128), lo = Neg192(DecLowerMag(d.dec_out128));
                U192 v = Add192(load3(g, i, 0), Neg192(load3(g, i, 3)));
                bool ovf = false;
                for (int e = t; e < kReplayBlock && (sk[e] >> 10) == (mine >> 10) && sk[e] != ~0ULL; ++e) {
                    const int64_t r = base + static_cast<int64_t>(sk[e] & 1023ULL);
                    if (NullBit(d.nulls, r)) continue;
                    unsigned long long vlo, vhi;
                    long long vtop;
                    DecInput(d, r, vlo, vhi, vtop);
                    v = Add192(v, U192{vlo, vhi, vtop});
                    if (Gt192(v, up) || Gt192(lo, v)) {
                        ovf = true;
                        break;
                    }
                }
                if (ovf) {
                    atomicOr(g.cflag(i), kDecOvf);
                    continue;
                }
                const bool neg = v.top < 0;
                const U192 z{0, 0, 0}, mag = neg ? Neg192(v) : v;
                store3(g, i, 0, neg ? z : mag);
                store3(g, i, 3, neg ? mag : z);
            }
        }
        __syncthreads();
    }
    // 3. clear CROSSED (the flags are read above for the last time)
    for (int64_t row = t; row < rows; row += kReplayBlock) {
        const unsigned int gid = row_gid[row];
        if (gid == 0xFFFFFFFFu) continue;
        GlobalRef g{args, gid};
        for (int i = 0; i < args.n; ++i) {
            if (args.a[i].kind == K_SUM_DEC_CHK && (LoadRelaxed32(g.cflag(i)) & kDecCrossed)) {
                atomicAnd(g.cflag(i), ~kDecCrossed);
            }
        }
    }
    __syncthreads();
    if (t == 0) *crossed = 0u;
}

constexpr int kLdsBlock = 1024;
constexpr int kLdsWaves = kLdsBlock / kWaveSize;
constexpr int kLdsMaxProbe = 32;
constexpr int kMaxLdsBlocks = kCursorSlots / kLdsWaves;

// LDS-local partial aggregation (single-word keys).  Each workgroup aggregates a contiguous chunk of rows
// in a private LDS table of L.S slots (linear probing, 64-bit LDS CAS), then merges every occupied slot into
// the global table with one global insert and one atomic per aggregate.  Rows whose key is NULL or the
// slot sentinel, or that find no LDS slot within kLdsMaxProbe steps, go straight to the global table.
__global__ void __launch_bounds__(kLdsBlock) agg_insert_lds(KeyDev kd, int64_t rows, int64_t rows_per_block,
                                                             Slot1 *slots, unsigned long long mask, GroupCtx g,
                                                             AggArgs args, LdsLayout L,
                                                             unsigned long long *spill_counter,
                                                             unsigned long long *cursors, ArrayMapDev am)
{
    OG_DYNAMIC_LDS(unsigned long long, lds_words);
    __shared__ unsigned long long wc[kLdsWaves][2];
    const int wslot = blockIdx.x * kLdsWaves + (WaveInBlock());
    LoadWaveChunk(wc[WaveInBlock()], cursors, wslot);
    g.wave = wc[WaveInBlock()];
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
            if (L.off[a][1] >= 0) reinterpret_cast<unsigned long long *>(base + L.off[a][1])[i] = InitWord1(kind);
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
            unsigned int gid = isNull ? SideGroup(g, 0, 0ULL, true, 1) : KeyGroup(am, slots, mask, key, g);
            if (gid != 0xFFFFFFFFu) UpdateAggs(args, gid, row);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < S; i += blockDim.x) {
        unsigned long long k = lkeys[i];
        if (k == kEmptyKey) continue;
        unsigned int gid = KeyGroup(am, slots, mask, k, g);  // LDS keys are never the slot sentinel
        if (gid != 0xFFFFFFFFu) MergeAggs(args, gid, LdsRef{base, L, i});
    }
    if (spilled) atomicAdd(spill_counter, static_cast<unsigned long long>(spilled));
    StoreWaveChunk(wc[WaveInBlock()], cursors, wslot);
}

// LDS-local partial aggregation of multi-word keys (several key columns or a DECIMAL128 key): as
// agg_insert_lds, with an LDS slot state per slot (0 empty, 1 being written, 2 ready) guarding the W key words.
// Hot keys of a skewed composite key are aggregated in LDS instead of contending on one global state row.
__global__ void __launch_bounds__(kLdsBlock) agg_insert_lds_multi(KeyDev kd, int64_t rows, int64_t rows_per_block,
                                                                   unsigned int *ctrl, unsigned long long *skeys,
                                                                   unsigned long long mask, GroupCtx g, AggArgs args,
                                                                   LdsLayout L, unsigned long long *spill_counter,
                                                                   unsigned long long *cursors)
{
    OG_DYNAMIC_LDS(unsigned long long, lds_words);
    __shared__ unsigned long long wc[kLdsWaves][2];
    const int wslot = blockIdx.x * kLdsWaves + WaveInBlock();
    LoadWaveChunk(wc[WaveInBlock()], cursors, wslot);
    g.wave = wc[WaveInBlock()];
    uint8_t *base = reinterpret_cast<uint8_t *>(lds_words);
    unsigned long long *lkeys = reinterpret_cast<unsigned long long *>(base + L.keys_off);
    unsigned int *lctrl = reinterpret_cast<unsigned int *>(base + L.ctrl_off);
    const int S = L.S, W = kd.W;
    for (int i = threadIdx.x; i < S; i += blockDim.x) {
        lctrl[i] = 0;
        for (int a = 0; a < args.n; ++a) {
            int kind = args.a[a].kind;
            if (Is32(kind)) {
                reinterpret_cast<int *>(base + L.off[a][0])[i] = static_cast<int>(InitWord(kind));
            } else {
                reinterpret_cast<unsigned long long *>(base + L.off[a][0])[i] = InitWord(kind);
            }
            if (L.off[a][1] >= 0) reinterpret_cast<unsigned long long *>(base + L.off[a][1])[i] = InitWord1(kind);
            if (L.off[a][2] >= 0) reinterpret_cast<unsigned long long *>(base + L.off[a][2])[i] = 0;
            if (L.flag_off[a] >= 0) base[L.flag_off[a] + i] = 0;
        }
    }
    __syncthreads();
    int64_t start = static_cast<int64_t>(blockIdx.x) * rows_per_block;
    int64_t end = start + rows_per_block < rows ? start + rows_per_block : rows;
    unsigned int spilled = 0;
    for (int64_t row = start + threadIdx.x; row < end; row += blockDim.x) {
        unsigned long long words[kMaxWords];
        EncodeKey(kd, row, words);
        const unsigned long long h = HashWords(words, W);
        int slot = -1;
        unsigned int pos = static_cast<unsigned int>(((h >> 32) * static_cast<unsigned long long>(S)) >> 32);
        int step = 0;
        while (step < kLdsMaxProbe) {
            unsigned int c = __hip_atomic_load(&lctrl[pos], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (c == 0) {
                unsigned int exp = 0;
                if (__hip_atomic_compare_exchange_strong(&lctrl[pos], &exp, 1u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_WORKGROUP)) {
                    for (int w = 0; w < W; ++w) lkeys[pos * W + w] = words[w];
                    __hip_atomic_store(&lctrl[pos], 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                    slot = static_cast<int>(pos);
                    break;
                }
                c = exp;
            }
            if (c == 1) continue;  // another lane is writing this slot's key words: read the state again
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
            ++spilled;
            unsigned int gid = InsertMulti(ctrl, skeys, mask, words, W, h, g);
            if (gid != 0xFFFFFFFFu) UpdateAggs(args, gid, row);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < S; i += blockDim.x) {
        if (lctrl[i] != 2u) continue;
        unsigned long long words[kMaxWords];
        for (int w = 0; w < W; ++w) words[w] = lkeys[i * W + w];
        unsigned int gid = InsertMulti(ctrl, skeys, mask, words, W, HashWords(words, W), g);
        if (gid != 0xFFFFFFFFu) MergeAggs(args, gid, LdsRef{base, L, i});
    }
    if (spilled) atomicAdd(spill_counter, static_cast<unsigned long long>(spilled));
    StoreWaveChunk(wc[WaveInBlock()], cursors, wslot);
}

// Rehash after growth: re-inserts the live groups of [0, n) under their ids (states do not move).
__global__ void agg_rehash_single(const unsigned long long *gkeys, const uint8_t *gvalid, int64_t n, Slot1 *slots,
                                  unsigned long long mask, const unsigned int *special)
{
    int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= n || !gvalid[gid]) return;
    if (special[0] == gid + 2 || special[1] == gid + 2) return;  // side groups are not in the table
    unsigned long long key = gkeys[gid];
    unsigned long long pos = Mix64(key) & mask;
    while (atomicCAS(&slots[pos].key, kEmptyKey, key) != kEmptyKey) pos = (pos + 1) & mask;
    slots[pos].gid1 = static_cast<unsigned int>(gid) + 1;
}

__global__ void agg_rehash_multi(const unsigned long long *gkeys, const uint8_t *gvalid, int64_t n, int W,
                                 unsigned int *ctrl, unsigned long long *skeys, unsigned long long mask)
{
    int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= n || !gvalid[gid]) return;
    unsigned long long words[kMaxWords];
    for (int w = 0; w < W; ++w) words[w] = gkeys[gid * W + w];
    unsigned long long pos = HashWords(words, W) & mask;
    while (atomicCAS(&ctrl[pos], 0u, static_cast<unsigned int>(gid) + 2) != 0u) pos = (pos + 1) & mask;
    for (int w = 0; w < W; ++w) skeys[pos * W + w] = words[w];
}

__global__ void fill_u32(unsigned int *p, unsigned int v, int64_t n)
{
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void fill_u8(uint8_t *p, uint8_t v, int64_t n)
{
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Writes the initial state row image `tmpl` (words_per_row 8-byte words) into rows [from, from + n).
__global__ void init_state_rows(unsigned long long *states, const unsigned long long *tmpl, int words_per_row,
                                int64_t from, int64_t n)
{
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n * words_per_row) return;
    states[from * words_per_row + i] = tmpl[i % words_per_row];
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

// ---- staging of many batches (omni_gpu_agg_add_batches) ------------------------------------------------------
constexpr int kStageCols = OMNI_GPU_MAX_KEYS + OMNI_GPU_MAX_AGGS;

struct StageSrc {
    const void *values;
    const uint8_t *nulls;  // LSB-first bitmap or nullptr
};

struct StageArgs {
    int nb;                              // batches
    int ncols;                           // staged columns
    long long rows;                      // total rows
    const long long *row_start;          // [nb + 1] first staged row of each batch
    const StageSrc *src;                 // [nb][ncols]
    int width[kStageCols];
    void *dst[kStageCols];               // staged values
    unsigned int *dst_nulls[kStageCols]; // staged bitmap, or nullptr when no batch has NULLs in the column
};

__device__ __forceinline__ int StageBatch(const StageArgs &a, long long row)
{
    int lo = 0, hi = a.nb - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (a.row_start[mid] <= row) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Column blockIdx.y, one row per thread: copies the value of staged row r from its batch.
__global__ void __launch_bounds__(kBlock) agg_stage_values(StageArgs a)
{
    const int c = blockIdx.y;
    const long long r = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (r >= a.rows) return;
    const int b = StageBatch(a, r);
    const long long i = r - a.row_start[b];
    const void *src = a.src[static_cast<long long>(b) * a.ncols + c].values;
    switch (a.width[c]) {
        case 8: static_cast<unsigned long long *>(a.dst[c])[r] = static_cast<const unsigned long long *>(src)[i]; break;
        case 4: static_cast<unsigned int *>(a.dst[c])[r] = static_cast<const unsigned int *>(src)[i]; break;
        case 2: static_cast<unsigned short *>(a.dst[c])[r] = static_cast<const unsigned short *>(src)[i]; break;
        case 16: {
            const unsigned long long *s2 = static_cast<const unsigned long long *>(src);
            unsigned long long *d2 = static_cast<unsigned long long *>(a.dst[c]);
            d2[2 * r] = s2[2 * i];
            d2[2 * r + 1] = s2[2 * i + 1];
            break;
        }
        default: static_cast<uint8_t *>(a.dst[c])[r] = static_cast<const uint8_t *>(src)[i]; break;
    }
}

// Column blockIdx.y, one 32-row bitmap word per thread (columns without NULLs return at once).
__global__ void __launch_bounds__(kBlock) agg_stage_nulls(StageArgs a)
{
    const int c = blockIdx.y;
    if (a.dst_nulls[c] == nullptr) return;
    const long long w = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    const long long r0 = w * 32;
    if (r0 >= a.rows) return;
    int b = StageBatch(a, r0);
    unsigned int word = 0;
    for (int k = 0; k < 32; ++k) {
        const long long r = r0 + k;
        if (r >= a.rows) break;
        while (a.row_start[b + 1] <= r) ++b;  // empty batches are skipped
        const uint8_t *nb = a.src[static_cast<long long>(b) * a.ncols + c].nulls;
        if (nb != nullptr && NullBit(nb, r - a.row_start[b])) word |= 1u << k;
    }
    a.dst_nulls[c][w] = word;
}


// Output rows [start, start + count) of the live groups, in the order of `index` (live group ids ascending).
__global__ void __launch_bounds__(256) agg_output(int64_t start, int64_t count, KeyDev kd, int single,
                                                  const unsigned int *index, const unsigned long long *gkeys,
                                                  const uint8_t *gnull, AggArgs args, OutArgs out,
                                                  int overflow_as_null, unsigned int *error_flag)
{
    int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    bool live = r < count;
    int64_t gid = live ? static_cast<int64_t>(index[start + r]) : 0;
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

}  // namespace omni_gpu
