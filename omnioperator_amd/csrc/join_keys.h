// Join keys other than one fixed-width column, translated into one exact int64 join key per row so that the
// single-key build and probe kernels serve them unchanged.
//
// HashBuilderOperatorFactory::InitVariant (core/src/operator/join/hash_builder.cpp:44-93) picks the key type of the
// reference's table from the key columns:
//   - several columns whose bit lengths (SHORT 16, INT 32, LONG / DECIMAL64 64; GetTypeLength :96-119) total at most
//     32, 64 or 128 bits: an int32, int64 or int128 key packed by ComputeMultiColKey
//     (join_hash_table_variants.cpp:105-146): key = key << bits | (KeyType)value for each column in order, the value
//     sign-extended to the key type, so a negative later value sets every earlier bit and distinct tuples can share
//     a key (they then match, as in the reference);
//   - a single DECIMAL128 column: a Decimal128 key (equality of both words);
//   - any other fixed-width combination: the columns' values serialised side by side (StringRef key, equality of
//     the bytes).
// An int32 / int64 packed key is the join key itself (int32 sign-extended).  A wider key (int128, Decimal128 or
// serialised) is first looked up in a device dictionary of the build side's distinct keys: its id -- the slot the
// key occupies -- is the join key, and a probe key absent from the dictionary becomes a NULL join key, which matches
// nothing exactly as an absent key does for every join type.  A row with a NULL in any key column has a NULL join key
// (EmplaceMultiKey... skips it on the build side, join_hash_table_variants.cpp:340-400; the probe treats it as a
// NULL key).  The reference never builds an array table for these keys (BuildHashTable tries one only for a single
// integer column, :695-746).
#pragma once

#include "omni_gpu_internal.h"

namespace omni_gpu {

enum JoinKeyMode : int {
    kKeySingle = 0,    // one fixed-width column, joined directly
    kKeyPacked32 = 1,  // ComputeMultiColKey into int32
    kKeyPacked64 = 2,  // ComputeMultiColKey into int64
    kKeyPacked128 = 3, // ComputeMultiColKey into int128 (dictionary)
    kKeyDecimal128 = 4,// one DECIMAL128 column (dictionary)
    kKeySerialized = 5,// the columns' values side by side (dictionary)
};

constexpr int kJoinKeyMaxWords = 2 * OMNI_GPU_MAX_KEYS;

struct JoinKeyCols {
    int n;                               // key columns
    int mode;                            // JoinKeyMode
    int W;                               // words of a dictionary key
    int type[OMNI_GPU_MAX_KEYS];
    int width[OMNI_GPU_MAX_KEYS];
    int bits[OMNI_GPU_MAX_KEYS];         // shift of each column in the packed modes
};

// One batch to translate: its key columns, rows, where its translated keys go, its first workgroup.
struct KeyBatch {
    const void *values[OMNI_GPU_MAX_KEYS];
    const uint8_t *nulls[OMNI_GPU_MAX_KEYS];
    int64_t rows;
    int64_t out_base;          // first translated row of the batch
    int64_t null_word_base;    // first 32-bit word of the batch's translated NULL bitmap (starts at its row 0)
    int64_t block_start;       // first workgroup of the translation launch
};

// Device dictionary of wide keys: record s is S = stride 64-bit words at recs + s * S, the key in words [0, W) and
// the state in word W (0 empty, 1 being written, 2 holds the key), so that a lookup reads one cache line; a key's
// id is its record index s.
struct KeyDict {
    unsigned long long *recs;
    unsigned long long mask;
    int W;
    int S;
};

// Words of a dictionary record for W key words: the next power of two above W (key words + state).
__host__ __device__ constexpr int DictStride(int W) { return W < 2 ? 2 : W < 4 ? 4 : W < 8 ? 8 : 16; }

// Stores v into w[o] for a runtime o with compile-time indices only, so that w stays in registers.
template <int MW>
__device__ __forceinline__ void PutWord(unsigned long long *w, int o, unsigned long long v)
{
#pragma unroll
    for (int q = 0; q < MW; ++q) {
        if (q == o) w[q] = v;
    }
}

template <int MW>
__device__ __forceinline__ unsigned long long KeyHash(const unsigned long long *w, int W)
{
    unsigned long long h = 0x2545F4914F6CDD1DULL;
#pragma unroll
    for (int k = 0; k < MW; ++k) {
        if (k < W) h = Mix64(h ^ (w[k] + 0x9E3779B97F4A7C15ULL * (k + 1)));
    }
    return h;
}

// Insert (build side) of a dictionary key from its home record; returns its id.  Device-scope relaxed atomics: a
// record's key words are stored, the store counter drained (vmcnt 0), and only then its state set to 2, so a lane
// that reads state 2 reads the finished key.
template <int MW>
__device__ __forceinline__ long long DictInsert(const KeyDict &d, const unsigned long long *w, unsigned long long pos)
{
    while (true) {
        unsigned long long *r = d.recs + pos * d.S;
        unsigned long long c = __hip_atomic_load(&r[d.W], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (c == 0) {
            c = atomicCAS(&r[d.W], 0ULL, 1ULL);
            if (c == 0) {
#pragma unroll
                for (int k = 0; k < MW; ++k) {
                    if (k < d.W) __hip_atomic_store(&r[k], w[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                __builtin_amdgcn_s_waitcnt(0x0F70);
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                __hip_atomic_store(&r[d.W], 2ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return static_cast<long long>(pos);
            }
        }
        if (c == 1) continue;  // another lane is writing this record's key
        bool same = true;
#pragma unroll
        for (int k = 0; k < MW; ++k) {
            if (k < d.W) same &= __hip_atomic_load(&r[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == w[k];
        }
        if (same) return static_cast<long long>(pos);
        pos = (pos + 1) & d.mask;
    }
}

// The probe side's lookup in the finished dictionary (the build side's launch came before), from record pos whose
// state is c; returns the key's id, or -1 when it is absent.
template <int MW>
__device__ __forceinline__ long long DictFind(const KeyDict &d, const unsigned long long *w, unsigned long long pos,
                                              unsigned long long c)
{
    while (c != 0) {
        const unsigned long long *r = d.recs + pos * d.S;
        bool same = true;
#pragma unroll
        for (int k = 0; k < MW; ++k) {
            if (k < d.W) same &= r[k] == w[k];
        }
        if (same) return static_cast<long long>(pos);
        pos = (pos + 1) & d.mask;
        c = d.recs[pos * d.S + d.W];
    }
    return -1;
}

// Last batch whose first workgroup is at or before blk (batches with no rows share the next one's start, so the
// last such batch is the one with rows): a 64-ary search by the whole wave over the ascending block starts, one load
// per lane and one ballot per round.
__device__ __forceinline__ int FindKeyBatch(const int64_t *starts, int nb, int64_t blk)
{
    int lo = 0, end = nb;  // starts[lo] <= blk, and the answer is in [lo, end)
    const int lane = LaneId();
    while (end - lo > 1) {
        const int step = (end - lo + 63) >> 6;
        const int idx = lo + lane * step;
        const unsigned long long m = __ballot(idx < end && starts[idx] <= blk);
        const int last = 63 - __builtin_clzll(m);  // lane 0 (idx = lo) is always set
        lo += last * step;
        end = min(end, lo + step);
    }
    return lo;
}

// Clears the state word of every dictionary record.
__global__ void join_dict_clear(unsigned long long *recs, int S, int W, int64_t records)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < records) recs[i * S + W] = 0;
}

constexpr int kKeyThreads = 256;         // lanes of a translation workgroup
constexpr int kKeyRowsPerThread = 4;     // rows per lane, kKeyThreads apart
constexpr int kKeyTile = kKeyThreads * kKeyRowsPerThread;  // rows per workgroup

// Translates the key columns of every batch into int64 join keys out[out_base + i] and a NULL bitmap per batch
// (32-bit words from null_word_base; each wave covers 64 consecutive rows of one batch per pass, so it writes whole
// words).  INSERT (build side of a dictionary mode) adds the keys to the dictionary; a lookup miss is a NULL join
// key.  starts[b] = batches[b].block_start; MW >= the key words W of the mode.
//
// Memory-level parallelism: a lane's rows are read column by column with every row's load issued before any is
// used (indices past the batch end are clamped to its last row and their results dropped), and on the probe side
// the home records of all its rows are loaded before any lookup resolves.
template <bool INSERT, int MW>
__global__ void __launch_bounds__(kKeyThreads) join_translate_keys(JoinKeyCols kc, const KeyBatch *batches,
                                                                   const int64_t *starts, int nb,
                                                                   int64_t uniform_blocks, KeyDict d,
                                                                   long long *out, unsigned int *null_words)
{
    constexpr int R = kKeyRowsPerThread;
    const int64_t blk = blockIdx.x;
    // batches of equal size (the last one may be shorter): the batch follows from the workgroup index
    const KeyBatch &b = batches[uniform_blocks > 0 ? static_cast<int>(min<int64_t>(blk / uniform_blocks, nb - 1))
                                                  : FindKeyBatch(starts, nb, blk)];
    const int64_t base = (blk - b.block_start) * kKeyTile + threadIdx.x;
    int64_t ci[R];
    bool live[R], isNull[R];
    unsigned long long w[R][MW];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int64_t i = base + k * kKeyThreads;
        live[k] = i < b.rows;
        ci[k] = live[k] ? i : b.rows - 1;  // a workgroup of a batch has at least one row
        isNull[k] = false;
#pragma unroll
        for (int q = 0; q < MW; ++q) w[k][q] = 0;
    }
    const bool dict = kc.mode != kKeyPacked32 && kc.mode != kKeyPacked64 && kc.mode != kKeyPacked128;
    int o = 0;  // next key word of a dictionary mode
#pragma unroll
    for (int c = 0; c < OMNI_GPU_MAX_KEYS; ++c) {
        if (c >= kc.n) break;
        const void *v = b.values[c];
        const uint8_t *nl = b.nulls[c];
        const int width = kc.width[c];
        unsigned long long x[R], y[R];  // value sign-extended (BYTE and BOOLEAN read as int8); y = high word
#pragma unroll
        for (int k = 0; k < R; ++k) {
            y[k] = 0;
            if (width == 8) x[k] = static_cast<const unsigned long long *>(v)[ci[k]];
            else if (width == 4) x[k] = static_cast<unsigned long long>(static_cast<long long>(static_cast<const int *>(v)[ci[k]]));
            else if (width == 2) x[k] = static_cast<unsigned long long>(static_cast<long long>(static_cast<const short *>(v)[ci[k]]));
            else if (width == 1) x[k] = static_cast<unsigned long long>(static_cast<long long>(static_cast<const int8_t *>(v)[ci[k]]));
            else {
                x[k] = static_cast<const unsigned long long *>(v)[2 * ci[k]];
                y[k] = static_cast<const unsigned long long *>(v)[2 * ci[k] + 1];
            }
        }
        if (nl) {
#pragma unroll
            for (int k = 0; k < R; ++k) isNull[k] |= (nl[ci[k] >> 3] >> (ci[k] & 7)) & 1;
        }
        const int bits = kc.bits[c];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (kc.mode == kKeyPacked32 || kc.mode == kKeyPacked64) {
                // key = key << bits | (KeyType)value; the int32 key keeps its low 32 bits (sign-extended at the end)
                w[k][0] = (w[k][0] << bits) | x[k];
            } else if (kc.mode == kKeyPacked128) {
                unsigned __int128 acc = (static_cast<unsigned __int128>(w[k][MW > 1 ? 1 : 0]) << 64) | w[k][0];
                acc = (acc << bits) | static_cast<unsigned __int128>(static_cast<__int128>(static_cast<long long>(x[k])));
                w[k][0] = static_cast<unsigned long long>(acc);
                if (MW > 1) w[k][MW > 1 ? 1 : 0] = static_cast<unsigned long long>(acc >> 64);
            } else if (width == 16) {
                PutWord<MW>(w[k], o, x[k]);
                PutWord<MW>(w[k], o + 1, y[k]);
            } else {  // the stored bytes, zero-extended
                PutWord<MW>(w[k], o, width == 8 ? x[k] : x[k] & ((1ULL << (8 * width)) - 1));
            }
        }
        o += width == 16 ? 2 : 1;
    }
    if (kc.mode == kKeyPacked32) {
#pragma unroll
        for (int k = 0; k < R; ++k) w[k][0] = static_cast<unsigned long long>(static_cast<long long>(static_cast<int>(w[k][0])));
    }
    const bool lookup = dict || kc.mode == kKeyPacked128;
    long long key[R];
    if (lookup) {
        unsigned long long home[R];
#pragma unroll
        for (int k = 0; k < R; ++k) home[k] = KeyHash<MW>(w[k], d.W) & d.mask;
        if (INSERT) {
#pragma unroll
            for (int k = 0; k < R; ++k) {
                key[k] = 0;
                if (live[k] && !isNull[k]) key[k] = DictInsert<MW>(d, w[k], home[k]);
            }
        } else {
            // every row's home record (state and the first key words) in flight before any compare
            unsigned long long st[R], r0[R], r1[R];
            const int sw = MW == 2 ? 2 : d.W;  // a mode of at most two key words has exactly two
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const unsigned long long *r = d.recs + home[k] * (MW == 2 ? 4 : d.S);
                r0[k] = r[0];
                r1[k] = r[1];
                st[k] = r[sw];
            }
#pragma unroll
            for (int k = 0; k < R; ++k) {
                key[k] = 0;
                if (!live[k] || isNull[k]) continue;
                long long id = -1;
                if (st[k] != 0) {
                    bool same = r0[k] == w[k][0] && (d.W < 2 || r1[k] == w[k][MW > 1 ? 1 : 0]);
                    if (same && d.W > 2) {
                        const unsigned long long *r = d.recs + home[k] * d.S;
#pragma unroll
                        for (int q = 2; q < MW; ++q) {
                            if (q < d.W) same &= r[q] == w[k][q];
                        }
                    }
                    id = same ? static_cast<long long>(home[k])
                              : DictFind<MW>(d, w[k], (home[k] + 1) & d.mask,
                                             d.recs[((home[k] + 1) & d.mask) * d.S + d.W]);
                }
                isNull[k] = id < 0;
                key[k] = id < 0 ? 0 : id;
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < R; ++k) key[k] = static_cast<long long>(w[k][0]);
    }
    const int lane = LaneId();
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int64_t i = base + k * kKeyThreads;
        if (live[k]) out[b.out_base + i] = isNull[k] ? 0 : key[k];
        // the wave covers batch rows [r0, r0 + 64), r0 a multiple of 64: word r0 / 32 and, if any of its rows
        // exist, word r0 / 32 + 1 (the batch's bitmap has ceil(rows / 32) words)
        const unsigned long long m = __ballot(live[k] && isNull[k]);
        const int64_t r0 = i - lane;
        if (lane == 0 && r0 < b.rows) null_words[b.null_word_base + (r0 >> 5)] = static_cast<unsigned int>(m);
        if (lane == 32 && r0 + 32 < b.rows) {
            null_words[b.null_word_base + (r0 >> 5) + 1] = static_cast<unsigned int>(m >> 32);
        }
    }
}

}  // namespace omni_gpu
