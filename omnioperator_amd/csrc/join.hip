// Inner hash join (single equality key) on gfx950.
//
// Build replaces HashBuilderOperator's end-of-input table build (core/src/operator/join/hash_builder.cpp:182-213,
// join_hash_table_variants.cpp:254-283, 695-808): the same choice between a direct-index array table (dense
// integer keys: range <= 8 * 2^ceil(log2(n / 0.75)) and <= UINT32_MAX) and a hash table, NULL build keys skipped.
// Probe replaces LookupJoinOperator::ProbeBatchForInnerJoin + the build-column gather
// (core/src/operator/join/lookup_join.cpp:924-1035, 1657-1779).
//
// Device layout:
//   build store   keys[n] (int64, sign-extended), key_null[n] (u8), payload values / NULL bytes per column
//   hash table    slots[cap] {i64 key; i32 head; u32 inline payload} (16 B, one cache-line read per probe step),
//                 linear probing at load <= 0.5; duplicate keys chain through next[n]
//   array table   heads[range] (i32 build row, -1 none) + next[n]
// Probe: one probe row per lane; matches are compacted with a wave-level exclusive scan and one atomic
// per wave on the output counter, so every wave writes its output rows contiguously.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

#include "omni_gpu_internal.h"

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
};

constexpr int kBlock = 256;
inline unsigned int Blocks(int64_t n) { return static_cast<unsigned int>((n + kBlock - 1) / kBlock); }

__global__ void join_load_keys(const void *src, int width, const uint8_t *nulls, int64_t rows, long long *dst,
                               uint8_t *dst_null)
{
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    bool isNull = NullBit(nulls, i);
    dst[i] = isNull ? 0 : static_cast<long long>(LoadWord(src, width, i));
    dst_null[i] = isNull ? 1 : 0;
}

__global__ void unpack_null_bytes(const uint8_t *bits, int64_t rows, uint8_t *dst)
{
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < rows) dst[i] = NullBit(bits, i) ? 1 : 0;
}

constexpr int kMinMaxRows = 16384;  // rows per workgroup of the min/max reduction

// Min/max of the non-NULL build keys: each workgroup reduces kMinMaxRows rows, then one atomic pair.
__global__ void __launch_bounds__(256) join_minmax(const long long *keys, const uint8_t *knull, int64_t n, long long *mm)
{
    __shared__ long long slo[4], shi[4];
    long long lo = std::numeric_limits<long long>::max(), hi = std::numeric_limits<long long>::min();
    int64_t start = static_cast<int64_t>(blockIdx.x) * kMinMaxRows;
    int64_t end = start + kMinMaxRows < n ? start + kMinMaxRows : n;
    for (int64_t i = start + threadIdx.x; i < end; i += 256) {
        if (!knull[i]) {
            lo = min(lo, keys[i]);
            hi = max(hi, keys[i]);
        }
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
        for (int w = 1; w < 4; ++w) {
            lo = min(lo, slo[w]);
            hi = max(hi, shi[w]);
        }
        atomicMin(&mm[0], lo);
        atomicMax(&mm[1], hi);
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
                                unsigned long long mask, int *next, int *sentinel_head, unsigned int *has_dup,
                                const void *inline_pay, int inline_width)
{
    int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (row >= n || knull[row]) return;
    long long key = keys[row];
    int old;
    if (static_cast<unsigned long long>(key) == kEmptyKey) {
        old = atomicExch(sentinel_head, static_cast<int>(row));
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

template <bool ARRAY>
__device__ __forceinline__ int Lookup(const TableDev &t, long long key, unsigned int &pay)
{
    if (ARRAY) {
        unsigned long long idx = static_cast<unsigned long long>(key) - static_cast<unsigned long long>(t.min_key);
        return idx < t.range ? t.heads[idx] : -1;
    }
    if (static_cast<unsigned long long>(key) == kEmptyKey) return *t.sentinel_head;
    unsigned long long pos = Mix64(static_cast<unsigned long long>(key)) & t.mask;
    while (true) {
        JoinSlot s = t.slots[pos];
        if (s.key == key) {
            pay = s.pay;
            return s.head;
        }
        if (static_cast<unsigned long long>(s.key) == kEmptyKey) return -1;
        pos = (pos + 1) & t.mask;
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
template <bool ARRAY, bool INLINE, bool DUPS>
__global__ void __launch_bounds__(kBlock) join_probe(ProbeArgs a, TableDev t)
{
    __shared__ unsigned long long sh[kBlock / 64 + 1];
    const ProbeBatch &b = a.nb == 1 ? a.single : a.batches[FindBatch(a, blockIdx.x)];
    const int64_t tile0 = (static_cast<int64_t>(blockIdx.x) - b.block_start) * kProbeTile;
    int head[kProbeRows];
    unsigned int pay[kProbeRows];
#pragma unroll
    for (int j = 0; j < kProbeRows; ++j) {
        int64_t row = tile0 + j * kBlock + threadIdx.x;
        head[j] = -1;
        pay[j] = 0;
        if (row < b.rows && !NullBit(b.nulls, row)) {
            head[j] = Lookup<ARRAY>(t, static_cast<long long>(LoadWord(b.keys, b.width, row)), pay[j]);
        }
    }
    const long long rowBase = b.row_base + tile0 + threadIdx.x;
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
                WriteMatch<INLINE>(a, static_cast<long long>(base + __popcll(m[j] & lt)), rowBase + j * kBlock, head[j],
                                   pay[j]);
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
        for (int h = head[j]; h >= 0; h = t.next[h]) WriteMatch<false>(a, static_cast<long long>(o++), rowBase + j * kBlock, h, 0);
    }
}

// ---- host side ----------------------------------------------------------------------------------------------

struct JoinHandle {
    omni_gpu_join_spec spec{};
    hipStream_t stream = nullptr;
    int64_t n = 0;             // build rows
    int64_t cap_rows = 0;      // allocated build-store rows
    DevBuf keys, knull;
    DevBuf pay[OMNI_GPU_MAX_AGGS], pay_null[OMNI_GPU_MAX_AGGS];
    bool pay_nullable[OMNI_GPU_MAX_AGGS] = {};
    bool finished = false;
    bool is_array = false;
    bool inline_pay = false;
    bool has_dup = false;
    int64_t max_dup = 0;
    uint64_t slots_n = 0;
    long long min_key = 0;
    unsigned long long range = 0;
    DevBuf slots, heads, next, scratch;  // scratch: [0] sentinel head i32, [4] has_dup u32, [8] max_len u32,
                                         //          [16] min/max i64[2], [32] match counter u64
    DevBuf batch_desc;
};

int EnsureBuildCapacity(JoinHandle &h, int64_t need)
{
    if (need <= h.cap_rows) return OMNI_GPU_OK;
    int64_t cap = std::max<int64_t>(need, std::max<int64_t>(h.cap_rows * 2, 65536));
    hipStream_t s = h.stream;
    OG_TRY(h.keys.Grow(sizeof(long long) * cap, sizeof(long long) * h.n, s));
    OG_TRY(h.knull.Grow(cap, h.n, s));
    for (int p = 0; p < h.spec.num_payloads; ++p) {
        int w = TypeWidth(h.spec.payload_types[p]);
        OG_TRY(h.pay[p].Grow(static_cast<size_t>(w) * cap, static_cast<size_t>(w) * h.n, s));
        OG_TRY(h.pay_null[p].Grow(cap, h.n, s));
    }
    h.cap_rows = cap;
    return OMNI_GPU_OK;
}

int RunProbe(JoinHandle &h, const ProbeBatch *batches, int nb, int64_t total_blocks, long long *out_rows,
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
    for (int p = 0; p < a.npay; ++p) {
        OG_CHECK(outs && outs[p].type_id == h.spec.payload_types[p], OMNI_GPU_ERR_INVALID, "payload output %d type mismatch", p);
        OG_CHECK(!h.pay_nullable[p] || cap == 0 || !outs[p].values || outs[p].nulls, OMNI_GPU_ERR_INVALID,
                 "payload %d has NULLs: its output needs a null bitmap", p);
        a.pay[p].width = TypeWidth(h.spec.payload_types[p]);
        a.pay[p].values = h.pay[p].p;
        a.pay[p].nulls = h.pay_nullable[p] ? h.pay_null[p].as<uint8_t>() : nullptr;
        a.pay[p].out = outs[p].values;
        a.pay[p].out_nulls = outs[p].nulls;
        if (outs[p].values && outs[p].nulls && cap > 0) {
            int64_t words = (cap + 31) / 32;
            KernelScope ks(s, "join_clear_nulls");
            join_fill_i32<<<Blocks(words), kBlock, 0, s>>>(reinterpret_cast<int *>(outs[p].nulls), 0, words);
            OG_TRY(CheckLaunch("join_fill_i32"));
        }
    }
    if (nb == 1) {
        a.single = batches[0];
        a.batches = nullptr;
    } else {
        OG_TRY(h.batch_desc.Reserve(sizeof(ProbeBatch) * nb, s));
        OG_HIP(hipMemcpyAsync(h.batch_desc.p, batches, sizeof(ProbeBatch) * nb, hipMemcpyHostToDevice, s));
        a.batches = h.batch_desc.as<ProbeBatch>();
    }
    OG_HIP(hipMemsetAsync(a.counter, 0, 8, s));
    TableDev t{};
    t.slots = h.slots.as<JoinSlot>();
    t.mask = h.slots_n ? h.slots_n - 1 : 0;
    t.heads = h.heads.as<int>();
    t.min_key = h.min_key;
    t.range = h.range;
    t.next = h.next.as<int>();
    t.sentinel_head = reinterpret_cast<const int *>(h.scratch.as<uint8_t>());
    if (total_blocks > 0 && h.n > 0) {
        KernelScope ks(s, "join_probe");
        dim3 grid(static_cast<unsigned int>(total_blocks));
        if (h.is_array) {
            if (h.has_dup) join_probe<true, false, true><<<grid, kBlock, 0, s>>>(a, t);
            else join_probe<true, false, false><<<grid, kBlock, 0, s>>>(a, t);
        } else if (h.has_dup) {
            join_probe<false, false, true><<<grid, kBlock, 0, s>>>(a, t);
        } else if (h.inline_pay) {
            join_probe<false, true, false><<<grid, kBlock, 0, s>>>(a, t);
        } else {
            join_probe<false, false, false><<<grid, kBlock, 0, s>>>(a, t);
        }
        OG_TRY(CheckLaunch("join_probe"));
    }
    unsigned long long m = 0;
    OG_HIP(hipMemcpyAsync(&m, a.counter, 8, hipMemcpyDeviceToHost, s));
    OG_HIP(hipStreamSynchronize(s));
    *num_matches = static_cast<int64_t>(m);
    if (static_cast<int64_t>(m) > cap) {
        return SetError(OMNI_GPU_ERR_OUTPUT_FULL, "join produced %llu matches for an output capacity of %lld", m,
                        static_cast<long long>(cap));
    }
    return OMNI_GPU_OK;
}

int FillBatch(ProbeBatch &b, const omni_gpu_column &c, int64_t rows, int64_t row_base, int64_t block_start,
              int32_t key_type)
{
    OG_CHECK(c.type_id == key_type, OMNI_GPU_ERR_INVALID, "probe key type %d != build key type %d", c.type_id, key_type);
    OG_CHECK(rows == 0 || c.values, OMNI_GPU_ERR_INVALID, "probe key has no values");
    OG_CHECK(rows >= 0, OMNI_GPU_ERR_INVALID, "negative row count");
    b.keys = c.values;
    b.nulls = c.nulls;
    b.rows = rows;
    b.row_base = row_base;
    b.block_start = block_start;
    b.width = TypeWidth(c.type_id);
    return OMNI_GPU_OK;
}

}  // namespace omni_gpu

using namespace omni_gpu;

extern "C" {

int omni_gpu_join_build_create(const omni_gpu_join_spec *spec, void **handle)
{
    OG_CHECK(spec && handle, OMNI_GPU_ERR_INVALID, "null argument");
    OG_CHECK(IsIntegerKeyType(spec->key_type), OMNI_GPU_ERR_INVALID,
             "join key type %d is not supported (integer-like single keys only)", spec->key_type);
    OG_CHECK(spec->num_payloads >= 0 && spec->num_payloads <= OMNI_GPU_MAX_AGGS, OMNI_GPU_ERR_INVALID,
             "num_payloads out of range");
    for (int p = 0; p < spec->num_payloads; ++p) {
        OG_CHECK(TypeWidth(spec->payload_types[p]) > 0, OMNI_GPU_ERR_INVALID, "payload type %d not supported",
                 spec->payload_types[p]);
    }
    std::unique_ptr<JoinHandle> h(new JoinHandle());
    h->spec = *spec;
    OG_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
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
    OG_CHECK(key->type_id == h.spec.key_type, OMNI_GPU_ERR_INVALID, "build key type mismatch");
    OG_CHECK(key->values, OMNI_GPU_ERR_INVALID, "build key has no values");
    OG_CHECK(h.spec.num_payloads == 0 || payloads, OMNI_GPU_ERR_INVALID, "null payload columns");
    OG_TRY(EnsureBuildCapacity(h, h.n + rows));
    hipStream_t s = h.stream;
    {
        KernelScope ks(s, "join_load_keys");
        join_load_keys<<<Blocks(rows), kBlock, 0, s>>>(key->values, TypeWidth(key->type_id), key->nulls, rows,
                                                       h.keys.as<long long>() + h.n, h.knull.as<uint8_t>() + h.n);
        OG_TRY(CheckLaunch("join_load_keys"));
    }
    for (int p = 0; p < h.spec.num_payloads; ++p) {
        OG_CHECK(payloads[p].type_id == h.spec.payload_types[p], OMNI_GPU_ERR_INVALID, "payload %d type mismatch", p);
        OG_CHECK(payloads[p].values, OMNI_GPU_ERR_INVALID, "payload %d has no values", p);
        int w = TypeWidth(h.spec.payload_types[p]);
        OG_HIP(hipMemcpyAsync(h.pay[p].as<uint8_t>() + static_cast<size_t>(w) * h.n, payloads[p].values,
                              static_cast<size_t>(w) * rows, hipMemcpyDeviceToDevice, s));
        if (payloads[p].nulls) {
            h.pay_nullable[p] = true;
            KernelScope ks(s, "join_unpack_nulls");
            unpack_null_bytes<<<Blocks(rows), kBlock, 0, s>>>(payloads[p].nulls, rows, h.pay_null[p].as<uint8_t>() + h.n);
            OG_TRY(CheckLaunch("unpack_null_bytes"));
        } else {
            OG_HIP(hipMemsetAsync(h.pay_null[p].as<uint8_t>() + h.n, 0, rows, s));
        }
    }
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
    int init_scratch[4] = {-1, 0, 0, 0};
    OG_HIP(hipMemcpyAsync(sc, init_scratch, sizeof(init_scratch), hipMemcpyHostToDevice, s));
    h.finished = true;
    if (h.n == 0) {
        h.max_dup = 0;
        return OMNI_GPU_OK;
    }
    OG_TRY(h.next.Reserve(sizeof(int) * h.n, s));
    // initDegree and the array-table rule of JoinHashTableVariants::BuildHashTable / TryToBuildArrayTable.
    bool arrayOk = false;
    if (h.spec.strategy == 0) {
        long long mm[2] = {std::numeric_limits<long long>::max(), std::numeric_limits<long long>::min()};
        long long *dmm = reinterpret_cast<long long *>(sc + 16);
        OG_HIP(hipMemcpyAsync(dmm, mm, sizeof(mm), hipMemcpyHostToDevice, s));
        {
            KernelScope ks(s, "join_minmax");
            join_minmax<<<static_cast<unsigned int>((h.n + kMinMaxRows - 1) / kMinMaxRows), 256, 0, s>>>(h.keys.as<long long>(), h.knull.as<uint8_t>(), h.n, dmm);
            OG_TRY(CheckLaunch("join_minmax"));
        }
        OG_HIP(hipMemcpyAsync(mm, dmm, sizeof(mm), hipMemcpyDeviceToHost, s));
        OG_HIP(hipStreamSynchronize(s));
        if (mm[1] >= mm[0]) {
            double degree = std::ceil(std::log2(static_cast<double>(h.n) / 0.75));
            long double len = std::pow(2.0L, static_cast<long double>(degree));
            long double range = static_cast<long double>(mm[1]) - static_cast<long double>(mm[0]);
            arrayOk = range <= 4294967295.0L && range <= 8.0L * len;
            if (arrayOk) {
                h.min_key = mm[0];
                h.range = static_cast<unsigned long long>(mm[1] - mm[0]) + 1;
            }
        }
    }
    unsigned int *has_dup = reinterpret_cast<unsigned int *>(sc + 4);
    unsigned int *max_len = reinterpret_cast<unsigned int *>(sc + 8);
    if (arrayOk) {
        h.is_array = true;
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
        h.slots_n = 1024;
        while (h.slots_n < static_cast<uint64_t>(2 * h.n)) h.slots_n <<= 1;
        OG_TRY(h.slots.Reserve(sizeof(JoinSlot) * h.slots_n, s));
        {
            KernelScope ks(s, "join_init_table");
            join_init_slots<<<Blocks(static_cast<int64_t>(h.slots_n)), kBlock, 0, s>>>(h.slots.as<JoinSlot>(),
                                                                                      static_cast<int64_t>(h.slots_n));
            OG_TRY(CheckLaunch("join_init_slots"));
        }
        h.inline_pay = h.spec.num_payloads == 1 && TypeWidth(h.spec.payload_types[0]) <= 4 && !h.pay_nullable[0];
        KernelScope ks(s, "join_build");
        join_build_hash<<<Blocks(h.n), kBlock, 0, s>>>(h.keys.as<long long>(), h.knull.as<uint8_t>(), h.n,
                                                       h.slots.as<JoinSlot>(), h.slots_n - 1, h.next.as<int>(),
                                                       reinterpret_cast<int *>(sc), has_dup,
                                                       h.inline_pay ? h.pay[0].p : nullptr,
                                                       h.inline_pay ? TypeWidth(h.spec.payload_types[0]) : 0);
        OG_TRY(CheckLaunch("join_build_hash"));
    }
    unsigned int flags[2] = {0, 0};
    OG_HIP(hipMemcpyAsync(flags, has_dup, 8, hipMemcpyDeviceToHost, s));
    OG_HIP(hipStreamSynchronize(s));
    h.has_dup = flags[0] != 0;
    if (h.has_dup) {
        h.inline_pay = false;
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
        OG_HIP(hipMemcpyAsync(flags, has_dup, 8, hipMemcpyDeviceToHost, s));
        OG_HIP(hipStreamSynchronize(s));
        h.max_dup = flags[1];
    } else {
        h.max_dup = 1;
    }
    return OMNI_GPU_OK;
}

int omni_gpu_join_build_info(void *handle, int32_t *is_array, int64_t *max_dup, int64_t *table_slots)
{
    OG_CHECK(handle, OMNI_GPU_ERR_INVALID, "null handle");
    JoinHandle &h = *static_cast<JoinHandle *>(handle);
    OG_CHECK(h.finished, OMNI_GPU_ERR_STATE, "build side not finished");
    if (is_array) *is_array = h.is_array ? 1 : 0;
    if (max_dup) *max_dup = h.max_dup;
    if (table_slots) *table_slots = h.is_array ? static_cast<int64_t>(h.range) : static_cast<int64_t>(h.slots_n);
    return OMNI_GPU_OK;
}

int omni_gpu_join_probe(void *handle, const omni_gpu_column *probe_key, int64_t rows, int64_t probe_row_base,
                        int64_t *out_probe_rows, omni_gpu_out_column *out_payloads, int64_t out_capacity,
                        int64_t *num_matches)
{
    OG_CHECK(handle && probe_key, OMNI_GPU_ERR_INVALID, "null argument");
    JoinHandle &h = *static_cast<JoinHandle *>(handle);
    ProbeBatch b{};
    OG_TRY(FillBatch(b, *probe_key, rows, probe_row_base, 0, h.spec.key_type));
    return RunProbe(h, &b, 1, (rows + kProbeTile - 1) / kProbeTile, reinterpret_cast<long long *>(out_probe_rows), out_payloads, out_capacity,
                    num_matches);
}

int omni_gpu_join_probe_batches(void *handle, const omni_gpu_column *keys, const int64_t *rows, int32_t num_batches,
                                int64_t *out_probe_rows, omni_gpu_out_column *out_payloads, int64_t out_capacity,
                                int64_t *num_matches)
{
    OG_CHECK(handle && keys && rows && num_batches > 0, OMNI_GPU_ERR_INVALID, "null argument or no batches");
    JoinHandle &h = *static_cast<JoinHandle *>(handle);
    std::vector<ProbeBatch> bs(num_batches);
    int64_t blocks = 0, base = 0;
    for (int i = 0; i < num_batches; ++i) {
        OG_TRY(FillBatch(bs[i], keys[i], rows[i], base, blocks, h.spec.key_type));
        blocks += (rows[i] + kProbeTile - 1) / kProbeTile;
        base += rows[i];
    }
    OG_CHECK(blocks < 0x7FFFFFFF, OMNI_GPU_ERR_INVALID, "too many probe rows for one launch");
    return RunProbe(h, bs.data(), num_batches, blocks, reinterpret_cast<long long *>(out_probe_rows), out_payloads,
                    out_capacity, num_matches);
}

int omni_gpu_join_destroy(void *handle)
{
    if (!handle) return OMNI_GPU_OK;
    JoinHandle *h = static_cast<JoinHandle *>(handle);
    hipStream_t s = h->stream;
    h->keys.Release(s);
    h->knull.Release(s);
    for (int p = 0; p < OMNI_GPU_MAX_AGGS; ++p) {
        h->pay[p].Release(s);
        h->pay_null[p].Release(s);
    }
    h->slots.Release(s);
    h->heads.Release(s);
    h->next.Release(s);
    h->scratch.Release(s);
    h->batch_desc.Release(s);
    if (s) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
    delete h;
    return OMNI_GPU_OK;
}

}  // extern "C"
