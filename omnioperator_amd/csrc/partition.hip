// Shuffle hash partitioning on gfx950: HashUtil::ComputePartitionIds (core/src/operator/hash_util.cpp:40-82),
// the partition id of each row for a hash exchange (PartitionedOutput).  Per row, a Murmur3 hash (seed 42,
// hash_util.h:56) is folded over the row's columns -- HashInt of a 4-byte value, HashLong of an 8-byte value
// (core/src/operator/util/mm3_util.h), a NULL leaving the hash unchanged -- and the id is
// Pmod(int32(hash), partitions).  The algorithm's steps and constants are the reference's, because the ids
// must equal the CPU operator's row for row.
//
// One streaming pass: each lane hashes kPartRows rows, column by column, with every load of a column issued
// before its hashes; reads 4 or 8 bytes per row per column (+ NULL bits) and writes 4 bytes per row.
#include "omni_gpu_internal.h"

namespace omni_gpu {

constexpr int kPartMaxCols = OMNI_GPU_MAX_AGGS;
constexpr int kPartBlock = 256;
constexpr int kPartRows = 4;  // rows per lane
constexpr unsigned int kMm3Seed = 42;

struct PartitionCols {
    int n;
    int width[kPartMaxCols];          // 1 (BOOLEAN), 4 or 8 bytes per value
    const void *values[kPartMaxCols];
    const uint8_t *nulls[kPartMaxCols];
};

__device__ __forceinline__ unsigned int Mm3Rotl(unsigned int i, unsigned int d) { return (i << d) | (i >> (32 - d)); }
__device__ __forceinline__ unsigned int Mm3MixK1(unsigned int k1) { return Mm3Rotl(k1 * 0xcc9e2d51u, 15) * 0x1b873593u; }
__device__ __forceinline__ unsigned int Mm3MixH1(unsigned int h1, unsigned int k1)
{
    return Mm3Rotl(h1 ^ k1, 13) * 5u + 0xe6546b64u;
}
__device__ __forceinline__ unsigned int Mm3Fmix(unsigned int h1, unsigned int length)
{
    h1 ^= length;
    h1 ^= h1 >> 16;
    h1 *= 0x85ebca6bu;
    h1 ^= h1 >> 13;
    h1 *= 0xc2b2ae35u;
    return h1 ^ (h1 >> 16);
}
__device__ __forceinline__ unsigned int Mm3HashInt(unsigned int v, unsigned int seed)
{
    return Mm3Fmix(Mm3MixH1(seed, Mm3MixK1(v)), 4);
}
__device__ __forceinline__ unsigned int Mm3HashLong(unsigned long long v, unsigned int seed)
{
    unsigned int h1 = Mm3MixH1(seed, Mm3MixK1(static_cast<unsigned int>(v)));
    h1 = Mm3MixH1(h1, Mm3MixK1(static_cast<unsigned int>(v >> 32)));
    return Mm3Fmix(h1, 8);
}

__global__ void __launch_bounds__(kPartBlock) partition_ids(PartitionCols c, int64_t rows, int partitions, int *out)
{
    const int64_t first = static_cast<int64_t>(blockIdx.x) * (kPartBlock * kPartRows) + threadIdx.x;
    unsigned int h[kPartRows];
    bool live[kPartRows];
    int64_t row[kPartRows];
#pragma unroll
    for (int j = 0; j < kPartRows; ++j) {
        row[j] = first + j * kPartBlock;
        live[j] = row[j] < rows;
        row[j] = live[j] ? row[j] : rows - 1;  // rows >= 1: out-of-range lanes re-read the last row
        h[j] = kMm3Seed;
    }
    for (int k = 0; k < c.n; ++k) {  // column type and NULL bitmap presence are uniform
        if (c.width[k] == 8) {
            const unsigned long long *v = static_cast<const unsigned long long *>(c.values[k]);
            unsigned long long x[kPartRows];
#pragma unroll
            for (int j = 0; j < kPartRows; ++j) x[j] = v[row[j]];
#pragma unroll
            for (int j = 0; j < kPartRows; ++j) {
                if (c.nulls[k] == nullptr || !NullBit(c.nulls[k], row[j])) h[j] = Mm3HashLong(x[j], h[j]);
            }
        } else if (c.width[k] == 4) {
            const unsigned int *v = static_cast<const unsigned int *>(c.values[k]);
            unsigned int x[kPartRows];
#pragma unroll
            for (int j = 0; j < kPartRows; ++j) x[j] = v[row[j]];
#pragma unroll
            for (int j = 0; j < kPartRows; ++j) {
                if (c.nulls[k] == nullptr || !NullBit(c.nulls[k], row[j])) h[j] = Mm3HashInt(x[j], h[j]);
            }
        } else {  // BOOLEAN: HashInt of 1 or 0 (Mm3Boolean)
            const uint8_t *v = static_cast<const uint8_t *>(c.values[k]);
            uint8_t x[kPartRows];
#pragma unroll
            for (int j = 0; j < kPartRows; ++j) x[j] = v[row[j]];
#pragma unroll
            for (int j = 0; j < kPartRows; ++j) {
                if (c.nulls[k] == nullptr || !NullBit(c.nulls[k], row[j])) h[j] = Mm3HashInt(x[j] != 0 ? 1u : 0u, h[j]);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < kPartRows; ++j) {
        if (!live[j]) continue;
        const int r = static_cast<int>(h[j]) % partitions;  // Pmod (mm3_util.h:413-420)
        out[row[j]] = r < 0 ? r + partitions : r;
    }
}

}  // namespace omni_gpu

using namespace omni_gpu;

extern "C" {

int omni_gpu_partition_ids(const omni_gpu_column *cols, int32_t num_cols, int64_t rows, int32_t num_partitions,
                           int32_t *out_ids)
{
    OG_CHECK(num_cols >= 1 && num_cols <= kPartMaxCols, OMNI_GPU_ERR_INVALID, "1 to %d columns", kPartMaxCols);
    OG_CHECK(cols, OMNI_GPU_ERR_INVALID, "null columns");
    OG_CHECK(num_partitions >= 1, OMNI_GPU_ERR_INVALID, "num_partitions must be at least 1");
    OG_CHECK(rows >= 0, OMNI_GPU_ERR_INVALID, "negative row count");
    OG_CHECK(rows == 0 || out_ids, OMNI_GPU_ERR_INVALID, "null out_ids");
    PartitionCols c{};
    c.n = num_cols;
    for (int k = 0; k < num_cols; ++k) {
        const int t = cols[k].type_id;
        switch (t) {
            case OMNI_TYPE_INT:
            case OMNI_TYPE_DATE32: c.width[k] = 4; break;
            case OMNI_TYPE_LONG:
            case OMNI_TYPE_TIMESTAMP:
            case OMNI_TYPE_DECIMAL64:
            case OMNI_TYPE_DOUBLE: c.width[k] = 8; break;
            case OMNI_TYPE_BOOLEAN: c.width[k] = 1; break;
            default:
                // SHORT / BYTE: Mm3Int reads them as 4-byte values (mm3_util.h); DECIMAL128 and strings hash byte
                // strings -- none is on the GPU path
                return SetError(OMNI_GPU_ERR_INVALID, "column %d: type %d is not supported by the GPU partitioner",
                                k, t);
        }
        OG_CHECK(rows == 0 || cols[k].values, OMNI_GPU_ERR_INVALID, "column %d has no values", k);
        c.values[k] = cols[k].values;
        c.nulls[k] = cols[k].nulls;
    }
    if (rows == 0) return OMNI_GPU_OK;
    const int64_t tile = static_cast<int64_t>(kPartBlock) * kPartRows;
    const int64_t blocks = (rows + tile - 1) / tile;
    OG_CHECK(blocks < 0x7FFFFFFF, OMNI_GPU_ERR_INVALID, "too many rows for one launch");
    hipStream_t s;
    OG_TRY(AcquireStream(&s));
    int rc = OMNI_GPU_OK;
    {
        KernelScope ks(s, "partition_ids");
        partition_ids<<<static_cast<unsigned int>(blocks), kPartBlock, 0, s>>>(c, rows, num_partitions, out_ids);
        rc = CheckLaunch("partition_ids");
    }
    if (rc == OMNI_GPU_OK) rc = SyncStream(s);
    ReleaseStream(s);
    return rc;
}

}  // extern "C"
