// Device-wide exclusive scans of u32 values (three launches: per-tile scans, a scan of the tile sums, the
// tile offsets added back), shared by the partitioned join probe and the aggregation's output compaction.
#pragma once

#include "omni_gpu_internal.h"

namespace omni_gpu {

constexpr int kScanBlock = 1024;                 // threads of the scan kernels
constexpr int kScanTile = 2 * kScanBlock;        // elements per scan workgroup

// Exclusive scan of kScanTile consecutive u32 values per workgroup (in place); the workgroup's sum goes to
// sums[blockIdx.x].
static __global__ void __launch_bounds__(kScanBlock) scan_tiles(unsigned int *v, long long n, unsigned int *sums)
{
    __shared__ unsigned int wsum[kScanBlock / kWaveSize];
    const long long i0 = static_cast<long long>(blockIdx.x) * kScanTile + 2 * threadIdx.x;
    unsigned int x0 = i0 < n ? v[i0] : 0u;
    unsigned int x1 = i0 + 1 < n ? v[i0 + 1] : 0u;
    unsigned int mine = x0 + x1;
    const int lane = LaneId(), wave = WaveInBlock();
    unsigned int incl = mine;
#pragma unroll
    for (int d = 1; d < kWaveSize; d <<= 1) {
        unsigned int up = __shfl_up(incl, d, kWaveSize);
        if (lane >= d) incl += up;
    }
    if (lane == kWaveSize - 1) wsum[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int run = 0;
        for (int w = 0; w < kScanBlock / kWaveSize; ++w) {
            unsigned int t = wsum[w];
            wsum[w] = run;
            run += t;
        }
        sums[blockIdx.x] = run;
    }
    __syncthreads();
    unsigned int excl = wsum[wave] + incl - mine;
    if (i0 < n) v[i0] = excl;
    if (i0 + 1 < n) v[i0 + 1] = excl + x0;
}

// Exclusive scan of n tile sums in one workgroup (kScanTile at a time, carrying the running total);
// *total receives their sum.
static __global__ void __launch_bounds__(kScanBlock) scan_sums(unsigned int *sums, int n, unsigned int *total)
{
    __shared__ unsigned int wsum[kScanBlock / kWaveSize];
    __shared__ unsigned int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < n; base += kScanTile) {
        const int i0 = base + 2 * threadIdx.x;
        unsigned int x0 = i0 < n ? sums[i0] : 0u;
        unsigned int x1 = i0 + 1 < n ? sums[i0 + 1] : 0u;
        unsigned int mine = x0 + x1;
        const int lane = LaneId(), wave = WaveInBlock();
        unsigned int incl = mine;
#pragma unroll
        for (int d = 1; d < kWaveSize; d <<= 1) {
            unsigned int up = __shfl_up(incl, d, kWaveSize);
            if (lane >= d) incl += up;
        }
        if (lane == kWaveSize - 1) wsum[wave] = incl;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned int run = carry;
            for (int w = 0; w < kScanBlock / kWaveSize; ++w) {
                unsigned int t = wsum[w];
                wsum[w] = run;
                run += t;
            }
            carry = run;
        }
        __syncthreads();
        unsigned int excl = wsum[wave] + incl - mine;
        if (i0 < n) sums[i0] = excl;
        if (i0 + 1 < n) sums[i0 + 1] = excl + x0;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

static __global__ void __launch_bounds__(kScanBlock) scan_add(unsigned int *v, long long n, const unsigned int *sums)
{
    const long long i0 = static_cast<long long>(blockIdx.x) * kScanTile + 2 * threadIdx.x;
    const unsigned int add = sums[blockIdx.x];
    if (i0 < n) v[i0] += add;
    if (i0 + 1 < n) v[i0 + 1] += add;
}

constexpr int kFlagItems = 16;                        // flags per thread of the compaction kernels
constexpr int kFlagTile = kScanBlock * kFlagItems;    // flags per workgroup

// Number of non-zero flags in each tile of kFlagTile flags.
static __global__ void __launch_bounds__(kScanBlock) flag_tile_counts(const uint8_t *flags, long long n, unsigned int *counts)
{
    __shared__ unsigned int wsum[kScanBlock / kWaveSize];
    const long long i0 = static_cast<long long>(blockIdx.x) * kFlagTile + static_cast<long long>(threadIdx.x) * kFlagItems;
    unsigned int c = 0;
    for (int j = 0; j < kFlagItems; ++j) c += (i0 + j < n && flags[i0 + j]) ? 1u : 0u;
#pragma unroll
    for (int d = kWaveSize / 2; d > 0; d >>= 1) c += __shfl_xor(c, d, kWaveSize);
    if (LaneId() == 0) wsum[WaveInBlock()] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int t = 0;
        for (int w = 0; w < kScanBlock / kWaveSize; ++w) t += wsum[w];
        counts[blockIdx.x] = t;
    }
}

// Writes the indices of the non-zero flags in increasing order: tile t's first one goes to out[offs[t]].
static __global__ void __launch_bounds__(kScanBlock) flag_compact(const uint8_t *flags, long long n, const unsigned int *offs,
                                                           unsigned int *out)
{
    __shared__ unsigned int wsum[kScanBlock / kWaveSize];
    const long long i0 = static_cast<long long>(blockIdx.x) * kFlagTile + static_cast<long long>(threadIdx.x) * kFlagItems;
    unsigned int c = 0;
    for (int j = 0; j < kFlagItems; ++j) c += (i0 + j < n && flags[i0 + j]) ? 1u : 0u;
    const int lane = LaneId(), wave = WaveInBlock();
    unsigned int incl = c;
#pragma unroll
    for (int d = 1; d < kWaveSize; d <<= 1) {
        unsigned int up = __shfl_up(incl, d, kWaveSize);
        if (lane >= d) incl += up;
    }
    if (lane == kWaveSize - 1) wsum[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int run = 0;
        for (int w = 0; w < kScanBlock / kWaveSize; ++w) {
            unsigned int t = wsum[w];
            wsum[w] = run;
            run += t;
        }
    }
    __syncthreads();
    unsigned int pos = offs[blockIdx.x] + wsum[wave] + incl - c;
    for (int j = 0; j < kFlagItems; ++j) {
        if (i0 + j < n && flags[i0 + j]) out[pos++] = static_cast<unsigned int>(i0 + j);
    }
}

}  // namespace omni_gpu
