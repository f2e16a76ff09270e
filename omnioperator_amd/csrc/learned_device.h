// Device code of the learned strategy/capacity advisor: the per-column key sketch and a 16-64-64-4 fp32 MLP
// (inference, gradient and Adam step), plus the seeded key generator the advisor is trained on.
// No reference counterpart: the reference sizes its tables by growth only (base_hash_map.h:255-292).
#pragma once

#include <limits>

#include "omni_gpu_internal.h"

namespace omni_gpu {

// ---- MLP ---------------------------------------------------------------------------------------------------
// Parameters, flattened: W1[16x64] b1[64] W2[64x64] b2[64] W3[64x4] b3[4] (row-major, y = x @ W + b).
// Output 0 regresses log2 of the group count; outputs 1..3 are the strategy logits.
// Loss = mean (y0 - t)^2 + mean cross-entropy(softmax(y1..3), c).
constexpr int kIn = 16, kH = 64, kOut = 4;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kIn * kH, kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kH * kH,
              kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + kH * kOut, kParams = kOffB3 + kOut;  // 5508
constexpr int kMlpFwdBlock = 64;  // samples per workgroup of the inference kernel
constexpr int kMlpBlock = 32;     // samples per workgroup of the gradient kernel (LDS: activations of 32 samples)

// Forward pass of one sample held by the calling thread; pre-activations go to the thread's LDS rows.
__device__ __forceinline__ void MlpForward(const float *P, const float *x, float *a1, float *a2, float *y)
{
    for (int j = 0; j < kH; ++j) {
        float s = P[kOffB1 + j];
        for (int i = 0; i < kIn; ++i) s += x[i] * P[kOffW1 + i * kH + j];
        a1[j] = s;
    }
    for (int j = 0; j < kH; ++j) {
        float s = P[kOffB2 + j];
        for (int i = 0; i < kH; ++i) s += fmaxf(a1[i], 0.f) * P[kOffW2 + i * kH + j];
        a2[j] = s;
    }
    for (int j = 0; j < kOut; ++j) {
        float s = P[kOffB3 + j];
        for (int i = 0; i < kH; ++i) s += fmaxf(a2[i], 0.f) * P[kOffW3 + i * kOut + j];
        y[j] = s;
    }
}

__global__ void __launch_bounds__(kMlpFwdBlock) mlp_forward(const float *__restrict__ params,
                                                            const float *__restrict__ X, int B, float *__restrict__ Y)
{
    __shared__ float P[kParams];
    __shared__ float A1[kMlpFwdBlock][kH + 1], A2[kMlpFwdBlock][kH + 1];
    for (int i = threadIdx.x; i < kParams; i += kMlpFwdBlock) P[i] = params[i];
    __syncthreads();
    const int b = blockIdx.x * kMlpFwdBlock + threadIdx.x;
    if (b >= B) return;
    float x[kIn], y[kOut];
    for (int i = 0; i < kIn; ++i) x[i] = X[b * kIn + i];
    MlpForward(P, x, A1[threadIdx.x], A2[threadIdx.x], y);
    for (int j = 0; j < kOut; ++j) Y[b * kOut + j] = y[j];
}

// Inference of one sample by one wave (many samples, latency-bound otherwise): lane j computes hidden unit j of
// both hidden layers, reading the previous layer's outputs lane by lane (shuffles), and the four outputs are
// reduced across the wave; lane 0 writes y[0..kOut).  Same network and parameter layout as MlpForward; P and x may
// be in LDS or global memory.
__device__ __forceinline__ void MlpWave(const float *__restrict__ P, const float *__restrict__ x, float *__restrict__ y)
{
    static_assert(kH == kWaveSize, "one lane per hidden unit");
    const int lane = LaneId();
    float s = P[kOffB1 + lane];
    for (int i = 0; i < kIn; ++i) s += x[i] * P[kOffW1 + i * kH + lane];
    const float h1 = fmaxf(s, 0.f);
    float t = P[kOffB2 + lane];
#pragma unroll 16
    for (int i = 0; i < kH; ++i) t += __shfl(h1, i, kWaveSize) * P[kOffW2 + i * kH + lane];
    const float h2 = fmaxf(t, 0.f);
#pragma unroll
    for (int o = 0; o < kOut; ++o) {
        float v = h2 * P[kOffW3 + lane * kOut + o];
#pragma unroll
        for (int d = kWaveSize / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWaveSize);
        if (lane == 0) y[o] = v + P[kOffB3 + o];
    }
}

// Forward + backward of up to kMlpBlock samples per workgroup (sample k of the batch is row idx[k] of X / T / C,
// or row k when idx is null); writes the workgroup's summed parameter gradient (already divided by B) to
// partial[blockIdx.x][kParams] and its share of the loss to loss[blockIdx.x].
__global__ void __launch_bounds__(kMlpBlock) mlp_backward(const float *__restrict__ params, const float *__restrict__ X,
                                                          const float *__restrict__ T, const int *__restrict__ C,
                                                          const int *__restrict__ idx, int B,
                                                          float *__restrict__ partial, float *__restrict__ loss)
{
    __shared__ float P[kParams];
    __shared__ float XS[kMlpBlock][kIn + 1];
    __shared__ float H1[kMlpBlock][kH + 1], H2[kMlpBlock][kH + 1];  // pre-activations, then relu of them
    __shared__ float D1[kMlpBlock][kH + 1], D2[kMlpBlock][kH + 1];  // dL/da1, dL/da2
    __shared__ float DY[kMlpBlock][kOut];
    __shared__ float LS[kMlpBlock];
    const int t = threadIdx.x;
    for (int i = t; i < kParams; i += kMlpBlock) P[i] = params[i];
    __syncthreads();
    const int b = blockIdx.x * kMlpBlock + t;
    const bool live = b < B;
    const int row = live ? (idx ? idx[b] : b) : 0;
    const float invB = 1.f / static_cast<float>(B);
    float x[kIn], y[kOut];
    for (int i = 0; i < kIn; ++i) {
        x[i] = live ? X[static_cast<size_t>(row) * kIn + i] : 0.f;
        XS[t][i] = x[i];
    }
    MlpForward(P, x, H1[t], H2[t], y);
    float l = 0.f;
    if (live) {
        const float tt = T[row];
        const int cc = C[row];
        const float d0 = y[0] - tt;
        l = d0 * d0;
        const float m = fmaxf(y[1], fmaxf(y[2], y[3]));
        const float e1 = expf(y[1] - m), e2 = expf(y[2] - m), e3 = expf(y[3] - m);
        const float sum = e1 + e2 + e3;
        const float zc = (cc == 0 ? y[1] : (cc == 1 ? y[2] : y[3])) - m;
        l += logf(sum) - zc;
        DY[t][0] = 2.f * d0 * invB;
        DY[t][1] = (e1 / sum - (cc == 0 ? 1.f : 0.f)) * invB;
        DY[t][2] = (e2 / sum - (cc == 1 ? 1.f : 0.f)) * invB;
        DY[t][3] = (e3 / sum - (cc == 2 ? 1.f : 0.f)) * invB;
    } else {
        for (int j = 0; j < kOut; ++j) DY[t][j] = 0.f;
    }
    LS[t] = l * invB;
    // dL/da2 = (dy @ W3^T) * relu'(a2); then h2 = relu(a2) in place
    for (int i = 0; i < kH; ++i) {
        float s = 0.f;
        for (int j = 0; j < kOut; ++j) s += DY[t][j] * P[kOffW3 + i * kOut + j];
        const float a = H2[t][i];
        D2[t][i] = a > 0.f ? s : 0.f;
        H2[t][i] = fmaxf(a, 0.f);
    }
    // dL/da1 = (dL/da2 @ W2^T) * relu'(a1); then h1 = relu(a1) in place
    for (int i = 0; i < kH; ++i) {
        float s = 0.f;
        for (int j = 0; j < kH; ++j) s += D2[t][j] * P[kOffW2 + i * kH + j];
        const float a = H1[t][i];
        D1[t][i] = a > 0.f ? s : 0.f;
        H1[t][i] = fmaxf(a, 0.f);
    }
    __syncthreads();
    // parameter gradients: each thread owns a strided set of parameters and sums over the workgroup's samples
    float *out = partial + static_cast<size_t>(blockIdx.x) * kParams;
    for (int p = t; p < kParams; p += kMlpBlock) {
        float g = 0.f;
        if (p < kOffB1) {                       // W1[i][j] = sum x_i dL/da1_j
            const int i = p / kH, j = p % kH;
            for (int k = 0; k < kMlpBlock; ++k) g += XS[k][i] * D1[k][j];
        } else if (p < kOffW2) {                // b1[j]
            const int j = p - kOffB1;
            for (int k = 0; k < kMlpBlock; ++k) g += D1[k][j];
        } else if (p < kOffB2) {                // W2[i][j] = sum h1_i dL/da2_j
            const int q = p - kOffW2, i = q / kH, j = q % kH;
            for (int k = 0; k < kMlpBlock; ++k) g += H1[k][i] * D2[k][j];
        } else if (p < kOffW3) {                // b2[j]
            const int j = p - kOffB2;
            for (int k = 0; k < kMlpBlock; ++k) g += D2[k][j];
        } else if (p < kOffB3) {                // W3[i][j] = sum h2_i dy_j
            const int q = p - kOffW3, i = q / kOut, j = q % kOut;
            for (int k = 0; k < kMlpBlock; ++k) g += H2[k][i] * DY[k][j];
        } else {                                // b3[j]
            const int j = p - kOffB3;
            for (int k = 0; k < kMlpBlock; ++k) g += DY[k][j];
        }
        out[p] = g;
    }
    if (t == 0) {
        float s = 0.f;
        for (int k = 0; k < kMlpBlock; ++k) s += LS[k];
        loss[blockIdx.x] = s;
    }
}

// Sums the workgroup partials in a fixed order (deterministic), optionally stores the gradient, and applies
// one Adam step (bias-corrected, step >= 1) when lr > 0.
__global__ void mlp_reduce_adam(const float *__restrict__ partial, int nparts, const float *__restrict__ loss_parts,
                                float *__restrict__ params, float *__restrict__ m, float *__restrict__ v,
                                float *__restrict__ grad_out, float *__restrict__ loss_out, float lr, float beta1,
                                float beta2, float eps, int step)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0 && loss_out) {
        float s = 0.f;
        for (int k = 0; k < nparts; ++k) s += loss_parts[k];
        *loss_out = s;
    }
    if (p >= kParams) return;
    float g = 0.f;
    for (int k = 0; k < nparts; ++k) g += partial[static_cast<size_t>(k) * kParams + p];
    if (grad_out) grad_out[p] = g;
    if (lr > 0.f) {
        float mm = beta1 * m[p] + (1.f - beta1) * g;
        float vv = beta2 * v[p] + (1.f - beta2) * g * g;
        m[p] = mm;
        v[p] = vv;
        float mh = mm / (1.f - powf(beta1, static_cast<float>(step)));
        float vh = vv / (1.f - powf(beta2, static_cast<float>(step)));
        params[p] -= lr * mh / (sqrtf(vh) + eps);
    }
}

// ---- key sketch --------------------------------------------------------------------------------------------
// A 1-in-`stride` row sample of the key columns is counted in a hash table of 64-bit key hashes (count per
// distinct sampled key); the reduction then gives the distinct count, the singleton / doubleton counts, the
// top key's frequency and the mass of keys seen at least 4 times, plus the sampled range and NULL count.

struct SketchStats {
    unsigned long long distinct, f1, f2, maxc, heavy, nulls, sampled;
    long long kmin, kmax;
};

struct SketchArgs {
    const void *values[2];
    const uint8_t *nulls[2];
    int width[2];
    int ncols;
    int64_t rows;
    int64_t stride;
    int64_t m;                    // sampled rows
    unsigned long long *hkeys;    // table of hashes (0 = empty; hash 0 is remapped to 1)
    unsigned int *counts;
    unsigned long long mask;
    SketchStats *stats;
};

__global__ void __launch_bounds__(256) sketch_insert(SketchArgs a)
{
    __shared__ long long smin[4], smax[4];
    __shared__ unsigned int snull[4];
    long long lo = std::numeric_limits<long long>::max(), hi = std::numeric_limits<long long>::min();
    unsigned int nn = 0;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i < a.m) {
        const int64_t row = i * a.stride;
        unsigned long long h = 0x243F6A8885A308D3ULL;
        bool anyNull = false;
        for (int c = 0; c < a.ncols; ++c) {
            bool isNull = NullBit(a.nulls[c], row);
            anyNull |= isNull;
            unsigned long long w = isNull ? 0x9E3779B97F4A7C15ULL : LoadWord(a.values[c], a.width[c], row);
            if (c == 0 && !isNull) {
                lo = static_cast<long long>(w);
                hi = static_cast<long long>(w);
            }
            h = Mix64(h ^ (w + 0x632BE59BD9B4E019ULL * (c + 1) + (isNull ? 1ULL : 0ULL)));
        }
        nn = anyNull ? 1u : 0u;
        if (h == 0) h = 1;
        unsigned long long pos = h & a.mask;
        while (true) {
            unsigned long long k = a.hkeys[pos];
            if (k == 0) {
                k = atomicCAS(&a.hkeys[pos], 0ULL, h);
                if (k == 0) k = h;
            }
            if (k == h) {
                atomicAdd(&a.counts[pos], 1u);
                break;
            }
            pos = (pos + 1) & a.mask;
        }
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        lo = min(lo, __shfl_xor(lo, d, 64));
        hi = max(hi, __shfl_xor(hi, d, 64));
        nn += __shfl_xor(nn, d, 64);
    }
    if (LaneId() == 0) {
        smin[threadIdx.x >> 6] = lo;
        smax[threadIdx.x >> 6] = hi;
        snull[threadIdx.x >> 6] = nn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int tn = 0;
        for (int w = 0; w < 4; ++w) {
            lo = min(lo, smin[w]);
            hi = max(hi, smax[w]);
            tn += snull[w];
        }
        if (lo <= hi) {
            atomicMin(&a.stats->kmin, lo);
            atomicMax(&a.stats->kmax, hi);
        }
        if (tn) atomicAdd(&a.stats->nulls, static_cast<unsigned long long>(tn));
    }
}

__global__ void __launch_bounds__(256) sketch_reduce(SketchArgs a, unsigned long long cap)
{
    __shared__ unsigned long long part[5][4];
    unsigned long long d = 0, f1 = 0, f2 = 0, mx = 0, heavy = 0;
    const unsigned long long i = static_cast<unsigned long long>(blockIdx.x) * 256 + threadIdx.x;
    if (i < cap) {
        unsigned int c = a.counts[i];
        d = c ? 1 : 0;
        f1 = c == 1 ? 1 : 0;
        f2 = c == 2 ? 1 : 0;
        mx = c;
        heavy = c >= 4 ? c : 0;
    }
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
        d += __shfl_xor(d, s, 64);
        f1 += __shfl_xor(f1, s, 64);
        f2 += __shfl_xor(f2, s, 64);
        heavy += __shfl_xor(heavy, s, 64);
        unsigned long long o = __shfl_xor(mx, s, 64);
        mx = o > mx ? o : mx;
    }
    if (LaneId() == 0) {
        const int w = threadIdx.x >> 6;
        part[0][w] = d;
        part[1][w] = f1;
        part[2][w] = f2;
        part[3][w] = mx;
        part[4][w] = heavy;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s[5] = {0, 0, 0, 0, 0};
        for (int w = 0; w < 4; ++w) {
            for (int k = 0; k < 5; ++k) {
                if (k == 3) s[3] = part[3][w] > s[3] ? part[3][w] : s[3];
                else s[k] += part[k][w];
            }
        }
        if (s[0]) atomicAdd(&a.stats->distinct, s[0]);
        if (s[1]) atomicAdd(&a.stats->f1, s[1]);
        if (s[2]) atomicAdd(&a.stats->f2, s[2]);
        if (s[4]) atomicAdd(&a.stats->heavy, s[4]);
        if (s[3]) atomicMax(&a.stats->maxc, s[3]);
    }
}

// Empties a sketch table of `cap` slots and seeds the statistics.  A kernel rather than memsets, so that the
// sketch's launch sequence holds no memset node when a caller captures it in a graph.
__global__ void __launch_bounds__(256) sketch_clear(unsigned long long *hkeys, unsigned int *counts,
                                                    unsigned long long cap, SketchStats *stats)
{
    const unsigned long long i = static_cast<unsigned long long>(blockIdx.x) * 256 + threadIdx.x;
    if (i < cap) {
        hkeys[i] = 0;
        counts[i] = 0;
    }
    if (i == 0) {
        SketchStats init{};
        init.kmin = std::numeric_limits<long long>::max();
        init.kmax = std::numeric_limits<long long>::min();
        *stats = init;
    }
}

// Copies the sketch of `rows` rows (m sampled) into a device-resident omni_gpu_sketch.
__global__ void sketch_publish(const SketchStats *st, const float *f, int64_t rows, int64_t m, omni_gpu_sketch *out)
{
    const int t = threadIdx.x;
    if (t < kIn) out->features[t] = f[t];
    if (t == 0) {
        out->rows = rows;
        out->sampled = m;
        out->distinct = static_cast<int64_t>(st->distinct);
        out->f1 = static_cast<int64_t>(st->f1);
        out->f2 = static_cast<int64_t>(st->f2);
        out->max_count = static_cast<int64_t>(st->maxc);
        out->heavy = static_cast<int64_t>(st->heavy);
        out->nulls = static_cast<int64_t>(st->nulls);
        out->key_min = st->kmin;
        out->key_max = st->kmax;
    }
}

// Feature i of the MLP's 16 input features from the sketch statistics (all roughly in [0, 1]): log2(a + 1) / den
// or a / den.  The cases only pick a and den, so that lanes computing different features of one sketch share the
// one logarithm and the one division.  Shared by the sketch entry point and the advisor so that training and
// inference see identical features.
__device__ __forceinline__ float SketchFeature(const SketchStats *st, int64_t rows, int64_t m, int key_bytes,
                                               int num_keys, int is_join, int i)
{
    const double md = m > 0 ? static_cast<double>(m) : 1.0;
    const double d = static_cast<double>(st->distinct), f1 = static_cast<double>(st->f1),
                 f2 = static_cast<double>(st->f2);
    double a = 1.0, den = 1.0;
    bool lg = false;
    switch (i) {
        case 0: a = static_cast<double>(rows), den = 32.0, lg = true; break;
        case 1: a = md, den = 32.0, lg = true; break;
        case 2: a = d, den = 32.0, lg = true; break;
        case 3: a = f1, den = md; break;
        case 4: a = f2, den = md; break;
        case 5: a = static_cast<double>(st->maxc), den = md; break;
        case 6: a = static_cast<double>(st->heavy), den = md; break;
        case 7: a = d, den = md; break;
        case 8:  // GEE estimate of the distinct keys
            a = sqrt(static_cast<double>(rows > 0 ? rows : 1) / md) * f1 + (d - f1), den = 32.0, lg = true;
            break;
        case 9:  // Chao1 estimate
            a = d + f1 * f1 / (2.0 * (f2 > 1.0 ? f2 : 1.0)), den = 32.0, lg = true;
            break;
        case 10:  // key range
            a = st->kmax >= st->kmin ? static_cast<double>(st->kmax) - static_cast<double>(st->kmin) : 0.0;
            den = 64.0, lg = true;
            break;
        case 11: a = static_cast<double>(st->nulls), den = md; break;
        case 12: a = static_cast<double>(key_bytes), den = 16.0; break;
        case 13: a = static_cast<double>(num_keys), den = 4.0; break;
        case 14: a = is_join ? 1.0 : 0.0; break;
        default: break;
    }
    return static_cast<float>((lg ? log2(a + 1.0) : a) / den);
}

// All 16 features in one thread.
__device__ __forceinline__ void SketchFeatures(const SketchStats *st, int64_t rows, int64_t m, int key_bytes,
                                               int num_keys, int is_join, float *f)
{
    for (int i = 0; i < kIn; ++i) f[i] = SketchFeature(st, rows, m, key_bytes, num_keys, is_join, i);
}

__global__ void sketch_features(const SketchStats *st, int64_t rows, int64_t m, int key_bytes, int num_keys,
                                int is_join, float *f)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    SketchFeatures(st, rows, m, key_bytes, num_keys, is_join, f);
}

// ---- per-batch sketch of a batched stream ------------------------------------------------------------------
// One workgroup per batch: a sample of at most kBatchSample rows of the batch's key column -- every row of a small
// batch, else 8 runs of 64 consecutive rows evenly spread over the batch (coalesced reads) -- is counted in an LDS
// table and reduced to the same statistics and 16 features as the column sketch (so the advisor network reads
// features of one kind), into X[b * kIn ..]; given the network's parameters, wave 0 then runs the network on them
// into Y[b * kOut ..].  Batches of at most min_rows rows get zero features and no outputs.
constexpr int kBatchSample = 512;
constexpr int kBatchSketchSlots = 2 * kBatchSample;
constexpr int kBatchSketchBlock = 256;

__device__ __forceinline__ unsigned long long WaveSum(unsigned long long v)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

__global__ void __launch_bounds__(kBatchSketchBlock) batch_features(const void *keys, const uint8_t *nulls, int width,
                                                                    const int64_t *offsets, float *X, int64_t min_rows,
                                                                    const float *params, float *Y)
{
    __shared__ unsigned long long hk[kBatchSketchSlots];
    __shared__ unsigned int cnt[kBatchSketchSlots];
    __shared__ SketchStats st;
    __shared__ float xs[kIn];
    const int b = blockIdx.x;
    const int64_t start = offsets[b], rows = offsets[b + 1] - offsets[b];
    if (rows <= min_rows) {  // the caller decides such batches without the network: zero features, no outputs
        if (threadIdx.x < kIn) X[static_cast<int64_t>(b) * kIn + threadIdx.x] = 0.f;
        return;
    }
    const int64_t m = rows < kBatchSample ? rows : kBatchSample;
    for (int i = threadIdx.x; i < kBatchSketchSlots; i += kBatchSketchBlock) {
        hk[i] = 0;
        cnt[i] = 0;
    }
    if (threadIdx.x == 0) {
        st = SketchStats{};
        st.kmin = std::numeric_limits<long long>::max();
        st.kmax = std::numeric_limits<long long>::min();
    }
    __syncthreads();
    long long lo = std::numeric_limits<long long>::max(), hi = std::numeric_limits<long long>::min();
    unsigned long long nn = 0;
    for (int64_t i = threadIdx.x; i < m; i += kBatchSketchBlock) {
        const int64_t r = rows <= kBatchSample ? i : (i >> 6) * (rows >> 3) + (i & 63);
        const int64_t row = start + r;
        const bool isNull = NullBit(nulls, row);
        const unsigned long long w = isNull ? 0x9E3779B97F4A7C15ULL : LoadWord(keys, width, row);
        unsigned long long h = Mix64(0x243F6A8885A308D3ULL ^ (w + 0x632BE59BD9B4E019ULL + (isNull ? 1ULL : 0ULL)));
        if (h == 0) h = 1;
        if (isNull) {
            ++nn;
        } else {
            lo = min(lo, static_cast<long long>(w));
            hi = max(hi, static_cast<long long>(w));
        }
        unsigned int pos = static_cast<unsigned int>(h) & (kBatchSketchSlots - 1);
        while (true) {
            unsigned long long k = hk[pos];
            if (k == 0) {
                k = atomicCAS(&hk[pos], 0ULL, h);
                if (k == 0) k = h;
            }
            if (k == h) {
                atomicAdd(&cnt[pos], 1u);
                break;
            }
            pos = (pos + 1) & (kBatchSketchSlots - 1);
        }
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        lo = min(lo, __shfl_xor(lo, d, 64));
        hi = max(hi, __shfl_xor(hi, d, 64));
    }
    nn = WaveSum(nn);
    if (LaneId() == 0) {
        if (lo <= hi) {
            atomicMin(&st.kmin, lo);
            atomicMax(&st.kmax, hi);
        }
        if (nn) atomicAdd(&st.nulls, nn);
    }
    __syncthreads();
    unsigned long long d = 0, f1 = 0, f2 = 0, heavy = 0, mx = 0;
    for (int i = threadIdx.x; i < kBatchSketchSlots; i += kBatchSketchBlock) {
        const unsigned int c = cnt[i];
        d += c ? 1 : 0;
        f1 += c == 1 ? 1 : 0;
        f2 += c == 2 ? 1 : 0;
        heavy += c >= 4 ? c : 0;
        mx = c > mx ? c : mx;
    }
    d = WaveSum(d);
    f1 = WaveSum(f1);
    f2 = WaveSum(f2);
    heavy = WaveSum(heavy);
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
        const unsigned long long o = __shfl_xor(mx, s, 64);
        mx = o > mx ? o : mx;
    }
    if (LaneId() == 0) {
        atomicAdd(&st.distinct, d);
        atomicAdd(&st.f1, f1);
        atomicAdd(&st.f2, f2);
        atomicAdd(&st.heavy, heavy);
        atomicMax(&st.maxc, mx);
    }
    __syncthreads();
    if (threadIdx.x < kIn) xs[threadIdx.x] = SketchFeature(&st, rows, m, width, 1, 0, threadIdx.x);
    __syncthreads();
    if (threadIdx.x < kIn) X[static_cast<int64_t>(b) * kIn + threadIdx.x] = xs[threadIdx.x];
    // the network's outputs for this batch (wave 0), when the caller passes its parameters
    if (Y && threadIdx.x < kWaveSize) MlpWave(params, xs, Y + static_cast<int64_t>(b) * kOut);
}

// Per-batch first phase from the network's outputs Y[b][4]: the predicted group count 2^(32 y0) - 1 against
// `limit` (a share of the per-batch LDS table's slots): at most that many start in the per-batch LDS table (0),
// more go straight to the hash-partitioned passes (2).
__global__ void batch_plan(const float *Y, int nb, float limit, int32_t *modes)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const float groups = exp2f(Y[static_cast<int64_t>(b) * kOut] * 32.f) - 1.f;
    modes[b] = groups > limit ? 2 : 0;
}

// ---- seeded key generator ----------------------------------------------------------------------------------
// Row i draws a rank r in [1, D] from a Zipf(s) law by inverting its continuous approximation (s = 0 is
// uniform), then maps r through splitmix64, a bijection, so the number of distinct keys is that of ranks.
// kind 0: int64 key; 1: int32 key (low 31 bits); 2: (int32, int64) composite whose pair is a function of r.
__device__ __forceinline__ unsigned long long SplitMix(unsigned long long x)
{
    unsigned long long z = x + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__global__ void gen_keys(int kind, int64_t rows, double D, double s, unsigned long long seed, long long *k64,
                         int *k32, long long *v64)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    const unsigned long long u0 = SplitMix(seed ^ (static_cast<unsigned long long>(i) * 0xD6E8FEB86659FD93ULL));
    const double u = (static_cast<double>(u0 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    double x;
    if (s < 1e-6) {
        x = 1.0 + u * D;
    } else if (fabs(s - 1.0) < 1e-6) {
        x = exp(u * log(D + 1.0));
    } else {
        const double e = 1.0 - s;
        x = pow(u * (pow(D + 1.0, e) - 1.0) + 1.0, 1.0 / e);
    }
    unsigned long long r = static_cast<unsigned long long>(x);
    if (r < 1) r = 1;
    if (r > static_cast<unsigned long long>(D)) r = static_cast<unsigned long long>(D);
    if (kind & 8) r = static_cast<unsigned long long>(i) + 1;  // sequential ranks: row i is rank i + 1
    unsigned long long in = r + (seed << 20);
    if ((kind & 16) && (u0 & 1023ULL) < 205ULL) in += 1ULL << 62;  // ~20% of rows outside the rank range
    const unsigned long long key = SplitMix(in);
    kind &= 7;
    if (kind == 0) {
        k64[i] = static_cast<long long>(key);
    } else if (kind == 1) {
        k32[i] = static_cast<int>(key & 0x7FFFFFFFULL);
    } else {
        k32[i] = static_cast<int>(key >> 48);
        k64[i] = static_cast<long long>(SplitMix(key ^ 0xA0761D6478BD642FULL));
    }
    if (v64) v64[i] = static_cast<long long>(u0 % 2000001ULL) - 1000000LL;
}

}  // namespace omni_gpu
