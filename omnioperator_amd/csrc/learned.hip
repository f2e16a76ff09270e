// Learned strategy / capacity advisor: key sketch, MLP inference / gradient / training, seeded key generator.
//
// The plain path sizes hash tables by growth (grow x4 and rehash, as BaseHashMap's Grower does,
// core/src/operator/hashmap/base_hash_map.h:255-292) and picks its table kind by a fixed rule.  The advisor
// predicts the group count and the fastest strategy from a sample sketch (1% of the rows, at least 4,096 and at
// most 65,536 of them), so the table can be sized once and
// the kernel chosen up front.  Results never depend on it: a prediction that is too small only costs a growth.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <map>
#include <mutex>
#include <numeric>
#include <optional>
#include <random>
#include <vector>

#include "learned_device.h"

namespace omni_gpu {

// Device buffers of one sketch: the sample's hash table (sized for the largest sample), statistics, features.
struct SketchBuffers {
    DevBuf hkeys, counts, stats, features;
    void Release(hipStream_t s)
    {
        hkeys.Release(s);
        counts.Release(s);
        stats.Release(s);
        features.Release(s);
    }
};

constexpr int64_t kMaxSample = 1 << 16;  // sampled rows at most: the sketch's cost stays flat
constexpr uint64_t kMaxSketchSlots = 2 * kMaxSample;  // table slots of the largest sample

// The sketch of one call: its arguments, table slots and the features' inputs.
struct SketchPlan {
    SketchArgs a{};
    uint64_t cap = 0;
    int key_bytes = 0;
};

// Validates the key columns and plans a 1-in-k sample of about 1% of the rows (at least 4096 or all of them, at
// most kMaxSample) counted in a hash table of cap >= 2 * m slots.
int PlanSketch(const omni_gpu_column *keys, int num_keys, int64_t rows, SketchPlan &p)
{
    OG_CHECK(keys, OMNI_GPU_ERR_INVALID, "null key columns");
    OG_CHECK(num_keys >= 1 && num_keys <= 2, OMNI_GPU_ERR_INVALID, "the sketch reads one or two key columns");
    OG_CHECK(rows >= 0, OMNI_GPU_ERR_INVALID, "negative row count");
    SketchArgs &a = p.a;
    a.ncols = num_keys;
    p.key_bytes = 0;
    for (int c = 0; c < num_keys; ++c) {
        a.width[c] = TypeWidth(keys[c].type_id);
        OG_CHECK(a.width[c] > 0 && a.width[c] <= 8 && (rows == 0 || keys[c].values), OMNI_GPU_ERR_INVALID,
                 "sketch key %d must be an integer-like column of <= 8 bytes", c);
        a.values[c] = keys[c].values;
        a.nulls[c] = keys[c].nulls;
        p.key_bytes += a.width[c];
    }
    a.rows = rows;
    int64_t m = std::min<int64_t>(rows, std::max<int64_t>(4096, rows / 100));
    m = std::min<int64_t>(m, kMaxSample);
    a.stride = m > 0 ? std::max<int64_t>(1, rows / m) : 1;
    a.m = m;
    p.cap = 1024;
    while (p.cap < static_cast<uint64_t>(2 * m)) p.cap <<= 1;
    return OMNI_GPU_OK;
}

// Launches the planned sketch on stream s into the given table and statistics: clear, count the sample, reduce,
// features.  `timed` brackets the launches with the library's kernel timers.
int LaunchSketch(hipStream_t s, SketchPlan &p, unsigned long long *hkeys, unsigned int *counts, SketchStats *stats,
                 float *features, int64_t rows, int num_keys, int is_join, bool timed)
{
    SketchArgs &a = p.a;
    a.hkeys = hkeys;
    a.counts = counts;
    a.mask = p.cap - 1;
    a.stats = stats;
    const unsigned int tableBlocks = static_cast<unsigned int>((p.cap + 255) / 256);
    sketch_clear<<<tableBlocks, 256, 0, s>>>(hkeys, counts, p.cap, stats);
    OG_TRY(CheckLaunch("sketch_clear"));
    if (a.m > 0) {
        std::optional<KernelScope> ks;
        if (timed) ks.emplace(s, "sketch_insert");
        sketch_insert<<<static_cast<unsigned int>((a.m + 255) / 256), 256, 0, s>>>(a);
        OG_TRY(CheckLaunch("sketch_insert"));
    }
    {
        std::optional<KernelScope> ks;
        if (timed) ks.emplace(s, "sketch_reduce");
        sketch_reduce<<<tableBlocks, 256, 0, s>>>(a, p.cap);
        OG_TRY(CheckLaunch("sketch_reduce"));
    }
    sketch_features<<<1, 64, 0, s>>>(stats, rows, a.m, p.key_bytes, num_keys, is_join, features);
    return CheckLaunch("sketch_features");
}

// Launches the sketch of `keys` on stream s into the buffers b (grown as needed).
int RunSketch(hipStream_t s, SketchBuffers &b, const omni_gpu_column *keys, int num_keys, int64_t rows, int is_join,
              int64_t *sampled)
{
    SketchPlan p;
    OG_TRY(PlanSketch(keys, num_keys, rows, p));
    *sampled = p.a.m;
    OG_TRY(b.hkeys.Reserve(8 * p.cap, s));
    OG_TRY(b.counts.Reserve(4 * p.cap, s));
    OG_TRY(b.stats.Reserve(sizeof(SketchStats), s));
    OG_TRY(b.features.Reserve(sizeof(float) * kIn, s));
    return LaunchSketch(s, p, b.hkeys.as<unsigned long long>(), b.counts.as<unsigned int>(), b.stats.as<SketchStats>(),
                        b.features.as<float>(), rows, num_keys, is_join, true);
}

// ---- the allocation-free sketch (omni_gpu_sketch_keys_noalloc) ------------------------------------------------
// Its table, counts, statistics and features live in `scratch`, a file-scope device array of fixed size that
// holds the largest sample's table (kMaxSketchSlots slots): 8 + 4 bytes per slot, then the statistics and the
// features.  Each GPU the library runs on holds its own copy until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in for the symbol): the device address of each GPU's
// copy comes from hipGetSymbolAddress, once per device, with that device current.
constexpr size_t kScratchStatsOff = 12 * kMaxSketchSlots;
constexpr size_t kScratchFeatOff = kScratchStatsOff + 256;
constexpr size_t kScratchBytes = kScratchFeatOff + 256;
static_assert(sizeof(SketchStats) <= 256 && sizeof(float) * kIn <= 256, "sketch scratch layout");
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_scratch_mu;
std::map<int, uint8_t *> g_scratch_ptr;  // device -> address of that device's scratch

int ScratchOf(int dev, uint8_t **p)
{
    std::lock_guard<std::mutex> g(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it != g_scratch_ptr.end()) {
        *p = it->second;
        return OMNI_GPU_OK;
    }
    void *ptr = nullptr;
    OG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    OG_CHECK(ptr != nullptr, OMNI_GPU_ERR_HIP, "no device address for the sketch scratch");
    g_scratch_ptr[dev] = static_cast<uint8_t *>(ptr);
    *p = static_cast<uint8_t *>(ptr);
    return OMNI_GPU_OK;
}

struct MlpHandle {
    hipStream_t stream = nullptr;
    DevBuf params, m, v, grad, partial, loss_parts, loss, X, T, C, idx, Y;
    SketchBuffers sketch;   // the advisor's sketch buffers, kept across calls
    int step = 0;
};

int UploadTraining(MlpHandle &h, const float *X, const float *T, const int32_t *C, int64_t n)
{
    hipStream_t s = h.stream;
    OG_TRY(h.X.Reserve(sizeof(float) * kIn * n, s));
    OG_TRY(h.T.Reserve(sizeof(float) * n, s));
    OG_TRY(h.C.Reserve(sizeof(int32_t) * n, s));
    OG_HIP(hipMemcpyAsync(h.X.p, X, sizeof(float) * kIn * n, hipMemcpyHostToDevice, s));
    OG_HIP(hipMemcpyAsync(h.T.p, T, sizeof(float) * n, hipMemcpyHostToDevice, s));
    OG_HIP(hipMemcpyAsync(h.C.p, C, sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
    return OMNI_GPU_OK;
}

// One gradient evaluation over `B` samples (rows idx[0..B) of the uploaded set, or rows [0, B) when idx is
// null), followed by an Adam step when lr > 0.
int GradStep(MlpHandle &h, const int *idx, int B, float lr, bool keep_grad)
{
    hipStream_t s = h.stream;
    const int nparts = (B + kMlpBlock - 1) / kMlpBlock;
    OG_TRY(h.partial.Reserve(sizeof(float) * kParams * nparts, s));
    OG_TRY(h.loss_parts.Reserve(sizeof(float) * nparts, s));
    {
        KernelScope ks(s, "mlp_backward");
        mlp_backward<<<nparts, kMlpBlock, 0, s>>>(h.params.as<float>(), h.X.as<float>(), h.T.as<float>(),
                                                  h.C.as<int>(), idx, B, h.partial.as<float>(),
                                                  h.loss_parts.as<float>());
        OG_TRY(CheckLaunch("mlp_backward"));
    }
    if (lr > 0.f) ++h.step;
    KernelScope ks(s, "mlp_adam");
    mlp_reduce_adam<<<(kParams + 255) / 256, 256, 0, s>>>(h.partial.as<float>(), nparts, h.loss_parts.as<float>(),
                                                         h.params.as<float>(), h.m.as<float>(), h.v.as<float>(),
                                                         keep_grad ? h.grad.as<float>() : nullptr,
                                                         h.loss.as<float>(), lr, 0.9f, 0.999f, 1e-8f,
                                                         std::max(1, h.step));
    return CheckLaunch("mlp_reduce_adam");
}

int LaunchBatchAdvisor(hipStream_t s, void *mlp, const void *keys, const uint8_t *nulls, int width,
                       const int64_t *offsets, int nb, float *X, float *Y, int32_t *modes, float limit, int64_t min_rows)
{
    if (nb <= 0) return OMNI_GPU_OK;
    // the sketch and the network's inference of each batch in one workgroup
    const float *params = mlp ? static_cast<const MlpHandle *>(mlp)->params.as<float>() : nullptr;
    batch_features<<<nb, kBatchSketchBlock, 0, s>>>(keys, nulls, width, offsets, X, min_rows, params,
                                                    mlp ? Y : nullptr);
    OG_TRY(CheckLaunch("batch_features"));
    if (!mlp || !modes) return OMNI_GPU_OK;
    batch_plan<<<(nb + 255) / 256, 256, 0, s>>>(Y, nb, limit, modes);
    return CheckLaunch("batch_plan");
}

}  // namespace omni_gpu

using namespace omni_gpu;

extern "C" {

int omni_gpu_batch_features(const omni_gpu_column *key, const int64_t *batch_offsets, int32_t num_batches,
                            float *features)
{
    OG_CHECK(key && batch_offsets && features && num_batches >= 0, OMNI_GPU_ERR_INVALID, "null argument");
    OG_CHECK(IsIntegerKeyType(key->type_id) && key->values, OMNI_GPU_ERR_INVALID, "key must be an integer column");
    OG_CHECK(batch_offsets[0] == 0, OMNI_GPU_ERR_INVALID, "batch_offsets[0] must be 0");
    for (int b = 0; b < num_batches; ++b) {
        OG_CHECK(batch_offsets[b + 1] >= batch_offsets[b], OMNI_GPU_ERR_INVALID, "batch %d has a negative size", b);
    }
    if (num_batches == 0) return OMNI_GPU_OK;
    hipStream_t s;
    OG_TRY(AcquireStream(&s));
    DevBuf off, X;
    int rc = off.Reserve(sizeof(int64_t) * (num_batches + 1), s);
    if (rc == OMNI_GPU_OK) rc = X.Reserve(sizeof(float) * kIn * num_batches, s);
    if (rc == OMNI_GPU_OK && hipMemcpyAsync(off.p, batch_offsets, sizeof(int64_t) * (num_batches + 1),
                                            hipMemcpyHostToDevice, s) != hipSuccess) {
        rc = SetError(OMNI_GPU_ERR_HIP, "offsets upload failed");
    }
    if (rc == OMNI_GPU_OK) {
        rc = LaunchBatchAdvisor(s, nullptr, key->values, key->nulls, TypeWidth(key->type_id), off.as<int64_t>(),
                                num_batches, X.as<float>(), nullptr, nullptr, 0.f, -1);
    }
    if (rc == OMNI_GPU_OK && (hipMemcpyAsync(features, X.p, sizeof(float) * kIn * num_batches, hipMemcpyDeviceToHost,
                                             s) != hipSuccess ||
                              hipStreamSynchronize(s) != hipSuccess)) {
        rc = SetError(OMNI_GPU_ERR_HIP, "feature readback failed");
    }
    off.Release(s);
    X.Release(s);
    ReleaseStream(s);
    return rc;
}

int omni_gpu_sketch_keys(const omni_gpu_column *keys, int32_t num_keys, int64_t rows, int32_t is_join,
                         omni_gpu_sketch *out)
{
    OG_CHECK(out, OMNI_GPU_ERR_INVALID, "null argument");
    std::memset(out, 0, sizeof(*out));
    hipStream_t s;
    OG_TRY(AcquireStream(&s));
    SketchBuffers b;
    int64_t m = 0;
    int rc = RunSketch(s, b, keys, num_keys, rows, is_join, &m);
    SketchStats st{};
    if (rc == OMNI_GPU_OK && (hipMemcpyAsync(&st, b.stats.p, sizeof(st), hipMemcpyDeviceToHost, s) != hipSuccess ||
                              hipMemcpyAsync(out->features, b.features.p, sizeof(float) * kIn, hipMemcpyDeviceToHost,
                                             s) != hipSuccess ||
                              hipStreamSynchronize(s) != hipSuccess)) {
        rc = SetError(OMNI_GPU_ERR_HIP, "sketch readback failed");
    }
    b.Release(s);
    ReleaseStream(s);
    OG_TRY(rc);
    out->rows = rows;
    out->sampled = m;
    out->distinct = static_cast<int64_t>(st.distinct);
    out->f1 = static_cast<int64_t>(st.f1);
    out->f2 = static_cast<int64_t>(st.f2);
    out->max_count = static_cast<int64_t>(st.maxc);
    out->heavy = static_cast<int64_t>(st.heavy);
    out->nulls = static_cast<int64_t>(st.nulls);
    out->key_min = st.kmin;
    out->key_max = st.kmax;
    return OMNI_GPU_OK;
}

int omni_gpu_sketch_keys_noalloc(const omni_gpu_column *keys, int32_t num_keys, int64_t rows, int32_t is_join,
                                 void *stream, omni_gpu_sketch *out)
{
    OG_CHECK(out, OMNI_GPU_ERR_INVALID, "null output");
    SketchPlan p;
    OG_TRY(PlanSketch(keys, num_keys, rows, p));
    OG_CHECK(p.cap <= kMaxSketchSlots, OMNI_GPU_ERR_INVALID,
             "sketch table of %llu slots exceeds the scratch's %llu slots", static_cast<unsigned long long>(p.cap),
             static_cast<unsigned long long>(kMaxSketchSlots));
    uint8_t *sc = nullptr;
    OG_TRY(ScratchOf(CurrentDevice(), &sc));
    hipStream_t s = static_cast<hipStream_t>(stream);
    unsigned long long *hkeys = reinterpret_cast<unsigned long long *>(sc);
    unsigned int *counts = reinterpret_cast<unsigned int *>(sc + 8 * kMaxSketchSlots);
    SketchStats *stats = reinterpret_cast<SketchStats *>(sc + kScratchStatsOff);
    float *features = reinterpret_cast<float *>(sc + kScratchFeatOff);
    OG_TRY(LaunchSketch(s, p, hkeys, counts, stats, features, rows, num_keys, is_join, false));
    sketch_publish<<<1, 64, 0, s>>>(stats, features, rows, p.a.m, out);
    return CheckLaunch("sketch_publish");
}

int omni_gpu_advise(void *mlp, const omni_gpu_column *keys, int32_t num_keys, int64_t rows, int32_t is_join,
                    float *y)
{
    OG_CHECK(mlp && y, OMNI_GPU_ERR_INVALID, "null argument");
    MlpHandle &h = *static_cast<MlpHandle *>(mlp);
    hipStream_t s = h.stream;
    int64_t m = 0;
    OG_TRY(RunSketch(s, h.sketch, keys, num_keys, rows, is_join, &m));
    OG_TRY(h.Y.Reserve(sizeof(float) * kOut, s));
    {
        KernelScope ks(s, "mlp_forward");
        mlp_forward<<<1, kMlpFwdBlock, 0, s>>>(h.params.as<float>(), h.sketch.features.as<float>(), 1, h.Y.as<float>());
        OG_TRY(CheckLaunch("mlp_forward"));
    }
    OG_TRY(CopyToHost(s, {{y, h.Y.p, sizeof(float) * kOut}}));
    return OMNI_GPU_OK;
}

int omni_gpu_mlp_create(uint64_t seed, void **handle)
{
    OG_CHECK(handle, OMNI_GPU_ERR_INVALID, "null argument");
    std::unique_ptr<MlpHandle> h(new MlpHandle());
    OG_TRY(AcquireStream(&h->stream));
    hipStream_t s = h->stream;
    OG_TRY(h->params.Reserve(sizeof(float) * kParams, s));
    OG_TRY(h->m.Reserve(sizeof(float) * kParams, s));
    OG_TRY(h->v.Reserve(sizeof(float) * kParams, s));
    OG_TRY(h->grad.Reserve(sizeof(float) * kParams, s));
    OG_TRY(h->loss.Reserve(sizeof(float), s));
    // He-uniform weights, zero biases
    std::vector<float> p(kParams, 0.f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int fan_in, int n) {
        const float lim = std::sqrt(6.0f / static_cast<float>(fan_in));
        std::uniform_real_distribution<float> u(-lim, lim);
        for (int i = 0; i < n; ++i) p[off + i] = u(rng);
    };
    fill(kOffW1, kIn, kIn * kH);
    fill(kOffW2, kH, kH * kH);
    fill(kOffW3, kH, kH * kOut);
    OG_HIP(hipMemcpyAsync(h->params.p, p.data(), sizeof(float) * kParams, hipMemcpyHostToDevice, s));
    OG_HIP(hipMemsetAsync(h->m.p, 0, sizeof(float) * kParams, s));
    OG_HIP(hipMemsetAsync(h->v.p, 0, sizeof(float) * kParams, s));
    OG_HIP(hipStreamSynchronize(s));
    *handle = h.release();
    return OMNI_GPU_OK;
}

int omni_gpu_mlp_num_params(void) { return kParams; }

int omni_gpu_mlp_set_params(void *handle, const float *params)
{
    OG_CHECK(handle && params, OMNI_GPU_ERR_INVALID, "null argument");
    MlpHandle &h = *static_cast<MlpHandle *>(handle);
    OG_HIP(hipMemcpyAsync(h.params.p, params, sizeof(float) * kParams, hipMemcpyHostToDevice, h.stream));
    OG_HIP(hipMemsetAsync(h.m.p, 0, sizeof(float) * kParams, h.stream));
    OG_HIP(hipMemsetAsync(h.v.p, 0, sizeof(float) * kParams, h.stream));
    h.step = 0;
    OG_HIP(hipStreamSynchronize(h.stream));
    return OMNI_GPU_OK;
}

int omni_gpu_mlp_get_params(void *handle, float *params)
{
    OG_CHECK(handle && params, OMNI_GPU_ERR_INVALID, "null argument");
    MlpHandle &h = *static_cast<MlpHandle *>(handle);
    OG_HIP(hipMemcpyAsync(params, h.params.p, sizeof(float) * kParams, hipMemcpyDeviceToHost, h.stream));
    OG_HIP(hipStreamSynchronize(h.stream));
    return OMNI_GPU_OK;
}

int omni_gpu_mlp_forward(void *handle, const float *X, int32_t n, float *Y)
{
    OG_CHECK(handle && X && Y && n >= 0, OMNI_GPU_ERR_INVALID, "bad argument");
    if (n == 0) return OMNI_GPU_OK;
    MlpHandle &h = *static_cast<MlpHandle *>(handle);
    hipStream_t s = h.stream;
    OG_TRY(h.X.Reserve(sizeof(float) * kIn * n, s));
    OG_TRY(h.Y.Reserve(sizeof(float) * kOut * n, s));
    OG_HIP(hipMemcpyAsync(h.X.p, X, sizeof(float) * kIn * n, hipMemcpyHostToDevice, s));
    {
        KernelScope ks(s, "mlp_forward");
        mlp_forward<<<(n + kMlpFwdBlock - 1) / kMlpFwdBlock, kMlpFwdBlock, 0, s>>>(h.params.as<float>(), h.X.as<float>(),
                                                                                  n, h.Y.as<float>());
        OG_TRY(CheckLaunch("mlp_forward"));
    }
    OG_HIP(hipMemcpyAsync(Y, h.Y.p, sizeof(float) * kOut * n, hipMemcpyDeviceToHost, s));
    OG_HIP(hipStreamSynchronize(s));
    return OMNI_GPU_OK;
}

int omni_gpu_mlp_gradient(void *handle, const float *X, const float *t, const int32_t *c, int32_t n, float *grad,
                          float *loss)
{
    OG_CHECK(handle && X && t && c && grad && n > 0, OMNI_GPU_ERR_INVALID, "bad argument");
    MlpHandle &h = *static_cast<MlpHandle *>(handle);
    for (int32_t i = 0; i < n; ++i) OG_CHECK(c[i] >= 0 && c[i] < 3, OMNI_GPU_ERR_INVALID, "class %d out of range", c[i]);
    OG_TRY(UploadTraining(h, X, t, c, n));
    OG_TRY(GradStep(h, nullptr, n, 0.f, true));
    OG_HIP(hipMemcpyAsync(grad, h.grad.p, sizeof(float) * kParams, hipMemcpyDeviceToHost, h.stream));
    float l = 0.f;
    OG_HIP(hipMemcpyAsync(&l, h.loss.p, sizeof(float), hipMemcpyDeviceToHost, h.stream));
    OG_HIP(hipStreamSynchronize(h.stream));
    if (loss) *loss = l;
    return OMNI_GPU_OK;
}

int omni_gpu_mlp_train(void *handle, const float *X, const float *t, const int32_t *c, int32_t n, int32_t batch,
                       int32_t epochs, float lr, uint64_t seed, float *epoch_loss)
{
    OG_CHECK(handle && X && t && c && n > 0 && batch > 0 && epochs > 0 && lr > 0.f, OMNI_GPU_ERR_INVALID,
             "bad argument");
    MlpHandle &h = *static_cast<MlpHandle *>(handle);
    for (int32_t i = 0; i < n; ++i) OG_CHECK(c[i] >= 0 && c[i] < 3, OMNI_GPU_ERR_INVALID, "class %d out of range", c[i]);
    hipStream_t s = h.stream;
    OG_TRY(UploadTraining(h, X, t, c, n));
    OG_TRY(h.idx.Reserve(sizeof(int) * n, s));
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937_64 rng(seed);
    for (int e = 0; e < epochs; ++e) {
        std::shuffle(perm.begin(), perm.end(), rng);
        OG_HIP(hipMemcpyAsync(h.idx.p, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
        // the copy reads `perm` asynchronously: wait before the next epoch reshuffles it
        double sum = 0.0;
        int steps = 0;
        for (int b0 = 0; b0 < n; b0 += batch) {
            const int B = std::min(batch, n - b0);
            OG_TRY(GradStep(h, h.idx.as<int>() + b0, B, lr, false));
            if (epoch_loss) {
                float l = 0.f;
                OG_HIP(hipMemcpyAsync(&l, h.loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
                OG_HIP(hipStreamSynchronize(s));
                sum += l * B;
                ++steps;
            }
        }
        OG_HIP(hipStreamSynchronize(s));
        if (epoch_loss) epoch_loss[e] = static_cast<float>(sum / n);
    }
    return OMNI_GPU_OK;
}

int omni_gpu_mlp_destroy(void *handle)
{
    if (!handle) return OMNI_GPU_OK;
    MlpHandle *h = static_cast<MlpHandle *>(handle);
    hipStream_t s = h->stream;
    for (DevBuf *b : {&h->params, &h->m, &h->v, &h->grad, &h->partial, &h->loss_parts, &h->loss, &h->X, &h->T, &h->C,
                      &h->idx, &h->Y}) {
        b->Release(s);
    }
    h->sketch.Release(s);
    ReleaseStream(s);
    delete h;
    return OMNI_GPU_OK;
}

int omni_gpu_generate_keys(int32_t kind, int64_t rows, double distinct, double zipf_s, uint64_t seed, void *k64,
                           void *k32, void *v64)
{
    const int32_t type = kind & 7;
    OG_CHECK(type <= 2 && (kind & ~31) == 0 && rows >= 0 && distinct >= 1.0 && zipf_s >= 0.0, OMNI_GPU_ERR_INVALID,
             "bad generator arguments");
    OG_CHECK((type == 1 || k64) && (type == 0 || k32), OMNI_GPU_ERR_INVALID, "missing output column");
    if (rows == 0) return OMNI_GPU_OK;
    hipStream_t s;
    OG_TRY(AcquireStream(&s));
    gen_keys<<<static_cast<unsigned int>((rows + 255) / 256), 256, 0, s>>>(
        kind, rows, distinct, zipf_s, seed, static_cast<long long *>(k64), static_cast<int *>(k32),
        static_cast<long long *>(v64));
    int rc = CheckLaunch("gen_keys");
    ReleaseStream(s);
    return rc;
}

}  // extern "C"
