// End-to-end timing of the GPU operators through the reference's Operator surface: host VectorBatches in, host
// VectorBatches out, every host<->device copy included (the way a query engine drives HashAggregationOperator and
// HashBuilderOperator + LookupJoinOperator).
//
//   omni_operator_bench agg  STEPS WARMUP   1M rows, int64 key over 1,000 values, int64 value, VectorBatches of 4,096;
//                                           SUM / COUNT(*) / MIN / MAX
//   omni_operator_bench join STEPS WARMUP   build 2M unique int64 keys + int32 payload, probe 64M int64 keys (80% hit),
//                                           VectorBatches of 65,536; output probe key + build payload
//
// Prints one JSON object: rows per step, milliseconds per step (wall clock), groups or matches (checked against the
// generator), and the library's own kernel time per step for the dominant kernels (HIP events).
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gpu_operators.h"

using namespace omniruntime;
using namespace omniruntime::op;
using namespace omniruntime::type;
using namespace omniruntime::vec;

namespace {

uint64_t SplitMix(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s;
    uint64_t Next() { return SplitMix(s++); }
    int64_t Range(int64_t lo, int64_t hi) { return lo + static_cast<int64_t>(Next() % static_cast<uint64_t>(hi - lo)); }
};

template <typename T> Vector<T> *Column(const T *vals, int32_t n, DataTypeId id)
{
    auto *v = new Vector<T>(n, id);
    std::memcpy(unsafe::UnsafeVector::GetRawValues(v), vals, sizeof(T) * static_cast<size_t>(n));
    return v;
}

double Now()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

double KernelMs(const char *prefix)
{
    int64_t n = 0;
    double ms = 0;
    omni_gpu_profile_get(prefix, &n, &ms);
    return ms;
}

int RunAgg(int steps, int warmup)
{
    const int32_t rows = 1000000, batch = 4096;
    Rng rng{0xC1};
    std::vector<int64_t> keys(rows), vals(rows);
    for (int32_t i = 0; i < rows; ++i) {
        keys[i] = rng.Range(0, 1000);
        vals[i] = rng.Range(-1000000, 1000001);
    }
    std::vector<uint32_t> groupCols = {0};
    std::vector<std::vector<uint32_t>> aggCols = {{1}, {1}, {1}};
    std::vector<DataTypes> inTypes(3, DataTypes({LongType()}));
    std::vector<DataTypes> outTypes(4, DataTypes({LongType()}));
    std::vector<uint32_t> funcs = {OMNI_AGG_SUM, OMNI_AGG_COUNT_ALL, OMNI_AGG_MIN, OMNI_AGG_MAX};
    std::vector<uint32_t> masks;
    // COUNT(*) takes no input column: aggsCols / aggInputTypes list the other three (group_aggregation.cpp:102-125)
    GpuHashAggregationOperatorFactory factory(groupCols, DataTypes({LongType()}), aggCols, inTypes, outTypes, funcs,
                                              masks, std::vector<bool>(4, true), std::vector<bool>(4, false), false);
    double total = 0, kernel = 0;
    int64_t groups = 0;
    for (int s = 0; s < warmup + steps; ++s) {
        std::vector<VectorBatch *> in;  // made before the clock starts: AddInput consumes them
        for (int32_t a = 0; a < rows; a += batch) {
            const int32_t n = std::min(batch, rows - a);
            auto *b = new VectorBatch(n);
            b->Append(Column(keys.data() + a, n, OMNI_LONG));
            b->Append(Column(vals.data() + a, n, OMNI_LONG));
            in.push_back(b);
        }
        omni_gpu_device_synchronize();
        if (s == warmup) {
            omni_gpu_profile_reset();
            omni_gpu_profile_enable(1);
        }
        const double t0 = Now();
        Operator *op = factory.CreateOperator();
        for (auto *b : in) op->AddInput(b);
        int64_t g = 0;
        for (;;) {
            VectorBatch *out = nullptr;
            op->GetOutput(&out);
            if (!out) break;
            g += out->GetRowCount();
            VectorHelper::FreeVecBatch(out);
        }
        Operator::DeleteOperator(op);
        const double t1 = Now();
        if (s >= warmup) total += t1 - t0;
        groups = g;
    }
    kernel = KernelMs("agg_insert");
    omni_gpu_profile_enable(0);
    std::printf("{\"rows\": %d, \"ms_per_step\": %.4f, \"groups\": %lld, \"groups_ok\": %s, \"kernel\": \"agg_insert\", "
                "\"kernel_ms_per_step\": %.4f}\n",
                rows, total / steps, static_cast<long long>(groups), groups == 1000 ? "true" : "false",
                kernel / steps);
    return groups == 1000 ? 0 : 1;
}

int RunJoin(int steps, int warmup)
{
    const int32_t nb = 2000000, np = 64000000, batch = 65536;
    std::vector<int64_t> bkeys(nb), pkeys(np);
    std::vector<int32_t> bpay(nb);
    Rng rng{0xC2};
    for (int32_t i = 0; i < nb; ++i) {
        bkeys[i] = static_cast<int64_t>(SplitMix(static_cast<uint64_t>(i)));
        bpay[i] = static_cast<int32_t>(rng.Next() & 0x7FFFFFFF);
    }
    int64_t hits = 0;
    for (int32_t i = 0; i < np; ++i) {
        const uint64_t r = rng.Next();
        if (r % 10 < 8) {
            pkeys[i] = bkeys[(r >> 8) % nb];
            ++hits;
        } else {
            pkeys[i] = static_cast<int64_t>(SplitMix(static_cast<uint64_t>(nb) + ((r >> 8) & ((1ULL << 40) - 1))));
        }
    }
    // misses are splitmix images of values >= nb, distinct from every build key's preimage: no false hit
    DataTypes buildTypes({LongType(), IntType()});
    DataTypes probeTypes({LongType()});
    DataTypes buildOutTypes({IntType()});
    int32_t hashCols[1] = {0}, probeOut[1] = {0}, probeHash[1] = {0}, buildOut[1] = {1};
    double total = 0;
    int64_t matches = 0;
    for (int s = 0; s < warmup + steps; ++s) {
        std::vector<VectorBatch *> bin, pin;
        for (int32_t a = 0; a < nb; a += batch) {
            const int32_t n = std::min(batch, nb - a);
            auto *b = new VectorBatch(n);
            b->Append(Column(bkeys.data() + a, n, OMNI_LONG));
            b->Append(Column(bpay.data() + a, n, OMNI_INT));
            bin.push_back(b);
        }
        for (int32_t a = 0; a < np; a += batch) {
            const int32_t n = std::min(batch, np - a);
            auto *b = new VectorBatch(n);
            b->Append(Column(pkeys.data() + a, n, OMNI_LONG));
            pin.push_back(b);
        }
        omni_gpu_device_synchronize();
        if (s == warmup) {
            omni_gpu_profile_reset();
            omni_gpu_profile_enable(1);
        }
        const double t0 = Now();
        auto *bf = GpuHashBuilderOperatorFactory::CreateHashBuilderOperatorFactory(OMNI_JOIN_TYPE_INNER, buildTypes,
                                                                                   hashCols, 1, 1);
        Operator *builder = bf->CreateOperator();
        for (auto *b : bin) builder->AddInput(b);
        VectorBatch *none = nullptr;
        builder->GetOutput(&none);
        auto *lf = GpuLookupJoinOperatorFactory::CreateLookupJoinOperatorFactory(
            probeTypes, probeOut, 1, probeHash, 1, buildOut, 1, buildOutTypes, reinterpret_cast<int64_t>(bf), nullptr,
            false, nullptr);
        Operator *lookup = lf->CreateOperator();
        // as LocalPlanner creates operators (local_planner.cpp:30-40): more input announced, noMoreInput() at the
        // end -- the lookup probes buffered runs of batches
        lookup->setNoMoreInput(false);
        int64_t m = 0;
        auto drain = [&]() {
            for (;;) {
                VectorBatch *out = nullptr;
                lookup->GetOutput(&out);
                if (!out) break;
                m += out->GetRowCount();
                VectorHelper::FreeVecBatch(out);
            }
        };
        for (auto *b : pin) {
            lookup->AddInput(b);
            drain();
        }
        lookup->noMoreInput();
        drain();
        Operator::DeleteOperator(lookup);
        Operator::DeleteOperator(builder);
        delete lf;
        delete bf;
        const double t1 = Now();
        if (s >= warmup) total += t1 - t0;
        matches = m;
    }
    const double probeMs = KernelMs("join_probe");
    const double buildMs = KernelMs("join_build") + KernelMs("join_gather_build");
    omni_gpu_profile_enable(0);
    std::printf("{\"rows\": %lld, \"build_rows\": %d, \"probe_rows\": %d, \"ms_per_step\": %.4f, \"matches\": %lld, "
                "\"matches_ok\": %s, \"kernel\": \"join_probe\", \"kernel_ms_per_step\": %.4f, "
                "\"build_kernel_ms_per_step\": %.4f}\n",
                static_cast<long long>(nb) + np, nb, np, total / steps, static_cast<long long>(matches),
                matches == hits ? "true" : "false", probeMs / steps, buildMs / steps);
    return matches == hits ? 0 : 1;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s agg|join [steps] [warmup]\n", argv[0]);
        return 2;
    }
    const std::string mode = argv[1];
    const int steps = argc > 2 ? std::max(1, std::atoi(argv[2])) : 3;
    const int warmup = argc > 3 ? std::max(0, std::atoi(argv[3])) : 1;
    try {
        if (omni_gpu_init(0) != OMNI_GPU_OK) {
            std::fprintf(stderr, "omni_gpu_init: %s\n", omni_gpu_last_error());
            return 1;
        }
        if (mode == "agg") return RunAgg(steps, warmup);
        if (mode == "join") return RunJoin(steps, warmup);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
    return 2;
}
