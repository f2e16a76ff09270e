// GPU-backed drop-ins for the reference's hash-aggregation and hash-join operators.
//
// The factories keep the reference constructors' argument meaning, so a caller (JNI, the native driver,
// tests) switches by constructing the Gpu* factory instead of the CPU one:
//   HashAggregationOperatorFactory  core/src/operator/aggregation/group_aggregation.h:233-245
//   HashBuilderOperatorFactory      core/src/operator/join/hash_builder.h:28-31
//   LookupJoinOperatorFactory       core/src/operator/join/lookup_join.h:148-152
//   LookupOuterJoinOperatorFactory  core/src/operator/join/lookup_outer_join.h
// Operators follow the reference's ownership rules: aggregation consumes and frees each input batch in
// AddInput; the builder takes its batches (their data moves to device memory); lookup keeps the current
// probe batch until all its matches are emitted; the caller owns every output batch.
// Errors reported by the C ABI become OmniException (the reference's error channel).
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "omni_gpu.h"
#include "omni_vector.h"

namespace omniruntime {
namespace expressions {
class Expr;  // core/src/expression/expressions.h:78 (join filters are not on the GPU path)
}
namespace op {
class OverflowConfig;  // core/src/operator/config/operator_config.h:174

// Owning device buffer allocated through the C ABI.
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    ~DeviceBuffer();
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    void Reserve(size_t bytes);  // contents not kept
    void *Get() const { return ptr_; }

private:
    void *ptr_ = nullptr;
    size_t bytes_ = 0;
};

// Owning page-locked host buffer allocated through the C ABI.
class HostBuffer {
public:
    HostBuffer() = default;
    ~HostBuffer();
    HostBuffer(const HostBuffer &) = delete;
    HostBuffer &operator=(const HostBuffer &) = delete;
    void Reserve(size_t bytes);  // contents not kept
    void *Get() const { return ptr_; }
    size_t Bytes() const { return bytes_; }
    void Swap(HostBuffer &o)
    {
        std::swap(ptr_, o.ptr_);
        std::swap(bytes_, o.bytes_);
    }

private:
    void *ptr_ = nullptr;
    size_t bytes_ = 0;
};

// Rows per output batch: ceil(1 MiB / row size) (core/src/operator/util/operator_util.h:111-115).
int32_t MaxRowCount(int32_t rowSize);

class GpuHashAggregationOperatorFactory : public OperatorFactory {
public:
    GpuHashAggregationOperatorFactory(std::vector<uint32_t> &groupByCol, const type::DataTypes &groupInputTypes,
                                      std::vector<std::vector<uint32_t>> &aggsCols,
                                      std::vector<type::DataTypes> &aggInputTypes,
                                      std::vector<type::DataTypes> &aggOutputTypes, std::vector<uint32_t> &aggFuncTypes,
                                      std::vector<uint32_t> &maskColsVector, std::vector<bool> inputRaws,
                                      std::vector<bool> outputPartials, bool overflowAsNull = false,
                                      bool isStatisticalAggregate = false);
    OmniStatus Init();
    Operator *CreateOperator() override;

    // Learned capacity hint (expected groups) forwarded to every operator this factory creates.
    void SetCapacityHint(int64_t groups) { spec_.capacity_hint = groups; }
    void SetStrategy(int32_t strategy) { spec_.strategy = strategy; }

    // How one aggregate of the factory (function + stage) reads the GPU aggregates of spec_ from `slot` on.
    enum AggRule {
        kDirect = 0,     // one GPU aggregate, one output column
        kAvgPartial = 1, // AVG with partial output: running sum (partial_out) + COUNT_COLUMN -> (sum, count)
        kAvgFinal = 2,   // AVG over partial (sum, count) input: SUM + SUM -> sum / count, NULL when count is 0
        kAvgMerge = 3,   // AVG over partial input with partial output: SUM + SUM -> (sum, count)
        // decimal SUM (SumSparkDecimalAggregator, partial state (sum, isEmpty), sum_spark_decimal_aggregator.h):
        kDecPartial = 4, // raw input, partial output: SUM (partial_out) + COUNT_COLUMN -> (sum, count == 0)
        kDecFinal = 5,   // partial input, final output: SUM over the non-empty sums + COUNT of NULL sums (a NULL
                         // partial sum marks an overflow, which sticks) -> sum, or the overflow when that count > 0
        kDecMerge = 6,   // partial input, partial output: kDecFinal's two + COUNT of the non-empty sums
    };
    // How a GPU aggregate's input column is derived from the batch (decimal SUM over partial input).
    enum InputMode {
        kPlain = 0,      // the column as it is
        kNonEmpty = 1,   // the sum column, NULL also where the isEmpty column is true
        kNullMarks = 2,  // non-NULL exactly where the sum column is NULL
    };
    struct AggPlan {
        AggRule rule;
        int32_t slot;
        std::vector<int32_t> outTypes;
        // the factory's aggregate, for AlignSchema: function, input column (-1: COUNT(*)), input type and
        // whether its input is raw and its output partial
        int32_t func = 0;
        int32_t col = -1;
        int32_t inType = 0;
        bool raw = true;
        bool partial = false;
    };

private:
    friend class GpuHashAggregationOperator;
    int32_t AddSlot(int32_t func, int32_t inType, int32_t outType, int32_t inScale, int32_t outScale, int32_t col,
                    bool partialOut, InputMode mode = kPlain, int32_t maskCol = -1);
    std::vector<uint32_t> groupByCols_;
    type::DataTypes groupTypes_;
    std::vector<int32_t> aggInputCol_;  // per GPU aggregate: input column, -1 for COUNT_ALL
    std::vector<int32_t> aggInputMode_; // per GPU aggregate: InputMode
    std::vector<int32_t> aggMaskCol_;   // per GPU aggregate: the isEmpty column of kNonEmpty, else -1
    std::vector<AggPlan> plan_;         // per aggregate of the factory
    omni_gpu_agg_spec spec_{};
};

class GpuHashAggregationOperator : public Operator {
public:
    explicit GpuHashAggregationOperator(const GpuHashAggregationOperatorFactory &f);
    ~GpuHashAggregationOperator() override;
    int32_t AddInput(vec::VectorBatch *vecBatch) override;
    int32_t GetOutput(vec::VectorBatch **result) override;
    OmniStatus Close() override;
    uint64_t GetHashMapUniqueKeys() override;
    // HashAggregationOperator::AlignSchema (group_aggregation.cpp:1182-1218): a raw input batch reshaped into the
    // partial output schema row by row (the group-by columns, then each aggregate's partial columns from that
    // row alone), for a partial stage that passes rows through instead of aggregating them.  Frees the input.
    vec::VectorBatch *AlignSchema(vec::VectorBatch *inputVecBatch) override;

    static constexpr int64_t kFlushRows = 1 << 20;  // staged rows that trigger a copy and an insert

private:
    // One input column's staged rows in page-locked host memory (col < 0: COUNT(*), nothing staged).
    struct StagedColumn {
        int32_t col = -1;
        int32_t mode = 0;   // GpuHashAggregationOperatorFactory::InputMode
        int32_t mask = -1;  // the isEmpty column of kNonEmpty
        int32_t type = 0;
        int32_t width = 0;
        bool anyNull = false;
        HostBuffer values, nulls;
    };
    void Flush();
    std::vector<uint32_t> groupByCols_;
    std::vector<int32_t> aggInputCol_;
    std::vector<int32_t> aggInputMode_;
    std::vector<int32_t> aggMaskCol_;
    std::vector<int32_t> outTypes_;  // GPU output columns: keys, then one per GPU aggregate
    std::vector<GpuHashAggregationOperatorFactory::AggPlan> plan_;
    omni_gpu_agg_spec spec_;
    void *handle_ = nullptr;
    bool hasInput_ = false;
    bool pending_ = false;  // a batch may still be reading the staging buffers
    int64_t groups_ = -1;
    int64_t emitted_ = 0;
    int32_t rowsPerBatch_ = 0;
    std::vector<std::unique_ptr<DeviceBuffer>> staging_;
    std::vector<std::unique_ptr<StagedColumn>> stage_;  // keys, then aggregate inputs
    int64_t stagedRows_ = 0;
    int64_t stageCap_ = kFlushRows;
};

class GpuHashBuilderOperatorFactory : public OperatorFactory {
public:
    // hash_builder.cpp:122-134: without a build side it is OMNI_BUILD_UNKNOWN (hash_builder.h:48)
    static GpuHashBuilderOperatorFactory *CreateHashBuilderOperatorFactory(JoinType joinType,
                                                                           const type::DataTypes &buildTypes,
                                                                           const int32_t *buildHashCols,
                                                                           int32_t buildHashColsCount,
                                                                           int32_t operatorCount);
    static GpuHashBuilderOperatorFactory *CreateHashBuilderOperatorFactory(JoinType joinType, BuildSide buildSide,
                                                                           const type::DataTypes &buildTypes,
                                                                           const int32_t *buildHashCols,
                                                                           int32_t buildHashColsCount,
                                                                           int32_t operatorCount);
    ~GpuHashBuilderOperatorFactory() override;
    Operator *CreateOperator() override;
    // One GPU table per builder partition (operatorCount, hash_builder.cpp:155-160): builder operator i fills
    // table i % operatorCount; a probe row looks only in the table of its hash partition.
    void *Table(int32_t partition = 0) const { return tables_.at(partition); }
    int32_t Partitions() const { return static_cast<int32_t>(tables_.size()); }
    const type::DataTypes &BuildTypes() const { return buildTypes_; }
    bool Finished() const { return finishedCount_ == Partitions(); }
    JoinType GetJoinType() const { return joinType_; }
    BuildSide GetBuildSide() const { return buildSide_; }
    const std::vector<int32_t> &KeyCols() const { return keyCols_; }

private:
    friend class GpuHashBuilderOperator;
    GpuHashBuilderOperatorFactory(JoinType joinType, BuildSide buildSide, const type::DataTypes &buildTypes,
                                  std::vector<int32_t> keyCols, int32_t operatorCount);
    JoinType joinType_;
    BuildSide buildSide_;
    type::DataTypes buildTypes_;
    std::vector<int32_t> keyCols_;  // buildHashCols
    std::vector<void *> tables_;
    std::vector<bool> partFinished_;
    int32_t finishedCount_ = 0;
    int32_t nextOperator_ = 0;
    // per partition: its build batches in device memory until its table is built
    std::vector<std::vector<std::unique_ptr<DeviceBuffer>>> retained_;
};

class GpuHashBuilderOperator : public Operator {
public:
    GpuHashBuilderOperator(GpuHashBuilderOperatorFactory &f, int32_t partition) : f_(f), partition_(partition) {}
    int32_t AddInput(vec::VectorBatch *vecBatch) override;
    int32_t GetOutput(vec::VectorBatch **result) override;

private:
    GpuHashBuilderOperatorFactory &f_;
    int32_t partition_;
};

class GpuLookupJoinOperatorFactory : public OperatorFactory {
public:
    // The signature of lookup_join.h:148-152.  filterExpr must be null (join filters are not on the GPU path);
    // isShuffleExchangeBuildPlan and overflowConfig are accepted and unused.
    static GpuLookupJoinOperatorFactory *CreateLookupJoinOperatorFactory(
        const type::DataTypes &probeTypes, int32_t *probeOutputCols, int32_t probeOutputColsCount,
        int32_t *probeHashCols, int32_t probeHashColsCount, int32_t *buildOutputCols, int32_t buildOutputColsCount,
        const type::DataTypes &buildOutputTypes, int64_t hashBuilderFactoryAddr, expressions::Expr *filterExpr,
        bool isShuffleExchangeBuildPlan, OverflowConfig *overflowConfig);
    Operator *CreateOperator() override;

private:
    friend class GpuLookupJoinOperator;
    type::DataTypes probeTypes_;
    std::vector<int32_t> probeOutputCols_;
    std::vector<int32_t> probeKeyCols_;  // probeHashCols
    std::vector<int32_t> buildOutputCols_;
    GpuHashBuilderOperatorFactory *builder_ = nullptr;
};

// Probes one batch at a time, as the reference does, unless the caller announced more input
// (setNoMoreInput(false), as LocalPlanner does for every operator, local_planner.cpp:30-40): then AddInput stages
// the probe batches' columns in page-locked host memory and every run of about half the table's slots (at least
// 2^20, at most 2^24 rows) -- or the rest, once noMoreInput() is called -- is probed in one call (the
// radix-partitioned chain for large tables), the probe output columns are gathered on the device, and GetOutput
// emits the run's rows before the operator takes more input (needsInput() is false meanwhile).  Outputs are the
// same rows either way; only their grouping into batches differs.  With several builder partitions each probe row
// probes the table of its hash partition, one batch at a time.
class GpuLookupJoinOperator : public Operator {
public:
    explicit GpuLookupJoinOperator(const GpuLookupJoinOperatorFactory &f) : f_(f) {}
    ~GpuLookupJoinOperator() override;
    int32_t AddInput(vec::VectorBatch *vecBatch) override;
    int32_t GetOutput(vec::VectorBatch **result) override;
    bool needsInput() override;

private:
    // One probe column's staged rows (key or probe output column).
    struct StagedProbeColumn {
        int32_t col = -1;
        int32_t type = 0;
        int32_t width = 0;
        bool anyNull = false;
        HostBuffer values, nulls;
    };
    int32_t ProbeOne(vec::VectorBatch *vecBatch);
    void ProbeTable(void *table, const std::vector<omni_gpu_column> &keys, int64_t rows, const int64_t *rowMap);
    void Stage(vec::VectorBatch *vecBatch);
    void FlushRun();
    vec::VectorBatch *EmitRun();
    const GpuLookupJoinOperatorFactory &f_;
    // buffered probe runs
    std::vector<vec::VectorBatch *> pending_;   // staged probe batches (freed once the run is emitted)
    std::vector<vec::VectorBatch *> emitting_;  // the batches of the run being emitted
    std::vector<std::unique_ptr<StagedProbeColumn>> stage_;  // key columns, then the other probe output columns
    std::vector<int32_t> outStage_;             // per probe output column: its index in stage_
    int64_t stagedRows_ = 0;
    int64_t stageCap_ = 0;
    int64_t flushRows_ = 0;
    bool inRun_ = false;                        // the rows of a flushed run are being emitted
    std::vector<std::unique_ptr<HostBuffer>> runVals_, runNulls_;  // probe output columns, then build output columns
    std::vector<bool> runAnyNull_;
    vec::VectorBatch *probe_ = nullptr;
    std::vector<int64_t> matchRows_;
    std::vector<std::vector<uint8_t>> matchPay_;       // per build output column: values
    std::vector<std::vector<uint8_t>> matchPayNulls_;  // per build output column: bitmap
    std::vector<uint8_t> exists_;                      // existence join: one flag per probe row
    int64_t matches_ = 0;
    int64_t emitted_ = 0;
    DeviceBuffer rowsBuf_;
    std::vector<std::unique_ptr<DeviceBuffer>> keyBufs_, keyNullBufs_;  // per probe key column
    std::vector<std::unique_ptr<DeviceBuffer>> payBufs_, payNullBufs_;
    std::vector<std::unique_ptr<DeviceBuffer>> gatherBufs_, gatherNullBufs_;  // gathered probe output columns
    int32_t OutputRowSize() const;
    int32_t EmitOne(vec::VectorBatch **result);
};

// LookupOuterJoinOperatorFactory (core/src/operator/join/lookup_outer_join.h): after every probe of a
// build-preserving outer join, emits the build rows no probe matched, probe output columns NULL.
class GpuLookupOuterJoinOperatorFactory : public OperatorFactory {
public:
    static GpuLookupOuterJoinOperatorFactory *CreateLookupOuterJoinOperatorFactory(
        const type::DataTypes &probeTypes, int32_t *probeOutputCols, int32_t probeOutputColsCount,
        int32_t *buildOutputCols, const type::DataTypes &buildOutputTypes, int64_t hashBuilderFactoryAddr);
    Operator *CreateOperator() override;

private:
    friend class GpuLookupOuterJoinOperator;
    type::DataTypes probeTypes_;
    std::vector<int32_t> probeOutputCols_;
    std::vector<int32_t> buildOutputCols_;
    GpuHashBuilderOperatorFactory *builder_ = nullptr;
};

class GpuLookupOuterJoinOperator : public Operator {
public:
    explicit GpuLookupOuterJoinOperator(const GpuLookupOuterJoinOperatorFactory &f) : f_(f) {}
    int32_t AddInput(vec::VectorBatch *vecBatch) override;
    int32_t GetOutput(vec::VectorBatch **result) override;

private:
    const GpuLookupOuterJoinOperatorFactory &f_;
    bool fetched_ = false;
    int64_t rows_ = 0;
    int64_t emitted_ = 0;
    std::vector<std::vector<uint8_t>> pay_, payNulls_;  // per build output column
};

}  // namespace op
}  // namespace omniruntime
