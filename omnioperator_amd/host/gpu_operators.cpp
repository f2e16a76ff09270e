// GPU-backed HashAggregation / HashBuilder / LookupJoin operators over the C ABI of include/omni_gpu.h.
#include "gpu_operators.h"

#include <algorithm>
#include <cstring>
#include <string>

namespace omniruntime {
namespace op {

namespace {
constexpr int32_t kMaxVecBatchBytes = 1024 * 1024;  // operator_util.h:33

void Check(int rc, const char *what)
{
    if (rc != OMNI_GPU_OK) {
        std::string code = rc == OMNI_GPU_ERR_OVERFLOW ? "OPERATOR_RUNTIME_ERROR" : "GPU_ERROR";
        throw OmniException(code, std::string(what) + ": " + omni_gpu_last_error());
    }
}

void EnsureInit()
{
    static bool done = false;
    if (!done) {
        Check(omni_gpu_init(0), "omni_gpu_init");
        done = true;
    }
}

// Storage address of row 0 of a flat vector (the slice's own row 0, UnsafeVector::GetRawValues); the storage type
// follows the type id as VectorHelper::CreateFlatVector lays it out.
void *FlatValues(vec::BaseVector *v)
{
    using vec::Vector;
    using vec::unsafe::UnsafeVector;
    switch (v->GetTypeId()) {
        case type::OMNI_DOUBLE: return UnsafeVector::GetRawValues(static_cast<Vector<double> *>(v));
        case type::OMNI_BOOLEAN: return UnsafeVector::GetRawValues(static_cast<Vector<uint8_t> *>(v));
        case type::OMNI_BYTE: return UnsafeVector::GetRawValues(static_cast<Vector<int8_t> *>(v));
        case type::OMNI_SHORT: return UnsafeVector::GetRawValues(static_cast<Vector<int16_t> *>(v));
        case type::OMNI_INT:
        case type::OMNI_DATE32: return UnsafeVector::GetRawValues(static_cast<Vector<int32_t> *>(v));
        case type::OMNI_DECIMAL128: return UnsafeVector::GetRawValues(static_cast<Vector<type::Decimal128> *>(v));
        default: return UnsafeVector::GetRawValues(static_cast<Vector<int64_t> *>(v));
    }
}

// Rows [0, rows) of a dictionary vector gathered through its ids into `out` (NULL rows zero).
template <typename T> void GatherDictionary(vec::BaseVector *v, int32_t rows, std::vector<uint8_t> &out)
{
    using vec::unsafe::UnsafeDictionaryVector;
    auto *dv = static_cast<vec::Vector<vec::DictionaryContainer<T>> *>(v);
    const T *dict = UnsafeDictionaryVector::GetDictionary(dv);
    const int32_t *ids = UnsafeDictionaryVector::GetIds(dv);
    out.assign(static_cast<size_t>(rows) * sizeof(T), 0);
    T *o = reinterpret_cast<T *>(out.data());
    for (int32_t r = 0; r < rows; ++r) {
        if (!v->IsNull(r)) o[r] = dict[ids[r]];
    }
}

// Host address of rows [0, rows) of a column as flat values: a flat vector (or slice) in place, a dictionary
// vector gathered into `tmp`; other encodings are not on the GPU path.
const void *ColumnValues(vec::BaseVector *v, int32_t rows, std::vector<uint8_t> &tmp)
{
    switch (v->GetEncoding()) {
        case vec::OMNI_FLAT: return FlatValues(v);
        case vec::OMNI_DICTIONARY:
            switch (v->GetTypeId()) {
                case type::OMNI_DOUBLE: GatherDictionary<double>(v, rows, tmp); break;
                case type::OMNI_BOOLEAN: GatherDictionary<uint8_t>(v, rows, tmp); break;
                case type::OMNI_BYTE: GatherDictionary<int8_t>(v, rows, tmp); break;
                case type::OMNI_SHORT: GatherDictionary<int16_t>(v, rows, tmp); break;
                case type::OMNI_INT:
                case type::OMNI_DATE32: GatherDictionary<int32_t>(v, rows, tmp); break;
                case type::OMNI_DECIMAL128: GatherDictionary<type::Decimal128>(v, rows, tmp); break;
                default: GatherDictionary<int64_t>(v, rows, tmp); break;
            }
            return tmp.data();
        default:
            throw OmniException("UNSUPPORTED_ERROR",
                                "vector encoding " + std::to_string(v->GetEncoding()) + " is not on the GPU path");
    }
}

// Copies a column to device memory: values into `vals`, the null bitmap into `nulls` when it has NULLs.
omni_gpu_column Upload(vec::BaseVector *v, int32_t rows, DeviceBuffer &vals, DeviceBuffer &nulls)
{
    omni_gpu_column c{};
    c.type_id = v->GetTypeId();
    std::vector<uint8_t> tmp;
    const void *src = ColumnValues(v, rows, tmp);
    size_t bytes = static_cast<size_t>(rows) * type::TypeWidth(v->GetTypeId());
    vals.Reserve(bytes);
    Check(omni_gpu_memcpy_h2d(vals.Get(), src, bytes), "upload values");
    c.values = vals.Get();
    if (v->HasNull()) {
        size_t nb = (rows + 7) / 8;
        nulls.Reserve(nb);
        Check(omni_gpu_memcpy_h2d(nulls.Get(), vec::unsafe::UnsafeBaseVector::GetNulls(v), nb), "upload nulls");
        c.nulls = static_cast<const uint8_t *>(nulls.Get());
    }
    return c;
}

// Copies a downloaded NULL bitmap of `rows` rows into a new flat vector and sets its NULL flag.
void SetNullsFrom(vec::BaseVector *v, const uint8_t *bits, int32_t rows)
{
    uint8_t *dst = vec::unsafe::UnsafeBaseVector::GetNulls(v);
    const size_t nb = (rows + 7) / 8;
    std::memcpy(dst, bits, nb);
    if (rows % 8) dst[nb - 1] &= static_cast<uint8_t>((1u << (rows % 8)) - 1);
    bool any = false;
    for (size_t i = 0; i < nb && !any; ++i) any = dst[i] != 0;
    v->SetNullFlag(any);
}
}  // namespace

HostBuffer::~HostBuffer()
{
    if (ptr_) omni_gpu_host_free(ptr_);
}

void HostBuffer::Reserve(size_t bytes)
{
    if (bytes <= bytes_ && ptr_) return;
    if (ptr_) omni_gpu_host_free(ptr_);
    ptr_ = nullptr;
    bytes_ = 0;
    Check(omni_gpu_host_alloc(&ptr_, bytes ? bytes : 1), "omni_gpu_host_alloc");
    bytes_ = bytes ? bytes : 1;
}

DeviceBuffer::~DeviceBuffer()
{
    if (ptr_) omni_gpu_free(ptr_);
}

void DeviceBuffer::Reserve(size_t bytes)
{
    if (bytes <= bytes_ && ptr_) return;
    if (ptr_) omni_gpu_free(ptr_);
    ptr_ = nullptr;
    bytes_ = 0;
    Check(omni_gpu_malloc(&ptr_, bytes ? bytes : 1), "omni_gpu_malloc");
    bytes_ = bytes ? bytes : 1;
}

int32_t MaxRowCount(int32_t rowSize) { return (kMaxVecBatchBytes + rowSize - 1) / rowSize; }

namespace {
// Value r of a flat vector of T.
template <typename T> T Cell(vec::BaseVector *v, int32_t r) { return static_cast<vec::Vector<T> *>(v)->GetValue(r); }
}  // namespace

// ================================================================================================
// Hash aggregation
// ================================================================================================

GpuHashAggregationOperatorFactory::GpuHashAggregationOperatorFactory(
    std::vector<uint32_t> &groupByCol, const type::DataTypes &groupInputTypes,
    std::vector<std::vector<uint32_t>> &aggsCols, std::vector<type::DataTypes> &aggInputTypes,
    std::vector<type::DataTypes> &aggOutputTypes, std::vector<uint32_t> &aggFuncTypes,
    std::vector<uint32_t> &maskColsVector, std::vector<bool> inputRaws, std::vector<bool> outputPartials,
    bool overflowAsNull, bool isStatisticalAggregate)
    : groupByCols_(groupByCol), groupTypes_(groupInputTypes)
{
    if (!maskColsVector.empty()) throw OmniException("UNSUPPORTED_ERROR", "aggregate masks are not on the GPU path");
    if (groupByCol.empty() || groupByCol.size() > OMNI_GPU_MAX_KEYS) {
        throw OmniException("UNSUPPORTED_ERROR", "1 to 4 group-by columns are supported on the GPU path");
    }
    if (aggFuncTypes.size() > OMNI_GPU_MAX_AGGS) throw OmniException("UNSUPPORTED_ERROR", "too many aggregates");
    spec_.num_keys = static_cast<int32_t>(groupByCol.size());
    for (size_t i = 0; i < groupByCol.size(); ++i) spec_.key_types[i] = groupInputTypes.GetType(i)->GetId();
    spec_.num_aggs = 0;
    spec_.overflow_as_null = overflowAsNull ? 1 : 0;
    // COUNT_ALL aggregates over raw input take no input column: aggsCols / aggInputTypes list only the others
    // (HashAggregationOperatorFactory::CreateOperator, group_aggregation.cpp:102-125).
    // Two-stage aggregation (inputRaw / outputPartial per aggregate): SUM, MIN and MAX have partial states equal
    // to their outputs, so a final stage adds / compares the partial values; a final COUNT adds the partial
    // counts; AVG's partial output is (double sum, long count) as two flat columns (the non-container layout of
    // AverageFlatIMAggregator::ExtractValuesBatch, average_flatim_aggregator.h:256-273) and its final stage adds
    // both and divides (:40-45, 275-290).
    size_t inputIdx = 0;
    for (size_t i = 0; i < aggFuncTypes.size(); ++i) {
        const bool raw = i < inputRaws.size() ? inputRaws[i] : true;
        const bool partial = i < outputPartials.size() ? outputPartials[i] : false;
        const int32_t func = static_cast<int32_t>(aggFuncTypes[i]);
        const type::DataTypes &outs = aggOutputTypes.at(i);
        const int32_t out0 = outs.GetType(0)->GetId();
        if (func == OMNI_AGG_COUNT_ALL && raw) {
            plan_.push_back({kDirect, AddSlot(func, type::OMNI_LONG, type::OMNI_LONG, 0, 0, -1, false), {out0}});
            plan_.back().func = func;
            plan_.back().partial = partial;
            continue;
        }
        const type::DataTypes &ins = aggInputTypes.at(inputIdx);
        const std::vector<uint32_t> &cols = aggsCols.at(inputIdx);
        ++inputIdx;
        const int32_t in0 = ins.GetType(0)->GetId();
        const int32_t col0 = static_cast<int32_t>(cols.at(0));
        const size_t planned = plan_.size();
        if (func == OMNI_AGG_COUNT_ALL || func == OMNI_AGG_COUNT_COLUMN) {
            const int32_t slot = raw ? AddSlot(OMNI_AGG_COUNT_COLUMN, in0, type::OMNI_LONG, 0, 0, col0, false)
                                     : AddSlot(OMNI_AGG_SUM, type::OMNI_LONG, type::OMNI_LONG, 0, 0, col0, false);
            plan_.push_back({kDirect, slot, {out0}});
        } else if (func == OMNI_AGG_AVG && raw && !partial) {
            plan_.push_back({kDirect, AddSlot(func, in0, out0, 0, 0, col0, false), {out0}});
        } else if (func == OMNI_AGG_AVG && raw) {
            if (outs.GetSize() != 2 || out0 != type::OMNI_DOUBLE || outs.GetType(1)->GetId() != type::OMNI_LONG) {
                throw OmniException("UNSUPPORTED_ERROR", "partial AVG output must be (DOUBLE sum, LONG count)");
            }
            const int32_t slot = AddSlot(OMNI_AGG_AVG, in0, type::OMNI_DOUBLE, 0, 0, col0, true);
            AddSlot(OMNI_AGG_COUNT_COLUMN, in0, type::OMNI_LONG, 0, 0, col0, false);
            plan_.push_back({kAvgPartial, slot, {type::OMNI_DOUBLE, type::OMNI_LONG}});
        } else if (func == OMNI_AGG_AVG) {
            if (cols.size() != 2 || ins.GetSize() != 2 || in0 != type::OMNI_DOUBLE ||
                ins.GetType(1)->GetId() != type::OMNI_LONG) {
                throw OmniException("UNSUPPORTED_ERROR", "AVG over partial input takes (DOUBLE sum, LONG count) columns");
            }
            const int32_t slot = AddSlot(OMNI_AGG_SUM, type::OMNI_DOUBLE, type::OMNI_DOUBLE, 0, 0, col0, false);
            AddSlot(OMNI_AGG_SUM, type::OMNI_LONG, type::OMNI_LONG, 0, 0, static_cast<int32_t>(cols[1]), false);
            if (partial) plan_.push_back({kAvgMerge, slot, {type::OMNI_DOUBLE, type::OMNI_LONG}});
            else plan_.push_back({kAvgFinal, slot, {type::OMNI_DOUBLE}});
        } else {
            const bool decimal = in0 == type::OMNI_DECIMAL64 || in0 == type::OMNI_DECIMAL128;
            const int32_t inScale = ins.GetType(0)->GetScale(), outScale = outs.GetType(0)->GetScale();
            if (decimal && func == OMNI_AGG_SUM && (!raw || partial)) {
                // SumSparkDecimalAggregator's partial state is (sum, isEmpty) (sum_spark_decimal_aggregator.h:68-72)
                if (partial && (outs.GetSize() != 2 || outs.GetType(1)->GetId() != type::OMNI_BOOLEAN)) {
                    throw OmniException("UNSUPPORTED_ERROR", "partial decimal SUM output must be (DECIMAL, BOOLEAN)");
                }
                if (!raw && (cols.size() != 2 || ins.GetSize() != 2 || ins.GetType(1)->GetId() != type::OMNI_BOOLEAN)) {
                    throw OmniException("UNSUPPORTED_ERROR", "decimal SUM over partial input takes (DECIMAL, BOOLEAN)");
                }
                if (raw) {
                    // ExtractValuesBatchInternal<OUTPUT_PARTIAL> (:246-301): the sum (0 when empty, NULL when
                    // overflowed) and isEmpty = no non-NULL input
                    const int32_t slot = AddSlot(func, in0, out0, inScale, outScale, col0, true);
                    AddSlot(OMNI_AGG_COUNT_COLUMN, in0, type::OMNI_LONG, 0, 0, col0, false);
                    plan_.push_back({kDecPartial, slot, {out0, type::OMNI_BOOLEAN}});
                } else {
                    // ProcessGroupInternalFinal / AddDecimalUseRowIndex (:16-66, 329-345): a NULL partial sum
                    // marks the group overflowed for good, an empty one is skipped, the others are added with the
                    // raw stage's checked add in row order
                    const int32_t empty = static_cast<int32_t>(cols[1]);
                    const int32_t slot = AddSlot(func, in0, out0, inScale, outScale, col0, partial, kNonEmpty, empty);
                    AddSlot(OMNI_AGG_COUNT_COLUMN, in0, type::OMNI_LONG, 0, 0, col0, false, kNullMarks);
                    if (partial) {
                        AddSlot(OMNI_AGG_COUNT_COLUMN, in0, type::OMNI_LONG, 0, 0, col0, false, kNonEmpty, empty);
                        plan_.push_back({kDecMerge, slot, {out0, type::OMNI_BOOLEAN}});
                    } else {
                        plan_.push_back({kDecFinal, slot, {out0}});
                    }
                }
            } else {
                plan_.push_back({kDirect, AddSlot(func, in0, out0, inScale, outScale, col0, false), {out0}});
            }
        }
        for (size_t k = planned; k < plan_.size(); ++k) {
            plan_[k].func = func;
            plan_[k].col = col0;
            plan_[k].inType = in0;
            plan_[k].raw = raw;
            plan_[k].partial = partial;
        }
    }
}

int32_t GpuHashAggregationOperatorFactory::AddSlot(int32_t func, int32_t inType, int32_t outType, int32_t inScale,
                                                   int32_t outScale, int32_t col, bool partialOut, InputMode mode,
                                                   int32_t maskCol)
{
    const int32_t n = spec_.num_aggs;
    if (n >= OMNI_GPU_MAX_AGGS) throw OmniException("UNSUPPORTED_ERROR", "too many aggregates");
    spec_.agg_funcs[n] = func;
    spec_.agg_in_types[n] = inType;
    spec_.agg_out_types[n] = outType;
    spec_.agg_in_scales[n] = inScale;
    spec_.agg_out_scales[n] = outScale;
    spec_.agg_partial_out[n] = partialOut ? 1 : 0;
    aggInputCol_.push_back(col);
    aggInputMode_.push_back(mode);
    aggMaskCol_.push_back(maskCol);
    spec_.num_aggs = n + 1;
    return n;
}

OmniStatus GpuHashAggregationOperatorFactory::Init() { return OMNI_STATUS_NORMAL; }

Operator *GpuHashAggregationOperatorFactory::CreateOperator() { return new GpuHashAggregationOperator(*this); }

GpuHashAggregationOperator::GpuHashAggregationOperator(const GpuHashAggregationOperatorFactory &f)
    : groupByCols_(f.groupByCols_), aggInputCol_(f.aggInputCol_), aggInputMode_(f.aggInputMode_),
      aggMaskCol_(f.aggMaskCol_), plan_(f.plan_), spec_(f.spec_)
{
    EnsureInit();
    Check(omni_gpu_agg_create(&spec_, &handle_), "omni_gpu_agg_create");
    int32_t rowSize = 0;  // of the output row (group_aggregation.cpp:414-429)
    for (int i = 0; i < spec_.num_keys; ++i) {
        outTypes_.push_back(spec_.key_types[i]);
        rowSize += type::TypeWidth(spec_.key_types[i]);
    }
    for (int i = 0; i < spec_.num_aggs; ++i) outTypes_.push_back(spec_.agg_out_types[i]);
    for (const auto &p : plan_) {
        for (int32_t t : p.outTypes) rowSize += type::TypeWidth(t);
    }
    rowsPerBatch_ = MaxRowCount(std::max(rowSize, 1));
}

GpuHashAggregationOperator::~GpuHashAggregationOperator() { Close(); }

OmniStatus GpuHashAggregationOperator::Close()
{
    if (handle_) {
        omni_gpu_agg_destroy(handle_);
        handle_ = nullptr;
    }
    return OMNI_STATUS_NORMAL;
}

// AddInput: group_aggregation.cpp:320-406 -- the batch is consumed and freed here.  Its columns are appended to
// page-locked host staging; once kFlushRows rows are staged (or at output), the run goes to the device in one copy
// per column and one omni_gpu_agg_add_batch, so many small batches cost one launch sequence and a few large copies
// instead of several small copies and launches each.  The groups do not depend on where the runs are cut.
int32_t GpuHashAggregationOperator::AddInput(vec::VectorBatch *vecBatch)
{
    hasInput_ = true;
    const int32_t rows = vecBatch->GetRowCount();
    if (rows <= 0) {
        vec::VectorHelper::FreeVecBatch(vecBatch);
        return 0;
    }
    const size_t nk = groupByCols_.size(), na = aggInputCol_.size();
    if (stage_.empty()) {
        for (size_t i = 0; i < nk + na; ++i) stage_.emplace_back(new StagedColumn());
        for (size_t i = 0; i < nk; ++i) stage_[i]->col = static_cast<int32_t>(groupByCols_[i]);
        for (size_t i = 0; i < na; ++i) {
            stage_[nk + i]->col = aggInputCol_[i];
            stage_[nk + i]->mode = aggInputMode_[i];
            stage_[nk + i]->mask = aggMaskCol_[i];
        }
    }
    if (stagedRows_ > 0 && stagedRows_ + rows > stageCap_) Flush();
    if (rows > stageCap_ || stage_[0]->values.Bytes() == 0) {
        stageCap_ = std::max<int64_t>(stageCap_, rows);
        for (auto &st : stage_) {
            if (st->col < 0) continue;
            const int32_t w = type::TypeWidth(vecBatch->Get(st->col)->GetTypeId());
            st->width = w;
            st->type = vecBatch->Get(st->col)->GetTypeId();
            st->values.Reserve(static_cast<size_t>(stageCap_) * w);
            st->nulls.Reserve(static_cast<size_t>(stageCap_ + 7) / 8 + 8);
            std::memset(st->nulls.Get(), 0, static_cast<size_t>(stageCap_ + 7) / 8 + 8);
        }
    }
    std::vector<uint8_t> tmp;
    for (auto &st : stage_) {
        if (st->col < 0) continue;
        vec::BaseVector *v = vecBatch->Get(st->col);
        const void *src = ColumnValues(v, rows, tmp);
        std::memcpy(static_cast<uint8_t *>(st->values.Get()) + static_cast<size_t>(stagedRows_) * st->width, src,
                    static_cast<size_t>(rows) * st->width);
        uint8_t *dst = static_cast<uint8_t *>(st->nulls.Get());
        if (st->mode == GpuHashAggregationOperatorFactory::kPlain) {
            if (v->HasNull()) {  // the batch's bits go to staged rows [stagedRows_, stagedRows_ + rows)
                const uint8_t *bits = vec::unsafe::UnsafeBaseVector::GetNulls(v);
                for (int32_t r = 0; r < rows; ++r) {
                    if ((bits[r >> 3] >> (r & 7)) & 1) {
                        const int64_t q = stagedRows_ + r;
                        dst[q >> 3] |= static_cast<uint8_t>(1u << (q & 7));
                        st->anyNull = true;
                    }
                }
            }
        } else {
            // a decimal partial sum column: NULL also where isEmpty is true (kNonEmpty), or NULL exactly where it
            // is not NULL (kNullMarks); the isEmpty column's raw values are read, as AddDecimalUseRowIndex does
            const uint8_t *bits = v->HasNull() ? vec::unsafe::UnsafeBaseVector::GetNulls(v) : nullptr;
            std::vector<uint8_t> tmpMask;
            const uint8_t *empty = st->mode == GpuHashAggregationOperatorFactory::kNonEmpty
                                       ? static_cast<const uint8_t *>(ColumnValues(vecBatch->Get(st->mask), rows, tmpMask))
                                       : nullptr;
            for (int32_t r = 0; r < rows; ++r) {
                const bool isNull = bits && ((bits[r >> 3] >> (r & 7)) & 1);
                const bool out = empty ? (isNull || empty[r] != 0) : !isNull;
                if (out) {
                    const int64_t q = stagedRows_ + r;
                    dst[q >> 3] |= static_cast<uint8_t>(1u << (q & 7));
                    st->anyNull = true;
                }
            }
        }
    }
    stagedRows_ += rows;
    vec::VectorHelper::FreeVecBatch(vecBatch);
    if (stagedRows_ >= kFlushRows) Flush();
    return 0;
}

// One copy per staged column and one insert for every batch staged since the last flush.
void GpuHashAggregationOperator::Flush()
{
    if (stagedRows_ == 0) return;
    if (pending_) Check(omni_gpu_agg_wait(handle_), "omni_gpu_agg_wait");  // the last insert still reads them
    const size_t nk = groupByCols_.size(), na = aggInputCol_.size();
    while (staging_.size() < 2 * (nk + na)) staging_.emplace_back(new DeviceBuffer());
    omni_gpu_column keys[OMNI_GPU_MAX_KEYS] = {};
    omni_gpu_column aggs[OMNI_GPU_MAX_AGGS] = {};
    for (size_t i = 0; i < nk + na; ++i) {
        StagedColumn &st = *stage_[i];
        omni_gpu_column &c = i < nk ? keys[i] : aggs[i - nk];
        if (st.col < 0) {  // COUNT(*): no input column
            c.type_id = type::OMNI_LONG;
            continue;
        }
        DeviceBuffer &dv = *staging_[2 * i], &dn = *staging_[2 * i + 1];
        const size_t bytes = static_cast<size_t>(stagedRows_) * st.width;
        dv.Reserve(bytes);
        Check(omni_gpu_memcpy_h2d(dv.Get(), st.values.Get(), bytes), "upload values");
        c.type_id = st.type;
        c.values = dv.Get();
        if (st.anyNull) {
            const size_t nb = static_cast<size_t>(stagedRows_ + 7) / 8;
            dn.Reserve(nb);
            Check(omni_gpu_memcpy_h2d(dn.Get(), st.nulls.Get(), nb), "upload nulls");
            c.nulls = static_cast<const uint8_t *>(dn.Get());
            std::memset(st.nulls.Get(), 0, nb);
            st.anyNull = false;
        }
    }
    Check(omni_gpu_agg_add_batch(handle_, keys, aggs, stagedRows_), "omni_gpu_agg_add_batch");
    pending_ = true;
    stagedRows_ = 0;
}

// GetOutput: group_aggregation.cpp:454-474, 1486-1538 -- batches of ceil(1 MiB / row size) rows.
int32_t GpuHashAggregationOperator::GetOutput(vec::VectorBatch **result)
{
    *result = nullptr;
    if (!hasInput_) return 0;
    if (groups_ < 0) {
        Flush();
        Check(omni_gpu_agg_num_groups(handle_, &groups_), "omni_gpu_agg_num_groups");
        pending_ = false;
    }
    if (emitted_ >= groups_) {
        SetStatus(OMNI_STATUS_FINISHED);
        return 0;
    }
    int32_t count = static_cast<int32_t>(std::min<int64_t>(rowsPerBatch_, groups_ - emitted_));
    size_t ncols = outTypes_.size();
    std::vector<std::unique_ptr<DeviceBuffer>> vals(ncols), nulls(ncols);
    std::vector<omni_gpu_out_column> outs(ncols);
    for (size_t c = 0; c < ncols; ++c) {
        vals[c].reset(new DeviceBuffer());
        nulls[c].reset(new DeviceBuffer());
        vals[c]->Reserve(static_cast<size_t>(count) * type::TypeWidth(outTypes_[c]));
        nulls[c]->Reserve((count + 7) / 8);
        outs[c].type_id = outTypes_[c];
        outs[c].values = vals[c]->Get();
        outs[c].nulls = static_cast<uint8_t *>(nulls[c]->Get());
    }
    Check(omni_gpu_agg_get_output(handle_, emitted_, count, outs.data()), "omni_gpu_agg_get_output");
    std::vector<std::unique_ptr<vec::BaseVector>> gpu(ncols);
    std::vector<uint8_t> bits((count + 7) / 8);
    for (size_t c = 0; c < ncols; ++c) {
        gpu[c].reset(vec::VectorHelper::CreateFlatVector(outTypes_[c], count));
        auto *v = gpu[c].get();
        Check(omni_gpu_memcpy_d2h(FlatValues(v), vals[c]->Get(), static_cast<size_t>(count) * type::TypeWidth(outTypes_[c])),
              "download");
        Check(omni_gpu_memcpy_d2h(bits.data(), nulls[c]->Get(), bits.size()), "download nulls");
        SetNullsFrom(v, bits.data(), count);
    }
    std::unique_ptr<vec::VectorBatch> batch(new vec::VectorBatch(count));  // freed if an overflow raises
    const size_t nk = static_cast<size_t>(spec_.num_keys);
    for (size_t c = 0; c < nk; ++c) batch->Append(gpu[c].release());
    for (const auto &p : plan_) {
        auto *first = gpu[nk + p.slot].get();
        auto *second = p.rule == GpuHashAggregationOperatorFactory::kDirect ? nullptr : gpu[nk + p.slot + 1].get();
        switch (p.rule) {
            case GpuHashAggregationOperatorFactory::kDirect: batch->Append(gpu[nk + p.slot].release()); break;
            case GpuHashAggregationOperatorFactory::kAvgPartial:
            case GpuHashAggregationOperatorFactory::kAvgMerge: {
                // (sum, count); a group without a non-NULL input is (0, 0), never NULL
                auto *os = new vec::Vector<double>(count, type::OMNI_DOUBLE);
                auto *oc = new vec::Vector<int64_t>(count, type::OMNI_LONG);
                for (int32_t r = 0; r < count; ++r) {
                    const int64_t c = second->IsNull(r) ? 0 : Cell<int64_t>(second, r);
                    os->SetValue(r, c == 0 || first->IsNull(r) ? 0.0 : Cell<double>(first, r));
                    oc->SetValue(r, c);
                }
                batch->Append(os);
                batch->Append(oc);
                break;
            }
            case GpuHashAggregationOperatorFactory::kAvgFinal: {
                auto *o = new vec::Vector<double>(count, type::OMNI_DOUBLE);
                for (int32_t r = 0; r < count; ++r) {
                    const int64_t c = second->IsNull(r) ? 0 : Cell<int64_t>(second, r);
                    if (c == 0) {
                        o->SetValue(r, 0.0);
                        o->SetNull(r);
                    } else {
                        o->SetValue(r, Cell<double>(first, r) / static_cast<double>(c));
                    }
                }
                batch->Append(o);
                break;
            }
            case GpuHashAggregationOperatorFactory::kDecPartial: {
                // (sum, isEmpty): the GPU's partial sum column as it is, isEmpty = no non-NULL input
                auto *e = new vec::Vector<uint8_t>(count, type::OMNI_BOOLEAN);
                for (int32_t r = 0; r < count; ++r) e->SetValue(r, Cell<int64_t>(second, r) == 0 ? 1 : 0);
                batch->Append(gpu[nk + p.slot].release());
                batch->Append(e);
                break;
            }
            case GpuHashAggregationOperatorFactory::kDecFinal:
            case GpuHashAggregationOperatorFactory::kDecMerge: {
                // a NULL partial sum among the group's inputs: the group is overflowed (SetNullOrThrowException in
                // the final output, NULL in a partial one)
                const bool merge = p.rule == GpuHashAggregationOperatorFactory::kDecMerge;
                vec::BaseVector *third = merge ? gpu[nk + p.slot + 2].get() : nullptr;
                vec::Vector<uint8_t> *e = merge ? new vec::Vector<uint8_t>(count, type::OMNI_BOOLEAN) : nullptr;
                std::unique_ptr<vec::Vector<uint8_t>> eOwn(e);
                for (int32_t r = 0; r < count; ++r) {
                    const int64_t marks = Cell<int64_t>(second, r);
                    if (marks > 0) {
                        if (!merge && !spec_.overflow_as_null) {
                            throw OmniException("OPERATOR_RUNTIME_ERROR", "decimal SUM overflow");
                        }
                        first->SetNull(r);
                    }
                    if (merge) e->SetValue(r, marks == 0 && Cell<int64_t>(third, r) == 0 ? 1 : 0);
                }
                batch->Append(gpu[nk + p.slot].release());
                if (merge) batch->Append(eOwn.release());
                break;
            }
        }
    }
    emitted_ += count;
    if (emitted_ >= groups_) SetStatus(OMNI_STATUS_FINISHED);
    *result = batch.release();
    return 1;
}

uint64_t GpuHashAggregationOperator::GetHashMapUniqueKeys()
{
    Flush();
    int64_t g = 0;
    Check(omni_gpu_agg_num_groups(handle_, &g), "omni_gpu_agg_num_groups");
    pending_ = false;
    return static_cast<uint64_t>(g);
}

namespace {
// Row r of a flat column's values (as ColumnValues returns them) widened to int64 or to double.
int64_t IntAt(const void *vals, int32_t typeId, int32_t r)
{
    switch (type::TypeWidth(typeId)) {
        case 1: return static_cast<const uint8_t *>(vals)[r];
        case 2: return static_cast<const int16_t *>(vals)[r];
        case 4: return static_cast<const int32_t *>(vals)[r];
        default: return static_cast<const int64_t *>(vals)[r];
    }
}
double DoubleAt(const void *vals, int32_t typeId, int32_t r)
{
    return typeId == type::OMNI_DOUBLE ? static_cast<const double *>(vals)[r]
                                       : static_cast<double>(IntAt(vals, typeId, r));
}
// Stores an int64 (or, for DOUBLE, a double) into row r of a new flat vector of typeId.
void StoreAt(vec::BaseVector *v, int32_t typeId, int32_t r, int64_t i, double d)
{
    switch (typeId) {
        case type::OMNI_DOUBLE: static_cast<vec::Vector<double> *>(v)->SetValue(r, d); break;
        case type::OMNI_BOOLEAN: static_cast<vec::Vector<uint8_t> *>(v)->SetValue(r, static_cast<uint8_t>(i)); break;
        case type::OMNI_BYTE: static_cast<vec::Vector<int8_t> *>(v)->SetValue(r, static_cast<int8_t>(i)); break;
        case type::OMNI_SHORT: static_cast<vec::Vector<int16_t> *>(v)->SetValue(r, static_cast<int16_t>(i)); break;
        case type::OMNI_INT:
        case type::OMNI_DATE32: static_cast<vec::Vector<int32_t> *>(v)->SetValue(r, static_cast<int32_t>(i)); break;
        case type::OMNI_DECIMAL128:
            static_cast<vec::Vector<type::Decimal128> *>(v)->SetValue(
                r, type::Decimal128(i < 0 ? -1 : 0, static_cast<uint64_t>(i)));
            break;
        default: static_cast<vec::Vector<int64_t> *>(v)->SetValue(r, i); break;
    }
}
}  // namespace

// Each aggregate's ProcessAlignAggSchema (typed_aggregator.h:243-249): COUNT(*) 1 (count_all_aggregator.h:38-47);
// COUNT(col) 0 / 1 (count_column_aggregator.cpp:186-205); SUM / MIN / MAX the value, NULL for a NULL row, as a
// slice when the output type is the input's (sum_flatim_aggregator.h:259-300, max_aggregator.cpp:197-240,
// min_aggregator.cpp); AVG (sum, count) = (value, 1), (0, 0) for a NULL row, the sum a slice when it is a DOUBLE
// input without NULLs (average_flatim_aggregator.h:292-347); decimal SUM (sum, isEmpty) = (value, NULL row), the
// sum a slice keeping the NULL rows when the output type is the input's, else 0 for a NULL row
// (sum_spark_decimal_aggregator.h:496-560).
vec::VectorBatch *GpuHashAggregationOperator::AlignSchema(vec::VectorBatch *inputVecBatch)
{
    const int32_t rows = inputVecBatch->GetRowCount();
    std::unique_ptr<vec::VectorBatch> in(inputVecBatch);
    std::unique_ptr<vec::VectorBatch> out(new vec::VectorBatch(rows));
    for (uint32_t c : groupByCols_) out->Append(vec::VectorHelper::SliceVector(in->Get(static_cast<int32_t>(c)), 0, rows));
    std::vector<uint8_t> tmp;
    for (const auto &p : plan_) {
        if (!p.raw || !p.partial) {
            throw OmniException("UNSUPPORTED_ERROR", "AlignSchema reshapes raw input into a partial output only");
        }
        vec::BaseVector *v = p.col >= 0 ? in->Get(p.col) : nullptr;
        const int32_t out0 = p.outTypes[0];
        if (p.func == OMNI_AGG_COUNT_ALL || p.func == OMNI_AGG_COUNT_COLUMN) {
            auto *c = new vec::Vector<int64_t>(rows, type::OMNI_LONG);
            for (int32_t r = 0; r < rows; ++r) c->SetValue(r, v == nullptr || !v->IsNull(r) ? 1 : 0);
            out->Append(c);
            continue;
        }
        const void *vals = ColumnValues(v, rows, tmp);
        if (p.func == OMNI_AGG_AVG) {
            if (p.inType == type::OMNI_DOUBLE && !v->HasNull()) {
                out->Append(vec::VectorHelper::SliceVector(v, 0, rows));
            } else {
                auto *sm = new vec::Vector<double>(rows, type::OMNI_DOUBLE);
                for (int32_t r = 0; r < rows; ++r) sm->SetValue(r, v->IsNull(r) ? 0.0 : DoubleAt(vals, p.inType, r));
                out->Append(sm);
            }
            auto *cn = new vec::Vector<int64_t>(rows, type::OMNI_LONG);
            for (int32_t r = 0; r < rows; ++r) cn->SetValue(r, v->IsNull(r) ? 0 : 1);
            out->Append(cn);
            continue;
        }
        const bool decimal = p.inType == type::OMNI_DECIMAL64 || p.inType == type::OMNI_DECIMAL128;
        if (decimal && p.func == OMNI_AGG_SUM) {
            if (out0 == p.inType) {
                out->Append(vec::VectorHelper::SliceVector(v, 0, rows));
            } else {  // DECIMAL64 into a DECIMAL128 sum
                auto *sm = vec::VectorHelper::CreateFlatVector(out0, rows);
                for (int32_t r = 0; r < rows; ++r) {
                    StoreAt(sm, out0, r, v->IsNull(r) ? 0 : IntAt(vals, p.inType, r), 0.0);
                }
                out->Append(sm);
            }
            auto *e = new vec::Vector<uint8_t>(rows, type::OMNI_BOOLEAN);
            for (int32_t r = 0; r < rows; ++r) e->SetValue(r, v->IsNull(r) ? 1 : 0);
            out->Append(e);
            continue;
        }
        // SUM, MIN, MAX
        if (out0 == p.inType) {
            out->Append(vec::VectorHelper::SliceVector(v, 0, rows));
            continue;
        }
        auto *o = vec::VectorHelper::CreateFlatVector(out0, rows);
        for (int32_t r = 0; r < rows; ++r) {
            if (v->IsNull(r)) {
                o->SetNull(r);
            } else {
                StoreAt(o, out0, r, IntAt(vals, p.inType, r), DoubleAt(vals, p.inType, r));
            }
        }
        out->Append(o);
    }
    return out.release();
}

// ================================================================================================
// Hash join: builder
// ================================================================================================

GpuHashBuilderOperatorFactory *GpuHashBuilderOperatorFactory::CreateHashBuilderOperatorFactory(
    JoinType joinType, const type::DataTypes &buildTypes, const int32_t *buildHashCols, int32_t buildHashColsCount,
    int32_t operatorCount)
{
    return CreateHashBuilderOperatorFactory(joinType, OMNI_BUILD_UNKNOWN, buildTypes, buildHashCols,
                                            buildHashColsCount, operatorCount);
}

GpuHashBuilderOperatorFactory *GpuHashBuilderOperatorFactory::CreateHashBuilderOperatorFactory(
    JoinType joinType, BuildSide buildSide, const type::DataTypes &buildTypes, const int32_t *buildHashCols,
    int32_t buildHashColsCount, int32_t operatorCount)
{
    if (operatorCount <= 0 || operatorCount > 10000) {  // hash_builder.cpp:33-35
        throw OmniException("OPERATOR_RUNTIME_ERROR", "operatorCount is not in the acceptable range [1, 10000].");
    }
    if (buildHashColsCount < 1 || buildHashColsCount > OMNI_GPU_MAX_KEYS) {
        throw OmniException("UNSUPPORTED_ERROR", "1 to " + std::to_string(OMNI_GPU_MAX_KEYS) +
                                                     " join key columns are supported");
    }
    if (operatorCount & (operatorCount - 1)) {
        // the probe masks its partition hash with operatorCount - 1 (lookup_join.cpp:496-498)
        throw OmniException("UNSUPPORTED_ERROR", "operatorCount must be a power of two on the GPU path");
    }
    for (int32_t i = 0; i < buildHashColsCount; ++i) {
        if (buildHashCols[i] < 0 || buildHashCols[i] >= buildTypes.GetSize()) {
            throw OmniException("INVALID_ARGUMENT", "join key column out of range");
        }
    }
    return new GpuHashBuilderOperatorFactory(joinType, buildSide, buildTypes,
                                             std::vector<int32_t>(buildHashCols, buildHashCols + buildHashColsCount),
                                             operatorCount);
}

// The key variant follows the key columns as in HashBuilderOperatorFactory::InitVariant (hash_builder.cpp:44-93);
// several columns or a DECIMAL128 column go through the library's key translation.
GpuHashBuilderOperatorFactory::GpuHashBuilderOperatorFactory(JoinType joinType, BuildSide buildSide,
                                                             const type::DataTypes &buildTypes,
                                                             std::vector<int32_t> keyCols, int32_t operatorCount)
    : joinType_(joinType), buildSide_(buildSide), buildTypes_(buildTypes), keyCols_(std::move(keyCols))
{
    EnsureInit();
    omni_gpu_join_spec spec{};
    spec.key_type = buildTypes.GetType(keyCols_[0])->GetId();
    spec.num_keys = static_cast<int32_t>(keyCols_.size());
    for (size_t i = 0; i < keyCols_.size(); ++i) spec.key_types[i] = buildTypes.GetType(keyCols_[i])->GetId();
    spec.num_payloads = buildTypes.GetSize();  // every build column is retained, like the reference's batches
    if (spec.num_payloads > OMNI_GPU_MAX_AGGS) throw OmniException("UNSUPPORTED_ERROR", "too many build columns");
    for (int32_t i = 0; i < spec.num_payloads; ++i) {
        spec.payload_types[i] = buildTypes.GetType(i)->GetId();
        if (type::TypeWidth(spec.payload_types[i]) == 0) {
            throw OmniException("UNSUPPORTED_ERROR", "variable-width build columns are not on the GPU path");
        }
    }
    spec.strategy = 0;
    tables_.assign(static_cast<size_t>(operatorCount), nullptr);
    partFinished_.assign(static_cast<size_t>(operatorCount), false);
    retained_.resize(static_cast<size_t>(operatorCount));
    for (auto &t : tables_) {
        Check(omni_gpu_join_build_create(&spec, &t), "omni_gpu_join_build_create");
        Check(omni_gpu_join_set_type(t, joinType, buildSide), "omni_gpu_join_set_type");
    }
}

GpuHashBuilderOperatorFactory::~GpuHashBuilderOperatorFactory()
{
    for (void *t : tables_) {
        if (t) omni_gpu_join_destroy(t);
    }
}

// HashBuilderOperatorFactory::CreateOperator (hash_builder.cpp:155-160): partitions round robin.
Operator *GpuHashBuilderOperatorFactory::CreateOperator()
{
    return new GpuHashBuilderOperator(*this, nextOperator_++ % Partitions());
}

// HashBuilderOperator::AddInput (hash_builder.cpp:169-180): the batch is retained -- here its columns move to
// device memory and the host batch is freed.
int32_t GpuHashBuilderOperator::AddInput(vec::VectorBatch *vecBatch)
{
    int32_t rows = vecBatch->GetRowCount();
    int32_t ncols = f_.buildTypes_.GetSize();
    if (rows > 0) {
        std::vector<omni_gpu_column> cols(ncols);
        auto &retained = f_.retained_[partition_];
        for (int32_t c = 0; c < ncols; ++c) {
            retained.emplace_back(new DeviceBuffer());
            retained.emplace_back(new DeviceBuffer());
            auto &vb = *retained[retained.size() - 2];
            auto &nb = *retained.back();
            cols[c] = Upload(vecBatch->Get(c), rows, vb, nb);
        }
        std::vector<omni_gpu_column> keys;
        for (int32_t k : f_.keyCols_) keys.push_back(cols[k]);
        Check(omni_gpu_join_build_add(f_.tables_[partition_], keys.data(), cols.data(), rows),
              "omni_gpu_join_build_add");
    }
    vec::VectorHelper::FreeVecBatch(vecBatch);
    return 0;
}

// HashBuilderOperator::GetOutput (hash_builder.cpp:182-213): builds the table once input is done.
int32_t GpuHashBuilderOperator::GetOutput(vec::VectorBatch **result)
{
    *result = nullptr;
    if (!f_.partFinished_[partition_]) {
        Check(omni_gpu_join_build_finish(f_.tables_[partition_]), "omni_gpu_join_build_finish");
        f_.partFinished_[partition_] = true;
        ++f_.finishedCount_;
        f_.retained_[partition_].clear();
    }
    SetStatus(OMNI_STATUS_FINISHED);
    return 0;
}

// ================================================================================================
// Hash join: lookup (inner)
// ================================================================================================

GpuLookupJoinOperatorFactory *GpuLookupJoinOperatorFactory::CreateLookupJoinOperatorFactory(
    const type::DataTypes &probeTypes, int32_t *probeOutputCols, int32_t probeOutputColsCount, int32_t *probeHashCols,
    int32_t probeHashColsCount, int32_t *buildOutputCols, int32_t buildOutputColsCount,
    const type::DataTypes &buildOutputTypes, int64_t hashBuilderFactoryAddr, expressions::Expr *filterExpr,
    bool isShuffleExchangeBuildPlan, OverflowConfig *overflowConfig)
{
    if (filterExpr != nullptr) throw OmniException("UNSUPPORTED_ERROR", "join filters are not on the GPU path");
    auto *builder = reinterpret_cast<GpuHashBuilderOperatorFactory *>(hashBuilderFactoryAddr);
    if (probeHashColsCount != static_cast<int32_t>(builder->KeyCols().size())) {
        throw OmniException("INVALID_ARGUMENT", "probe and build sides have different numbers of join keys");
    }
    for (int32_t i = 0; i < probeHashColsCount; ++i) {
        // the probe's keys are translated as the build's: their types must agree column by column
        if (probeHashCols[i] < 0 || probeHashCols[i] >= probeTypes.GetSize() ||
            probeTypes.GetType(probeHashCols[i])->GetId() !=
                builder->BuildTypes().GetType(builder->KeyCols()[i])->GetId()) {
            throw OmniException("UNSUPPORTED_ERROR", "probe join key types must equal the build join key types");
        }
    }
    auto *f = new GpuLookupJoinOperatorFactory();
    f->probeTypes_ = probeTypes;
    f->probeOutputCols_.assign(probeOutputCols, probeOutputCols + probeOutputColsCount);
    f->probeKeyCols_.assign(probeHashCols, probeHashCols + probeHashColsCount);
    f->buildOutputCols_.assign(buildOutputCols, buildOutputCols + buildOutputColsCount);
    // the build side is shared by raw factory address (lookup_join.cpp:84-95)
    f->builder_ = builder;
    for (int32_t i = 0; i < buildOutputColsCount; ++i) {
        if (buildOutputTypes.GetType(i)->GetId() != f->builder_->BuildTypes().GetType(buildOutputCols[i])->GetId()) {
            throw OmniException("INVALID_ARGUMENT", "build output type does not match the build column");
        }
    }
    if (f->builder_->GetJoinType() == OMNI_JOIN_TYPE_LEFT_SEMI && buildOutputColsCount > 0) {
        // the reference's semi join would emit its first matching build row's columns
        throw OmniException("UNSUPPORTED_ERROR", "semi joins with build output columns are not on the GPU path");
    }
    return f;
}

Operator *GpuLookupJoinOperatorFactory::CreateOperator() { return new GpuLookupJoinOperator(*this); }

GpuLookupJoinOperator::~GpuLookupJoinOperator()
{
    if (probe_) vec::VectorHelper::FreeVecBatch(probe_);
    for (auto *b : pending_) vec::VectorHelper::FreeVecBatch(b);
    for (auto *b : emitting_) vec::VectorHelper::FreeVecBatch(b);
}

bool GpuLookupJoinOperator::needsInput()
{
    return GetStatus() != OMNI_STATUS_FINISHED && !noMoreInput_ && !inRun_ && probe_ == nullptr;
}

// Bytes of one output row: ceil(1 MiB / this) rows per output batch (lookup_join.cpp:211-244).
int32_t GpuLookupJoinOperator::OutputRowSize() const
{
    const auto &btypes = f_.builder_->BuildTypes();
    int32_t rowSize = 0;
    for (int32_t c : f_.probeOutputCols_) rowSize += type::TypeWidth(f_.probeTypes_.GetType(c)->GetId());
    for (int32_t c : f_.buildOutputCols_) rowSize += type::TypeWidth(btypes.GetType(c)->GetId());
    if (f_.builder_->GetJoinType() == OMNI_JOIN_TYPE_EXISTENCE) rowSize += 1;
    return std::max(rowSize, 1);
}

int32_t GpuLookupJoinOperator::AddInput(vec::VectorBatch *vecBatch)
{
    if (!f_.builder_->Finished()) throw OmniException("OPERATOR_RUNTIME_ERROR", "probe before the build finished");
    if (noMoreInput_ || f_.builder_->GetJoinType() == OMNI_JOIN_TYPE_EXISTENCE || f_.builder_->Partitions() > 1) {
        return ProbeOne(vecBatch);
    }
    SetStatus(OMNI_STATUS_NORMAL);
    if (vecBatch->GetRowCount() <= 0) {
        vec::VectorHelper::FreeVecBatch(vecBatch);
        return 0;
    }
    Stage(vecBatch);
    return 0;
}

// Appends a probe batch's key and probe output columns to the staged run (values, and NULL bits at the run's
// row positions); the batch is kept until the run's rows are emitted.
void GpuLookupJoinOperator::Stage(vec::VectorBatch *vecBatch)
{
    const int32_t rows = vecBatch->GetRowCount();
    if (stage_.empty()) {
        int32_t is_array = 0;
        int64_t max_dup = 0, slots = 0;
        Check(omni_gpu_join_build_info(f_.builder_->Table(), &is_array, &max_dup, &slots), "omni_gpu_join_build_info");
        flushRows_ = std::min<int64_t>(std::max<int64_t>(slots / 2, 1 << 20), 1 << 24);
        for (int32_t c : f_.probeKeyCols_) {
            stage_.emplace_back(new StagedProbeColumn());
            stage_.back()->col = c;
        }
        for (int32_t c : f_.probeOutputCols_) {
            int32_t at = -1;
            for (size_t i = 0; i < stage_.size(); ++i) {
                if (stage_[i]->col == c) at = static_cast<int32_t>(i);
            }
            if (at < 0) {
                at = static_cast<int32_t>(stage_.size());
                stage_.emplace_back(new StagedProbeColumn());
                stage_.back()->col = c;
            }
            outStage_.push_back(at);
        }
        for (auto &st : stage_) {
            st->type = vecBatch->Get(st->col)->GetTypeId();
            st->width = type::TypeWidth(st->type);
        }
    }
    if (stagedRows_ + rows > stageCap_) {
        // grow (a run ends at the first batch past flushRows_): new buffers, the staged rows copied over
        const int64_t cap = std::max<int64_t>(flushRows_ + (1 << 17), stagedRows_ + rows);
        for (auto &st : stage_) {
            HostBuffer v, n;
            v.Reserve(static_cast<size_t>(cap) * st->width);
            n.Reserve(static_cast<size_t>(cap + 7) / 8 + 8);
            std::memset(n.Get(), 0, static_cast<size_t>(cap + 7) / 8 + 8);
            if (stagedRows_ > 0) {
                std::memcpy(v.Get(), st->values.Get(), static_cast<size_t>(stagedRows_) * st->width);
                std::memcpy(n.Get(), st->nulls.Get(), static_cast<size_t>(stagedRows_ + 7) / 8);
            }
            st->values.Swap(v);
            st->nulls.Swap(n);
        }
        stageCap_ = cap;
    }
    std::vector<uint8_t> tmp;
    for (auto &st : stage_) {
        vec::BaseVector *v = vecBatch->Get(st->col);
        const void *src = ColumnValues(v, rows, tmp);
        std::memcpy(static_cast<uint8_t *>(st->values.Get()) + static_cast<size_t>(stagedRows_) * st->width, src,
                    static_cast<size_t>(rows) * st->width);
        if (v->HasNull()) {
            const uint8_t *bits = vec::unsafe::UnsafeBaseVector::GetNulls(v);
            uint8_t *dst = static_cast<uint8_t *>(st->nulls.Get());
            for (int32_t r = 0; r < rows; ++r) {
                if ((bits[r >> 3] >> (r & 7)) & 1) {
                    const int64_t q = stagedRows_ + r;
                    dst[q >> 3] |= static_cast<uint8_t>(1u << (q & 7));
                    st->anyNull = true;
                }
            }
        }
    }
    stagedRows_ += rows;
    pending_.push_back(vecBatch);
}

// Probes the staged run in one call, gathers its probe output columns on the device and downloads the output
// columns into page-locked host memory; the run's batches move to emitting_.
void GpuLookupJoinOperator::FlushRun()
{
    const int64_t rows = stagedRows_;
    const size_t ns = stage_.size(), nk = f_.probeKeyCols_.size(), npo = f_.probeOutputCols_.size();
    while (keyBufs_.size() < ns) {
        keyBufs_.emplace_back(new DeviceBuffer());
        keyNullBufs_.emplace_back(new DeviceBuffer());
    }
    std::vector<omni_gpu_column> cols(ns);
    for (size_t i = 0; i < ns; ++i) {
        StagedProbeColumn &st = *stage_[i];
        keyBufs_[i]->Reserve(static_cast<size_t>(rows) * st.width);
        Check(omni_gpu_memcpy_h2d(keyBufs_[i]->Get(), st.values.Get(), static_cast<size_t>(rows) * st.width),
              "upload probe column");
        cols[i].type_id = st.type;
        cols[i].values = keyBufs_[i]->Get();
        if (st.anyNull) {
            const size_t nb = static_cast<size_t>(rows + 7) / 8;
            keyNullBufs_[i]->Reserve(nb);
            Check(omni_gpu_memcpy_h2d(keyNullBufs_[i]->Get(), st.nulls.Get(), nb), "upload probe nulls");
            cols[i].nulls = static_cast<const uint8_t *>(keyNullBufs_[i]->Get());
            std::memset(st.nulls.Get(), 0, nb);
            st.anyNull = false;
        }
    }
    int32_t is_array = 0;
    int64_t max_dup = 0, slots = 0;
    Check(omni_gpu_join_build_info(f_.builder_->Table(), &is_array, &max_dup, &slots), "omni_gpu_join_build_info");
    // outer joins emit a row per unmatched probe row as well
    const int64_t cap = std::max<int64_t>(rows, rows * std::max<int64_t>(1, max_dup));
    rowsBuf_.Reserve(sizeof(int64_t) * cap);
    const auto &btypes = f_.builder_->BuildTypes();
    const int32_t nb = btypes.GetSize();
    std::vector<omni_gpu_out_column> outs(nb);
    while (payBufs_.size() < static_cast<size_t>(nb)) {
        payBufs_.emplace_back(new DeviceBuffer());
        payNullBufs_.emplace_back(new DeviceBuffer());
    }
    for (int32_t c = 0; c < nb; ++c) outs[c].type_id = btypes.GetType(c)->GetId();
    for (int32_t col : f_.buildOutputCols_) {
        payBufs_[col]->Reserve(static_cast<size_t>(cap) * type::TypeWidth(btypes.GetType(col)->GetId()));
        payNullBufs_[col]->Reserve(4 * ((cap + 31) / 32));
        outs[col].values = payBufs_[col]->Get();
        outs[col].nulls = static_cast<uint8_t *>(payNullBufs_[col]->Get());
    }
    Check(omni_gpu_join_probe(f_.builder_->Table(), cols.data(), rows, 0, static_cast<int64_t *>(rowsBuf_.Get()),
                              outs.data(), cap, &matches_),
          "omni_gpu_join_probe");
    emitted_ = 0;
    // the probe output columns of the output rows, gathered on the device by probe row
    while (gatherBufs_.size() < npo) {
        gatherBufs_.emplace_back(new DeviceBuffer());
        gatherNullBufs_.emplace_back(new DeviceBuffer());
    }
    std::vector<omni_gpu_column> src(npo);
    std::vector<omni_gpu_out_column> gout(npo);
    for (size_t i = 0; i < npo; ++i) {
        src[i] = cols[outStage_[i]];
        gatherBufs_[i]->Reserve(static_cast<size_t>(std::max<int64_t>(matches_, 1)) * stage_[outStage_[i]]->width);
        gatherNullBufs_[i]->Reserve(4 * ((std::max<int64_t>(matches_, 1) + 31) / 32));
        gout[i].type_id = src[i].type_id;
        gout[i].values = gatherBufs_[i]->Get();
        gout[i].nulls = static_cast<uint8_t *>(gatherNullBufs_[i]->Get());
    }
    if (matches_ > 0) {
        Check(omni_gpu_join_gather_rows(src.data(), static_cast<int32_t>(npo), static_cast<const int64_t *>(rowsBuf_.Get()),
                                        matches_, gout.data()),
              "omni_gpu_join_gather_rows");
    }
    // download: probe output columns, then build output columns
    const size_t nout = npo + f_.buildOutputCols_.size();
    while (runVals_.size() < nout) {
        runVals_.emplace_back(new HostBuffer());
        runNulls_.emplace_back(new HostBuffer());
    }
    runAnyNull_.assign(nout, false);
    const size_t nbytes = static_cast<size_t>(matches_ + 7) / 8;
    for (size_t i = 0; i < nout; ++i) {
        const bool probeSide = i < npo;
        const int32_t col = probeSide ? 0 : f_.buildOutputCols_[i - npo];
        const int w = probeSide ? stage_[outStage_[i]]->width : type::TypeWidth(btypes.GetType(col)->GetId());
        const void *dv = probeSide ? gatherBufs_[i]->Get() : payBufs_[col]->Get();
        const void *dn = probeSide ? gatherNullBufs_[i]->Get() : payNullBufs_[col]->Get();
        runVals_[i]->Reserve(static_cast<size_t>(std::max<int64_t>(matches_, 1)) * w);
        runNulls_[i]->Reserve(nbytes + 8);
        if (matches_ == 0) continue;
        Check(omni_gpu_memcpy_d2h(runVals_[i]->Get(), dv, static_cast<size_t>(matches_) * w), "download output");
        Check(omni_gpu_memcpy_d2h(runNulls_[i]->Get(), dn, nbytes), "download output nulls");
        const uint8_t *bits = static_cast<const uint8_t *>(runNulls_[i]->Get());
        bool any = false;
        for (size_t b = 0; b < nbytes && !any; ++b) {
            uint8_t m = bits[b];
            if (b == nbytes - 1 && (matches_ & 7)) m &= static_cast<uint8_t>((1u << (matches_ & 7)) - 1);
            any = m != 0;
        }
        runAnyNull_[i] = any;
    }
    emitting_.swap(pending_);
    stagedRows_ = 0;
    inRun_ = true;
}

// The next output batch of the flushed run: copies out of the downloaded columns.
vec::VectorBatch *GpuLookupJoinOperator::EmitRun()
{
    const auto &btypes = f_.builder_->BuildTypes();
    const int32_t count = static_cast<int32_t>(std::min<int64_t>(MaxRowCount(OutputRowSize()), matches_ - emitted_));
    std::unique_ptr<vec::VectorBatch> out(new vec::VectorBatch(count));
    const size_t npo = f_.probeOutputCols_.size();
    auto column = [&](size_t i, int32_t tid, bool allNull) {
        auto *v = vec::VectorHelper::CreateFlatVector(tid, count);
        const int w = type::TypeWidth(tid);
        if (allNull) {
            for (int32_t r = 0; r < count; ++r) v->SetNull(r);
            return v;
        }
        std::memcpy(FlatValues(v), static_cast<const uint8_t *>(runVals_[i]->Get()) + static_cast<size_t>(w) * emitted_,
                    static_cast<size_t>(w) * count);
        if (runAnyNull_[i]) {
            const uint8_t *bits = static_cast<const uint8_t *>(runNulls_[i]->Get());
            for (int32_t r = 0; r < count; ++r) {
                const int64_t m = emitted_ + r;
                if ((bits[m >> 3] >> (m & 7)) & 1) v->SetNull(r);
            }
        }
        return v;
    };
    for (size_t i = 0; i < npo; ++i) out->Append(column(i, stage_[outStage_[i]]->type, false));
    const JoinType jt = f_.builder_->GetJoinType();
    for (size_t j = 0; j < f_.buildOutputCols_.size(); ++j) {
        // LEFT_ANTI: AppendRow(probePosition, nullptr, 0), build columns NULL
        out->Append(column(npo + j, btypes.GetType(f_.buildOutputCols_[j])->GetId(), jt == OMNI_JOIN_TYPE_LEFT_ANTI));
    }
    if (f_.builder_->GetBuildSide() == OMNI_BUILD_LEFT) {  // lookup_join.cpp:1959-1962: build columns first
        out->RotateLeft(static_cast<int32_t>(npo));
    }
    emitted_ += count;
    return out.release();
}

int32_t GpuLookupJoinOperator::GetOutput(vec::VectorBatch **result)
{
    *result = nullptr;
    if (probe_) return EmitOne(result);
    if (inRun_ && emitted_ >= matches_) {
        for (auto *b : emitting_) vec::VectorHelper::FreeVecBatch(b);
        emitting_.clear();
        inRun_ = false;
    }
    if (!inRun_ && stagedRows_ > 0 && (noMoreInput_ || stagedRows_ >= flushRows_)) FlushRun();
    if (inRun_ && emitted_ < matches_) {
        *result = EmitRun();
        return 0;
    }
    if (inRun_) {  // a run without output rows
        for (auto *b : emitting_) vec::VectorHelper::FreeVecBatch(b);
        emitting_.clear();
        inRun_ = false;
    }
    if (noMoreInput_ && stagedRows_ == 0) SetStatus(OMNI_STATUS_FINISHED);
    return 0;
}

namespace {
// Appends bits [0, n) of src after the first `at` bits of dst.
void AppendBits(std::vector<uint8_t> &dst, int64_t at, const uint8_t *src, int64_t n)
{
    dst.resize(static_cast<size_t>((at + n + 7) / 8 + 1), 0);
    for (int64_t i = 0; i < n; ++i) {
        if ((src[i >> 3] >> (i & 7)) & 1) dst[(at + i) >> 3] |= static_cast<uint8_t>(1u << ((at + i) & 7));
    }
}
}  // namespace

// Probes `rows` rows of device key columns against one table and appends the output rows to matchRows_ (probe
// row ids, through rowMap when it is given) and matchPay_ / matchPayNulls_; EXISTENCE sets exists_ instead.
void GpuLookupJoinOperator::ProbeTable(void *table, const std::vector<omni_gpu_column> &keys, int64_t rows,
                                       const int64_t *rowMap)
{
    int32_t is_array = 0;
    int64_t max_dup = 0, slots = 0;
    Check(omni_gpu_join_build_info(table, &is_array, &max_dup, &slots), "omni_gpu_join_build_info");
    if (f_.builder_->GetJoinType() == OMNI_JOIN_TYPE_EXISTENCE) {  // one output row per probe row
        rowsBuf_.Reserve(static_cast<size_t>(rows));
        Check(omni_gpu_join_probe_exists(table, keys.data(), rows, static_cast<uint8_t *>(rowsBuf_.Get())),
              "omni_gpu_join_probe_exists");
        std::vector<uint8_t> flags(static_cast<size_t>(rows));
        Check(omni_gpu_memcpy_d2h(flags.data(), rowsBuf_.Get(), static_cast<size_t>(rows)), "download flags");
        for (int64_t r = 0; r < rows; ++r) exists_[rowMap ? rowMap[r] : r] = flags[r];
        return;
    }
    // outer joins emit a row per unmatched probe row as well
    const int64_t cap = std::max<int64_t>(std::max<int64_t>(1, rows), rows * std::max<int64_t>(1, max_dup));
    rowsBuf_.Reserve(sizeof(int64_t) * cap);
    const auto &btypes = f_.builder_->BuildTypes();
    int32_t nb = btypes.GetSize();
    std::vector<omni_gpu_out_column> outs(nb);
    while (payBufs_.size() < static_cast<size_t>(nb)) {
        payBufs_.emplace_back(new DeviceBuffer());
        payNullBufs_.emplace_back(new DeviceBuffer());
    }
    for (int32_t c = 0; c < nb; ++c) outs[c].type_id = btypes.GetType(c)->GetId();
    for (int32_t col : f_.buildOutputCols_) {
        payBufs_[col]->Reserve(static_cast<size_t>(cap) * type::TypeWidth(btypes.GetType(col)->GetId()));
        payNullBufs_[col]->Reserve(4 * ((cap + 31) / 32));
        outs[col].values = payBufs_[col]->Get();
        outs[col].nulls = static_cast<uint8_t *>(payNullBufs_[col]->Get());
    }
    int64_t m = 0;
    Check(omni_gpu_join_probe(table, keys.data(), rows, 0, static_cast<int64_t *>(rowsBuf_.Get()), outs.data(), cap, &m),
          "omni_gpu_join_probe");
    const int64_t at = matches_;
    matchRows_.resize(static_cast<size_t>(at + m));
    Check(omni_gpu_memcpy_d2h(matchRows_.data() + at, rowsBuf_.Get(), sizeof(int64_t) * m), "download rows");
    if (rowMap) {
        for (int64_t k = at; k < at + m; ++k) matchRows_[k] = rowMap[matchRows_[k]];
    }
    std::vector<uint8_t> bits;
    for (size_t i = 0; i < f_.buildOutputCols_.size(); ++i) {
        int32_t col = f_.buildOutputCols_[i];
        int w = type::TypeWidth(btypes.GetType(col)->GetId());
        matchPay_[i].resize(static_cast<size_t>(at + m) * w);
        Check(omni_gpu_memcpy_d2h(matchPay_[i].data() + static_cast<size_t>(at) * w, payBufs_[col]->Get(),
                                  static_cast<size_t>(m) * w),
              "download payload");
        bits.assign(static_cast<size_t>((m + 7) / 8 + 1), 0);
        if (m > 0) {
            Check(omni_gpu_memcpy_d2h(bits.data(), payNullBufs_[col]->Get(), (m + 7) / 8), "download payload nulls");
        }
        AppendBits(matchPayNulls_[i], at, bits.data(), m);
    }
    matches_ = at + m;
}

// One batch probed on its own (the reference's flow, lookup_join.cpp:420-475).  With several builder partitions
// each row probes the table of its hash partition (ProbeBatchForInnerJoin with partitionMask, :950-1000): the
// rows are grouped by partition, each group's keys gathered on the device and probed against its table.
int32_t GpuLookupJoinOperator::ProbeOne(vec::VectorBatch *vecBatch)
{
    if (probe_) vec::VectorHelper::FreeVecBatch(probe_);
    probe_ = vecBatch;
    SetStatus(OMNI_STATUS_NORMAL);
    int32_t rows = vecBatch->GetRowCount();
    matches_ = emitted_ = 0;
    matchRows_.clear();
    matchPay_.assign(f_.buildOutputCols_.size(), {});
    matchPayNulls_.assign(f_.buildOutputCols_.size(), {});
    if (rows <= 0) return 0;
    const size_t nk = f_.probeKeyCols_.size();
    std::vector<omni_gpu_column> keys(nk);
    while (keyBufs_.size() < 2 * nk) {
        keyBufs_.emplace_back(new DeviceBuffer());
        keyNullBufs_.emplace_back(new DeviceBuffer());
    }
    for (size_t k = 0; k < nk; ++k) {
        keys[k] = Upload(vecBatch->Get(f_.probeKeyCols_[k]), rows, *keyBufs_[k], *keyNullBufs_[k]);
    }
    const bool existence = f_.builder_->GetJoinType() == OMNI_JOIN_TYPE_EXISTENCE;
    if (existence) exists_.assign(static_cast<size_t>(rows), 0);
    const int32_t P = f_.builder_->Partitions();
    if (P == 1) {
        ProbeTable(f_.builder_->Table(), keys, rows, nullptr);
    } else {
        DeviceBuffer parts;
        parts.Reserve(sizeof(uint32_t) * rows);
        Check(omni_gpu_join_partition_ids(keys.data(), static_cast<int32_t>(nk), rows, static_cast<uint32_t>(P - 1),
                                          static_cast<uint32_t *>(parts.Get())),
              "omni_gpu_join_partition_ids");
        std::vector<uint32_t> part(static_cast<size_t>(rows));
        Check(omni_gpu_memcpy_d2h(part.data(), parts.Get(), sizeof(uint32_t) * rows), "download partitions");
        std::vector<int64_t> start(static_cast<size_t>(P) + 1, 0), idx(static_cast<size_t>(rows));
        for (uint32_t q : part) ++start[q + 1];
        for (int32_t q = 0; q < P; ++q) start[q + 1] += start[q];
        std::vector<int64_t> fill(start.begin(), start.end() - 1);
        for (int32_t r = 0; r < rows; ++r) idx[fill[part[r]]++] = r;
        DeviceBuffer didx;
        didx.Reserve(sizeof(int64_t) * rows);
        Check(omni_gpu_memcpy_h2d(didx.Get(), idx.data(), sizeof(int64_t) * rows), "upload row groups");
        std::vector<omni_gpu_column> sub(nk);
        std::vector<omni_gpu_out_column> gout(nk);
        for (int32_t q = 0; q < P; ++q) {
            const int64_t n = start[q + 1] - start[q];
            if (n == 0) continue;
            for (size_t k = 0; k < nk; ++k) {
                const int w = type::TypeWidth(keys[k].type_id);
                keyBufs_[nk + k]->Reserve(static_cast<size_t>(n) * w);
                keyNullBufs_[nk + k]->Reserve(4 * ((n + 31) / 32));
                gout[k].type_id = keys[k].type_id;
                gout[k].values = keyBufs_[nk + k]->Get();
                gout[k].nulls = static_cast<uint8_t *>(keyNullBufs_[nk + k]->Get());
                sub[k] = omni_gpu_column{keys[k].type_id, 0, keyBufs_[nk + k]->Get(),
                                         static_cast<const uint8_t *>(keyNullBufs_[nk + k]->Get())};
            }
            Check(omni_gpu_join_gather_rows(keys.data(), static_cast<int32_t>(nk),
                                            static_cast<const int64_t *>(didx.Get()) + start[q], n, gout.data()),
                  "omni_gpu_join_gather_rows");
            ProbeTable(f_.builder_->Table(q), sub, n, idx.data() + start[q]);
        }
    }
    if (existence) {
        matchRows_.resize(rows);
        for (int32_t r = 0; r < rows; ++r) matchRows_[r] = r;
        matches_ = rows;
    }
    return 0;
}

// Output batch = probe output columns (gathered by probe row) ++ build output columns
// (LookupJoinOutputBuilder::BuildOutput, lookup_join.cpp:1924-1985), at most ceil(1 MiB / row size) rows.
int32_t GpuLookupJoinOperator::EmitOne(vec::VectorBatch **result)
{
    *result = nullptr;
    if (!probe_) return 0;
    if (emitted_ >= matches_) {
        vec::VectorHelper::FreeVecBatch(probe_);
        probe_ = nullptr;
        if (noMoreInput_) SetStatus(OMNI_STATUS_FINISHED);
        return 0;
    }
    const auto &btypes = f_.builder_->BuildTypes();
    int32_t count = static_cast<int32_t>(std::min<int64_t>(MaxRowCount(OutputRowSize()), matches_ - emitted_));
    auto *out = new vec::VectorBatch(count);
    std::vector<uint8_t> tmp;
    for (int32_t c : f_.probeOutputCols_) {
        vec::BaseVector *src = probe_->Get(c);
        const auto *sv = static_cast<const uint8_t *>(ColumnValues(src, probe_->GetRowCount(), tmp));
        auto *v = vec::VectorHelper::CreateFlatVector(src->GetTypeId(), count);
        auto *dv = static_cast<uint8_t *>(FlatValues(v));
        const int w = type::TypeWidth(src->GetTypeId());
        for (int32_t r = 0; r < count; ++r) {
            int64_t pr = matchRows_[emitted_ + r];
            std::memcpy(dv + static_cast<size_t>(w) * r, sv + static_cast<size_t>(w) * pr, w);
            if (src->IsNull(static_cast<int32_t>(pr))) v->SetNull(r);
        }
        out->Append(v);
    }
    const JoinType jt = f_.builder_->GetJoinType();
    if (jt == OMNI_JOIN_TYPE_EXISTENCE) {  // ConstructExistenceColumn (lookup_join.cpp:1997-2004)
        auto *v = vec::VectorHelper::CreateFlatVector(type::OMNI_BOOLEAN, count);
        std::memcpy(FlatValues(v), exists_.data() + emitted_, static_cast<size_t>(count));
        out->Append(v);
    } else {
        for (size_t i = 0; i < f_.buildOutputCols_.size(); ++i) {
            int32_t tid = btypes.GetType(f_.buildOutputCols_[i])->GetId();
            auto *v = vec::VectorHelper::CreateFlatVector(tid, count);
            const int w = type::TypeWidth(tid);
            if (jt == OMNI_JOIN_TYPE_LEFT_ANTI) {  // AppendRow(probePosition, nullptr, 0): build columns NULL
                std::memset(FlatValues(v), 0, static_cast<size_t>(w) * count);
                for (int32_t r = 0; r < count; ++r) v->SetNull(r);
                out->Append(v);
                continue;
            }
            std::memcpy(FlatValues(v), matchPay_[i].data() + static_cast<size_t>(w) * emitted_,
                        static_cast<size_t>(w) * count);
            for (int32_t r = 0; r < count; ++r) {
                int64_t m = emitted_ + r;
                if ((matchPayNulls_[i][m >> 3] >> (m & 7)) & 1) v->SetNull(r);
            }
            out->Append(v);
        }
        if (f_.builder_->GetBuildSide() == OMNI_BUILD_LEFT) {  // lookup_join.cpp:1959-1962: build columns first
            out->RotateLeft(static_cast<int32_t>(f_.probeOutputCols_.size()));
        }
    }
    emitted_ += count;
    *result = out;
    return 0;
}

// ================================================================================================
// Hash join: outer pass over unmatched build rows
// ================================================================================================

GpuLookupOuterJoinOperatorFactory *GpuLookupOuterJoinOperatorFactory::CreateLookupOuterJoinOperatorFactory(
    const type::DataTypes &probeTypes, int32_t *probeOutputCols, int32_t probeOutputColsCount,
    int32_t *buildOutputCols, const type::DataTypes &buildOutputTypes, int64_t hashBuilderFactoryAddr)
{
    auto *f = new GpuLookupOuterJoinOperatorFactory();
    f->probeTypes_ = probeTypes;
    f->probeOutputCols_.assign(probeOutputCols, probeOutputCols + probeOutputColsCount);
    f->builder_ = reinterpret_cast<GpuHashBuilderOperatorFactory *>(hashBuilderFactoryAddr);
    f->buildOutputCols_.assign(buildOutputCols, buildOutputCols + buildOutputTypes.GetSize());
    for (int32_t i = 0; i < buildOutputTypes.GetSize(); ++i) {
        if (buildOutputTypes.GetType(i)->GetId() != f->builder_->BuildTypes().GetType(buildOutputCols[i])->GetId()) {
            throw OmniException("INVALID_ARGUMENT", "build output type does not match the build column");
        }
    }
    return f;
}

Operator *GpuLookupOuterJoinOperatorFactory::CreateOperator() { return new GpuLookupOuterJoinOperator(*this); }

// lookup_outer_join.cpp:151: the outer operator takes no input.
int32_t GpuLookupOuterJoinOperator::AddInput(vec::VectorBatch *vecBatch)
{
    vec::VectorHelper::FreeVecBatch(vecBatch);
    return 0;
}

// LookupOuterJoinOperator::GetOutput / BuildVecBatch (lookup_outer_join.cpp:155-215): batches of the unmatched
// build rows, probe output columns NULL, build output columns first when the build side is LEFT.
int32_t GpuLookupOuterJoinOperator::GetOutput(vec::VectorBatch **result)
{
    *result = nullptr;
    if (!f_.builder_->Finished()) throw OmniException("OPERATOR_RUNTIME_ERROR", "outer output before the build finished");
    const auto &btypes = f_.builder_->BuildTypes();
    if (!fetched_) {
        // every builder partition's unmatched build rows, partition after partition
        fetched_ = true;
        rows_ = 0;
        pay_.assign(f_.buildOutputCols_.size(), {});
        payNulls_.assign(f_.buildOutputCols_.size(), {});
        int32_t nb = btypes.GetSize();
        std::vector<uint8_t> bits;
        for (int32_t q = 0; q < f_.builder_->Partitions(); ++q) {
            void *table = f_.builder_->Table(q);
            std::vector<omni_gpu_out_column> outs(nb);
            for (int32_t c = 0; c < nb; ++c) outs[c].type_id = btypes.GetType(c)->GetId();
            // first call with no capacity returns the row count
            int64_t rows = 0;
            int rc = omni_gpu_join_outer_output(table, nullptr, outs.data(), 0, &rows);
            if (rc != OMNI_GPU_OK && rc != OMNI_GPU_ERR_OUTPUT_FULL) Check(rc, "omni_gpu_join_outer_output");
            std::vector<std::unique_ptr<DeviceBuffer>> vals, nulls;
            for (int32_t col : f_.buildOutputCols_) {
                vals.emplace_back(new DeviceBuffer());
                nulls.emplace_back(new DeviceBuffer());
                vals.back()->Reserve(std::max<size_t>(1, static_cast<size_t>(rows) * type::TypeWidth(outs[col].type_id)));
                nulls.back()->Reserve(4 * ((rows + 31) / 32) + 4);
                outs[col].values = vals.back()->Get();
                outs[col].nulls = static_cast<uint8_t *>(nulls.back()->Get());
            }
            if (rows > 0) {
                Check(omni_gpu_join_outer_output(table, nullptr, outs.data(), rows, &rows), "omni_gpu_join_outer_output");
            }
            for (size_t i = 0; i < f_.buildOutputCols_.size(); ++i) {
                int w = type::TypeWidth(outs[f_.buildOutputCols_[i]].type_id);
                pay_[i].resize(static_cast<size_t>(rows_ + rows) * w);
                bits.assign(static_cast<size_t>((rows + 7) / 8 + 1), 0);
                if (rows > 0) {
                    Check(omni_gpu_memcpy_d2h(pay_[i].data() + static_cast<size_t>(rows_) * w, vals[i]->Get(),
                                              static_cast<size_t>(rows) * w),
                          "download payload");
                    Check(omni_gpu_memcpy_d2h(bits.data(), nulls[i]->Get(), (rows + 7) / 8), "download nulls");
                }
                AppendBits(payNulls_[i], rows_, bits.data(), rows);
            }
            rows_ += rows;
        }
    }
    if (emitted_ >= rows_) {
        SetStatus(OMNI_STATUS_FINISHED);
        return 0;
    }
    int32_t rowSize = 0;
    for (int32_t c : f_.probeOutputCols_) rowSize += type::TypeWidth(f_.probeTypes_.GetType(c)->GetId());
    for (int32_t c : f_.buildOutputCols_) rowSize += type::TypeWidth(btypes.GetType(c)->GetId());
    int32_t count = static_cast<int32_t>(std::min<int64_t>(MaxRowCount(std::max(rowSize, 1)), rows_ - emitted_));
    auto *out = new vec::VectorBatch(count);
    for (int32_t c : f_.probeOutputCols_) {
        const int32_t tid = f_.probeTypes_.GetType(c)->GetId();
        auto *v = vec::VectorHelper::CreateFlatVector(tid, count);
        std::memset(FlatValues(v), 0, static_cast<size_t>(type::TypeWidth(tid)) * count);
        for (int32_t r = 0; r < count; ++r) v->SetNull(r);
        out->Append(v);
    }
    for (size_t i = 0; i < f_.buildOutputCols_.size(); ++i) {
        const int32_t tid = btypes.GetType(f_.buildOutputCols_[i])->GetId();
        auto *v = vec::VectorHelper::CreateFlatVector(tid, count);
        const int w = type::TypeWidth(tid);
        std::memcpy(FlatValues(v), pay_[i].data() + static_cast<size_t>(w) * emitted_, static_cast<size_t>(w) * count);
        for (int32_t r = 0; r < count; ++r) {
            int64_t m = emitted_ + r;
            if ((payNulls_[i][m >> 3] >> (m & 7)) & 1) v->SetNull(r);
        }
        out->Append(v);
    }
    if (f_.builder_->GetBuildSide() == OMNI_BUILD_LEFT) out->RotateLeft(static_cast<int32_t>(f_.probeOutputCols_.size()));
    emitted_ += count;
    if (emitted_ >= rows_) SetStatus(OMNI_STATUS_FINISHED);
    *result = out;
    return 0;
}

}  // namespace op
}  // namespace omniruntime
