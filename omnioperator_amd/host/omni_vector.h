// Minimal OmniVec-compatible columnar containers and the Operator / OperatorFactory surface of the
// reference (core/src/vector/vector_batch.h:10-66, core/src/vector/vector.h:202-426,
// core/src/vector/nulls_buffer.h:16-117, core/src/type/data_type.h:40-65, core/src/type/decimal128.h:21-148,
// core/src/operator/operator.h:22-264, core/src/operator/operator_factory.h:12-27, core/src/operator/status.h:9-13).
//
// Only the subset the hash-aggregation / hash-join path needs: flat vectors of fixed-width types with an
// LSB-first null bitmap (bit = 1 means NULL), batches of such vectors, and the push/pull operator API.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace omniruntime {
namespace type {

enum DataTypeId {
    OMNI_NONE = 0,
    OMNI_INT = 1,
    OMNI_LONG = 2,
    OMNI_DOUBLE = 3,
    OMNI_BOOLEAN = 4,
    OMNI_SHORT = 5,
    OMNI_DECIMAL64 = 6,
    OMNI_DECIMAL128 = 7,
    OMNI_DATE32 = 8,
    OMNI_DATE64 = 9,
    OMNI_TIMESTAMP = 12,
    OMNI_VARCHAR = 15,
};

// Decimal128 memory layout of the reference: little-endian {uint64 low; int64 high}.
struct Decimal128 {
    uint64_t low = 0;
    int64_t high = 0;
    Decimal128() = default;
    Decimal128(int64_t hi, uint64_t lo) : low(lo), high(hi) {}
    bool operator==(const Decimal128 &o) const { return low == o.low && high == o.high; }
};

class DataType {
public:
    explicit DataType(DataTypeId id, int32_t precision = 0, int32_t scale = 0)
        : id_(id), precision_(precision), scale_(scale) {}
    DataTypeId GetId() const { return id_; }
    int32_t GetPrecision() const { return precision_; }
    int32_t GetScale() const { return scale_; }

private:
    DataTypeId id_;
    int32_t precision_;
    int32_t scale_;
};
using DataTypePtr = std::shared_ptr<DataType>;

inline DataTypePtr IntType() { return std::make_shared<DataType>(OMNI_INT); }
inline DataTypePtr LongType() { return std::make_shared<DataType>(OMNI_LONG); }
inline DataTypePtr ShortType() { return std::make_shared<DataType>(OMNI_SHORT); }
inline DataTypePtr DoubleType() { return std::make_shared<DataType>(OMNI_DOUBLE); }
inline DataTypePtr Date32Type() { return std::make_shared<DataType>(OMNI_DATE32); }
inline DataTypePtr TimestampType() { return std::make_shared<DataType>(OMNI_TIMESTAMP); }
inline DataTypePtr Decimal64Type(int32_t p = 18, int32_t s = 0) { return std::make_shared<DataType>(OMNI_DECIMAL64, p, s); }
inline DataTypePtr Decimal128Type(int32_t p = 38, int32_t s = 0) { return std::make_shared<DataType>(OMNI_DECIMAL128, p, s); }

class DataTypes {
public:
    DataTypes() = default;
    explicit DataTypes(std::vector<DataTypePtr> types) : types_(std::move(types)) {}
    int32_t GetSize() const { return static_cast<int32_t>(types_.size()); }
    const DataTypePtr &GetType(int32_t i) const { return types_.at(i); }
    const std::vector<DataTypePtr> &Get() const { return types_; }
    std::vector<int32_t> GetIds() const
    {
        std::vector<int32_t> ids;
        for (auto &t : types_) ids.push_back(t->GetId());
        return ids;
    }

private:
    std::vector<DataTypePtr> types_;
};

// Value width in bytes of a fixed-width type (0 for variable-width types).
inline int32_t TypeWidth(int32_t id)
{
    switch (id) {
        case OMNI_SHORT: return 2;
        case OMNI_INT:
        case OMNI_DATE32: return 4;
        case OMNI_LONG:
        case OMNI_TIMESTAMP:
        case OMNI_DECIMAL64:
        case OMNI_DOUBLE:
        case OMNI_DATE64: return 8;
        case OMNI_DECIMAL128: return 16;
        case OMNI_BOOLEAN: return 1;
        default: return 0;
    }
}
}  // namespace type

namespace vec {
using type::DataTypeId;

enum Encoding { OMNI_FLAT = 0, OMNI_DICTIONARY = 1 };

// A flat vector: contiguous fixed-width values plus an LSB-first null bitmap.
class BaseVector {
public:
    BaseVector(DataTypeId id, int32_t size)
        : id_(id), size_(size), width_(type::TypeWidth(id)), values_(static_cast<size_t>(size) * type::TypeWidth(id)),
          nulls_((size + 7) / 8 + 8, 0) {}
    virtual ~BaseVector() = default;
    DataTypeId GetTypeId() const { return id_; }
    int32_t GetSize() const { return size_; }
    Encoding GetEncoding() const { return OMNI_FLAT; }
    bool IsNull(int32_t i) const { return (nulls_[i >> 3] >> (i & 7)) & 1; }
    void SetNull(int32_t i) { nulls_[i >> 3] |= static_cast<uint8_t>(1u << (i & 7)); }
    void SetNotNull(int32_t i) { nulls_[i >> 3] &= static_cast<uint8_t>(~(1u << (i & 7))); }
    bool HasNull() const
    {
        for (int32_t i = 0; i < size_; ++i)
            if (IsNull(i)) return true;
        return false;
    }
    uint8_t *RawValues() { return values_.data(); }
    const uint8_t *RawValues() const { return values_.data(); }
    uint8_t *RawNulls() { return nulls_.data(); }
    const uint8_t *RawNulls() const { return nulls_.data(); }
    int32_t Width() const { return width_; }

protected:
    DataTypeId id_;
    int32_t size_;
    int32_t width_;
    std::vector<uint8_t> values_;
    std::vector<uint8_t> nulls_;
};

template <typename T> class Vector : public BaseVector {
public:
    Vector(int32_t size, DataTypeId id) : BaseVector(id, size) {}
    T GetValue(int32_t i) const
    {
        T v;
        std::memcpy(&v, values_.data() + sizeof(T) * i, sizeof(T));
        return v;
    }
    void SetValue(int32_t i, T v)
    {
        std::memcpy(values_.data() + sizeof(T) * i, &v, sizeof(T));
        SetNotNull(i);
    }
};

class VectorBatch {
public:
    explicit VectorBatch(int32_t rowCount) : rowCount_(rowCount) {}
    ~VectorBatch() = default;
    void Append(BaseVector *v) { vectors_.emplace_back(v); }
    BaseVector *Get(int32_t i) const { return vectors_.at(i).get(); }
    int32_t GetVectorCount() const { return static_cast<int32_t>(vectors_.size()); }
    int32_t GetRowCount() const { return rowCount_; }
    // Moves the first k columns to the end (the reference rotates GetVectors(), lookup_join.cpp:1959-1962).
    void RotateLeft(int32_t k) { std::rotate(vectors_.begin(), vectors_.begin() + k, vectors_.end()); }

private:
    int32_t rowCount_;
    std::vector<std::unique_ptr<BaseVector>> vectors_;
};

struct VectorHelper {
    static void FreeVecBatch(VectorBatch *b) { delete b; }
    static BaseVector *CreateFlatVector(int32_t typeId, int32_t size)
    {
        return new BaseVector(static_cast<DataTypeId>(typeId), size);
    }
};
}  // namespace vec

namespace op {
enum OmniStatus { OMNI_STATUS_NORMAL = 0, OMNI_STATUS_FINISHED = 1, OMNI_STATUS_ERROR = -1 };

// core/src/plannode/planNode.h:406-414 and :28-32
enum JoinType {
    OMNI_JOIN_TYPE_INNER = 0,
    OMNI_JOIN_TYPE_LEFT,
    OMNI_JOIN_TYPE_RIGHT,
    OMNI_JOIN_TYPE_FULL,
    OMNI_JOIN_TYPE_LEFT_SEMI,
    OMNI_JOIN_TYPE_LEFT_ANTI,
    OMNI_JOIN_TYPE_EXISTENCE
};
enum BuildSide { OMNI_BUILD_UNKNOWN = 0, OMNI_BUILD_LEFT, OMNI_BUILD_RIGHT };

class OmniException : public std::runtime_error {
public:
    OmniException(const std::string &code, const std::string &msg) : std::runtime_error(code + ": " + msg), code_(code) {}
    const std::string &Code() const { return code_; }

private:
    std::string code_;
};

// Push/pull operator surface of the reference (operator.h:22-264).
class Operator {
public:
    virtual ~Operator() = default;
    virtual int32_t AddInput(vec::VectorBatch *vecBatch) = 0;
    virtual int32_t GetOutput(vec::VectorBatch **result) = 0;
    virtual OmniStatus Init() { return OMNI_STATUS_NORMAL; }
    virtual OmniStatus Close() { return OMNI_STATUS_NORMAL; }
    virtual uint64_t GetHashMapUniqueKeys() { return 0; }
    static void DeleteOperator(Operator *op)
    {
        op->Close();
        delete op;
    }
    OmniStatus GetStatus() const { return status_; }
    void SetStatus(OmniStatus s) { status_ = s; }
    virtual void noMoreInput() { noMoreInput_ = true; }
    virtual bool needsInput() { return status_ != OMNI_STATUS_FINISHED && !noMoreInput_; }
    bool isFinished() const { return status_ == OMNI_STATUS_FINISHED; }

protected:
    OmniStatus status_ = OMNI_STATUS_NORMAL;
    bool noMoreInput_ = true;
};

class OperatorFactory {
public:
    virtual ~OperatorFactory() = default;
    virtual Operator *CreateOperator() { return nullptr; }
};
}  // namespace op
}  // namespace omniruntime
