// Minimal OmniVec-compatible columnar containers and the Operator / OperatorFactory surface of the
// reference (core/src/vector/vector_batch.h:10-66, core/src/vector/vector.h:202-426,
// core/src/vector/nulls_buffer.h:16-117, core/src/type/data_type.h:40-65, core/src/type/decimal128.h:21-148,
// core/src/operator/operator.h:22-264, core/src/operator/operator_factory.h:12-27, core/src/operator/status.h:9-13).
//
// Only the subset the hash-aggregation / hash-join path needs, under the reference's names: flat and dictionary
// vectors of fixed-width types with slices (an offset into shared buffers), an LSB-first null bitmap (bit = 1 means
// NULL), the unsafe raw accessors (vector/unsafe_vector.h:17-90), batches of vectors, and the push/pull operator
// API.  The reference's own headers cannot be compiled in this image (vector.h includes jemalloc), so this subset
// stands in for them; the GPU operators use only accessors the reference's headers have under the same names.
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
    OMNI_BYTE = 18,
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
inline DataTypePtr BooleanType() { return std::make_shared<DataType>(OMNI_BOOLEAN); }
inline DataTypePtr ByteType() { return std::make_shared<DataType>(OMNI_BYTE); }

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
        case OMNI_BOOLEAN:
        case OMNI_BYTE: return 1;
        default: return 0;
    }
}
}  // namespace type

namespace vec {
using type::DataTypeId;

// vector.h:34-41
enum Encoding { OMNI_FLAT = 0, OMNI_DICTIONARY = 1, OMNI_ENCODING_CONTAINER = 2 };

class BaseVector;
template <typename T> class Vector;
template <typename T> class DictionaryContainer;
namespace unsafe {
class UnsafeBaseVector;
class UnsafeVector;
class UnsafeDictionaryVector;
}  // namespace unsafe

// LSB-first NULL bitmap (bit = 1 is NULL) of `size` rows (nulls_buffer.h:16-80).  A slice shares its parent's
// bitmap when the slice starts on a byte boundary and copies its bits otherwise, so GetNulls() always addresses
// the slice's own row 0.
class NullsBuffer {
public:
    explicit NullsBuffer(int32_t size, NullsBuffer *input = nullptr, int32_t sliceOffset = 0) : size_(size)
    {
        if (input != nullptr && sliceOffset % 8 == 0) {
            buf_ = input->buf_;
            nulls_ = input->nulls_ + sliceOffset / 8;
            hasNull_ = input->hasNull_;
            return;
        }
        buf_ = std::make_shared<std::vector<uint8_t>>(static_cast<size_t>((size + 7) / 8 + 8), 0);
        nulls_ = buf_->data();
        if (input != nullptr) {
            for (int32_t i = 0; i < size; ++i) SetNull(i, input->IsNull(sliceOffset + i));
        }
    }
    bool IsNull(int32_t i) const { return (nulls_[i >> 3] >> (i & 7)) & 1; }
    void SetNull(int32_t i, bool isNull)
    {
        if (isNull) {
            nulls_[i >> 3] |= static_cast<uint8_t>(1u << (i & 7));
            hasNull_ = true;
        } else {
            nulls_[i >> 3] &= static_cast<uint8_t>(~(1u << (i & 7)));
        }
    }
    bool HasNull() const { return hasNull_; }
    void SetNullFlag(bool hasNull) { hasNull_ = hasNull; }
    uint8_t *GetNulls() const { return nulls_; }

private:
    int32_t size_;
    std::shared_ptr<std::vector<uint8_t>> buf_;
    uint8_t *nulls_ = nullptr;
    bool hasNull_ = false;
};

// vector.h:45-200: size, encoding, type, slice offset and the NULL bitmap.
class BaseVector {
public:
    BaseVector(int32_t size, Encoding encoding, DataTypeId typeId, NullsBuffer *nulls = nullptr, int32_t sliceOffset = 0)
        : size(size), encoding(encoding), dataTypeId(typeId),
          nullsBuffer(std::make_shared<NullsBuffer>(size, nulls, sliceOffset))
    {
    }
    virtual ~BaseVector() = default;
    DataTypeId GetTypeId() const { return dataTypeId; }
    int32_t GetSize() const { return size; }
    Encoding GetEncoding() const { return encoding; }
    int32_t GetOffset() const { return offset; }
    bool IsNull(int32_t i) const { return nullsBuffer->IsNull(i); }
    void SetNull(int32_t i) { nullsBuffer->SetNull(i, true); }
    void SetNotNull(int32_t i) { nullsBuffer->SetNull(i, false); }
    bool HasNull() const { return nullsBuffer->HasNull(); }
    // Whether the vector may hold NULLs, for a bitmap written through UnsafeBaseVector::GetNulls (vector.h:117-120).
    void SetNullFlag(bool hasNull) { nullsBuffer->SetNullFlag(hasNull); }
    // A view of rows [positionOffset, positionOffset + length) sharing this vector's buffers (vector.h:152-155).
    virtual BaseVector *Slice(int32_t positionOffset, int32_t length) = 0;

protected:
    friend class unsafe::UnsafeBaseVector;
    int32_t size;
    Encoding encoding;
    DataTypeId dataTypeId;
    int32_t offset = 0;
    std::shared_ptr<NullsBuffer> nullsBuffer;
};

// A flat vector of T (vector.h:202-426): values[offset + i] is row i.
template <typename T> class Vector : public BaseVector {
public:
    Vector(int32_t size, DataTypeId typeId)
        : BaseVector(size, OMNI_FLAT, typeId), valuesBuffer(std::make_shared<std::vector<T>>(static_cast<size_t>(size)))
    {
        values = valuesBuffer->data();
    }
    T GetValue(int32_t i) const { return values[offset + i]; }
    void SetValue(int32_t i, T v)
    {
        values[offset + i] = v;
        SetNotNull(i);
    }
    BaseVector *Slice(int32_t positionOffset, int32_t length) override
    {
        auto *v = new Vector<T>(*this, positionOffset, length);
        return v;
    }

private:
    friend class unsafe::UnsafeVector;
    Vector(const Vector<T> &parent, int32_t positionOffset, int32_t length)
        : BaseVector(length, OMNI_FLAT, parent.dataTypeId, parent.nullsBuffer.get(), positionOffset),
          valuesBuffer(parent.valuesBuffer), values(parent.values)
    {
        offset = parent.offset + positionOffset;
    }
    std::shared_ptr<std::vector<T>> valuesBuffer;
    T *values = nullptr;
};

// Row i of a dictionary vector is dictionary[ids[offset + i]] (dictionary_container.h, vector.h:434-555).
template <typename T> class DictionaryContainer {
public:
    DictionaryContainer(std::vector<int32_t> ids, std::shared_ptr<std::vector<T>> dictionary)
        : ids(std::move(ids)), dictionary(std::move(dictionary))
    {
    }
    std::vector<int32_t> ids;
    std::shared_ptr<std::vector<T>> dictionary;
};

template <typename T> class Vector<DictionaryContainer<T>> : public BaseVector {
public:
    Vector(int32_t size, std::shared_ptr<DictionaryContainer<T>> container, DataTypeId typeId)
        : BaseVector(size, OMNI_DICTIONARY, typeId), container(std::move(container))
    {
    }
    T GetValue(int32_t i) const { return (*container->dictionary)[container->ids[offset + i]]; }
    BaseVector *Slice(int32_t positionOffset, int32_t length) override
    {
        auto *v = new Vector<DictionaryContainer<T>>(length, container, dataTypeId);
        v->offset = offset + positionOffset;
        v->nullsBuffer = std::make_shared<NullsBuffer>(length, nullsBuffer.get(), positionOffset);
        return v;
    }

private:
    friend class unsafe::UnsafeDictionaryVector;
    std::shared_ptr<DictionaryContainer<T>> container;
};

// The raw accessors the reference's operators use (unsafe_vector.h:17-90).
namespace unsafe {
class UnsafeBaseVector {
public:
    static uint8_t *GetNulls(BaseVector *vector) { return vector->nullsBuffer->GetNulls(); }
};
class UnsafeVector {
public:
    template <typename T> static T *GetRawValues(Vector<T> *vector) { return vector->values + vector->offset; }
};
class UnsafeDictionaryVector {
public:
    template <typename T> static T *GetDictionary(Vector<DictionaryContainer<T>> *vector)
    {
        return vector->container->dictionary->data();
    }
    template <typename T> static int32_t *GetIds(Vector<DictionaryContainer<T>> *vector)
    {
        return vector->container->ids.data() + vector->offset;
    }
};
}  // namespace unsafe

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

// A flat vector of the type's storage (vector_helper.h CreateFlatVector): double, a byte for BOOLEAN, int16, int32,
// Decimal128, else int64.
struct VectorHelper {
    static void FreeVecBatch(VectorBatch *b) { delete b; }
    // vector_helper.h SliceVector: a view of rows [offset, offset + length) sharing the vector's buffers
    static BaseVector *SliceVector(BaseVector *v, int32_t offset, int32_t length) { return v->Slice(offset, length); }
    static BaseVector *CreateFlatVector(int32_t typeId, int32_t size)
    {
        const auto id = static_cast<DataTypeId>(typeId);
        switch (typeId) {
            case type::OMNI_DOUBLE: return new Vector<double>(size, id);
            case type::OMNI_BOOLEAN: return new Vector<uint8_t>(size, id);
            case type::OMNI_BYTE: return new Vector<int8_t>(size, id);
            case type::OMNI_SHORT: return new Vector<int16_t>(size, id);
            case type::OMNI_INT:
            case type::OMNI_DATE32: return new Vector<int32_t>(size, id);
            case type::OMNI_DECIMAL128: return new Vector<type::Decimal128>(size, id);
            default: return new Vector<int64_t>(size, id);
        }
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
    // operator.h:92-95: the identity unless the operator reshapes raw input into its output schema
    virtual vec::VectorBatch *AlignSchema(vec::VectorBatch *inputVecBatch) { return inputVecBatch; }
    static void DeleteOperator(Operator *op)
    {
        op->Close();
        delete op;
    }
    OmniStatus GetStatus() const { return status_; }
    void SetStatus(OmniStatus s) { status_ = s; }
    virtual void noMoreInput() { noMoreInput_ = true; }
    virtual void setNoMoreInput(bool noMoreInput) { noMoreInput_ = noMoreInput; }
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
