// CPU restatement of OmniOperator's hash-aggregation and hash-join hot path.
//
// TEST INFRASTRUCTURE ONLY.  This is the oracle the GPU path is checked against and the CPU
// baseline bench.py times.  Nothing under omnioperator_amd/ links or calls it.
//
// The reference cannot be compiled in this image (aarch64-only asm/NEON, jemalloc, securec, LLVM-15),
// so this file restates its algorithm on x86-64, function by function, citing the reference file:line
// each part follows (paths relative to the reference repository root):
//   - hashes:       core/src/operator/hashmap/crc_hasher.h, crc32c.h, group_hasher.h,
//                   core/src/operator/hash_util.h
//   - Swiss table:  core/src/operator/hashmap/base_hash_map.h (BaseHashMap, Grower, ProbeSeq)
//   - key handling: core/src/operator/hashmap/column_marshaller.h, vector_marshaller.cpp
//   - aggregation:  core/src/operator/aggregation/group_aggregation.cpp and aggregator/*.h
//   - join:         core/src/operator/join/join_hash_table_variants.cpp, lookup_join.cpp, row_ref.h
// Parity of each part is pinned by tests/test_oracle_*.py (known-answer vectors from the reference's
// own tests, closed forms and a brute-force sort-based oracle).
#include <nmmintrin.h>
#include <emmintrin.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../include/omni_gpu.h"

using int128 = __int128;

namespace {

thread_local std::string g_err;

// ================================================================================================
// Hashes
// ================================================================================================

// HashUint64: crc_hasher.h:21-26.  One aarch64 `crc32cx w, w, x` step with the accumulator
// preset to all ones and no final inversion; x86 SSE4.2 `crc32 r64` computes the same CRC-32C step.
inline uint64_t HashUint64(uint64_t x) { return static_cast<uint32_t>(_mm_crc32_u64(0xFFFFFFFFu, x)); }

// Bitwise (reflected polynomial 0x82F63B78) form of the same step, used to cross-check the SSE4.2 one.
inline uint32_t Crc32cStepSoftware(uint32_t crc, uint64_t x)
{
    for (int byte = 0; byte < 8; ++byte) {
        crc ^= static_cast<uint32_t>((x >> (8 * byte)) & 0xFF);
        for (int bit = 0; bit < 8; ++bit) {
            crc = (crc >> 1) ^ (0x82F63B78u & (0u - (crc & 1u)));
        }
    }
    return crc;
}

// CRC32HasherForInt: crc_hasher.h:44-54 -- the key is zero-extended into a uint64 (a union whose
// 64-bit member is cleared first), then HashUint64.
template <typename T> inline uint64_t CRC32HasherForInt(T key)
{
    uint64_t u = 0;
    std::memcpy(&u, &key, sizeof(T));
    return HashUint64(u);
}

// HashCRC32<Decimal128>: crc_hasher.h:75-86 -- two chained steps over low then high word.
inline uint64_t HashCrcDecimal128(uint64_t lo, uint64_t hi)
{
    uint32_t c = static_cast<uint32_t>(_mm_crc32_u64(0xFFFFFFFFu, lo));
    return static_cast<uint32_t>(_mm_crc32_u64(c, hi));
}

// Extend (crc32c.h:106-145): standard CRC-32C of a byte string (pre- and post-inverted).
inline uint32_t Crc32cExtend(uint32_t crc, const uint8_t *p, size_t n)
{
    uint64_t c = ~crc & 0xFFFFFFFFu;
    while (n >= 8) {
        uint64_t v;
        std::memcpy(&v, p, 8);
        c = _mm_crc32_u64(c, v);
        p += 8;
        n -= 8;
    }
    uint32_t c32 = static_cast<uint32_t>(c);
    while (n--) c32 = _mm_crc32_u8(c32, *p++);
    return ~c32;
}

constexpr uint64_t P1 = 0x9E3779B185EBCA87ULL;  // hash_util.h:42-46
constexpr uint64_t P2 = 0xC2B2AE3D27D4EB4FULL;
constexpr uint64_t P3 = 0x165667B19E3779F9ULL;
constexpr uint64_t P4 = 0x85EBCA77C2B2AE63ULL;
constexpr uint64_t P5 = 0x27D4EB2F165667C5ULL;

inline uint64_t Rotl(uint64_t v, unsigned r) { return (v << r) | (v >> (64 - r)); }

// HashUtil::HashValue(int64) / (int32): hash_util.h:74-82.
inline int64_t HashValueI64(int64_t v) { return static_cast<int64_t>(Rotl(static_cast<uint64_t>(v) * P2, 31) * P1); }

// xxHash64-style helpers: hash_util.h:160-268.
inline uint64_t XxMix(uint64_t acc, uint64_t v) { return Rotl(acc + v * P2, 31) * P1; }
inline uint64_t XxTail64(uint64_t h, uint64_t v) { return Rotl(h ^ XxMix(0, v), 27) * P1 + P4; }
inline uint64_t XxTail32(uint64_t h, uint32_t v) { return Rotl(h ^ (static_cast<uint64_t>(v) * P1), 23) * P2 + P3; }
inline uint64_t XxTail8(uint64_t h, uint8_t v) { return Rotl(h ^ (static_cast<uint64_t>(v) * P5), 11) * P1; }
inline uint64_t XxFinal(uint64_t h)
{
    h ^= h >> 33;
    h *= P2;
    h ^= h >> 29;
    h *= P3;
    h ^= h >> 32;
    return h;
}
inline uint64_t XxMerge(uint64_t h, uint64_t v) { return (h ^ XxMix(0, v)) * P1 + P4; }

// XxHash64Hash(seed=0, data, 0, len): hash_util.h:255-268 with UpdateBody/UpdateTail.
inline int64_t XxHash64(const uint8_t *p, int32_t len)
{
    uint64_t h;
    if (len >= 32) {
        uint64_t v1 = P1 + P2, v2 = P2, v3 = 0, v4 = 0 - P1;
        int32_t remaining = len;
        const uint8_t *q = p;
        while (remaining >= 32) {
            uint64_t a, b, c, d;
            std::memcpy(&a, q, 8);
            std::memcpy(&b, q + 8, 8);
            std::memcpy(&c, q + 16, 8);
            std::memcpy(&d, q + 24, 8);
            v1 = XxMix(v1, a);
            v2 = XxMix(v2, b);
            v3 = XxMix(v3, c);
            v4 = XxMix(v4, d);
            q += 32;
            remaining -= 32;
        }
        h = Rotl(v1, 1) + Rotl(v2, 7) + Rotl(v3, 12) + Rotl(v4, 18);
        h = XxMerge(h, v1);
        h = XxMerge(h, v2);
        h = XxMerge(h, v3);
        h = XxMerge(h, v4);
    } else {
        h = P5;
    }
    h += static_cast<uint64_t>(len);
    int32_t i = len & ~31;
    while (i <= len - 8) {
        uint64_t v;
        std::memcpy(&v, p + i, 8);
        h = XxTail64(h, v);
        i += 8;
    }
    if (i <= len - 4) {
        uint32_t v;
        std::memcpy(&v, p + i, 4);
        h = XxTail32(h, v);
        i += 4;
    }
    while (i < len) {
        h = XxTail8(h, p[i]);
        ++i;
    }
    return static_cast<int64_t>(XxFinal(h));
}

// GetRawHashPartition: hash_util.h:146-154 (bit-reverse, then the xxhash single-long finaliser).
inline uint32_t RawHashPartition(int64_t rawHash, uint32_t mask)
{
    uint64_t r = 0, x = static_cast<uint64_t>(rawHash);
    for (int i = 0; i < 64; ++i) r |= ((x >> i) & 1ULL) << (63 - i);
    uint64_t h = P5 + 8;
    h = XxTail64(h, r);
    h = XxFinal(h);
    return static_cast<uint32_t>(h) & mask;
}

// GetRawHashPosition: hash_util.h:131-141 (murmur3 fmix64).
inline uint32_t RawHashPosition(int64_t rawHash, uint64_t mask)
{
    uint64_t h = static_cast<uint64_t>(rawHash);
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdULL;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ULL;
    h ^= h >> 33;
    return static_cast<uint32_t>(h & mask);
}

// ================================================================================================
// Columns
// ================================================================================================

struct HostColumn {
    int32_t type_id;
    int32_t scale;
    const void *values;
    const uint8_t *nulls;  // LSB-first bitmap, bit = 1 means NULL
};

inline bool IsNull(const HostColumn &c, int64_t i) { return c.nulls && ((c.nulls[i >> 3] >> (i & 7)) & 1); }

inline int TypeWidth(int32_t t)
{
    switch (t) {
        case OMNI_TYPE_SHORT: return 2;
        case OMNI_TYPE_INT:
        case OMNI_TYPE_DATE32: return 4;
        case OMNI_TYPE_LONG:
        case OMNI_TYPE_TIMESTAMP:
        case OMNI_TYPE_DECIMAL64:
        case OMNI_TYPE_DOUBLE: return 8;
        case OMNI_TYPE_DECIMAL128: return 16;
        case OMNI_TYPE_BOOLEAN:
        case OMNI_TYPE_BYTE: return 1;
        default: return 0;
    }
}

inline int64_t LoadInt(const HostColumn &c, int64_t i)
{
    switch (c.type_id) {
        case OMNI_TYPE_SHORT: return static_cast<const int16_t *>(c.values)[i];
        case OMNI_TYPE_INT:
        case OMNI_TYPE_DATE32: return static_cast<const int32_t *>(c.values)[i];
        case OMNI_TYPE_BOOLEAN: return static_cast<const uint8_t *>(c.values)[i] ? 1 : 0;
        case OMNI_TYPE_BYTE: return static_cast<const int8_t *>(c.values)[i];
        default: return static_cast<const int64_t *>(c.values)[i];
    }
}

inline int128 LoadInt128(const HostColumn &c, int64_t i)
{
    if (c.type_id == OMNI_TYPE_DECIMAL128) {
        const uint64_t *w = static_cast<const uint64_t *>(c.values) + 2 * i;
        return static_cast<int128>((static_cast<unsigned __int128>(w[1]) << 64) | w[0]);
    }
    return LoadInt(c, i);
}

// ================================================================================================
// Swiss table: BaseHashMap (base_hash_map.h:455-887)
// ================================================================================================

constexpr int8_t kEmpty = -128;   // base_hash_map.h:402-405
constexpr int8_t kSentinel = -1;
constexpr int kGroupWidth = 16;   // base_hash_map.h:407-411
constexpr double kMaxLoad = 0.9;  // base_hash_map.h:94

// Grower: base_hash_map.h:255-292 -- +2 degrees below 23, +1 from 23 on; threshold = ceil(cap*0.9).
struct Grower {
    uint8_t degree;
    uint64_t threshold = 0;
    explicit Grower(uint8_t d) : degree(d) { Calc(); }
    void Calc() { threshold = static_cast<uint64_t>(std::ceil(static_cast<double>(1ULL << degree) * kMaxLoad)); }
    uint64_t Grow()
    {
        degree += (degree >= 23 ? 1 : 2);
        Calc();
        return 1ULL << degree;
    }
    uint64_t Size() const { return 1ULL << degree; }
};

template <typename K> struct Slot {
    K key{};
    uint32_t value = 0;
    uint64_t hash = 0;
};

// K: key type; Hasher(K) -> uint64 hash; equality = cached hash equal and key ==.
template <typename K, typename Hasher> class SwissMap {
public:
    explicit SwissMap(uint8_t degree) : grower_(degree) { InitSlots(grower_.Size()); }

    // Emplace: base_hash_map.h:528-547 (rehash check first, then find-or-insert).
    uint32_t *Emplace(const K &key, bool &inserted)
    {
        if (elements_ > grower_.threshold) Rehash();
        return EmplaceNoRehash(key, hasher_(key), inserted);
    }

    // EmplaceNotNullKey(key, hash): base_hash_map.h:549-576 (no rehash check).
    uint32_t *EmplaceNoRehash(const K &key, uint64_t hash, bool &inserted)
    {
        inserted = false;
        uint64_t pos = FindPosition(key, hash, inserted);
        if (inserted) {
            slots_[pos].key = key;
            slots_[pos].value = 0;
            ctrl_[pos] = static_cast<int8_t>(hash & 0x7F);
            ++elements_;
        }
        slots_[pos].hash = hash;
        return &slots_[pos].value;
    }

    // EmplaceNullValue: base_hash_map.h:579-589 -- one out-of-table slot, counted in the size.
    uint32_t *EmplaceNull(bool &inserted)
    {
        inserted = !hasNull_;
        if (!hasNull_) {
            hasNull_ = true;
            ++elements_;
        }
        return &nullValue_;
    }

    // FindMatchPosition: base_hash_map.h:507-519.
    const uint32_t *Find(const K &key) const
    {
        bool notFound = false;
        uint64_t pos = const_cast<SwissMap *>(this)->FindPosition(key, hasher_(key), notFound);
        return notFound ? nullptr : &slots_[pos].value;
    }

    bool HasNull() const { return hasNull_; }
    uint32_t NullValue() const { return nullValue_; }
    uint64_t Size() const { return elements_; }
    uint64_t Capacity() const { return capacity_; }

    // Slot-order traversal (HashmapIteratorOutput::HandleElements, base_hash_map.h:687-743):
    // occupied slots in index order, the NULL slot last.
    template <typename F> void ForEachInSlotOrder(F &&f) const
    {
        for (uint64_t i = 0; i < capacity_; ++i) {
            if (ctrl_[i] >= 0) f(slots_[i].key, slots_[i].value, false);
        }
        if (hasNull_) f(K{}, nullValue_, true);
    }

private:
    void InitSlots(uint64_t cap)  // base_hash_map.h:472-483
    {
        capacity_ = cap;
        ctrl_.assign(cap + kGroupWidth, kEmpty);
        std::fill(ctrl_.begin() + cap, ctrl_.end(), kSentinel);
        slots_.assign(cap, Slot<K>());
    }

    // ProbeSeq + FindPosition: base_hash_map.h:411-445, 807-835.  The 16 control bytes starting at
    // the (unaligned) probe offset are compared against H2; slot indices wrap with the mask.
    uint64_t FindPosition(const K &key, uint64_t hash, bool &inserted)
    {
        const uint64_t mask = capacity_ - 1;
        uint64_t offset = (hash >> 7) & mask;
        uint64_t index = 0;
        const __m128i h2 = _mm_set1_epi8(static_cast<char>(hash & 0x7F));
        const __m128i empty = _mm_set1_epi8(static_cast<char>(kEmpty));
        while (ctrl_[offset] != kEmpty) {
            __m128i g = _mm_loadu_si128(reinterpret_cast<const __m128i *>(ctrl_.data() + offset));
            uint32_t match = static_cast<uint32_t>(_mm_movemask_epi8(_mm_cmpeq_epi8(g, h2)));
            while (match) {
                uint32_t v = static_cast<uint32_t>(__builtin_ctz(match));
                match &= match - 1;
                const Slot<K> &s = slots_[(offset + v) & mask];
                if (s.hash == hash && s.key == key) return (offset + v) & mask;
            }
            uint32_t em = static_cast<uint32_t>(_mm_movemask_epi8(_mm_cmpeq_epi8(g, empty)));
            if (em) {
                inserted = true;
                return (offset + static_cast<uint32_t>(__builtin_ctz(em))) & mask;
            }
            index += kGroupWidth;
            offset = (offset + index) & mask;
            if (index > capacity_) break;
        }
        inserted = true;
        return offset;
    }

    // Rehash: base_hash_map.h:758-789 -- grow, then reinsert in old slot order.
    void Rehash()
    {
        std::vector<int8_t> oldCtrl;
        std::vector<Slot<K>> oldSlots;
        oldCtrl.swap(ctrl_);
        oldSlots.swap(slots_);
        uint64_t oldCap = capacity_;
        InitSlots(grower_.Grow());
        for (uint64_t i = 0; i < oldCap; ++i) {
            if (oldCtrl[i] < 0) continue;
            bool dummy = false;
            uint64_t pos = FindPosition(oldSlots[i].key, oldSlots[i].hash, dummy);
            ctrl_[pos] = static_cast<int8_t>(oldSlots[i].hash & 0x7F);
            slots_[pos] = oldSlots[i];
        }
    }

    Grower grower_;
    Hasher hasher_;
    std::vector<int8_t> ctrl_;
    std::vector<Slot<K>> slots_;
    uint64_t capacity_ = 0;
    uint64_t elements_ = 0;
    bool hasNull_ = false;
    uint32_t nullValue_ = 0;
};

template <typename T> struct GroupbyCrcHasher {  // group_hasher.h:28-48
    uint64_t operator()(T k) const { return CRC32HasherForInt<T>(k); }
};

struct ByteKey {
    std::string bytes;
    bool operator==(const ByteKey &o) const { return bytes == o.bytes; }
};
struct ByteKeyHasher {  // group_hasher.h:21-26 -> HashUtil::HashValue(int8*, len) = XxHash64Hash(0, ...)
    uint64_t operator()(const ByteKey &k) const
    {
        return static_cast<uint64_t>(XxHash64(reinterpret_cast<const uint8_t *>(k.bytes.data()),
                                              static_cast<int32_t>(k.bytes.size())));
    }
};

template <typename T> struct JoinIntHasher {  // HashCRC32<T> for integers: crc_hasher.h:28-42 (identity)
    uint64_t operator()(T k) const { return static_cast<uint64_t>(static_cast<int64_t>(k)); }
};

// ================================================================================================
// Composite-key serialisation: SerializeValueIntoArena / FixedLenTypeSerializer
// (vector_marshaller.cpp:155-232).  Per column: [1-byte size][value truncated to size bytes];
// NULL = [0].  Integers use the smallest power-of-two byte count that keeps the sign bit.
// ================================================================================================

inline uint8_t FixedIntSize(int64_t v)
{
    uint64_t t = static_cast<uint64_t>(v < 0 ? ~v : v);
    uint32_t bytes = t == 0 ? 0 : static_cast<uint32_t>((64 - __builtin_clzll(t)) / 8);
    // aarch64 `clz` of 0 is 32, so the reference's `1 << (32 - clz(bytes))` yields 1 for bytes == 0.
    uint32_t lead = bytes == 0 ? 32 : static_cast<uint32_t>(__builtin_clz(bytes));
    return static_cast<uint8_t>(1u << (32 - lead));
}

inline void SerializeColumnValue(const HostColumn &c, int64_t row, std::string &out)
{
    if (IsNull(c, row)) {
        out.push_back('\0');
        return;
    }
    switch (c.type_id) {
        case OMNI_TYPE_DOUBLE: {
            out.push_back(8);
            out.append(static_cast<const char *>(c.values) + 8 * row, 8);
            return;
        }
        case OMNI_TYPE_BOOLEAN: {
            out.push_back(1);
            out.push_back(static_cast<const uint8_t *>(c.values)[row] ? 1 : 0);
            return;
        }
        case OMNI_TYPE_DECIMAL128: {  // Decimal128Serializer: vector_marshaller.cpp:144-152
            out.push_back(16);
            out.append(static_cast<const char *>(c.values) + 16 * row, 16);
            return;
        }
        default: {
            int64_t v = LoadInt(c, row);
            uint8_t sz = FixedIntSize(v);
            out.push_back(static_cast<char>(sz));
            out.append(reinterpret_cast<const char *>(&v), sz);  // little-endian truncation
            return;
        }
    }
}

// ================================================================================================
// Aggregators (default policy: SupportContainerVecRule::NOT_SUPPORT, util/policy.h:162-167)
// ================================================================================================

enum class VState : int8_t { EMPTY = 0, NORMAL = 1, OVERFLOWED = 2 };  // state_flag_operation.h:12-16

constexpr int128 kDec128Max = (static_cast<int128>(5421010862427522170LL) << 64) + 687399551400673279LL;  // 10^38-1

inline bool AddChecked128(int128 a, int128 b, int128 &r)  // base_operations.h:23-28
{
    bool o = __builtin_add_overflow(a, b, &r);
    return o || r > kDec128Max || r < -kDec128Max;
}
inline bool MulChecked128(int128 a, int128 b, int128 &r)  // base_operations.h:30-35
{
    bool o = __builtin_mul_overflow(a, b, &r);
    return o || r > kDec128Max || r < -kDec128Max;
}

struct AggDesc {
    int32_t func, inType, outType, inScale, outScale;
    bool partialOut = false;  // partial output: AVG its sum column; decimal SUM 0 when empty, NULL on overflow
    size_t offset = 0;  // state offset inside the per-group blob
    size_t size = 0;
};

template <typename T> inline T Ld(const uint8_t *p)
{
    T v;
    std::memcpy(&v, p, sizeof(T));
    return v;
}
template <typename T> inline void St(uint8_t *p, T v) { std::memcpy(p, &v, sizeof(T)); }

inline bool IsDecimal(int32_t t) { return t == OMNI_TYPE_DECIMAL64 || t == OMNI_TYPE_DECIMAL128; }

// State layouts are the reference's packed structs (#pragma pack(1)):
//   SUM int/long: {int64 value; int8 state}    sum_flatim_aggregator.h:18-44
//   SUM double:   {double value; int8 state}
//   SUM decimal:  {int64|int128 value; int8 state}  sum_spark_decimal_aggregator.h
//   COUNT:        {int64 count}                count_column_aggregator.h
//   AVG:          {double value; int64 count}  average_flatim_aggregator.h:20-44
//   MIN/MAX:      {ResultType value; int8 state}  max_aggregator.h:64-88 (int16 promoted to int32)
size_t StateSize(const AggDesc &a)
{
    switch (a.func) {
        case OMNI_AGG_COUNT_ALL:
        case OMNI_AGG_COUNT_COLUMN: return 8;
        case OMNI_AGG_AVG: return 16;
        case OMNI_AGG_SUM:
            if (IsDecimal(a.inType)) return (a.outType == OMNI_TYPE_DECIMAL128 ? 16 : 8) + 1;
            return 9;
        case OMNI_AGG_MIN:
        case OMNI_AGG_MAX:
            if (a.inType == OMNI_TYPE_DECIMAL128) return 17;
            if (a.inType == OMNI_TYPE_SHORT || a.inType == OMNI_TYPE_INT || a.inType == OMNI_TYPE_DATE32) return 5;
            return 9;
        default: return 0;
    }
}

void InitState(const AggDesc &a, uint8_t *s)
{
    std::memset(s, 0, a.size);
    if (a.func == OMNI_AGG_MIN || a.func == OMNI_AGG_MAX) {
        bool mn = a.func == OMNI_AGG_MIN;  // MinAggregator::InitStates: min_aggregator.cpp:81-86
        switch (a.inType) {
            case OMNI_TYPE_SHORT:
            case OMNI_TYPE_INT:
            case OMNI_TYPE_DATE32:
                St<int32_t>(s, mn ? std::numeric_limits<int32_t>::max() : std::numeric_limits<int32_t>::min());
                break;
            case OMNI_TYPE_DOUBLE:
                St<double>(s, mn ? std::numeric_limits<double>::max() : std::numeric_limits<double>::lowest());
                break;
            case OMNI_TYPE_DECIMAL128:
                St<int128>(s, mn ? kDec128Max : -kDec128Max);
                break;
            default:
                St<int64_t>(s, mn ? std::numeric_limits<int64_t>::max() : std::numeric_limits<int64_t>::min());
        }
    }
}

// Row-ordered update of one aggregator over a batch: AddUseRowIndex (operations_hash_aggregator.h:16-32)
// with the per-function update rules (SumOp sum_aggregator.h:12-54; MaxOp max_aggregator.h:39-56;
// AddDecimalUseRowIndex sum_spark_decimal_aggregator.h:15-66; AVG average_flatim_aggregator.h:30-35).
void UpdateStates(const AggDesc &a, const HostColumn *col, uint8_t *const *rowStates, int64_t rows)
{
    for (int64_t i = 0; i < rows; ++i) {
        uint8_t *s = rowStates[i] + a.offset;
        if (a.func == OMNI_AGG_COUNT_ALL) {
            St<int64_t>(s, Ld<int64_t>(s) + 1);
            continue;
        }
        if (IsNull(*col, i)) continue;
        switch (a.func) {
            case OMNI_AGG_COUNT_COLUMN: St<int64_t>(s, Ld<int64_t>(s) + 1); break;
            case OMNI_AGG_AVG: {
                double v = col->type_id == OMNI_TYPE_DOUBLE ? static_cast<const double *>(col->values)[i]
                                                            : static_cast<double>(LoadInt(*col, i));
                St<double>(s, Ld<double>(s) + v);
                St<int64_t>(s + 8, Ld<int64_t>(s + 8) + 1);
                break;
            }
            case OMNI_AGG_SUM: {
                if (IsDecimal(a.inType)) {
                    uint8_t *flag = s + a.size - 1;
                    if (*flag == static_cast<uint8_t>(VState::OVERFLOWED)) break;
                    bool ovf;
                    if (a.outType == OMNI_TYPE_DECIMAL128) {
                        int128 r;
                        ovf = AddChecked128(Ld<int128>(s), LoadInt128(*col, i), r);
                        St<int128>(s, r);
                    } else {
                        int64_t r;
                        ovf = __builtin_add_overflow(Ld<int64_t>(s), LoadInt(*col, i), &r);
                        St<int64_t>(s, r);
                    }
                    *flag = static_cast<uint8_t>(ovf ? VState::OVERFLOWED : VState::NORMAL);
                } else if (a.inType == OMNI_TYPE_DOUBLE) {
                    St<double>(s, Ld<double>(s) + static_cast<const double *>(col->values)[i]);
                    s[8] = static_cast<uint8_t>(VState::NORMAL);
                } else {  // wrapping two's-complement add (CareOverFlow = false)
                    uint64_t r = static_cast<uint64_t>(Ld<int64_t>(s)) + static_cast<uint64_t>(LoadInt(*col, i));
                    St<int64_t>(s, static_cast<int64_t>(r));
                    s[8] = static_cast<uint8_t>(VState::NORMAL);
                }
                break;
            }
            case OMNI_AGG_MIN:
            case OMNI_AGG_MAX: {
                bool mn = a.func == OMNI_AGG_MIN;
                switch (a.inType) {
                    case OMNI_TYPE_SHORT:
                    case OMNI_TYPE_INT:
                    case OMNI_TYPE_DATE32: {
                        int32_t v = static_cast<int32_t>(LoadInt(*col, i)), c = Ld<int32_t>(s);
                        St<int32_t>(s, mn ? std::min(c, v) : std::max(c, v));
                        s[4] = static_cast<uint8_t>(VState::NORMAL);
                        break;
                    }
                    case OMNI_TYPE_DOUBLE: {
                        double v = static_cast<const double *>(col->values)[i], c = Ld<double>(s);
                        St<double>(s, mn ? std::min(c, v) : std::max(c, v));
                        s[8] = static_cast<uint8_t>(VState::NORMAL);
                        break;
                    }
                    case OMNI_TYPE_DECIMAL128: {
                        int128 v = LoadInt128(*col, i), c = Ld<int128>(s);
                        St<int128>(s, mn ? std::min(c, v) : std::max(c, v));
                        s[16] = static_cast<uint8_t>(VState::NORMAL);
                        break;
                    }
                    default: {
                        int64_t v = LoadInt(*col, i), c = Ld<int64_t>(s);
                        St<int64_t>(s, mn ? std::min(c, v) : std::max(c, v));
                        s[8] = static_cast<uint8_t>(VState::NORMAL);
                    }
                }
                break;
            }
            default: break;
        }
    }
}

struct OutCol {
    int32_t type_id;
    int32_t reserved;
    void *values;
    uint8_t *nulls;  // LSB-first bitmap
};

inline void SetNullBit(OutCol &o, int64_t r) { o.nulls[r >> 3] |= static_cast<uint8_t>(1u << (r & 7)); }

// ExtractValuesBatch of each aggregator (sum_flatim_aggregator.h:243-257, sum_spark_decimal_aggregator.h:246-301,
// max_aggregator.cpp:34-55, average_flatim_aggregator.h:275-290).  Returns false on an overflow that
// must raise (overflowAsNull == false).
bool ExtractState(const AggDesc &a, const uint8_t *s, OutCol &o, int64_t r, bool overflowAsNull)
{
    switch (a.func) {
        case OMNI_AGG_COUNT_ALL:
        case OMNI_AGG_COUNT_COLUMN: static_cast<int64_t *>(o.values)[r] = Ld<int64_t>(s); return true;
        case OMNI_AGG_AVG: {
            int64_t cnt = Ld<int64_t>(s + 8);
            if (a.partialOut) {  // partial output: the sum, 0 for an empty group (average_flatim_aggregator.h:260-273)
                static_cast<double *>(o.values)[r] = cnt == 0 ? 0.0 : Ld<double>(s);
                return true;
            }
            if (cnt == 0) {
                SetNullBit(o, r);
                static_cast<double *>(o.values)[r] = 0;
            } else {
                static_cast<double *>(o.values)[r] = Ld<double>(s) / static_cast<double>(cnt);
            }
            return true;
        }
        case OMNI_AGG_SUM: {
            uint8_t flag = s[a.size - 1];
            if (!IsDecimal(a.inType)) {
                if (flag != static_cast<uint8_t>(VState::NORMAL)) {
                    SetNullBit(o, r);
                    std::memset(static_cast<uint8_t *>(o.values) + 8 * r, 0, 8);
                } else {
                    std::memcpy(static_cast<uint8_t *>(o.values) + 8 * r, s, 8);
                }
                return true;
            }
            // ExtractValuesBatchInternal<OUTPUT_PARTIAL> (sum_spark_decimal_aggregator.h:246-301): the partial
            // output is 0 for an empty group and NULL for an overflowed one, without raising
            int width = a.outType == OMNI_TYPE_DECIMAL128 ? 16 : 8;
            std::memset(static_cast<uint8_t *>(o.values) + width * r, 0, width);
            if (flag == static_cast<uint8_t>(VState::EMPTY)) {
                if (!a.partialOut) SetNullBit(o, r);
                return true;
            }
            int128 v = a.outType == OMNI_TYPE_DECIMAL128 ? Ld<int128>(s) : static_cast<int128>(Ld<int64_t>(s));
            bool ovf = flag == static_cast<uint8_t>(VState::OVERFLOWED);
            int128 scaled = 0;
            if (!ovf) {
                // MulCheckedOverflow(value, 10^(outScale - inScale)) against the DECIMAL128 limits only, for both
                // output types (sum_spark_decimal_aggregator.h:284-296)
                int128 p = 1;
                for (int k = 0; k < a.outScale - a.inScale; ++k) p *= 10;
                ovf = MulChecked128(v, p, scaled);
            }
            if (ovf) {
                SetNullBit(o, r);
                return overflowAsNull || a.partialOut;
            }
            if (width == 8) {
                // SetValToVector: static_cast<int64_t> of the rescaled int128 (:580-589), i.e. its low 64 bits
                static_cast<int64_t *>(o.values)[r] = static_cast<int64_t>(scaled);
            } else {
                std::memcpy(static_cast<uint8_t *>(o.values) + 16 * r, &scaled, 16);
            }
            return true;
        }
        case OMNI_AGG_MIN:
        case OMNI_AGG_MAX: {
            int w = TypeWidth(a.outType);
            uint8_t *dst = static_cast<uint8_t *>(o.values) + static_cast<int64_t>(w) * r;
            std::memset(dst, 0, w);
            if (s[a.size - 1] == static_cast<uint8_t>(VState::EMPTY)) {
                SetNullBit(o, r);
                return true;
            }
            if (a.inType == OMNI_TYPE_SHORT) {
                St<int16_t>(dst, static_cast<int16_t>(Ld<int32_t>(s)));
            } else {
                std::memcpy(dst, s, w);
            }
            return true;
        }
        default: return true;
    }
}

// ================================================================================================
// HashAggregationOperator (group_aggregation.cpp), single stage
// ================================================================================================

enum class HandleType { fixedInt16, fixedInt32, fixedInt64, serialize };

class HashAgg {
public:
    explicit HashAgg(const omni_gpu_agg_spec &spec) : spec_(spec)
    {
        // ChooseGroupByType: group_aggregation.cpp:145-164.
        handle_ = HandleType::serialize;
        if (spec.num_keys == 1) {
            int32_t t = spec.key_types[0];
            if (t == OMNI_TYPE_INT || t == OMNI_TYPE_DATE32) handle_ = HandleType::fixedInt32;
            else if (t == OMNI_TYPE_LONG || t == OMNI_TYPE_TIMESTAMP || t == OMNI_TYPE_DECIMAL64) handle_ = HandleType::fixedInt64;
            else if (t == OMNI_TYPE_SHORT) handle_ = HandleType::fixedInt16;
        }
        size_t off = 0;
        for (int i = 0; i < spec.num_aggs; ++i) {
            AggDesc a{spec.agg_funcs[i], spec.agg_in_types[i], spec.agg_out_types[i], spec.agg_in_scales[i],
                      spec.agg_out_scales[i]};
            a.partialOut = (spec.agg_funcs[i] == OMNI_AGG_AVG ||
                            (spec.agg_funcs[i] == OMNI_AGG_SUM && IsDecimal(spec.agg_in_types[i]))) &&
                           spec.agg_partial_out[i] != 0;
            a.size = StateSize(a);
            a.offset = off;
            off += a.size;
            aggs_.push_back(a);
        }
        blob_ = off == 0 ? 1 : off;
        // Default degrees: BaseHashMap 15 (base_hash_map.h:885); ColumnSerializeHandler 16 (column_marshaller.h:34).
        if (handle_ == HandleType::fixedInt16) m16_.reset(new SwissMap<int16_t, GroupbyCrcHasher<int16_t>>(15));
        if (handle_ == HandleType::fixedInt32) m32_.reset(new SwissMap<int32_t, GroupbyCrcHasher<int32_t>>(15));
        if (handle_ == HandleType::fixedInt64) m64_.reset(new SwissMap<int64_t, GroupbyCrcHasher<int64_t>>(15));
        if (handle_ == HandleType::serialize) ms_.reset(new SwissMap<ByteKey, ByteKeyHasher>(16));
    }

    // AddInput + Emplace<>: group_aggregation.cpp:320-406, 811-890.
    void AddBatch(const HostColumn *keys, const HostColumn *vals, int64_t rows)
    {
        if (rows <= 0) return;
        rowStates_.resize(rows);
        for (int64_t r = 0; r < rows; ++r) {
            bool inserted = false;
            uint32_t *v = EmplaceRow(keys, r, inserted);
            if (inserted) {
                uint32_t gid = static_cast<uint32_t>(numGroups_++);
                *v = gid;
                states_.resize(static_cast<size_t>(numGroups_) * blob_);
                for (auto &a : aggs_) InitState(a, states_.data() + static_cast<size_t>(gid) * blob_ + a.offset);
                StoreKeys(keys, r, gid);
            }
            rowStates_[r] = reinterpret_cast<uint8_t *>(static_cast<uintptr_t>(*v));  // gid for now
        }
        // Pointers are materialised after all inserts of the batch (states_ may have grown).
        for (int64_t r = 0; r < rows; ++r) {
            rowStates_[r] = states_.data() + static_cast<size_t>(reinterpret_cast<uintptr_t>(rowStates_[r])) * blob_;
        }
        for (size_t i = 0; i < aggs_.size(); ++i) UpdateStates(aggs_[i], &vals[i], rowStates_.data(), rows);
    }

    int64_t NumGroups() const { return numGroups_; }

    // Output<> / TraverseHashmapToGetOneResult: group_aggregation.cpp:1486-1538, 892-955 -- all groups in
    // slot order (NULL group last).  outs = keys then aggregates.
    bool Output(OutCol *outs)
    {
        std::vector<uint32_t> order;
        order.reserve(numGroups_);
        auto visit = [&](const auto &, uint32_t gid, bool) { order.push_back(gid); };
        if (m16_) m16_->ForEachInSlotOrder(visit);
        if (m32_) m32_->ForEachInSlotOrder(visit);
        if (m64_) m64_->ForEachInSlotOrder(visit);
        if (ms_) ms_->ForEachInSlotOrder(visit);
        int nk = spec_.num_keys;
        for (int c = 0; c < nk + static_cast<int>(aggs_.size()); ++c) {
            std::memset(outs[c].nulls, 0, (order.size() + 7) / 8);
        }
        bool ok = true;
        for (size_t r = 0; r < order.size(); ++r) {
            uint32_t gid = order[r];
            for (int k = 0; k < nk; ++k) {
                int w = TypeWidth(spec_.key_types[k]);
                std::memcpy(static_cast<uint8_t *>(outs[k].values) + w * r, &keyVals_[k][static_cast<size_t>(gid) * w], w);
                if (keyNull_[k][gid]) SetNullBit(outs[k], static_cast<int64_t>(r));
            }
            const uint8_t *blob = states_.data() + static_cast<size_t>(gid) * blob_;
            for (size_t a = 0; a < aggs_.size(); ++a) {
                ok &= ExtractState(aggs_[a], blob + aggs_[a].offset, outs[nk + a], static_cast<int64_t>(r),
                                   spec_.overflow_as_null != 0);
            }
        }
        return ok;
    }

private:
    uint32_t *EmplaceRow(const HostColumn *keys, int64_t r, bool &inserted)
    {
        switch (handle_) {
            // GroupbySingleFixHandler::InsertValueToHashmap: column_marshaller.h:225-234.
            case HandleType::fixedInt16:
                if (IsNull(keys[0], r)) return m16_->EmplaceNull(inserted);
                return m16_->Emplace(static_cast<const int16_t *>(keys[0].values)[r], inserted);
            case HandleType::fixedInt32:
                if (IsNull(keys[0], r)) return m32_->EmplaceNull(inserted);
                return m32_->Emplace(static_cast<const int32_t *>(keys[0].values)[r], inserted);
            case HandleType::fixedInt64:
                if (IsNull(keys[0], r)) return m64_->EmplaceNull(inserted);
                return m64_->Emplace(static_cast<const int64_t *>(keys[0].values)[r], inserted);
            default: {
                // ColumnSerializeHandler::InsertValueToHashmap: column_marshaller.h:36-46.
                ByteKey k;
                for (int c = 0; c < spec_.num_keys; ++c) SerializeColumnValue(keys[c], r, k.bytes);
                return ms_->Emplace(k, inserted);
            }
        }
    }

    void StoreKeys(const HostColumn *keys, int64_t r, uint32_t gid)
    {
        if (keyVals_.empty()) {
            keyVals_.resize(spec_.num_keys);
            keyNull_.resize(spec_.num_keys);
        }
        for (int k = 0; k < spec_.num_keys; ++k) {
            int w = TypeWidth(keys[k].type_id);
            auto &kv = keyVals_[k];
            kv.resize(static_cast<size_t>(gid + 1) * w);
            bool isNull = IsNull(keys[k], r);
            if (isNull) {
                std::memset(&kv[static_cast<size_t>(gid) * w], 0, w);
            } else {
                std::memcpy(&kv[static_cast<size_t>(gid) * w], static_cast<const uint8_t *>(keys[k].values) + w * r, w);
            }
            keyNull_[k].push_back(isNull ? 1 : 0);
        }
    }

    omni_gpu_agg_spec spec_;
    HandleType handle_;
    std::vector<AggDesc> aggs_;
    size_t blob_ = 1;
    std::unique_ptr<SwissMap<int16_t, GroupbyCrcHasher<int16_t>>> m16_;
    std::unique_ptr<SwissMap<int32_t, GroupbyCrcHasher<int32_t>>> m32_;
    std::unique_ptr<SwissMap<int64_t, GroupbyCrcHasher<int64_t>>> m64_;
    std::unique_ptr<SwissMap<ByteKey, ByteKeyHasher>> ms_;
    std::vector<uint8_t> states_;
    std::vector<uint8_t *> rowStates_;
    std::vector<std::vector<uint8_t>> keyVals_;
    std::vector<std::vector<uint8_t>> keyNull_;
    int64_t numGroups_ = 0;
};

// ================================================================================================
// Shuffle hash partitioning: HashUtil::ComputePartitionIds (core/src/operator/hash_util.cpp:40-82) with the
// Murmur3 steps of core/src/operator/util/mm3_util.h (MixK1, MixH1, Fmix, HashInt, HashLong, Pmod).
// ================================================================================================

inline uint32_t Mm3Rotl(uint32_t i, uint32_t d) { return (i << d) | (i >> (32 - d)); }
inline uint32_t Mm3MixK1(uint32_t k1) { return Mm3Rotl(k1 * 0xcc9e2d51u, 15) * 0x1b873593u; }
inline uint32_t Mm3MixH1(uint32_t h1, uint32_t k1) { return Mm3Rotl(h1 ^ k1, 13) * 5u + 0xe6546b64u; }
inline uint32_t Mm3Fmix(uint32_t h1, uint32_t length)
{
    h1 ^= length;
    h1 ^= h1 >> 16;
    h1 *= 0x85ebca6bu;
    h1 ^= h1 >> 13;
    h1 *= 0xc2b2ae35u;
    return h1 ^ (h1 >> 16);
}
inline uint32_t Mm3HashInt(uint32_t v, uint32_t seed) { return Mm3Fmix(Mm3MixH1(seed, Mm3MixK1(v)), 4); }
inline uint32_t Mm3HashLong(uint64_t v, uint32_t seed)
{
    uint32_t h1 = Mm3MixH1(seed, Mm3MixK1(static_cast<uint32_t>(v)));
    h1 = Mm3MixH1(h1, Mm3MixK1(static_cast<uint32_t>(v >> 32)));
    return Mm3Fmix(h1, 8);
}
inline int32_t Mm3Pmod(int32_t x, int32_t y)
{
    int32_t r = x % y;
    return r < 0 ? r + y : r;
}

// Columns: INT / DATE32 (HashInt of the value), LONG / TIMESTAMP / DECIMAL64 / DOUBLE (HashLong of the 8 raw
// bytes), BOOLEAN (HashInt of 0 or 1).  A NULL leaves the running hash unchanged; the seed is 42
// (MM3HASH_SEED, hash_util.h:56).
void PartitionIds(const HostColumn *cols, int32_t ncols, int64_t rows, int32_t partitions, int32_t *out)
{
    for (int64_t r = 0; r < rows; ++r) {
        uint32_t h = 42;
        for (int32_t c = 0; c < ncols; ++c) {
            const HostColumn &col = cols[c];
            if (IsNull(col, r)) continue;
            switch (col.type_id) {
                case OMNI_TYPE_BOOLEAN:
                    h = Mm3HashInt(static_cast<const uint8_t *>(col.values)[r] != 0 ? 1u : 0u, h);
                    break;
                case OMNI_TYPE_INT:
                case OMNI_TYPE_DATE32:
                    h = Mm3HashInt(static_cast<const uint32_t *>(col.values)[r], h);
                    break;
                default:
                    h = Mm3HashLong(static_cast<const uint64_t *>(col.values)[r], h);
                    break;
            }
        }
        out[r] = Mm3Pmod(static_cast<int32_t>(h), partitions);
    }
}

// ================================================================================================
// Inner hash join: HashBuilderOperator + JoinHashTableVariants + LookupJoinOperator
// ================================================================================================

// Join keys other than one fixed-width column: the key the reference's table compares, as an int64 join key.
// HashBuilderOperatorFactory::InitVariant (hash_builder.cpp:44-93) with GetTypeLength (:96-119): several SHORT (16
// bits) / INT (32) / LONG, DECIMAL64 (64) columns of at most 32, 64 or 128 bits in total are packed by
// ComputeMultiColKey (join_hash_table_variants.cpp:105-146) into int32 / int64 / int128; a single DECIMAL128 column
// is a Decimal128 key; other combinations are serialised (a StringRef key of the values' bytes).  A row with a NULL
// in any key column has no key (EmplaceMultiKey..., :340-400).  int32 / int64 keys are the join key itself; wider
// keys are numbered by a dictionary of the build side's keys, and a probe key outside it has no key.
class JoinKeyCodec {
public:
    // GetTypeLength (hash_builder.cpp:96-119): the bits a key column takes in a packed key, 0 when the type
    // makes the key serialised.
    static int PackBits(int32_t t)
    {
        switch (t) {
            case OMNI_TYPE_BYTE: return 8;
            case OMNI_TYPE_SHORT: return 16;
            case OMNI_TYPE_INT: return 32;
            case OMNI_TYPE_LONG:
            case OMNI_TYPE_DECIMAL64: return 64;
            default: return 0;
        }
    }

    explicit JoinKeyCodec(const std::vector<int32_t> &types) : types_(types)
    {
        const size_t n = types.size();
        if (n == 1) {
            kind_ = Kind::bytes;  // Decimal128: equality of both words
            return;
        }
        int bits = 0;
        bool fixedInt = true;
        for (int32_t t : types) {
            const int b = PackBits(t);
            if (b == 0) fixedInt = false;
            bits += b;
        }
        if (fixedInt && bits <= 32) kind_ = Kind::int32;
        else if (fixedInt && bits <= 64) kind_ = Kind::int64;
        else if (fixedInt && bits <= 128) kind_ = Kind::int128;
        else kind_ = Kind::bytes;
    }

    // The join key of row r; false when the row has no key (a NULL column, or a probe key not in the build side).
    bool Encode(const HostColumn *cols, int64_t r, bool insert, int64_t &key)
    {
        for (size_t c = 0; c < types_.size(); ++c) {
            if (IsNull(cols[c], r)) return false;
        }
        std::string bytes;
        switch (kind_) {
            case Kind::int32: {  // key = key << bits | (int32_t)value
                int32_t k = 0;
                for (size_t c = 0; c < types_.size(); ++c) {
                    const int shift = PackBits(types_[c]);
                    k = static_cast<int32_t>((static_cast<uint32_t>(k) << shift) |
                                             static_cast<uint32_t>(static_cast<int32_t>(LoadInt(cols[c], r))));
                }
                key = k;
                return true;
            }
            case Kind::int64: {
                int64_t k = 0;
                for (size_t c = 0; c < types_.size(); ++c) {
                    const int shift = PackBits(types_[c]);
                    k = static_cast<int64_t>((static_cast<uint64_t>(k) << shift) |
                                             static_cast<uint64_t>(LoadInt(cols[c], r)));
                }
                key = k;
                return true;
            }
            case Kind::int128: {
                int128 k = 0;
                for (size_t c = 0; c < types_.size(); ++c) {
                    const int shift = PackBits(types_[c]);
                    k = static_cast<int128>((static_cast<unsigned __int128>(k) << shift) |
                                            static_cast<unsigned __int128>(static_cast<int128>(LoadInt(cols[c], r))));
                }
                bytes.assign(reinterpret_cast<const char *>(&k), sizeof(k));
                break;
            }
            case Kind::bytes:
                for (size_t c = 0; c < types_.size(); ++c) {
                    const int w = TypeWidth(types_[c]);
                    bytes.append(static_cast<const char *>(cols[c].values) + static_cast<int64_t>(w) * r, w);
                }
                break;
        }
        auto it = dict_.find(bytes);
        if (it != dict_.end()) {
            key = it->second;
            return true;
        }
        if (!insert) return false;
        key = static_cast<int64_t>(dict_.size());
        dict_.emplace(std::move(bytes), key);
        return true;
    }

private:
    enum class Kind { int32, int64, int128, bytes };
    std::vector<int32_t> types_;
    Kind kind_ = Kind::bytes;
    std::unordered_map<std::string, int64_t> dict_;
};

class HashJoin {
public:
    explicit HashJoin(int32_t keyType) : keyType_(keyType) {}
    explicit HashJoin(const std::vector<int32_t> &keyTypes) : keyType_(OMNI_TYPE_LONG), codec_(new JoinKeyCodec(keyTypes))
    {
    }

    // The key columns of a batch (several columns or one DECIMAL128) as one int64 key column with its bitmap.
    void Translate(const HostColumn *cols, int64_t rows, bool insert, int64_t *out, uint8_t *bitmap)
    {
        std::memset(bitmap, 0, static_cast<size_t>(rows + 7) / 8);
        for (int64_t r = 0; r < rows; ++r) {
            int64_t k = 0;
            if (!codec_->Encode(cols, r, insert, k)) bitmap[r >> 3] |= static_cast<uint8_t>(1u << (r & 7));
            out[r] = k;
        }
    }

    // HashBuilderOperator::AddInput retains the batch (hash_builder.cpp:169-180).
    void AddBatch(const HostColumn &key, const HostColumn *payloads, int32_t npay, int64_t rows)
    {
        if (payTypes_.empty()) {
            for (int p = 0; p < npay; ++p) payTypes_.push_back(payloads[p].type_id);
            pay_.resize(npay);
            payNull_.resize(npay);
        }
        for (int64_t r = 0; r < rows; ++r) {
            keyNull_.push_back(IsNull(key, r) ? 1 : 0);
            keys_.push_back(IsNull(key, r) ? 0 : LoadInt(key, r));
        }
        for (int p = 0; p < npay; ++p) {
            int w = TypeWidth(payTypes_[p]);
            const uint8_t *src = static_cast<const uint8_t *>(payloads[p].values);
            pay_[p].insert(pay_[p].end(), src, src + w * rows);
            for (int64_t r = 0; r < rows; ++r) payNull_[p].push_back(IsNull(payloads[p], r) ? 1 : 0);
        }
    }

    // BuildHashTable: join_hash_table_variants.cpp:695-746; TryToBuildArrayTable :748-808;
    // EmplaceSingleNotNullKeyToNormalHashTable :254-283 (NULL keys are skipped).
    void Finish()
    {
        int64_t n = static_cast<int64_t>(keys_.size());
        uint8_t degree = static_cast<uint8_t>(std::ceil(std::log2(static_cast<double>(n) / 0.75)));
        if (n == 0) degree = 0;
        int64_t len = static_cast<int64_t>(std::pow(2.0, degree));
        next_.assign(n, -1);
        // Array table decision (single integer key): range must be <= 8 * 2^initDegree and <= UINT32_MAX.
        int64_t mn = std::numeric_limits<int64_t>::max(), mx = std::numeric_limits<int64_t>::min();
        for (int64_t i = 0; i < n; ++i) {
            if (keyNull_[i]) continue;
            mn = std::min(mn, keys_[i]);
            mx = std::max(mx, keys_[i]);
        }
        // Only INT / DATE32 / SHORT / LONG / TIMESTAMP keys try it (join_hash_table_variants.cpp:700-736), and
        // max - min must not overflow the key type (TryToBuildArrayTable :792-794).
        const bool candidate = !codec_ && (keyType_ == OMNI_TYPE_INT || keyType_ == OMNI_TYPE_DATE32 ||
                               keyType_ == OMNI_TYPE_SHORT || keyType_ == OMNI_TYPE_LONG ||
                               keyType_ == OMNI_TYPE_TIMESTAMP);
        bool arrayOk = candidate && mx >= mn;
        if (arrayOk) {
            int128 range = static_cast<int128>(mx) - mn;
            const int128 typeMax = keyType_ == OMNI_TYPE_SHORT ? 32767
                                   : (keyType_ == OMNI_TYPE_INT || keyType_ == OMNI_TYPE_DATE32)
                                       ? 2147483647
                                       : static_cast<int128>(std::numeric_limits<int64_t>::max());
            arrayOk = range <= typeMax && range <= 0xFFFFFFFFLL && range <= static_cast<int128>(8) * len;
        }
        if (arrayOk) {
            isArray_ = true;
            min_ = mn;
            heads_.assign(static_cast<size_t>(mx - mn + 1), -1);
            tails_.assign(heads_.size(), -1);
            for (int64_t i = 0; i < n; ++i) {
                if (keyNull_[i]) continue;
                Link(heads_[keys_[i] - mn], tails_[keys_[i] - mn], i);
            }
            return;
        }
        if (degree < 5) degree = 5;
        map_.reset(new SwissMap<int64_t, JoinIntHasher<int64_t>>(degree));
        for (int64_t i = 0; i < n; ++i) {
            if (keyNull_[i]) continue;
            bool inserted = false;
            uint32_t *v = map_->EmplaceNoRehash(keys_[i], static_cast<uint64_t>(keys_[i]), inserted);
            if (inserted) {
                *v = static_cast<uint32_t>(listHeads_.size());
                listHeads_.push_back(-1);
                listTails_.push_back(-1);
            }
            Link(listHeads_[*v], listTails_[*v], i);
        }
    }

    bool IsArray() const { return isArray_; }

    // ProbeBatchForInnerJoin: lookup_join.cpp:924-1035 (NULL probe keys skipped, each match appended).
    int64_t Probe(const HostColumn &key, int64_t rows, int64_t rowBase, int64_t *outProbe, int64_t *outBuild,
                  int64_t cap)
    {
        int64_t m = 0;
        for (int64_t r = 0; r < rows; ++r) {
            if (IsNull(key, r)) continue;
            int64_t k = LoadInt(key, r);
            int64_t head = -1;
            if (isArray_) {
                int128 idx = static_cast<int128>(k) - min_;
                if (idx >= 0 && idx < static_cast<int128>(heads_.size())) head = heads_[static_cast<size_t>(idx)];
            } else if (map_) {
                const uint32_t *v = map_->Find(k);
                if (v) head = listHeads_[*v];
            }
            for (int64_t b = head; b >= 0; b = next_[b]) {
                if (m < cap) {
                    outProbe[m] = rowBase + r;
                    outBuild[m] = b;
                }
                ++m;
            }
        }
        return m;
    }

    // Join types other than inner, restated from the per-type probe loops (no join filter, single table):
    //   mode 1 ProbeBatchForOppositeSideOuterJoin / ProbeBatchForFullJoin (lookup_join.cpp:1037-1108, 1178-1240):
    //          NULL key -> one row without a build row; each match; no match -> one row without a build row;
    //   mode 2 ProbeBatchForLeftSemiJoin (:1242-1296): one row at the first match, NULL keys skipped;
    //   mode 3 ProbeBatchForLeftAntiJoin (:1298-1358): one row when nothing matches, NULL keys included;
    //   mode 0 ProbeBatchForSameSideOuterJoin (:1110-1176): each match, NULL keys skipped.
    // outBuild = -1 marks a row without a build row (its build columns are NULL).  markVisited records each
    // matched build row (RowRefListWithFlags::visited via PositionVisited, join_hash_table_variants.h:129-137).
    int64_t ProbeTyped(const HostColumn &key, int64_t rows, int64_t rowBase, int32_t mode, bool markVisited,
                       int64_t *outProbe, int64_t *outBuild, int64_t cap)
    {
        if (markVisited && visited_.size() != keys_.size()) visited_.assign(keys_.size(), 0);
        int64_t m = 0;
        auto emit = [&](int64_t probe, int64_t build) {
            if (m < cap) {
                outProbe[m] = probe;
                outBuild[m] = build;
            }
            ++m;
        };
        for (int64_t r = 0; r < rows; ++r) {
            const int64_t pr = rowBase + r;
            if (IsNull(key, r)) {
                if (mode == 1 || mode == 3) emit(pr, -1);
                continue;
            }
            const int64_t head = Lookup(LoadInt(key, r));
            if (mode == 2) {
                if (head >= 0) emit(pr, -1);
                continue;
            }
            if (mode == 3) {
                if (head < 0) emit(pr, -1);
                continue;
            }
            for (int64_t b = head; b >= 0; b = next_[b]) {
                emit(pr, b);
                if (markVisited) visited_[b] = 1;
            }
            if (mode == 1 && head < 0) emit(pr, -1);
        }
        return m;
    }

    // ProbeBatchForExistenceJoin (lookup_join.cpp:1393-1438): AppendExistenceRow<matched> per probe row.
    void Exists(const HostColumn &key, int64_t rows, uint8_t *out) const
    {
        for (int64_t r = 0; r < rows; ++r) out[r] = (!IsNull(key, r) && Lookup(LoadInt(key, r)) >= 0) ? 1 : 0;
    }

    // LookupOuterJoinOperator::GetOutput (lookup_outer_join.cpp:155-320, 290-325): the build rows never visited,
    // including the NULL-key rows, which the build-preserving types keep in the null-key slot
    // (join_hash_table_variants.cpp:80-82, 286-330) where no probe reaches them.  Build-row order.
    int64_t Outer(int64_t *outBuild, int64_t cap)
    {
        if (visited_.size() != keys_.size()) visited_.assign(keys_.size(), 0);
        int64_t m = 0;
        for (size_t b = 0; b < keys_.size(); ++b) {
            if (visited_[b]) continue;
            if (m < cap) outBuild[m] = static_cast<int64_t>(b);
            ++m;
        }
        return m;
    }

    // ConstructBuildColumn gather (lookup_join.cpp:1667-1779); build row -1 gathers a NULL.
    void GatherPayload(int32_t p, const int64_t *buildRows, int64_t m, void *out, uint8_t *nulls) const
    {
        int w = TypeWidth(payTypes_[p]);
        std::memset(nulls, 0, (m + 7) / 8);
        for (int64_t i = 0; i < m; ++i) {
            if (buildRows[i] < 0) {
                std::memset(static_cast<uint8_t *>(out) + w * i, 0, w);
                nulls[i >> 3] |= static_cast<uint8_t>(1u << (i & 7));
                continue;
            }
            std::memcpy(static_cast<uint8_t *>(out) + w * i, &pay_[p][static_cast<size_t>(buildRows[i]) * w], w);
            if (payNull_[p][buildRows[i]]) nulls[i >> 3] |= static_cast<uint8_t>(1u << (i & 7));
        }
    }

private:
    // First build row of a non-NULL key, or -1 (array table or Swiss table, as Finish chose).
    int64_t Lookup(int64_t k) const
    {
        if (isArray_) {
            int128 idx = static_cast<int128>(k) - min_;
            if (idx >= 0 && idx < static_cast<int128>(heads_.size())) return heads_[static_cast<size_t>(idx)];
            return -1;
        }
        if (map_) {
            const uint32_t *v = map_->Find(k);
            if (v) return listHeads_[*v];
        }
        return -1;
    }

    void Link(int64_t &head, int64_t &tail, int64_t row)
    {
        if (head < 0) head = row;
        else next_[tail] = row;
        tail = row;
    }

    int32_t keyType_;
    std::unique_ptr<JoinKeyCodec> codec_;
    std::vector<int64_t> keys_;
    std::vector<uint8_t> keyNull_;
    std::vector<int32_t> payTypes_;
    std::vector<std::vector<uint8_t>> pay_;
    std::vector<std::vector<uint8_t>> payNull_;
    bool isArray_ = false;
    int64_t min_ = 0;
    std::vector<int64_t> heads_, tails_, next_;
    std::unique_ptr<SwissMap<int64_t, JoinIntHasher<int64_t>>> map_;
    std::vector<uint8_t> visited_;
    std::vector<int64_t> listHeads_, listTails_;
};

}  // namespace

// ================================================================================================
// C ABI for the Python tests and the CPU baseline (host pointers only)
// ================================================================================================
extern "C" {

typedef struct oracle_column {
    int32_t type_id;
    int32_t scale;
    const void *values;
    const uint8_t *nulls;
} oracle_column;

const char *oracle_last_error(void) { return g_err.c_str(); }

uint64_t oracle_hash_uint64(uint64_t x) { return HashUint64(x); }
uint32_t oracle_crc32c_step_software(uint32_t crc, uint64_t x) { return Crc32cStepSoftware(crc, x); }
uint64_t oracle_crc_hash_int(int64_t v, int32_t type_id)
{
    switch (type_id) {
        case OMNI_TYPE_SHORT: return CRC32HasherForInt<int16_t>(static_cast<int16_t>(v));
        case OMNI_TYPE_INT:
        case OMNI_TYPE_DATE32: return CRC32HasherForInt<int32_t>(static_cast<int32_t>(v));
        default: return CRC32HasherForInt<int64_t>(v);
    }
}
uint64_t oracle_crc_hash_decimal128(uint64_t lo, uint64_t hi) { return HashCrcDecimal128(lo, hi); }
uint32_t oracle_crc32c_extend(uint32_t crc, const uint8_t *data, size_t n) { return Crc32cExtend(crc, data, n); }
uint64_t oracle_join_hash_int(int64_t v) { return JoinIntHasher<int64_t>()(v); }
int64_t oracle_hash_value_i64(int64_t v) { return HashValueI64(v); }
int64_t oracle_hash_value_double(double v)  // HashValue(double) via DoubleToLongBits: hash_util.h:84-87, 270-279
{
    uint64_t bits;
    std::memcpy(&bits, &v, 8);
    if ((bits & 0x7FF0000000000000ULL) == 0x7FF0000000000000ULL && (bits & 0x000FFFFFFFFFFFFFULL) != 0) {
        bits = 0x7FF8000000000000ULL;
    }
    return HashValueI64(static_cast<int64_t>(bits));
}
int64_t oracle_hash_value_decimal128(int64_t low, int64_t high)  // hash_util.h:99-103
{
    auto single = [](uint64_t v) { return XxFinal(XxTail64(P5 + 8, v)); };
    return static_cast<int64_t>(single(static_cast<uint64_t>(low)) ^
                                single(static_cast<uint64_t>(high) & 0x7FFFFFFFFFFFFFFFULL));
}
int64_t oracle_xxhash64(const uint8_t *data, int32_t len) { return XxHash64(data, len); }
uint32_t oracle_raw_hash_partition(int64_t h, uint32_t mask) { return RawHashPartition(h, mask); }
uint32_t oracle_raw_hash_position(int64_t h, uint64_t mask) { return RawHashPosition(h, mask); }

int32_t oracle_serialize_row(const oracle_column *cols, int32_t ncols, int64_t row, uint8_t *out, int32_t cap)
{
    std::string s;
    for (int c = 0; c < ncols; ++c) {
        HostColumn hc{cols[c].type_id, cols[c].scale, cols[c].values, cols[c].nulls};
        SerializeColumnValue(hc, row, s);
    }
    if (static_cast<int32_t>(s.size()) <= cap) std::memcpy(out, s.data(), s.size());
    return static_cast<int32_t>(s.size());
}

// --- BaseHashMap<int64> probe used by the hash-map tests ---------------------------------------------
void *oracle_map_create(int32_t degree) { return new SwissMap<int64_t, GroupbyCrcHasher<int64_t>>(static_cast<uint8_t>(degree)); }
int32_t oracle_map_emplace(void *h, int64_t key, int32_t is_null, uint32_t value)
{
    auto *m = static_cast<SwissMap<int64_t, GroupbyCrcHasher<int64_t>> *>(h);
    bool inserted = false;
    uint32_t *v = is_null ? m->EmplaceNull(inserted) : m->Emplace(key, inserted);
    if (inserted) *v = value;
    return inserted ? 1 : 0;
}
int64_t oracle_map_size(void *h) { return static_cast<int64_t>(static_cast<SwissMap<int64_t, GroupbyCrcHasher<int64_t>> *>(h)->Size()); }
int64_t oracle_map_capacity(void *h) { return static_cast<int64_t>(static_cast<SwissMap<int64_t, GroupbyCrcHasher<int64_t>> *>(h)->Capacity()); }
int64_t oracle_map_values_in_order(void *h, uint32_t *out)
{
    int64_t n = 0;
    static_cast<SwissMap<int64_t, GroupbyCrcHasher<int64_t>> *>(h)->ForEachInSlotOrder(
        [&](int64_t, uint32_t v, bool) { out[n++] = v; });
    return n;
}
void oracle_map_destroy(void *h) { delete static_cast<SwissMap<int64_t, GroupbyCrcHasher<int64_t>> *>(h); }

// --- aggregation --------------------------------------------------------------------------------------
void *oracle_agg_create(const omni_gpu_agg_spec *spec) { return new HashAgg(*spec); }
int32_t oracle_agg_add_batch(void *h, const oracle_column *keys, const oracle_column *aggs, int64_t rows)
{
    auto *a = static_cast<HashAgg *>(h);
    std::vector<HostColumn> k, v;
    for (int i = 0; i < OMNI_GPU_MAX_KEYS; ++i) k.push_back({keys[i].type_id, keys[i].scale, keys[i].values, keys[i].nulls});
    for (int i = 0; i < OMNI_GPU_MAX_AGGS; ++i) v.push_back({aggs[i].type_id, aggs[i].scale, aggs[i].values, aggs[i].nulls});
    a->AddBatch(k.data(), v.data(), rows);
    return 0;
}
int64_t oracle_agg_num_groups(void *h) { return static_cast<HashAgg *>(h)->NumGroups(); }
int32_t oracle_agg_get_output(void *h, omni_gpu_out_column *outs)
{
    bool ok = static_cast<HashAgg *>(h)->Output(reinterpret_cast<OutCol *>(outs));
    if (!ok) {
        g_err = "decimal overflow";
        return OMNI_GPU_ERR_OVERFLOW;
    }
    return 0;
}
void oracle_agg_destroy(void *h) { delete static_cast<HashAgg *>(h); }

// A stream of independent group-bys: for each VectorBatch b (rows [offsets[b], offsets[b + 1]) of the columns)
// one HashAggregationOperator takes the batch (AddInput) and emits its groups (GetOutput), as an engine runs one
// operator per batch (group_aggregation.cpp:320-406, 454-474).  `threads` host threads take batches from a shared
// counter (threads = 1: the reference's one operator per thread).  Writes each batch's group count; returns the
// total, or -1 on a decimal overflow that raises.
int64_t oracle_agg_stream(const omni_gpu_agg_spec *spec, const oracle_column *keys, const oracle_column *aggs,
                          const int64_t *offsets, int32_t nb, int32_t threads, int64_t *groups_per_batch)
{
    std::atomic<int32_t> next{0};
    std::atomic<bool> failed{false};
    auto shifted = [](const oracle_column &c, int64_t start, int64_t rows, std::vector<uint8_t> &bits) {
        HostColumn h{c.type_id, c.scale, nullptr, nullptr};
        if (c.values) h.values = static_cast<const uint8_t *>(c.values) + start * TypeWidth(c.type_id);
        if (c.nulls) {
            bits.assign(static_cast<size_t>(rows + 7) / 8 + 1, 0);
            for (int64_t r = 0; r < rows; ++r) {
                if ((c.nulls[(start + r) >> 3] >> ((start + r) & 7)) & 1) bits[r >> 3] |= 1u << (r & 7);
            }
            h.nulls = bits.data();
        }
        return h;
    };
    auto work = [&]() {
        std::vector<std::vector<uint8_t>> kbits(OMNI_GPU_MAX_KEYS), abits(OMNI_GPU_MAX_AGGS);
        std::vector<std::vector<uint8_t>> vals, nulls;
        while (true) {
            const int32_t b = next.fetch_add(1);
            if (b >= nb || failed.load()) break;
            const int64_t start = offsets[b], rows = offsets[b + 1] - offsets[b];
            HostColumn k[OMNI_GPU_MAX_KEYS], v[OMNI_GPU_MAX_AGGS];
            for (int i = 0; i < OMNI_GPU_MAX_KEYS; ++i) k[i] = shifted(keys[i], start, rows, kbits[i]);
            for (int i = 0; i < OMNI_GPU_MAX_AGGS; ++i) v[i] = shifted(aggs[i], start, rows, abits[i]);
            HashAgg a(*spec);
            a.AddBatch(k, v, rows);
            const int64_t g = a.NumGroups();
            const int ncols = spec->num_keys + spec->num_aggs;
            vals.resize(ncols);
            nulls.resize(ncols);
            OutCol outs[OMNI_GPU_MAX_KEYS + OMNI_GPU_MAX_AGGS];
            for (int c = 0; c < ncols; ++c) {
                const int32_t t = c < spec->num_keys ? spec->key_types[c] : spec->agg_out_types[c - spec->num_keys];
                vals[c].resize(static_cast<size_t>(std::max<int64_t>(1, g)) * TypeWidth(t));
                nulls[c].assign(static_cast<size_t>(g + 7) / 8 + 1, 0);
                outs[c] = OutCol{t, 0, vals[c].data(), nulls[c].data()};
            }
            if (!a.Output(outs)) failed.store(true);
            groups_per_batch[b] = g;
        }
    };
    const int nt = std::max(1, threads);
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
    if (failed.load()) {
        g_err = "decimal overflow";
        return -1;
    }
    int64_t total = 0;
    for (int32_t b = 0; b < nb; ++b) total += groups_per_batch[b];
    return total;
}

// --- join -----------------------------------------------------------------------------------------------
void *oracle_join_create(int32_t key_type) { return new HashJoin(key_type); }
int32_t oracle_join_add_batch(void *h, const oracle_column *key, const oracle_column *payloads, int32_t npay,
                              int64_t rows)
{
    HostColumn k{key->type_id, key->scale, key->values, key->nulls};
    std::vector<HostColumn> p;
    for (int i = 0; i < npay; ++i) p.push_back({payloads[i].type_id, payloads[i].scale, payloads[i].values, payloads[i].nulls});
    static_cast<HashJoin *>(h)->AddBatch(k, p.data(), npay, rows);
    return 0;
}
void *oracle_join_create_multi(const int32_t *types, int32_t n)
{
    return new HashJoin(std::vector<int32_t>(types, types + n));
}
// Key columns keys[0..n) of `rows` rows into int64 join keys and their bitmap (insert: the build side numbers its
// wide keys; the probe side only looks them up).
void oracle_join_translate(void *h, const oracle_column *keys, int32_t n, int64_t rows, int32_t insert, int64_t *out,
                           uint8_t *bitmap)
{
    std::vector<HostColumn> k;
    for (int i = 0; i < n; ++i) k.push_back({keys[i].type_id, keys[i].scale, keys[i].values, keys[i].nulls});
    static_cast<HashJoin *>(h)->Translate(k.data(), rows, insert != 0, out, bitmap);
}
int32_t oracle_join_finish(void *h)
{
    static_cast<HashJoin *>(h)->Finish();
    return 0;
}
int32_t oracle_join_is_array(void *h) { return static_cast<HashJoin *>(h)->IsArray() ? 1 : 0; }
int64_t oracle_join_probe(void *h, const oracle_column *key, int64_t rows, int64_t row_base, int64_t *out_probe,
                          int64_t *out_build, int64_t cap)
{
    HostColumn k{key->type_id, key->scale, key->values, key->nulls};
    return static_cast<HashJoin *>(h)->Probe(k, rows, row_base, out_probe, out_build, cap);
}
void oracle_join_gather(void *h, int32_t p, const int64_t *build_rows, int64_t m, void *out, uint8_t *nulls)
{
    static_cast<HashJoin *>(h)->GatherPayload(p, build_rows, m, out, nulls);
}
int64_t oracle_join_probe_typed(void *h, const oracle_column *key, int64_t rows, int64_t row_base, int32_t mode,
                                int32_t mark_visited, int64_t *out_probe, int64_t *out_build, int64_t cap)
{
    HostColumn k{key->type_id, key->scale, key->values, key->nulls};
    return static_cast<HashJoin *>(h)->ProbeTyped(k, rows, row_base, mode, mark_visited != 0, out_probe, out_build,
                                                  cap);
}
void oracle_join_exists(void *h, const oracle_column *key, int64_t rows, uint8_t *out)
{
    HostColumn k{key->type_id, key->scale, key->values, key->nulls};
    static_cast<HashJoin *>(h)->Exists(k, rows, out);
}
int64_t oracle_join_outer(void *h, int64_t *out_build, int64_t cap)
{
    return static_cast<HashJoin *>(h)->Outer(out_build, cap);
}
void oracle_join_destroy(void *h) { delete static_cast<HashJoin *>(h); }

// --- shuffle partitioning ------------------------------------------------------------------------------
uint32_t oracle_mm3_hash_int(uint32_t v, uint32_t seed) { return Mm3HashInt(v, seed); }
uint32_t oracle_mm3_hash_long(uint64_t v, uint32_t seed) { return Mm3HashLong(v, seed); }
void oracle_partition_ids(const oracle_column *cols, int32_t ncols, int64_t rows, int32_t partitions, int32_t *out)
{
    std::vector<HostColumn> c;
    for (int i = 0; i < ncols; ++i) c.push_back({cols[i].type_id, cols[i].scale, cols[i].values, cols[i].nulls});
    PartitionIds(c.data(), ncols, rows, partitions, out);
}

}  // extern "C"
