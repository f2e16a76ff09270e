// Operator-level tests of the GPU drop-ins, written like the reference's gtest cases
// (core/test/operator/join_test.cpp, aggregation_test.cpp): build VectorBatches, push them through
// AddInput, pull GetOutput, compare as row sets (VecBatchMatchIgnoreOrder semantics).
// Exit status 0 when every case passes.
#include <algorithm>
#include <array>
#include <climits>
#include <cstdio>
#include <map>
#include <cstring>
#include <functional>
#include <string>
#include <tuple>
#include <vector>

#include "gpu_operators.h"

using namespace omniruntime;
using namespace omniruntime::op;
using namespace omniruntime::type;
using namespace omniruntime::vec;

static int g_failures = 0;
#define EXPECT_TRUE(c)                                                               \
    do {                                                                             \
        if (!(c)) {                                                                  \
            std::printf("  expectation failed: %s (%s:%d)\n", #c, __FILE__, __LINE__); \
            ++g_failures;                                                            \
        }                                                                            \
    } while (0)

// A cell rendered as text ("null" for NULL) so rows of mixed types compare as tuples.
static std::string Cell(const BaseVector *v, int32_t r)
{
    if (v->IsNull(r)) return "null";
    char buf[64];
    switch (v->GetTypeId()) {
        case OMNI_SHORT: std::snprintf(buf, sizeof buf, "%d", static_cast<const Vector<int16_t> *>(v)->GetValue(r)); break;
        case OMNI_INT:
        case OMNI_DATE32: std::snprintf(buf, sizeof buf, "%d", static_cast<const Vector<int32_t> *>(v)->GetValue(r)); break;
        case OMNI_DOUBLE: std::snprintf(buf, sizeof buf, "%.12g", static_cast<const Vector<double> *>(v)->GetValue(r)); break;
        case OMNI_BOOLEAN: std::snprintf(buf, sizeof buf, "%s", static_cast<const Vector<uint8_t> *>(v)->GetValue(r) ? "true" : "false"); break;
        case OMNI_BYTE: std::snprintf(buf, sizeof buf, "%d", static_cast<int>(static_cast<const Vector<int8_t> *>(v)->GetValue(r))); break;
        case OMNI_DECIMAL128: {
            const Decimal128 d = static_cast<const Vector<Decimal128> *>(v)->GetValue(r);
            __int128 x = static_cast<__int128>((static_cast<unsigned __int128>(static_cast<uint64_t>(d.high)) << 64) | d.low);
            std::string out;
            const bool neg = x < 0;
            unsigned __int128 u = neg ? -static_cast<unsigned __int128>(x) : static_cast<unsigned __int128>(x);
            do {
                out.insert(out.begin(), static_cast<char>('0' + static_cast<int>(u % 10)));
                u /= 10;
            } while (u != 0);
            return neg ? "-" + out : out;
        }
        default: std::snprintf(buf, sizeof buf, "%lld", static_cast<long long>(static_cast<const Vector<int64_t> *>(v)->GetValue(r)));
    }
    return buf;
}

using Rows = std::vector<std::vector<std::string>>;

static void AppendRows(Rows &rows, const VectorBatch *b)
{
    for (int32_t r = 0; r < b->GetRowCount(); ++r) {
        std::vector<std::string> row;
        for (int32_t c = 0; c < b->GetVectorCount(); ++c) row.push_back(Cell(b->Get(c), r));
        rows.push_back(row);
    }
}

static bool SameRowSet(Rows a, Rows b)
{
    std::sort(a.begin(), a.end());
    std::sort(b.begin(), b.end());
    return a == b;
}

template <typename T> static Vector<T> *MakeVector(DataTypeId id, const std::vector<T> &vals)
{
    auto *v = new Vector<T>(static_cast<int32_t>(vals.size()), id);
    for (size_t i = 0; i < vals.size(); ++i) v->SetValue(static_cast<int32_t>(i), vals[i]);
    return v;
}

static Rows Drain(Operator *op)
{
    Rows rows;
    for (int guard = 0; guard < 100000; ++guard) {
        VectorBatch *out = nullptr;
        op->GetOutput(&out);
        if (!out) break;
        AppendRows(rows, out);
        VectorHelper::FreeVecBatch(out);
    }
    return rows;
}

// join_test.cpp:283-308 (TestInnerEqualityJoinWithOneBuildOp)
static void TestInnerJoinOneBuildOp()
{
    DataTypes buildTypes({LongType(), LongType()});
    auto *build = new VectorBatch(10);
    build->Append(MakeVector<int64_t>(OMNI_LONG, {1, 2, 1, 2, 3, 4, 5, 6, 7, 1}));
    build->Append(MakeVector<int64_t>(OMNI_LONG, {79, 79, 70, 70, 70, 70, 70, 70, 70, 70}));
    int32_t hashCols[1] = {0};
    auto *bf = GpuHashBuilderOperatorFactory::CreateHashBuilderOperatorFactory(OMNI_JOIN_TYPE_INNER, buildTypes, hashCols, 1, 1);
    Operator *builder = bf->CreateOperator();
    builder->AddInput(build);
    VectorBatch *none = nullptr;
    builder->GetOutput(&none);
    EXPECT_TRUE(none == nullptr && builder->GetStatus() == OMNI_STATUS_FINISHED);

    DataTypes probeTypes({LongType(), LongType()});
    auto *probe = new VectorBatch(10);
    probe->Append(MakeVector<int64_t>(OMNI_LONG, {1, 2, 3, 4, 5, 6, 1, 1, 2, 3}));
    probe->Append(MakeVector<int64_t>(OMNI_LONG, {78, 78, 78, 78, 78, 78, 78, 82, 82, 65}));
    int32_t probeOut[1] = {1}, probeHash[1] = {0}, buildOut[1] = {1};
    DataTypes buildOutTypes({LongType()});
    auto *lf = GpuLookupJoinOperatorFactory::CreateLookupJoinOperatorFactory(
        probeTypes, probeOut, 1, probeHash, 1, buildOut, 1, buildOutTypes, reinterpret_cast<int64_t>(bf), nullptr,
        false, nullptr);
    Operator *lookup = lf->CreateOperator();
    lookup->AddInput(probe);
    Rows got = Drain(lookup);
    std::vector<int64_t> e0 = {78, 78, 78, 78, 78, 78, 78, 82, 82, 65};
    std::vector<int> ids = {0, 0, 0, 1, 1, 2, 3, 4, 5, 6, 6, 6, 7, 7, 7, 8, 8, 9};
    std::vector<int64_t> e1 = {79, 70, 70, 79, 70, 70, 70, 70, 70, 79, 70, 70, 79, 70, 70, 79, 70, 70};
    Rows want;
    for (size_t i = 0; i < ids.size(); ++i) want.push_back({std::to_string(e0[ids[i]]), std::to_string(e1[i])});
    EXPECT_TRUE(got.size() == 18);
    EXPECT_TRUE(SameRowSet(got, want));
    Operator::DeleteOperator(builder);
    Operator::DeleteOperator(lookup);
    delete lf;
    delete bf;
}

// join_test.cpp:4265-4321 (TestInnerEqualityJoinOfArrayTable)
static void TestInnerJoinArrayTable()
{
    DataTypes types({ShortType(), ShortType()});
    auto *build = new VectorBatch(6);
    build->Append(MakeVector<int16_t>(OMNI_SHORT, {1, 2, 3, 4, 1, 2}));
    build->Append(MakeVector<int16_t>(OMNI_SHORT, {111, 11, 333, 33, 20, 16}));
    int32_t hashCols[1] = {0};
    auto *bf = GpuHashBuilderOperatorFactory::CreateHashBuilderOperatorFactory(OMNI_JOIN_TYPE_INNER, types, hashCols, 1, 1);
    Operator *builder = bf->CreateOperator();
    builder->AddInput(build);
    VectorBatch *none = nullptr;
    builder->GetOutput(&none);
    auto *probe = new VectorBatch(4);
    probe->Append(MakeVector<int16_t>(OMNI_SHORT, {1, 2, 3, 4}));
    probe->Append(MakeVector<int16_t>(OMNI_SHORT, {11, 22, 33, 44}));
    int32_t probeOut[2] = {0, 1}, probeHash[1] = {0}, buildOut[2] = {0, 1};
    auto *lf = GpuLookupJoinOperatorFactory::CreateLookupJoinOperatorFactory(
        types, probeOut, 2, probeHash, 1, buildOut, 2, types, reinterpret_cast<int64_t>(bf), nullptr, false, nullptr);
    Operator *lookup = lf->CreateOperator();
    lookup->AddInput(probe);
    Rows got = Drain(lookup);
    Rows want = {{"1", "11", "1", "111"}, {"1", "11", "1", "20"}, {"2", "22", "2", "11"},
                 {"2", "22", "2", "16"},  {"3", "33", "3", "333"}, {"4", "44", "4", "33"}};
    EXPECT_TRUE(SameRowSet(got, want));
    int32_t is_array = 0;
    int64_t dup = 0, slots = 0;
    omni_gpu_join_build_info(bf->Table(), &is_array, &dup, &slots);
    EXPECT_TRUE(is_array == 1 && dup == 2);
    Operator::DeleteOperator(builder);
    Operator::DeleteOperator(lookup);
    delete lf;
    delete bf;
}

static GpuHashAggregationOperatorFactory *StdAggFactory(std::vector<DataTypePtr> groupTypes, uint32_t firstAggCol)
{
    std::vector<uint32_t> groupCols;
    for (uint32_t i = 0; i < groupTypes.size(); ++i) groupCols.push_back(i);
    std::vector<std::vector<uint32_t>> aggCols;
    std::vector<DataTypes> inTypes, outTypes;
    std::vector<uint32_t> funcs = {OMNI_AGG_SUM, OMNI_AGG_AVG, OMNI_AGG_COUNT_COLUMN, OMNI_AGG_MIN, OMNI_AGG_MAX};
    std::vector<DataTypePtr> outs = {LongType(), DoubleType(), LongType(), LongType(), LongType()};
    for (uint32_t i = 0; i < 5; ++i) {
        aggCols.push_back({firstAggCol + i});
        inTypes.push_back(DataTypes({LongType()}));
        outTypes.push_back(DataTypes({outs[i]}));
    }
    std::vector<uint32_t> masks;
    return new GpuHashAggregationOperatorFactory(groupCols, DataTypes(groupTypes), aggCols, inTypes, outTypes, funcs, masks,
                                                 std::vector<bool>(5, true), std::vector<bool>(5, false), false);
}

// aggregation_test.cpp:544-630 (verify_correctness), single stage
static void TestAggVerifyCorrectness()
{
    auto *f = StdAggFactory({LongType(), LongType()}, 2);
    Operator *op = f->CreateOperator();
    for (int b = 0; b < 20; ++b) {
        auto *batch = new VectorBatch(2000);
        std::vector<int64_t> k(2000), one(2000, 1);
        for (int j = 0; j < 2000; ++j) k[j] = j % 10;
        batch->Append(MakeVector<int64_t>(OMNI_LONG, k));
        batch->Append(MakeVector<int64_t>(OMNI_LONG, k));
        for (int a = 0; a < 5; ++a) batch->Append(MakeVector<int64_t>(OMNI_LONG, one));
        op->AddInput(batch);
    }
    EXPECT_TRUE(op->GetHashMapUniqueKeys() == 10);
    Rows got = Drain(op);
    Rows want;
    for (int k = 0; k < 10; ++k) want.push_back({std::to_string(k), std::to_string(k), "4000", "1", "4000", "1", "1"});
    EXPECT_TRUE(SameRowSet(got, want));
    EXPECT_TRUE(op->GetStatus() == OMNI_STATUS_FINISHED);
    Operator::DeleteOperator(op);
    delete f;
}

// Pushes the batches through an operator and returns its output batches (the caller frees them).
static std::vector<VectorBatch *> RunAgg(Operator *op, std::vector<VectorBatch *> in)
{
    for (auto *b : in) op->AddInput(b);
    std::vector<VectorBatch *> out;
    for (int guard = 0; guard < 100000; ++guard) {
        VectorBatch *o = nullptr;
        op->GetOutput(&o);
        if (!o) break;
        out.push_back(o);
    }
    return out;
}

// aggregation_test.cpp:544-630 (verify_correctness) as written there: two partial operators over 10 batches each,
// then a final operator over their outputs.  Partial AVG is (DOUBLE sum, LONG count) as two flat columns.
static void TestAggTwoStageVerifyCorrectness()
{
    std::vector<uint32_t> funcs = {OMNI_AGG_SUM, OMNI_AGG_AVG, OMNI_AGG_COUNT_COLUMN, OMNI_AGG_MIN, OMNI_AGG_MAX};
    std::vector<uint32_t> groupCols = {0, 1}, masks;
    std::vector<std::vector<uint32_t>> pcols = {{2}, {3}, {4}, {5}, {6}};
    std::vector<DataTypes> pin(5, DataTypes({LongType()}));
    std::vector<DataTypes> pout = {DataTypes({LongType()}), DataTypes({DoubleType(), LongType()}), DataTypes({LongType()}),
                                   DataTypes({LongType()}), DataTypes({LongType()})};
    GpuHashAggregationOperatorFactory pf(groupCols, DataTypes({LongType(), LongType()}), pcols, pin, pout, funcs, masks,
                                         std::vector<bool>(5, true), std::vector<bool>(5, true), false);
    std::vector<VectorBatch *> partials;
    for (int opi = 0; opi < 2; ++opi) {
        std::vector<VectorBatch *> in;
        for (int b = 0; b < 10; ++b) {
            auto *batch = new VectorBatch(2000);
            std::vector<int64_t> k(2000), one(2000, 1);
            for (int j = 0; j < 2000; ++j) k[j] = j % 10;
            batch->Append(MakeVector<int64_t>(OMNI_LONG, k));
            batch->Append(MakeVector<int64_t>(OMNI_LONG, k));
            for (int a = 0; a < 5; ++a) batch->Append(MakeVector<int64_t>(OMNI_LONG, one));
            in.push_back(batch);
        }
        Operator *op = pf.CreateOperator();
        auto out = RunAgg(op, in);
        EXPECT_TRUE(out.size() == 1 && out[0]->GetVectorCount() == 8);
        Rows rows;
        for (auto *o : out) AppendRows(rows, o);
        Rows want;
        for (int k = 0; k < 10; ++k) {
            want.push_back({std::to_string(k), std::to_string(k), "2000", "2000", "2000", "2000", "1", "1"});
        }
        EXPECT_TRUE(SameRowSet(rows, want));
        partials.insert(partials.end(), out.begin(), out.end());
        Operator::DeleteOperator(op);
    }
    std::vector<std::vector<uint32_t>> fcols = {{2}, {3, 4}, {5}, {6}, {7}};
    std::vector<DataTypes> fin = {DataTypes({LongType()}), DataTypes({DoubleType(), LongType()}), DataTypes({LongType()}),
                                  DataTypes({LongType()}), DataTypes({LongType()})};
    std::vector<DataTypes> fout = {DataTypes({LongType()}), DataTypes({DoubleType()}), DataTypes({LongType()}),
                                   DataTypes({LongType()}), DataTypes({LongType()})};
    GpuHashAggregationOperatorFactory ff(groupCols, DataTypes({LongType(), LongType()}), fcols, fin, fout, funcs, masks,
                                         std::vector<bool>(5, false), std::vector<bool>(5, false), false);
    Operator *fop = ff.CreateOperator();
    auto out = RunAgg(fop, partials);
    Rows got;
    for (auto *o : out) {
        EXPECT_TRUE(o->GetVectorCount() == 7);
        AppendRows(got, o);
        VectorHelper::FreeVecBatch(o);
    }
    Rows want;
    for (int k = 0; k < 10; ++k) want.push_back({std::to_string(k), std::to_string(k), "4000", "1", "4000", "1", "1"});
    EXPECT_TRUE(SameRowSet(got, want));
    Operator::DeleteOperator(fop);
}

// Two stages equal one stage: SUM, AVG(long), COUNT(col), COUNT(*), MIN, MAX, AVG(double) over seeded rows with NULL
// keys, NULL inputs and one group whose AVG inputs are all NULL (partial (0, 0), final NULL).  Three partial
// operators take a third of the batches each; the final operator merges their outputs.  AVG inputs are integral,
// so every double sum is exact in any order and the comparison is exact.
static void TestAggTwoStageMatchesSingleStage()
{
    const int kRows = 30000, kBatch = 4096;
    uint64_t st = 0x2545F4914F6CDD1DULL;
    auto rnd = [&st]() {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        return st;
    };
    std::vector<int64_t> key(kRows), v1(kRows), v2(kRows);
    std::vector<double> v3(kRows);
    std::vector<bool> nk(kRows), n1(kRows), n2(kRows), n3(kRows);
    for (int i = 0; i < kRows; ++i) {
        key[i] = static_cast<int64_t>(rnd() % 500);
        nk[i] = rnd() % 50 == 0;
        v1[i] = static_cast<int64_t>(rnd() % 2000001) - 1000000;
        n1[i] = rnd() % 20 == 0;
        v2[i] = static_cast<int64_t>(rnd() % 1000);
        n2[i] = rnd() % 10 == 0;
        v3[i] = static_cast<double>(static_cast<int64_t>(rnd() % 2001) - 1000);
        n3[i] = rnd() % 20 == 0;
        if (i % 97 == 0) {  // group 777: every AVG input NULL
            key[i] = 777;
            nk[i] = false;
            n1[i] = n3[i] = true;
        }
    }
    auto makeBatches = [&](int lo, int hi) {
        std::vector<VectorBatch *> in;
        for (int a = lo; a < hi; a += kBatch) {
            int b = std::min(hi, a + kBatch), n = b - a;
            auto *batch = new VectorBatch(n);
            auto *k = MakeVector<int64_t>(OMNI_LONG, std::vector<int64_t>(key.begin() + a, key.begin() + b));
            auto *c1 = MakeVector<int64_t>(OMNI_LONG, std::vector<int64_t>(v1.begin() + a, v1.begin() + b));
            auto *c2 = MakeVector<int64_t>(OMNI_LONG, std::vector<int64_t>(v2.begin() + a, v2.begin() + b));
            auto *c3 = MakeVector<double>(OMNI_DOUBLE, std::vector<double>(v3.begin() + a, v3.begin() + b));
            for (int r = 0; r < n; ++r) {
                if (nk[a + r]) k->SetNull(r);
                if (n1[a + r]) c1->SetNull(r);
                if (n2[a + r]) c2->SetNull(r);
                if (n3[a + r]) c3->SetNull(r);
            }
            batch->Append(k);
            batch->Append(c1);
            batch->Append(c2);
            batch->Append(c3);
            in.push_back(batch);
        }
        return in;
    };
    std::vector<uint32_t> funcs = {OMNI_AGG_SUM, OMNI_AGG_AVG, OMNI_AGG_COUNT_COLUMN, OMNI_AGG_COUNT_ALL,
                                   OMNI_AGG_MIN, OMNI_AGG_MAX, OMNI_AGG_AVG};
    std::vector<uint32_t> groupCols = {0}, masks;
    DataTypes keyTypes({LongType()});
    // raw input columns (COUNT_ALL over raw input takes none)
    std::vector<std::vector<uint32_t>> rcols = {{1}, {1}, {2}, {1}, {2}, {3}};
    std::vector<DataTypes> rin = {DataTypes({LongType()}), DataTypes({LongType()}), DataTypes({LongType()}),
                                  DataTypes({LongType()}), DataTypes({LongType()}), DataTypes({DoubleType()})};
    std::vector<DataTypes> single = {DataTypes({LongType()}), DataTypes({DoubleType()}), DataTypes({LongType()}),
                                     DataTypes({LongType()}), DataTypes({LongType()}), DataTypes({LongType()}),
                                     DataTypes({DoubleType()})};
    std::vector<DataTypes> part = {DataTypes({LongType()}), DataTypes({DoubleType(), LongType()}),
                                   DataTypes({LongType()}), DataTypes({LongType()}), DataTypes({LongType()}),
                                   DataTypes({LongType()}), DataTypes({DoubleType(), LongType()})};
    GpuHashAggregationOperatorFactory sf(groupCols, keyTypes, rcols, rin, single, funcs, masks,
                                         std::vector<bool>(7, true), std::vector<bool>(7, false), false);
    GpuHashAggregationOperatorFactory pf(groupCols, keyTypes, rcols, rin, part, funcs, masks,
                                         std::vector<bool>(7, true), std::vector<bool>(7, true), false);
    // partial output: key, sum, avg (sum, count), count, count(*), min, max, avg (sum, count) = columns 0..9
    std::vector<std::vector<uint32_t>> fcols = {{1}, {2, 3}, {4}, {5}, {6}, {7}, {8, 9}};
    std::vector<DataTypes> fin = {DataTypes({LongType()}), DataTypes({DoubleType(), LongType()}), DataTypes({LongType()}),
                                  DataTypes({LongType()}), DataTypes({LongType()}), DataTypes({LongType()}),
                                  DataTypes({DoubleType(), LongType()})};
    GpuHashAggregationOperatorFactory ff(groupCols, keyTypes, fcols, fin, single, funcs, masks,
                                         std::vector<bool>(7, false), std::vector<bool>(7, false), false);
    Operator *sop = sf.CreateOperator();
    Rows want;
    for (auto *o : RunAgg(sop, makeBatches(0, kRows))) {
        AppendRows(want, o);
        VectorHelper::FreeVecBatch(o);
    }
    Operator::DeleteOperator(sop);
    std::vector<VectorBatch *> partials;
    const int cut[4] = {0, 3 * kBatch, 5 * kBatch, kRows};
    for (int p = 0; p < 3; ++p) {
        Operator *op = pf.CreateOperator();
        auto out = RunAgg(op, makeBatches(cut[p], cut[p + 1]));
        partials.insert(partials.end(), out.begin(), out.end());
        Operator::DeleteOperator(op);
    }
    Operator *fop = ff.CreateOperator();
    Rows got;
    for (auto *o : RunAgg(fop, partials)) {
        AppendRows(got, o);
        VectorHelper::FreeVecBatch(o);
    }
    Operator::DeleteOperator(fop);
    EXPECT_TRUE(want.size() == 502);  // 500 keys, the NULL key, key 777
    EXPECT_TRUE(SameRowSet(got, want));
    bool sawNullAvg = false;
    for (const auto &r : got) sawNullAvg = sawNullAvg || (r[0] == "777" && r[2] == "null" && r[7] == "null");
    EXPECT_TRUE(sawNullAvg);
}

// aggregation_test.cpp:632-740 (NULL keys form their own group), LONG key
static void TestAggNullKeys()
{
    auto *f = StdAggFactory({LongType()}, 1);
    Operator *op = f->CreateOperator();
    std::vector<std::vector<int>> keyNulls = {{0, 11}, {0, 11, 20}};
    for (int b = 0; b < 2; ++b) {
        auto *batch = new VectorBatch(21);
        std::vector<int64_t> k(21), one(21, 1);
        for (int j = 0; j < 21; ++j) k[j] = j % 3;
        auto *kv = MakeVector<int64_t>(OMNI_LONG, k);
        for (int r : keyNulls[b]) kv->SetNull(r);
        batch->Append(kv);
        for (int a = 0; a < 5; ++a) {
            auto *v = MakeVector<int64_t>(OMNI_LONG, one);
            if (a == 1)
                for (int r = 0; r < 5; ++r) v->SetNull(r);
            if (a >= 2) {
                v->SetNull(0);
                v->SetNull(1);
            }
            batch->Append(v);
        }
        op->AddInput(batch);
    }
    Rows got = Drain(op);
    Rows want = {{"null", "5", "1", "3", "1", "1"}, {"0", "12", "1", "12", "1", "1"},
                 {"1", "14", "1", "12", "1", "1"}, {"2", "11", "1", "11", "1", "1"}};
    EXPECT_TRUE(SameRowSet(got, want));
    Operator::DeleteOperator(op);
    delete f;
}

// Output is cut into batches of ceil(1 MiB / row size) rows (operator_util.h:111-115).
static void TestAggOutputBatching()
{
    auto *f = StdAggFactory({LongType()}, 1);
    Operator *op = f->CreateOperator();
    const int n = 100000;
    auto *batch = new VectorBatch(n);
    std::vector<int64_t> k(n), v(n);
    for (int j = 0; j < n; ++j) {
        k[j] = static_cast<int64_t>(j) * 2654435761LL;
        v[j] = j;
    }
    batch->Append(MakeVector<int64_t>(OMNI_LONG, k));
    for (int a = 0; a < 5; ++a) batch->Append(MakeVector<int64_t>(OMNI_LONG, v));
    op->AddInput(batch);
    int batches = 0;
    int64_t rows = 0, sum = 0;
    while (true) {
        VectorBatch *out = nullptr;
        op->GetOutput(&out);
        if (!out) break;
        ++batches;
        EXPECT_TRUE(out->GetRowCount() <= MaxRowCount(48));
        rows += out->GetRowCount();
        for (int r = 0; r < out->GetRowCount(); ++r) sum += static_cast<Vector<int64_t> *>(out->Get(1))->GetValue(r);
        VectorHelper::FreeVecBatch(out);
    }
    EXPECT_TRUE(rows == n);
    EXPECT_TRUE(batches == (n + MaxRowCount(48) - 1) / MaxRowCount(48));
    EXPECT_TRUE(sum == static_cast<int64_t>(n) * (n - 1) / 2);
    Operator::DeleteOperator(op);
    delete f;
}

// A decimal SUM overflow raises OmniException when overflowAsNull is false.
static void TestAggDecimalOverflowThrows()
{
    std::vector<uint32_t> groupCols = {0};
    std::vector<std::vector<uint32_t>> aggCols = {{1}};
    std::vector<DataTypes> inTypes = {DataTypes({Decimal64Type(18, 0)})}, outTypes = {DataTypes({Decimal64Type(18, 0)})};
    std::vector<uint32_t> funcs = {OMNI_AGG_SUM}, masks;
    GpuHashAggregationOperatorFactory f(groupCols, DataTypes({LongType()}), aggCols, inTypes, outTypes, funcs, masks,
                                        {true}, {false}, false);
    Operator *op = f.CreateOperator();
    auto *batch = new VectorBatch(2);
    batch->Append(MakeVector<int64_t>(OMNI_LONG, {1, 1}));
    batch->Append(MakeVector<int64_t>(OMNI_DECIMAL64, {INT64_MAX, 1}));
    op->AddInput(batch);
    bool threw = false;
    try {
        Drain(op);
    } catch (const OmniException &e) {
        threw = e.Code() == "OPERATOR_RUNTIME_ERROR";
    }
    EXPECT_TRUE(threw);
    Operator::DeleteOperator(op);
}

// Two-column LONG build and probe batches (NULL where `nulls` lists the row of the first key column), key columns
// `keyCols` on both sides, every column output; returns the lookup operator's rows and, when `outer`, the outer
// operator's rows.
static std::pair<Rows, Rows> RunTypedJoin(JoinType jt, BuildSide bs, std::vector<int32_t> keyCols,
                                          const std::vector<std::vector<int64_t>> &bcols, std::vector<int> bnulls,
                                          const std::vector<std::vector<int64_t>> &pcols, bool outer)
{
    DataTypes types({LongType(), LongType()});
    auto *build = new VectorBatch(static_cast<int32_t>(bcols[0].size()));
    for (int c = 0; c < 2; ++c) build->Append(MakeVector<int64_t>(OMNI_LONG, bcols[c]));
    for (int r : bnulls) build->Get(keyCols[0])->SetNull(r);
    const int32_t nk = static_cast<int32_t>(keyCols.size());
    auto *bf = GpuHashBuilderOperatorFactory::CreateHashBuilderOperatorFactory(jt, bs, types, keyCols.data(), nk, 1);
    Operator *builder = bf->CreateOperator();
    builder->AddInput(build);
    VectorBatch *none = nullptr;
    builder->GetOutput(&none);
    auto *probe = new VectorBatch(static_cast<int32_t>(pcols[0].size()));
    for (int c = 0; c < 2; ++c) probe->Append(MakeVector<int64_t>(OMNI_LONG, pcols[c]));
    const bool probeOnly = jt == OMNI_JOIN_TYPE_LEFT_SEMI || jt == OMNI_JOIN_TYPE_LEFT_ANTI ||
                           jt == OMNI_JOIN_TYPE_EXISTENCE;
    int32_t probeOut[2] = {0, 1}, buildOut[2] = {0, 1};
    DataTypes buildOutTypes = probeOnly ? DataTypes() : DataTypes({LongType(), LongType()});
    auto *lf = GpuLookupJoinOperatorFactory::CreateLookupJoinOperatorFactory(
        types, probeOut, jt == OMNI_JOIN_TYPE_EXISTENCE ? 0 : 2, keyCols.data(), nk, buildOut, probeOnly ? 0 : 2,
        buildOutTypes, reinterpret_cast<int64_t>(bf), nullptr, false, nullptr);
    Operator *lookup = lf->CreateOperator();
    lookup->AddInput(probe);
    std::pair<Rows, Rows> res;
    res.first = Drain(lookup);
    GpuLookupOuterJoinOperatorFactory *of = nullptr;
    Operator *outerOp = nullptr;
    if (outer) {
        of = GpuLookupOuterJoinOperatorFactory::CreateLookupOuterJoinOperatorFactory(
            types, probeOut, 2, buildOut, buildOutTypes, reinterpret_cast<int64_t>(bf));
        outerOp = of->CreateOperator();
        res.second = Drain(outerOp);
        EXPECT_TRUE(outerOp->GetStatus() == OMNI_STATUS_FINISHED);
        Operator::DeleteOperator(outerOp);
        delete of;
    }
    Operator::DeleteOperator(builder);
    Operator::DeleteOperator(lookup);
    delete lf;
    delete bf;
    return res;
}

static const std::vector<std::vector<int64_t>> kBuild4 = {{1, 2, 3, 4}, {111, 11, 333, 33}};
static const std::vector<std::vector<int64_t>> kProbe4 = {{1, 2, 3, 4}, {11, 22, 33, 44}};

// join_test.cpp:773-833 (TestLeftEqualityJoin)
static void TestLeftJoin()
{
    auto r = RunTypedJoin(OMNI_JOIN_TYPE_LEFT, OMNI_BUILD_UNKNOWN, {1}, kBuild4, {}, kProbe4, false);
    Rows want = {{"1", "11", "2", "11"}, {"2", "22", "null", "null"}, {"3", "33", "4", "33"}, {"4", "44", "null", "null"}};
    EXPECT_TRUE(SameRowSet(r.first, want));
}

// join_test.cpp:835-918 (TestLeftEqualityJoinAndBuildLeft): build columns come first, the outer operator emits
// the unmatched build rows
static void TestLeftJoinBuildLeft()
{
    auto r = RunTypedJoin(OMNI_JOIN_TYPE_LEFT, OMNI_BUILD_LEFT, {1}, kBuild4, {}, kProbe4, true);
    Rows want = {{"2", "11", "1", "11"}, {"4", "33", "3", "33"}};  // (build col0, build col1, probe col0, probe col1)
    EXPECT_TRUE(SameRowSet(r.first, want));
    Rows wantOuter = {{"3", "333", "null", "null"}, {"1", "111", "null", "null"}};
    EXPECT_TRUE(SameRowSet(r.second, wantOuter));
}

// join_test.cpp:2881-2974 (TestOneKeyFullEqualityJoinWithBuildNull)
static void TestFullJoinBuildNull()
{
    auto r = RunTypedJoin(OMNI_JOIN_TYPE_FULL, OMNI_BUILD_UNKNOWN, {1}, {{1, 2, 3, 4, 5, 6}, {111, 11, 333, 33, 11, 11}},
                          {0, 2}, kProbe4, true);
    Rows want = {{"1", "11", "2", "11"}, {"1", "11", "5", "11"}, {"1", "11", "6", "11"},
                 {"2", "22", "null", "null"}, {"3", "33", "4", "33"}, {"4", "44", "null", "null"}};
    EXPECT_TRUE(SameRowSet(r.first, want));
    Rows wantOuter = {{"null", "null", "1", "null"}, {"null", "null", "3", "null"}};
    EXPECT_TRUE(SameRowSet(r.second, wantOuter));
}

// join_test.cpp:4157-4208 (TestLeftAntiEqualityJoin)
static void TestLeftAntiJoin()
{
    auto r = RunTypedJoin(OMNI_JOIN_TYPE_LEFT_ANTI, OMNI_BUILD_UNKNOWN, {1}, kBuild4, {},
                          {{1, 2, 3, 4, 3, 6}, {11, 22, 33, 44, 22, 44}}, false);
    Rows want = {{"2", "22"}, {"4", "44"}, {"3", "22"}, {"6", "44"}};
    EXPECT_TRUE(SameRowSet(r.first, want));
}

// join_test.cpp:4584-4636 (TestExistenceEqualityJoin): no probe output columns, one BOOLEAN per probe row
static void TestExistenceJoin()
{
    auto r = RunTypedJoin(OMNI_JOIN_TYPE_EXISTENCE, OMNI_BUILD_UNKNOWN, {1}, kBuild4, {},
                          {{1, 2, 3, 4, 3, 6}, {11, 22, 33, 44, 22, 44}}, false);
    Rows want = {{"true"}, {"false"}, {"true"}, {"false"}, {"false"}, {"false"}};
    EXPECT_TRUE(r.first == want);
}

// join_test.cpp:2975-3072 (TestFixedKeysFullEqualityJoinWithBuildNull): two LONG keys (an int128 key), NULL in the
// second key column of build rows 0 and 2
static void TestFixedKeysFullJoinBuildNull()
{
    std::vector<std::vector<int64_t>> bcols = {{1, 1, 3, 3, 5, 6}, {111, 11, 333, 33, 11, 11}};
    DataTypes types({LongType(), LongType()});
    auto *build = new VectorBatch(6);
    for (int c = 0; c < 2; ++c) build->Append(MakeVector<int64_t>(OMNI_LONG, bcols[c]));
    build->Get(1)->SetNull(0);
    build->Get(1)->SetNull(2);
    int32_t keys[2] = {0, 1}, out[2] = {0, 1};
    auto *bf = GpuHashBuilderOperatorFactory::CreateHashBuilderOperatorFactory(OMNI_JOIN_TYPE_FULL, types, keys, 2, 1);
    Operator *builder = bf->CreateOperator();
    builder->AddInput(build);
    VectorBatch *none = nullptr;
    builder->GetOutput(&none);
    auto *probe = new VectorBatch(4);
    probe->Append(MakeVector<int64_t>(OMNI_LONG, {1, 2, 3, 4}));
    probe->Append(MakeVector<int64_t>(OMNI_LONG, {11, 22, 33, 44}));
    auto *lf = GpuLookupJoinOperatorFactory::CreateLookupJoinOperatorFactory(
        types, out, 2, keys, 2, out, 2, types, reinterpret_cast<int64_t>(bf), nullptr, false, nullptr);
    Operator *lookup = lf->CreateOperator();
    lookup->AddInput(probe);
    Rows want = {{"1", "11", "1", "11"}, {"2", "22", "null", "null"}, {"3", "33", "3", "33"},
                 {"4", "44", "null", "null"}};
    EXPECT_TRUE(SameRowSet(Drain(lookup), want));
    auto *of = GpuLookupOuterJoinOperatorFactory::CreateLookupOuterJoinOperatorFactory(types, out, 2, out, types,
                                                                                      reinterpret_cast<int64_t>(bf));
    Operator *outerOp = of->CreateOperator();
    Rows wantOuter = {{"null", "null", "1", "null"}, {"null", "null", "3", "null"}, {"null", "null", "5", "11"},
                      {"null", "null", "6", "11"}};
    EXPECT_TRUE(SameRowSet(Drain(outerOp), wantOuter));
    Operator::DeleteOperator(outerOp);
    Operator::DeleteOperator(builder);
    Operator::DeleteOperator(lookup);
    delete of;
    delete lf;
    delete bf;
}

// join_test.cpp:4637-4705 (TestMultiKeysFullEqualityJoinInteger): SHORT, INT, LONG keys (112 bits, an int128 key)
static void TestMultiKeysFullJoinInteger()
{
    DataTypes types({ShortType(), IntType(), LongType()});
    auto *build = new VectorBatch(6);
    build->Append(MakeVector<int16_t>(OMNI_SHORT, {1, 2, 3, 4, 5, 6}));
    build->Append(MakeVector<int32_t>(OMNI_INT, {11, 22, 33, 44, 55, 66}));
    build->Append(MakeVector<int64_t>(OMNI_LONG, {111, 222, 333, 444, 555, 666}));
    int32_t keys[3] = {0, 1, 2};
    auto *bf = GpuHashBuilderOperatorFactory::CreateHashBuilderOperatorFactory(OMNI_JOIN_TYPE_FULL, types, keys, 3, 1);
    Operator *builder = bf->CreateOperator();
    builder->AddInput(build);
    VectorBatch *none = nullptr;
    builder->GetOutput(&none);
    auto *probe = new VectorBatch(4);
    probe->Append(MakeVector<int16_t>(OMNI_SHORT, {1, 3, 5, 7}));
    probe->Append(MakeVector<int32_t>(OMNI_INT, {11, 33, 55, 77}));
    probe->Append(MakeVector<int64_t>(OMNI_LONG, {111, 333, 555, 777}));
    auto *lf = GpuLookupJoinOperatorFactory::CreateLookupJoinOperatorFactory(
        types, keys, 3, keys, 3, keys, 3, types, reinterpret_cast<int64_t>(bf), nullptr, false, nullptr);
    Operator *lookup = lf->CreateOperator();
    lookup->AddInput(probe);
    Rows want = {{"1", "11", "111", "1", "11", "111"}, {"3", "33", "333", "3", "33", "333"},
                 {"5", "55", "555", "5", "55", "555"}, {"7", "77", "777", "null", "null", "null"}};
    EXPECT_TRUE(SameRowSet(Drain(lookup), want));
    Operator::DeleteOperator(builder);
    Operator::DeleteOperator(lookup);
    delete lf;
    delete bf;
}

// A dictionary vector of int64 values: row i is dict[ids[i]]; NULL where nullAt(i).
static BaseVector *MakeDictVector(const std::vector<int32_t> &ids, const std::vector<int64_t> &dict,
                                  const std::function<bool(int)> &nullAt)
{
    auto container = std::make_shared<DictionaryContainer<int64_t>>(ids, std::make_shared<std::vector<int64_t>>(dict));
    auto *v = new Vector<DictionaryContainer<int64_t>>(static_cast<int32_t>(ids.size()), container, OMNI_LONG);
    for (size_t i = 0; i < ids.size(); ++i) {
        if (nullAt(static_cast<int>(i))) v->SetNull(static_cast<int32_t>(i));
    }
    return v;
}

// A flat int64 vector of `n + off` rows sliced to its rows [off, off + n): row i of the slice is vals[i], NULL where
// nullAt(i) (the slice's NULL bitmap is a bit-shifted copy unless off is a multiple of 8).
static BaseVector *MakeSlicedVector(const std::vector<int64_t> &vals, int32_t off, const std::function<bool(int)> &nullAt)
{
    const int32_t n = static_cast<int32_t>(vals.size());
    Vector<int64_t> parent(n + off + 5, OMNI_LONG);
    for (int32_t i = 0; i < n + off + 5; ++i) parent.SetValue(i, -777);  // rows outside the slice
    for (int32_t i = 0; i < off; ++i) parent.SetNull(i);
    for (int32_t i = 0; i < n; ++i) {
        parent.SetValue(off + i, vals[i]);
        if (nullAt(i)) parent.SetNull(off + i);
    }
    return parent.Slice(off, n);
}

// SUM / COUNT(col) / MIN / MAX over int64 values grouped by an int64 key, as a std::map (NULL key -> "null").
static Rows ExpectedAgg(const std::vector<int64_t> &k, const std::vector<bool> &kn, const std::vector<int64_t> &v,
                        const std::vector<bool> &vn)
{
    struct Acc {
        int64_t sum = 0, cnt = 0, mn = 0, mx = 0;
        bool any = false;
    };
    std::map<std::string, Acc> m;
    for (size_t i = 0; i < k.size(); ++i) {
        Acc &a = m[kn[i] ? "null" : std::to_string(k[i])];
        if (vn[i]) continue;
        a.sum = static_cast<int64_t>(static_cast<uint64_t>(a.sum) + static_cast<uint64_t>(v[i]));
        a.mn = a.any ? std::min(a.mn, v[i]) : v[i];
        a.mx = a.any ? std::max(a.mx, v[i]) : v[i];
        a.cnt += 1;
        a.any = true;
    }
    Rows want;
    for (auto &e : m) {
        const Acc &a = e.second;
        auto s = [&](int64_t x) { return a.any ? std::to_string(x) : std::string("null"); };
        want.push_back({e.first, s(a.sum), std::to_string(a.cnt), s(a.mn), s(a.mx)});
    }
    return want;
}

static GpuHashAggregationOperatorFactory *SumCountMinMaxFactory()
{
    std::vector<uint32_t> groupCols = {0};
    std::vector<std::vector<uint32_t>> aggCols = {{1}, {1}, {1}, {1}};
    std::vector<DataTypes> inTypes(4, DataTypes({LongType()})), outTypes(4, DataTypes({LongType()}));
    std::vector<uint32_t> funcs = {OMNI_AGG_SUM, OMNI_AGG_COUNT_COLUMN, OMNI_AGG_MIN, OMNI_AGG_MAX};
    std::vector<uint32_t> masks;
    return new GpuHashAggregationOperatorFactory(groupCols, DataTypes({LongType()}), aggCols, inTypes, outTypes, funcs,
                                                 masks, std::vector<bool>(4, true), std::vector<bool>(4, false), false);
}

// Sliced key columns (byte-aligned and not) and dictionary-encoded value columns read through the reference's
// accessors (GetRawValues / GetNulls of the slice, GetIds / GetDictionary) give the groups of their logical rows.
static void TestAggSlicedAndDictionaryInputs()
{
    auto *f = SumCountMinMaxFactory();
    Operator *op = f->CreateOperator();
    std::vector<int64_t> allK, allV;
    std::vector<bool> allKn, allVn;
    const int32_t offsets[3] = {3, 8, 13};
    for (int b = 0; b < 3; ++b) {
        const int32_t n = 1000 + 37 * b;
        std::vector<int64_t> k(n), dict = {-5, 7, 1LL << 40, -(1LL << 50), 0};
        std::vector<int32_t> ids(n);
        for (int32_t i = 0; i < n; ++i) {
            k[i] = (i * 7 + b) % 23;
            ids[i] = (i * 3 + b) % 5;
        }
        auto kNull = [b](int i) { return (i + b) % 17 == 0; };
        auto vNull = [b](int i) { return (i + 2 * b) % 11 == 0; };
        auto *batch = new VectorBatch(n);
        batch->Append(MakeSlicedVector(k, offsets[b], kNull));
        batch->Append(MakeDictVector(ids, dict, vNull));
        for (int32_t i = 0; i < n; ++i) {
            allK.push_back(k[i]);
            allKn.push_back(kNull(i));
            allV.push_back(dict[ids[i]]);
            allVn.push_back(vNull(i));
        }
        op->AddInput(batch);
    }
    Rows got = Drain(op);
    EXPECT_TRUE(SameRowSet(got, ExpectedAgg(allK, allKn, allV, allVn)));
    Operator::DeleteOperator(op);
    delete f;
}

// 300 batches of 4,097 rows (1.2M rows: the host staging flushes to the device part-way, and batch boundaries fall
// inside NULL-bitmap bytes) give the groups of adding them one by one.
static void TestAggStagingAcrossFlushes()
{
    auto *f = SumCountMinMaxFactory();
    Operator *op = f->CreateOperator();
    std::vector<int64_t> allK, allV;
    std::vector<bool> allKn, allVn;
    for (int b = 0; b < 300; ++b) {
        const int32_t n = 4097;
        std::vector<int64_t> k(n), v(n);
        auto *kv = new Vector<int64_t>(n, OMNI_LONG);
        auto *vv = new Vector<int64_t>(n, OMNI_LONG);
        for (int32_t i = 0; i < n; ++i) {
            const int64_t row = static_cast<int64_t>(b) * n + i;
            k[i] = (row * 2654435761LL) % 5003;
            v[i] = (row % 1000) - 500;
            kv->SetValue(i, k[i]);
            vv->SetValue(i, v[i]);
            const bool kn = row % 1009 == 0, vn = row % 13 == 5;
            if (kn) kv->SetNull(i);
            if (vn) vv->SetNull(i);
            allK.push_back(k[i]);
            allKn.push_back(kn);
            allV.push_back(v[i]);
            allVn.push_back(vn);
        }
        auto *batch = new VectorBatch(n);
        batch->Append(kv);
        batch->Append(vv);
        op->AddInput(batch);
    }
    Rows got = Drain(op);
    EXPECT_TRUE(SameRowSet(got, ExpectedAgg(allK, allKn, allV, allVn)));
    Operator::DeleteOperator(op);
    delete f;
}

// An encoding the GPU path does not read (the AVG partial container) is refused with OmniException.
class ContainerVector : public BaseVector {
public:
    explicit ContainerVector(int32_t n) : BaseVector(n, OMNI_ENCODING_CONTAINER, OMNI_LONG) {}
    BaseVector *Slice(int32_t, int32_t) override { return nullptr; }
};

static void TestUnsupportedEncodingThrows()
{
    auto *f = SumCountMinMaxFactory();
    Operator *op = f->CreateOperator();
    auto *batch = new VectorBatch(4);
    batch->Append(MakeVector<int64_t>(OMNI_LONG, {1, 2, 3, 4}));
    batch->Append(new ContainerVector(4));
    bool threw = false;
    try {
        op->AddInput(batch);
    } catch (const OmniException &e) {
        threw = e.Code() == "UNSUPPORTED_ERROR";
    }
    EXPECT_TRUE(threw);
    Operator::DeleteOperator(op);
    delete f;
}

// Inner join whose build key column is dictionary-encoded and whose probe columns are a slice (key) and a
// dictionary vector (output), against a nested-loop join of the logical rows.
static void TestJoinSlicedAndDictionaryInputs()
{
    const int32_t nb = 300, np = 1000;
    std::vector<int32_t> bids(nb);
    std::vector<int64_t> bdict, bpay(nb);
    for (int i = 0; i < 60; ++i) bdict.push_back(i * 1000003LL);
    for (int32_t i = 0; i < nb; ++i) {
        bids[i] = (i * 7) % 60;
        bpay[i] = 5000 + i;
    }
    auto bKeyNull = [](int i) { return i % 29 == 3; };
    DataTypes buildTypes({LongType(), LongType()});
    auto *build = new VectorBatch(nb);
    build->Append(MakeDictVector(bids, bdict, bKeyNull));
    build->Append(MakeVector<int64_t>(OMNI_LONG, bpay));
    int32_t hashCols[1] = {0};
    auto *bf = GpuHashBuilderOperatorFactory::CreateHashBuilderOperatorFactory(OMNI_JOIN_TYPE_INNER, buildTypes, hashCols, 1, 1);
    Operator *builder = bf->CreateOperator();
    builder->AddInput(build);
    VectorBatch *none = nullptr;
    builder->GetOutput(&none);

    std::vector<int64_t> pk(np), pdictv = {11, 22, 33};
    std::vector<int32_t> pids(np);
    for (int32_t i = 0; i < np; ++i) {
        pk[i] = ((i * 13) % 80) * 1000003LL;
        pids[i] = i % 3;
    }
    auto pKeyNull = [](int i) { return i % 31 == 7; };
    auto pOutNull = [](int i) { return i % 10 == 4; };
    DataTypes probeTypes({LongType(), LongType()});
    auto *probe = new VectorBatch(np);
    probe->Append(MakeSlicedVector(pk, 5, pKeyNull));
    probe->Append(MakeDictVector(pids, pdictv, pOutNull));
    int32_t probeOut[1] = {1}, probeHash[1] = {0}, buildOut[1] = {1};
    DataTypes buildOutTypes({LongType()});
    auto *lf = GpuLookupJoinOperatorFactory::CreateLookupJoinOperatorFactory(
        probeTypes, probeOut, 1, probeHash, 1, buildOut, 1, buildOutTypes, reinterpret_cast<int64_t>(bf), nullptr,
        false, nullptr);
    Operator *lookup = lf->CreateOperator();
    lookup->AddInput(probe);
    Rows got = Drain(lookup);
    Rows want;
    for (int32_t p = 0; p < np; ++p) {
        if (pKeyNull(p)) continue;
        for (int32_t b = 0; b < nb; ++b) {
            if (bKeyNull(b) || bdict[bids[b]] != pk[p]) continue;
            want.push_back({pOutNull(p) ? "null" : std::to_string(pdictv[pids[p]]), std::to_string(bpay[b])});
        }
    }
    EXPECT_TRUE(!want.empty());
    EXPECT_TRUE(SameRowSet(got, want));
    Operator::DeleteOperator(builder);
    Operator::DeleteOperator(lookup);
    delete lf;
    delete bf;
}


// ---- two-stage decimal SUM (SumSparkDecimalAggregator: partial state (sum, isEmpty)) --------------------------

static Decimal128 D128(__int128 v)
{
    return Decimal128(static_cast<int64_t>(v >> 64), static_cast<uint64_t>(static_cast<unsigned __int128>(v)));
}
static const __int128 kMax38 = static_cast<__int128>(5421010862427522170LL) * (static_cast<__int128>(1) << 64) +
                               static_cast<__int128>(687399551400673279ULL);  // 10^38 - 1

// A one-key (LONG) decimal SUM operator factory: raw or partial (sum, isEmpty) input, final or partial output.
static GpuHashAggregationOperatorFactory *DecSumFactory(bool raw, bool partial, DataTypePtr inT, DataTypePtr outT,
                                                        bool overflowAsNull)
{
    std::vector<uint32_t> groupCols = {0};
    std::vector<std::vector<uint32_t>> aggCols = {raw ? std::vector<uint32_t>{1} : std::vector<uint32_t>{1, 2}};
    std::vector<DataTypes> inTypes = {raw ? DataTypes({inT}) : DataTypes({inT, BooleanType()})};
    std::vector<DataTypes> outTypes = {partial ? DataTypes({outT, BooleanType()}) : DataTypes({outT})};
    std::vector<uint32_t> funcs = {OMNI_AGG_SUM}, masks;
    return new GpuHashAggregationOperatorFactory(groupCols, DataTypes({LongType()}), aggCols, inTypes, outTypes, funcs,
                                                 masks, {raw}, {partial}, overflowAsNull);
}

// A batch (key, decimal sum[, isEmpty]); sumNull lists the NULL sums.
static VectorBatch *DecBatch(const std::vector<int64_t> &keys, BaseVector *sums, const std::vector<int> &sumNull,
                             const std::vector<uint8_t> *empty)
{
    auto *b = new VectorBatch(static_cast<int32_t>(keys.size()));
    b->Append(MakeVector<int64_t>(OMNI_LONG, keys));
    for (int r : sumNull) sums->SetNull(r);
    b->Append(sums);
    if (empty) b->Append(MakeVector<uint8_t>(OMNI_BOOLEAN, *empty));
    return b;
}

static Rows RunOne(GpuHashAggregationOperatorFactory *f, std::vector<VectorBatch *> in)
{
    Operator *op = f->CreateOperator();
    for (auto *b : in) op->AddInput(b);
    Rows rows = Drain(op);
    Operator::DeleteOperator(op);
    delete f;
    return rows;
}

// aggregation_test.cpp:3680-3734 (spark_sum_decimal64_normal): DECIMAL64(18, 6) into DECIMAL128(28, 6); the
// partial output of one row of 999999999999.999999 is that value with isEmpty false, and the three rows sum to
// 2999999999999.999997 -- in one stage, and in two (two partial operators, then a final one).
static void TestAggDecimalTwoStageKnownAnswers()
{
    const int64_t v = 999999999999999999LL;
    Rows p1 = RunOne(DecSumFactory(true, true, Decimal64Type(18, 6), Decimal128Type(28, 6), false),
                     {DecBatch({7}, MakeVector<int64_t>(OMNI_DECIMAL64, {v}), {}, nullptr)});
    EXPECT_TRUE(SameRowSet(p1, {{"7", "999999999999999999", "false"}}));
    Rows one = RunOne(DecSumFactory(true, false, Decimal64Type(18, 6), Decimal128Type(28, 6), false),
                      {DecBatch({7, 7, 7}, MakeVector<int64_t>(OMNI_DECIMAL64, {v, v, v}), {}, nullptr)});
    EXPECT_TRUE(SameRowSet(one, {{"7", "2999999999999999997"}}));
    Rows p2 = RunOne(DecSumFactory(true, true, Decimal64Type(18, 6), Decimal128Type(28, 6), false),
                     {DecBatch({7, 7}, MakeVector<int64_t>(OMNI_DECIMAL64, {v, v}), {}, nullptr)});
    EXPECT_TRUE(SameRowSet(p2, {{"7", "1999999999999999998", "false"}}));
    const std::vector<uint8_t> no = {0, 0};
    Rows fin = RunOne(DecSumFactory(false, false, Decimal128Type(28, 6), Decimal128Type(28, 6), false),
                      {DecBatch({7, 7}, MakeVector<Decimal128>(OMNI_DECIMAL128, {D128(v), D128(2 * static_cast<__int128>(v))}),
                                {}, &no)});
    EXPECT_TRUE(SameRowSet(fin, {{"7", "2999999999999999997"}}));
}

// aggregation_test.cpp:3790-3902 (spark_sum_decimal128_overflow_*): DECIMAL128(38, 0); the partial output of
// 10^38 - 1 is itself, and a final stage adding three such partial sums overflows: an error when overflowAsNull is
// false, a NULL when it is true.
static void TestAggDecimalTwoStageOverflow()
{
    Rows p = RunOne(DecSumFactory(true, true, Decimal128Type(38, 0), Decimal128Type(38, 0), false),
                    {DecBatch({1}, MakeVector<Decimal128>(OMNI_DECIMAL128, {D128(kMax38)}), {}, nullptr)});
    EXPECT_TRUE(SameRowSet(p, {{"1", "99999999999999999999999999999999999999", "false"}}));
    const std::vector<uint8_t> no = {0, 0, 0};
    auto batch = [&]() {
        return DecBatch({1, 1, 1}, MakeVector<Decimal128>(OMNI_DECIMAL128, {D128(kMax38), D128(kMax38), D128(kMax38)}),
                        {}, &no);
    };
    bool threw = false;
    try {
        RunOne(DecSumFactory(false, false, Decimal128Type(38, 0), Decimal128Type(38, 0), false), {batch()});
    } catch (const OmniException &e) {
        threw = e.Code() == "OPERATOR_RUNTIME_ERROR";
    }
    EXPECT_TRUE(threw);
    Rows n = RunOne(DecSumFactory(false, false, Decimal128Type(38, 0), Decimal128Type(38, 0), true), {batch()});
    EXPECT_TRUE(SameRowSet(n, {{"1", "null"}}));
}

// The partial state's rules (sum_spark_decimal_aggregator.h:16-66, 246-301): a raw partial stage outputs 0 with
// isEmpty true for a group of NULL inputs, and NULL with isEmpty false for an overflowed group (never an error);
// over partial input, a NULL sum overflows the group whatever follows (even with isEmpty set), empty inputs are
// skipped, and a group of empty inputs only is NULL (final) or (0, true) (partial).
static void TestAggDecimalPartialStateRules()
{
    Rows p = RunOne(DecSumFactory(true, true, Decimal64Type(18, 0), Decimal64Type(18, 0), false),
                    {DecBatch({1, 1, 2, 2, 3}, MakeVector<int64_t>(OMNI_DECIMAL64, {INT64_MAX, 1, 0, 0, 5}), {2, 3},
                              nullptr)});
    EXPECT_TRUE(SameRowSet(p, {{"1", "null", "false"}, {"2", "0", "true"}, {"3", "5", "false"}}));
    // keys: 1 NULL sum then 5; 2 empty only; 3 a sum and an empty; 4 a NULL sum flagged empty; 5 two sums
    const std::vector<int64_t> keys = {1, 1, 2, 2, 3, 3, 4, 5, 5};
    const std::vector<uint8_t> empty = {0, 0, 1, 1, 0, 1, 1, 0, 0};
    auto batch = [&]() {
        return DecBatch(keys, MakeVector<int64_t>(OMNI_DECIMAL64, {0, 5, 0, 0, 3, 0, 0, 3, 4}), {0, 6}, &empty);
    };
    Rows fin = RunOne(DecSumFactory(false, false, Decimal64Type(18, 0), Decimal64Type(18, 0), true), {batch()});
    EXPECT_TRUE(SameRowSet(fin, {{"1", "null"}, {"2", "null"}, {"3", "3"}, {"4", "null"}, {"5", "7"}}));
    Rows merge = RunOne(DecSumFactory(false, true, Decimal64Type(18, 0), Decimal64Type(18, 0), false), {batch()});
    EXPECT_TRUE(SameRowSet(merge, {{"1", "null", "false"}, {"2", "0", "true"}, {"3", "3", "false"},
                                   {"4", "null", "false"}, {"5", "7", "false"}}));
    bool threw = false;
    try {
        RunOne(DecSumFactory(false, false, Decimal64Type(18, 0), Decimal64Type(18, 0), false), {batch()});
    } catch (const OmniException &e) {
        threw = e.Code() == "OPERATOR_RUNTIME_ERROR";
    }
    EXPECT_TRUE(threw);
}

// Two stages equal one on seeded rows that do not overflow: DECIMAL64(18, 2) into DECIMAL128(38, 2), keys over 50
// values, 20% NULL inputs, split over three partial operators, their outputs into one final operator.
static void TestAggDecimalTwoStageMatchesSingleStage()
{
    uint64_t st = 0xDEC2;
    auto next = [&]() {
        st += 0x9E3779B97F4A7C15ULL;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    };
    std::vector<VectorBatch *> one, parts[3];
    for (int b = 0; b < 6; ++b) {
        std::vector<int64_t> k(1000), v(1000);
        std::vector<int> nulls;
        for (int r = 0; r < 1000; ++r) {
            k[r] = static_cast<int64_t>(next() % 50);
            v[r] = static_cast<int64_t>(next() % 2000000001ULL) - 1000000000LL;
            if (next() % 5 == 0) nulls.push_back(r);
        }
        one.push_back(DecBatch(k, MakeVector<int64_t>(OMNI_DECIMAL64, v), nulls, nullptr));
        parts[b % 3].push_back(DecBatch(k, MakeVector<int64_t>(OMNI_DECIMAL64, v), nulls, nullptr));
    }
    // a key whose inputs are all NULL: NULL in one stage, (0, true) partials and NULL again in two
    one.push_back(DecBatch({99, 99}, MakeVector<int64_t>(OMNI_DECIMAL64, {1, 2}), {0, 1}, nullptr));
    parts[0].push_back(DecBatch({99, 99}, MakeVector<int64_t>(OMNI_DECIMAL64, {1, 2}), {0, 1}, nullptr));
    Rows single = RunOne(DecSumFactory(true, false, Decimal64Type(18, 2), Decimal128Type(38, 2), false), one);
    auto *ff = DecSumFactory(false, false, Decimal128Type(38, 2), Decimal128Type(38, 2), false);
    Operator *fin = ff->CreateOperator();
    for (auto &ps : parts) {
        auto *pf = DecSumFactory(true, true, Decimal64Type(18, 2), Decimal128Type(38, 2), false);
        Operator *op = pf->CreateOperator();
        for (auto *o : RunAgg(op, ps)) fin->AddInput(o);
        Operator::DeleteOperator(op);
        delete pf;
    }
    Rows two = Drain(fin);
    Operator::DeleteOperator(fin);
    delete ff;
    EXPECT_TRUE(single.size() == 51);
    EXPECT_TRUE(SameRowSet(single, two));
}

// HashAggregationOperator::AlignSchema (group_aggregation.cpp:1182-1218): raw rows reshaped into the partial schema
// one by one; a final stage over aligned batches equals one stage over the raw rows (SUM / COUNT(col) / AVG / MAX
// of LONG, COUNT(*), SUM of DECIMAL64 into DECIMAL128), and the aligned columns follow each aggregator's rule.
static void TestAggAlignSchema()
{
    std::vector<uint32_t> groupCols = {0};
    std::vector<std::vector<uint32_t>> aggCols = {{1}, {1}, {1}, {1}, {2}};
    std::vector<DataTypes> inTypes = {DataTypes({LongType()}), DataTypes({LongType()}), DataTypes({LongType()}),
                                      DataTypes({LongType()}), DataTypes({Decimal64Type(18, 2)})};
    std::vector<DataTypes> partOut = {DataTypes({LongType()}), DataTypes({LongType()}),
                                      DataTypes({DoubleType(), LongType()}), DataTypes({LongType()}),
                                      DataTypes({LongType()}), DataTypes({Decimal128Type(38, 2), BooleanType()})};
    std::vector<uint32_t> funcs = {OMNI_AGG_SUM, OMNI_AGG_COUNT_COLUMN, OMNI_AGG_AVG, OMNI_AGG_MAX,
                                   OMNI_AGG_COUNT_ALL, OMNI_AGG_SUM}, masks;
    // COUNT(*) over raw input takes no input column: aggCols / inTypes list the other aggregates only
    GpuHashAggregationOperatorFactory pf(groupCols, DataTypes({LongType()}), aggCols, inTypes, partOut, funcs, masks,
                                         {true, true, true, true, true, true}, {true, true, true, true, true, true});
    auto raw = [](int salt) {
        auto *b = new VectorBatch(5);
        b->Append(MakeVector<int64_t>(OMNI_LONG, {1, 2, 1, 3, 2}));
        auto *v = MakeVector<int64_t>(OMNI_LONG, {10 + salt, 20, 30, 40, 50});
        v->SetNull(3);
        b->Append(v);
        auto *d = MakeVector<int64_t>(OMNI_DECIMAL64, {100, 200, 300 + salt, 400, 500});
        d->SetNull(1);
        b->Append(d);
        return b;
    };
    Operator *op = pf.CreateOperator();
    VectorBatch *aligned = op->AlignSchema(raw(0));
    Rows got;
    AppendRows(got, aligned);
    EXPECT_TRUE(SameRowSet(got, {{"1", "10", "1", "10", "1", "10", "1", "100", "false"},
                                 {"2", "20", "1", "20", "1", "20", "1", "0", "true"},
                                 {"1", "30", "1", "30", "1", "30", "1", "300", "false"},
                                 {"3", "null", "0", "0", "0", "null", "1", "400", "false"},
                                 {"2", "50", "1", "50", "1", "50", "1", "500", "false"}}));
    // the final stage over two aligned batches equals one stage over the raw rows
    std::vector<std::vector<uint32_t>> finCols = {{1}, {2}, {3, 4}, {5}, {6}, {7, 8}};
    std::vector<DataTypes> finIn = {DataTypes({LongType()}), DataTypes({LongType()}),
                                    DataTypes({DoubleType(), LongType()}), DataTypes({LongType()}),
                                    DataTypes({LongType()}), DataTypes({Decimal128Type(38, 2), BooleanType()})};
    std::vector<DataTypes> finOut = {DataTypes({LongType()}), DataTypes({LongType()}), DataTypes({DoubleType()}),
                                     DataTypes({LongType()}), DataTypes({LongType()}),
                                     DataTypes({Decimal128Type(38, 2)})};
    GpuHashAggregationOperatorFactory ff(groupCols, DataTypes({LongType()}), finCols, finIn, finOut, funcs, masks,
                                         {false, false, false, false, false, false},
                                         {false, false, false, false, false, false});
    Operator *fin = ff.CreateOperator();
    fin->AddInput(aligned);
    fin->AddInput(op->AlignSchema(raw(1)));
    Rows two = Drain(fin);
    std::vector<std::vector<uint32_t>> oneCols = {{1}, {1}, {1}, {1}, {2}};
    std::vector<DataTypes> oneOut = finOut;
    GpuHashAggregationOperatorFactory of(groupCols, DataTypes({LongType()}), oneCols, inTypes, oneOut, funcs, masks,
                                         {true, true, true, true, true, true},
                                         {false, false, false, false, false, false});
    Operator *single = of.CreateOperator();
    single->AddInput(raw(0));
    single->AddInput(raw(1));
    Rows expect = Drain(single);
    EXPECT_TRUE(expect.size() == 3);
    EXPECT_TRUE(SameRowSet(two, expect));
    Operator::DeleteOperator(single);
    Operator::DeleteOperator(fin);
    Operator::DeleteOperator(op);
}


// ---- buffered probe runs (setNoMoreInput(false) ... noMoreInput()) -------------------------------------------

// A join of `probeRows` probe rows (batches of 65,536, a LONG key with NULLs and a LONG payload column with NULLs)
// against 3,000 build rows (keys 0..999 three times each for the inner / left joins: duplicate keys), run once
// batch by batch and once as buffered runs; both give the same rows, the buffered operator takes no input while
// it emits a run and finishes after noMoreInput().
static void BufferedVersusPerBatch(JoinType jt, int64_t probeRows)
{
    DataTypes btypes({LongType(), LongType()});
    std::vector<int64_t> bk(3000), bv(3000);
    for (int i = 0; i < 3000; ++i) {
        bk[i] = i % 1000;
        bv[i] = 7 * i;
    }
    int32_t keyCols[1] = {0};
    auto *bf = GpuHashBuilderOperatorFactory::CreateHashBuilderOperatorFactory(jt, OMNI_BUILD_UNKNOWN, btypes, keyCols, 1, 1);
    Operator *builder = bf->CreateOperator();
    auto *build = new VectorBatch(3000);
    build->Append(MakeVector<int64_t>(OMNI_LONG, bk));
    build->Append(MakeVector<int64_t>(OMNI_LONG, bv));
    builder->AddInput(build);
    VectorBatch *none = nullptr;
    builder->GetOutput(&none);
    const bool probeOnly = jt == OMNI_JOIN_TYPE_LEFT_ANTI;
    DataTypes ptypes({LongType(), LongType()});
    int32_t probeOut[2] = {0, 1}, buildOut[1] = {1};
    DataTypes buildOutTypes = probeOnly ? DataTypes() : DataTypes({LongType()});
    auto *lf = GpuLookupJoinOperatorFactory::CreateLookupJoinOperatorFactory(
        ptypes, probeOut, 2, keyCols, 1, buildOut, probeOnly ? 0 : 1, buildOutTypes, reinterpret_cast<int64_t>(bf),
        nullptr, false, nullptr);
    auto batches = [&]() {
        std::vector<VectorBatch *> out;
        uint64_t st = 0xB0F;
        for (int64_t a = 0; a < probeRows; a += 65536) {
            const int32_t n = static_cast<int32_t>(std::min<int64_t>(65536, probeRows - a));
            std::vector<int64_t> k(n), v(n);
            for (int32_t r = 0; r < n; ++r) {
                st = st * 6364136223846793005ULL + 1442695040888963407ULL;
                k[r] = static_cast<int64_t>((st >> 33) % 1500);  // a third of the keys miss
                v[r] = a + r;
            }
            auto *b = new VectorBatch(n);
            auto *kv = MakeVector<int64_t>(OMNI_LONG, k);
            auto *vv = MakeVector<int64_t>(OMNI_LONG, v);
            for (int32_t r = 0; r < n; r += 97) kv->SetNull(r);
            for (int32_t r = 5; r < n; r += 89) vv->SetNull(r);
            b->Append(kv);
            b->Append(vv);
            out.push_back(b);
        }
        return out;
    };
    // rows as int64 triples (INT64_MIN for NULL), sorted: cheaper than text rows at a million rows
    using Row3 = std::array<int64_t, 3>;
    auto collect = [](std::vector<Row3> &rows, VectorBatch *o) {
        for (int32_t r = 0; r < o->GetRowCount(); ++r) {
            Row3 t = {INT64_MIN, INT64_MIN, INT64_MIN};
            for (int32_t c = 0; c < o->GetVectorCount() && c < 3; ++c) {
                const BaseVector *v = o->Get(c);
                if (!v->IsNull(r)) t[c] = static_cast<const Vector<int64_t> *>(v)->GetValue(r);
            }
            rows.push_back(t);
        }
        VectorHelper::FreeVecBatch(o);
    };
    auto drain = [&](Operator *op, std::vector<Row3> &rows) {
        int n = 0;
        for (;;) {
            VectorBatch *o = nullptr;
            op->GetOutput(&o);
            if (!o) break;
            collect(rows, o);
            ++n;
        }
        return n;
    };
    Operator *one = lf->CreateOperator();
    std::vector<Row3> perBatch;
    for (auto *b : batches()) {
        one->AddInput(b);
        drain(one, perBatch);
    }
    Operator *buf = lf->CreateOperator();
    buf->setNoMoreInput(false);
    std::vector<Row3> buffered;
    int runsSeen = 0;
    for (auto *b : batches()) {
        EXPECT_TRUE(buf->needsInput());
        buf->AddInput(b);
        VectorBatch *o = nullptr;
        buf->GetOutput(&o);
        if (o) {
            ++runsSeen;
            EXPECT_TRUE(!buf->needsInput());  // a run is being emitted
            collect(buffered, o);
            drain(buf, buffered);
            EXPECT_TRUE(buf->needsInput());
        }
    }
    buf->noMoreInput();
    drain(buf, buffered);
    EXPECT_TRUE(buf->GetStatus() == OMNI_STATUS_FINISHED);
    EXPECT_TRUE(!perBatch.empty());
    std::sort(perBatch.begin(), perBatch.end());
    std::sort(buffered.begin(), buffered.end());
    EXPECT_TRUE(perBatch == buffered);
    if (probeRows > (1 << 20)) EXPECT_TRUE(runsSeen >= 1);  // a run flushed before noMoreInput()
    Operator::DeleteOperator(one);
    Operator::DeleteOperator(buf);
    Operator::DeleteOperator(builder);
    delete lf;
    delete bf;
}

static void TestJoinBufferedRunsInner() { BufferedVersusPerBatch(OMNI_JOIN_TYPE_INNER, 1300000); }
static void TestJoinBufferedRunsLeft() { BufferedVersusPerBatch(OMNI_JOIN_TYPE_LEFT, 200000); }
static void TestJoinBufferedRunsAnti() { BufferedVersusPerBatch(OMNI_JOIN_TYPE_LEFT_ANTI, 200000); }


// ---- several builder partitions (operatorCount > 1) ------------------------------------------------------------

// Hash-table partitions of LONG keys (NULL where null[i]) from the library (pinned against the oracle's HashUtil
// restatement by tests/test_gpu_join.py::test_join_partition_ids_match_oracle).
static std::vector<uint32_t> KeyPartitions(const std::vector<int64_t> &keys, const std::vector<bool> &null, uint32_t mask)
{
    const int64_t n = static_cast<int64_t>(keys.size());
    void *dv = nullptr, *dn = nullptr, *dp = nullptr;
    std::vector<uint8_t> bits((n + 7) / 8 + 1, 0);
    for (int64_t i = 0; i < n; ++i) {
        if (null[i]) bits[i >> 3] |= static_cast<uint8_t>(1u << (i & 7));
    }
    omni_gpu_malloc(&dv, 8 * n);
    omni_gpu_malloc(&dn, bits.size());
    omni_gpu_malloc(&dp, 4 * n);
    omni_gpu_memcpy_h2d(dv, keys.data(), 8 * n);
    omni_gpu_memcpy_h2d(dn, bits.data(), bits.size());
    omni_gpu_column c{OMNI_LONG, 0, dv, static_cast<const uint8_t *>(dn)};
    std::vector<uint32_t> out(n);
    EXPECT_TRUE(omni_gpu_join_partition_ids(&c, 1, n, mask, static_cast<uint32_t *>(dp)) == 0);
    omni_gpu_memcpy_d2h(out.data(), dp, 4 * n);
    omni_gpu_free(dv);
    omni_gpu_free(dn);
    omni_gpu_free(dp);
    return out;
}

// A LONG-key join over P builder operators (HashBuilderOperatorFactory with operatorCount = P): build rows
// either routed to the builder of their key's hash partition or dealt out by batch regardless of the key; every
// probe row only sees the build rows of its own partition's builder (lookup_join.cpp:950-1000), checked against
// a brute force over the same assignment, and -- when routed -- against the one-table join.
static Rows MultiPartitionJoin(JoinType jt, int32_t P, bool routed, Rows *expected)
{
    const int32_t nb = 900, np = 2500;
    std::vector<int64_t> bk(nb), bv(nb), pk(np), pv(np);
    std::vector<bool> bnull(nb, false), pnull(np, false);
    for (int32_t i = 0; i < nb; ++i) {
        bk[i] = (i * 37) % 400;  // keys with two or three build rows each
        bv[i] = 1000 + i;
        bnull[i] = i % 53 == 0;
    }
    for (int32_t i = 0; i < np; ++i) {
        pk[i] = (i * 7919) % 600;
        pv[i] = i;
        pnull[i] = i % 41 == 0;
    }
    std::vector<int32_t> owner(nb);
    if (routed) {
        std::vector<uint32_t> q = KeyPartitions(bk, bnull, static_cast<uint32_t>(P - 1));
        for (int32_t i = 0; i < nb; ++i) owner[i] = static_cast<int32_t>(q[i]);
    } else {
        for (int32_t i = 0; i < nb; ++i) owner[i] = (i / 50) % P;
    }
    DataTypes types({LongType(), LongType()});
    int32_t keyCols[1] = {0};
    auto *bf = GpuHashBuilderOperatorFactory::CreateHashBuilderOperatorFactory(jt, types, keyCols, 1, P);
    std::vector<Operator *> builders;
    for (int32_t q = 0; q < P; ++q) builders.push_back(bf->CreateOperator());  // partitions 0..P-1
    for (int32_t q = 0; q < P; ++q) {
        std::vector<int64_t> k, v;
        std::vector<int> nulls;
        for (int32_t i = 0; i < nb; ++i) {
            if (owner[i] != q) continue;
            if (bnull[i]) nulls.push_back(static_cast<int>(k.size()));
            k.push_back(bk[i]);
            v.push_back(bv[i]);
        }
        if (!k.empty()) {
            auto *b = new VectorBatch(static_cast<int32_t>(k.size()));
            auto *kv = MakeVector<int64_t>(OMNI_LONG, k);
            for (int r : nulls) kv->SetNull(r);
            b->Append(kv);
            b->Append(MakeVector<int64_t>(OMNI_LONG, v));
            builders[q]->AddInput(b);
        }
        VectorBatch *none = nullptr;
        EXPECT_TRUE(!bf->Finished() || q == P - 1);
        builders[q]->GetOutput(&none);
    }
    EXPECT_TRUE(bf->Finished());
    int32_t probeOut[2] = {0, 1}, buildOut[1] = {1};
    DataTypes buildOutTypes({LongType()});
    auto *lf = GpuLookupJoinOperatorFactory::CreateLookupJoinOperatorFactory(
        types, probeOut, 2, keyCols, 1, buildOut, 1, buildOutTypes, reinterpret_cast<int64_t>(bf), nullptr, false,
        nullptr);
    Operator *lookup = lf->CreateOperator();
    Rows got;
    for (int32_t a = 0; a < np; a += 1000) {
        const int32_t n = std::min(1000, np - a);
        auto *b = new VectorBatch(n);
        auto *kv = MakeVector<int64_t>(OMNI_LONG, std::vector<int64_t>(pk.begin() + a, pk.begin() + a + n));
        for (int32_t r = 0; r < n; ++r) {
            if (pnull[a + r]) kv->SetNull(r);
        }
        b->Append(kv);
        b->Append(MakeVector<int64_t>(OMNI_LONG, std::vector<int64_t>(pv.begin() + a, pv.begin() + a + n)));
        lookup->AddInput(b);
        Rows r = Drain(lookup);
        got.insert(got.end(), r.begin(), r.end());
    }
    if (expected) {
        std::vector<uint32_t> pq = KeyPartitions(pk, pnull, static_cast<uint32_t>(P - 1));
        expected->clear();
        for (int32_t i = 0; i < np; ++i) {
            bool any = false;
            for (int32_t j = 0; j < nb && !pnull[i]; ++j) {
                if (!bnull[j] && bk[j] == pk[i] && owner[j] == static_cast<int32_t>(pq[i])) {
                    expected->push_back({std::to_string(pk[i]), std::to_string(pv[i]), std::to_string(bv[j])});
                    any = true;
                }
            }
            if (!any && jt == OMNI_JOIN_TYPE_LEFT) {
                expected->push_back({pnull[i] ? "null" : std::to_string(pk[i]), std::to_string(pv[i]), "null"});
            }
        }
    }
    Operator::DeleteOperator(lookup);
    for (auto *b : builders) Operator::DeleteOperator(b);
    delete lf;
    delete bf;
    return got;
}

static void TestJoinMultiPartitionRouted()
{
    Rows want, one = MultiPartitionJoin(OMNI_JOIN_TYPE_INNER, 1, false, nullptr);
    for (int32_t P : {2, 4}) {
        Rows got = MultiPartitionJoin(OMNI_JOIN_TYPE_INNER, P, true, &want);
        EXPECT_TRUE(SameRowSet(got, want));
        EXPECT_TRUE(SameRowSet(got, one));  // rows routed by the probe's own partition function: one table's rows
    }
}

// BYTE (int8) keys through the drop-in operators: a group-by over negative and extreme values, and an inner join
// on (BYTE, INT) keys, whose packed key (8 + 32 bits) lets (3, -1) and (8, -1) match as ComputeMultiColKey's
// sign extension makes them (join_hash_table_variants.cpp:105-146).
static void TestByteKeysThroughOperators()
{
    auto *f = StdAggFactory({ByteType()}, 1);
    Operator *op = f->CreateOperator();
    auto *batch = new VectorBatch(6);
    batch->Append(MakeVector<int8_t>(OMNI_BYTE, {-1, -128, 127, -1, 127, -1}));
    for (int a = 0; a < 5; ++a) batch->Append(MakeVector<int64_t>(OMNI_LONG, {2, 5, 7, 4, 9, 6}));
    op->AddInput(batch);
    Rows want = {{"-1", "12", "4", "3", "2", "6"}, {"-128", "5", "5", "1", "5", "5"}, {"127", "16", "8", "2", "7", "9"}};
    EXPECT_TRUE(SameRowSet(Drain(op), want));
    Operator::DeleteOperator(op);
    delete f;

    DataTypes types({ByteType(), IntType()});
    auto *build = new VectorBatch(3);
    build->Append(MakeVector<int8_t>(OMNI_BYTE, {3, -7, 100}));
    build->Append(MakeVector<int32_t>(OMNI_INT, {-1, 9, 10}));
    int32_t keys[2] = {0, 1};
    auto *bf = GpuHashBuilderOperatorFactory::CreateHashBuilderOperatorFactory(OMNI_JOIN_TYPE_INNER, types, keys, 2, 1);
    Operator *builder = bf->CreateOperator();
    builder->AddInput(build);
    VectorBatch *none = nullptr;
    builder->GetOutput(&none);
    auto *probe = new VectorBatch(3);
    probe->Append(MakeVector<int8_t>(OMNI_BYTE, {8, -7, 101}));
    probe->Append(MakeVector<int32_t>(OMNI_INT, {-1, 9, 10}));
    auto *lf = GpuLookupJoinOperatorFactory::CreateLookupJoinOperatorFactory(
        types, keys, 2, keys, 2, keys, 2, types, reinterpret_cast<int64_t>(bf), nullptr, false, nullptr);
    Operator *lookup = lf->CreateOperator();
    lookup->AddInput(probe);
    Rows jwant = {{"8", "-1", "3", "-1"}, {"-7", "9", "-7", "9"}};
    EXPECT_TRUE(SameRowSet(Drain(lookup), jwant));
    Operator::DeleteOperator(builder);
    Operator::DeleteOperator(lookup);
    delete lf;
    delete bf;
}

static void TestJoinMultiPartitionUnrouted()
{
    for (JoinType jt : {OMNI_JOIN_TYPE_INNER, OMNI_JOIN_TYPE_LEFT}) {
        Rows want;
        Rows got = MultiPartitionJoin(jt, 4, false, &want);
        EXPECT_TRUE(!want.empty());
        EXPECT_TRUE(SameRowSet(got, want));
    }
}

int main()
{
    std::vector<std::pair<const char *, std::function<void()>>> tests = {
        {"InnerJoinOneBuildOp", TestInnerJoinOneBuildOp},
        {"InnerJoinArrayTable", TestInnerJoinArrayTable},
        {"LeftJoin", TestLeftJoin},
        {"LeftJoinBuildLeft", TestLeftJoinBuildLeft},
        {"FullJoinBuildNull", TestFullJoinBuildNull},
        {"LeftAntiJoin", TestLeftAntiJoin},
        {"ExistenceJoin", TestExistenceJoin},
        {"FixedKeysFullJoinBuildNull", TestFixedKeysFullJoinBuildNull},
        {"MultiKeysFullJoinInteger", TestMultiKeysFullJoinInteger},
        {"AggVerifyCorrectness", TestAggVerifyCorrectness},
        {"AggTwoStageVerifyCorrectness", TestAggTwoStageVerifyCorrectness},
        {"AggTwoStageMatchesSingleStage", TestAggTwoStageMatchesSingleStage},
        {"AggNullKeys", TestAggNullKeys},
        {"AggOutputBatching", TestAggOutputBatching},
        {"AggDecimalOverflowThrows", TestAggDecimalOverflowThrows},
        {"AggSlicedAndDictionaryInputs", TestAggSlicedAndDictionaryInputs},
        {"AggStagingAcrossFlushes", TestAggStagingAcrossFlushes},
        {"UnsupportedEncodingThrows", TestUnsupportedEncodingThrows},
        {"JoinSlicedAndDictionaryInputs", TestJoinSlicedAndDictionaryInputs},
        {"AggDecimalTwoStageKnownAnswers", TestAggDecimalTwoStageKnownAnswers},
        {"AggDecimalTwoStageOverflow", TestAggDecimalTwoStageOverflow},
        {"AggDecimalPartialStateRules", TestAggDecimalPartialStateRules},
        {"AggDecimalTwoStageMatchesSingleStage", TestAggDecimalTwoStageMatchesSingleStage},
        {"AggAlignSchema", TestAggAlignSchema},
        {"JoinBufferedRunsInner", TestJoinBufferedRunsInner},
        {"JoinBufferedRunsLeft", TestJoinBufferedRunsLeft},
        {"JoinBufferedRunsAnti", TestJoinBufferedRunsAnti},
        {"JoinMultiPartitionRouted", TestJoinMultiPartitionRouted},
        {"JoinMultiPartitionUnrouted", TestJoinMultiPartitionUnrouted},
        {"ByteKeysThroughOperators", TestByteKeysThroughOperators},
    };
    for (auto &t : tests) {
        int before = g_failures;
        std::printf("[ RUN  ] %s\n", t.first);
        std::fflush(stdout);
        try {
            t.second();
        } catch (const std::exception &e) {
            std::printf("  exception: %s\n", e.what());
            ++g_failures;
        }
        std::printf("[ %s ] %s\n", g_failures == before ? " OK " : "FAIL", t.first);
        std::fflush(stdout);
    }
    std::printf("%d failure(s)\n", g_failures);
    return g_failures == 0 ? 0 : 1;
}
