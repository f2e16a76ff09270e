// Host-side helpers that map an aggregation spec onto device descriptors and table layouts.
#pragma once

#include <cstring>
#include <vector>

#include "agg_device.h"

namespace omni_gpu {

// Maps aggregate i of the spec onto a device aggregate kind, rejecting combinations the GPU path lacks.
inline int ClassifyAgg(const omni_gpu_agg_spec &sp, int i, AggDev &d)
{
    int f = sp.agg_funcs[i], ti = sp.agg_in_types[i], to = sp.agg_out_types[i];
    d.in_width = TypeWidth(ti);
    d.in_double = ti == OMNI_TYPE_DOUBLE;
    const bool decIn = ti == OMNI_TYPE_DECIMAL64 || ti == OMNI_TYPE_DECIMAL128;
    d.partial_out = ((f == OMNI_AGG_AVG || (f == OMNI_AGG_SUM && decIn)) && sp.agg_partial_out[i] != 0) ? 1 : 0;
    d.out_width = TypeWidth(to);
    d.dec_scale_up = 0;
    d.dec_out128 = 0;
    bool intIn = IsIntegerKeyType(ti) && ti != OMNI_TYPE_DECIMAL64;
    switch (f) {
        case OMNI_AGG_COUNT_ALL:
            OG_CHECK(to == OMNI_TYPE_LONG, OMNI_GPU_ERR_INVALID, "COUNT output must be LONG");
            d.kind = K_COUNT_ALL;
            return OMNI_GPU_OK;
        case OMNI_AGG_COUNT_COLUMN:
            OG_CHECK(to == OMNI_TYPE_LONG, OMNI_GPU_ERR_INVALID, "COUNT output must be LONG");
            d.kind = K_COUNT_COL;
            return OMNI_GPU_OK;
        case OMNI_AGG_SUM:
            if (ti == OMNI_TYPE_DECIMAL64 || ti == OMNI_TYPE_DECIMAL128) {
                OG_CHECK(to == OMNI_TYPE_DECIMAL64 || to == OMNI_TYPE_DECIMAL128, OMNI_GPU_ERR_INVALID,
                         "decimal SUM output must be decimal");
                OG_CHECK(sp.agg_out_scales[i] >= sp.agg_in_scales[i] && sp.agg_out_scales[i] - sp.agg_in_scales[i] <= 38,
                         OMNI_GPU_ERR_INVALID, "decimal SUM output scale must be >= input scale");
                // SumSparkAggregatorFactory (aggregator_factory.cpp:104-116): DECIMAL128 input is summed into a
                // DECIMAL128 state only
                OG_CHECK(!(ti == OMNI_TYPE_DECIMAL128 && to == OMNI_TYPE_DECIMAL64), OMNI_GPU_ERR_INVALID,
                         "SUM(DECIMAL128) output must be DECIMAL128");
                // A DECIMAL64 -> DECIMAL128 sum never leaves +-(10^38 - 1) (at most 2^63 per row), so its exact total
                // is its row-order result; the other two can overflow part-way and are checked in row order.
                d.kind = (ti == OMNI_TYPE_DECIMAL64 && to == OMNI_TYPE_DECIMAL128) ? K_SUM_DEC : K_SUM_DEC_CHK;
                d.dec_scale_up = sp.agg_out_scales[i] - sp.agg_in_scales[i];
                d.dec_out128 = to == OMNI_TYPE_DECIMAL128;
                return OMNI_GPU_OK;
            }
            if (ti == OMNI_TYPE_DOUBLE) {
                OG_CHECK(to == OMNI_TYPE_DOUBLE, OMNI_GPU_ERR_INVALID, "SUM(double) output must be DOUBLE");
                d.kind = K_SUM_F64;
                return OMNI_GPU_OK;
            }
            OG_CHECK(intIn && to == OMNI_TYPE_LONG, OMNI_GPU_ERR_INVALID, "SUM(int) output must be LONG");
            d.kind = K_SUM_I64;
            return OMNI_GPU_OK;
        case OMNI_AGG_AVG:
            OG_CHECK((intIn || ti == OMNI_TYPE_DOUBLE) && to == OMNI_TYPE_DOUBLE, OMNI_GPU_ERR_INVALID,
                     "AVG supports integer/double input with DOUBLE output");
            d.kind = K_AVG;
            return OMNI_GPU_OK;
        case OMNI_AGG_MIN:
        case OMNI_AGG_MAX: {
            bool mn = f == OMNI_AGG_MIN;
            OG_CHECK(ti == to, OMNI_GPU_ERR_INVALID, "MIN/MAX output type must equal input type");
            if (ti == OMNI_TYPE_SHORT || ti == OMNI_TYPE_INT || ti == OMNI_TYPE_DATE32) d.kind = mn ? K_MIN_I32 : K_MAX_I32;
            else if (ti == OMNI_TYPE_LONG || ti == OMNI_TYPE_TIMESTAMP || ti == OMNI_TYPE_DECIMAL64) d.kind = mn ? K_MIN_I64 : K_MAX_I64;
            else if (ti == OMNI_TYPE_DOUBLE) d.kind = mn ? K_MIN_F64 : K_MAX_F64;
            else return SetError(OMNI_GPU_ERR_INVALID, "MIN/MAX over type %d is not supported on the GPU path", ti);
            return OMNI_GPU_OK;
        }
        default: return SetError(OMNI_GPU_ERR_INVALID, "aggregate function %d is not supported", f);
    }
}


// Group state row layout (array of structures, one row per group id): every aggregate's 8-byte words first,
// then the 4-byte MIN/MAX words, then the flag bytes, so each atomic word is naturally aligned; the row is
// padded to a multiple of 8 bytes.  Fills args.a[i].off / flag_off and args.row_bytes.
inline void AssignRowLayout(AggArgs &args)
{
    int off = 0;
    for (int i = 0; i < args.n; ++i) {
        AggDev &d = args.a[i];
        int k = d.kind;
        for (int w = 0; w < 6; ++w) d.off[w] = -1;
        d.flag_off = -1;
        d.cflag_off = -1;
        if (!Is32(k)) {
            d.off[0] = off;
            off += 8;
        }
        const int extra = k == K_SUM_DEC_CHK ? 5 : ExtraWords(k);
        for (int w = 1; w <= extra; ++w) {
            d.off[w] = off;
            off += 8;
        }
    }
    for (int i = 0; i < args.n; ++i) {
        if (Is32(args.a[i].kind)) {
            args.a[i].off[0] = off;
            off += 4;
        }
        if (args.a[i].kind == K_SUM_DEC_CHK) {
            args.a[i].cflag_off = off;
            off += 4;
        }
    }
    for (int i = 0; i < args.n; ++i) {
        int k = args.a[i].kind;
        if (k != K_COUNT_ALL && k != K_COUNT_COL && k != K_AVG) args.a[i].flag_off = off++;
    }
    args.row_bytes = (off + 7) & ~7;
    if (args.row_bytes == 0) args.row_bytes = 8;
}

// Initial image of one state row (MIN/MAX at the type's extreme, everything else zero).
inline std::vector<unsigned long long> RowTemplate(const AggArgs &args)
{
    std::vector<unsigned long long> w(static_cast<size_t>(args.row_bytes / 8), 0ULL);
    uint8_t *b = reinterpret_cast<uint8_t *>(w.data());
    for (int i = 0; i < args.n; ++i) {
        const AggDev &d = args.a[i];
        unsigned long long init = InitWord(d.kind);
        if (Is32(d.kind)) {
            unsigned int v = static_cast<unsigned int>(init);
            std::memcpy(b + d.off[0], &v, 4);
        } else {
            std::memcpy(b + d.off[0], &init, 8);
        }
        if (ExtraWords(d.kind) >= 1 && d.off[1] >= 0) {
            unsigned long long w1 = InitWord1(d.kind);
            std::memcpy(b + d.off[1], &w1, 8);
        }
    }
    return w;
}

// LDS table layout for the aggregates of a handle within `budget` bytes: per slot `key_words` u64 key words (and a
// u32 slot state for multi-word keys) plus each aggregate's state words (4 bytes for 32-bit MIN/MAX) and a flag
// byte where the aggregate needs one.  A checked decimal SUM (K_SUM_DEC_CHK) is never aggregated in LDS: a handle
// that has one takes the global table.
inline LdsLayout MakeLdsLayout(const AggArgs &args, int budget, int key_words = 1)
{
    LdsLayout L{};
    int perSlot = 8 * key_words + (key_words > 1 ? 4 : 0);
    for (int i = 0; i < args.n; ++i) {
        int k = args.a[i].kind;
        perSlot += (Is32(k) ? 4 : 8) + 8 * ExtraWords(k);
        if (k != K_COUNT_ALL && k != K_COUNT_COL && k != K_AVG) perSlot += 1;
    }
    int S = (budget - 64 * (2 * OMNI_GPU_MAX_AGGS + 1)) / perSlot;  // room for 8-byte alignment padding
    S = S - S % 64;
    L.S = S;
    int off = 0;
    auto take = [&](int bytes) {
        int o = off;
        off += (bytes + 7) & ~7;
        return o;
    };
    L.keys_off = take(8 * S * key_words);
    L.ctrl_off = key_words > 1 ? take(4 * S) : -1;
    for (int i = 0; i < args.n; ++i) {
        int k = args.a[i].kind;
        L.off[i][0] = take((Is32(k) ? 4 : 8) * S);
        L.off[i][1] = ExtraWords(k) >= 1 ? take(8 * S) : -1;
        L.off[i][2] = ExtraWords(k) >= 2 ? take(8 * S) : -1;
        L.flag_off[i] = (k != K_COUNT_ALL && k != K_COUNT_COL && k != K_AVG) ? take(S) : -1;
    }
    L.bytes = off;
    return L;
}



}  // namespace omni_gpu
