"""Pins the CPU oracle's hash aggregation: the reference's own fixtures, closed forms, and an
independent sort-based brute force (oracle.brute) on seeded random inputs."""
import numpy as np
import pytest

import oracle
from oracle import brute
from oracle import (AVG, COUNT_ALL, COUNT_COLUMN, DATE32, DECIMAL64, DECIMAL128, DOUBLE, INT, LONG, MAX, MIN, SHORT,
                    SUM, TIMESTAMP, Col)
from tests.helpers import assert_row_sets_equal, load_golden, rows_of

FIX = load_golden("operator_fixtures.json")
TYPES = {"SHORT": SHORT, "INT": INT, "DATE32": DATE32, "LONG": LONG, "TIMESTAMP": TIMESTAMP, "DECIMAL64": DECIMAL64}
STD_AGGS = [(SUM, LONG, LONG), (AVG, LONG, DOUBLE), (COUNT_COLUMN, LONG, LONG), (MIN, LONG, LONG), (MAX, LONG, LONG)]


def cols_rows(outs):
    return rows_of([(c.type_id, c.values, c.nulls) for c in outs])


def test_verify_correctness_fixture():
    f = FIX["agg_verify_correctness"]
    n = f["batches"] * f["rows_per_batch"]
    j = np.tile(np.arange(f["rows_per_batch"], dtype=np.int64) % f["cardinality"], f["batches"])
    keys = [Col(LONG, j), Col(LONG, j.copy())]
    aggs = [Col(LONG, np.ones(n, np.int64)) for _ in range(5)]
    spec = oracle.make_agg_spec([LONG, LONG], STD_AGGS)
    out = oracle.hash_agg(spec, keys, aggs, f["rows_per_batch"])
    e = f["expected"]
    want = [(k, k, e["sum"], e["avg"], e["count"], e["min"], e["max"]) for k in e["keys"]]
    assert_row_sets_equal(cols_rows(out), want, float_cols=[3])


@pytest.mark.parametrize("key_type", ["SHORT", "INT", "DATE32", "LONG", "TIMESTAMP", "DECIMAL64"])
def test_null_key_fixture(key_type):
    f = FIX["agg_null_keys"]
    t = TYPES[key_type]
    r = f["rows_per_batch"]
    keys, knull = [], []
    aggs_v = [[] for _ in range(5)]
    aggs_n = [[] for _ in range(5)]
    for b in range(2):
        k = np.arange(r) % f["cardinality"]
        nl = np.zeros(r, bool)
        nl[f["key_null_rows"][b]] = True
        keys.append(k)
        knull.append(nl)
        for a in range(5):
            aggs_v[a].append(np.ones(r, np.int64))
            m = np.zeros(r, bool)
            m[f["agg_null_rows"][str(a + 1)][b]] = True
            aggs_n[a].append(m)
    kc = Col(t, np.concatenate(keys).astype(oracle.np_dtype(t)), np.concatenate(knull))
    ac = [Col(LONG, np.concatenate(aggs_v[a]), np.concatenate(aggs_n[a])) for a in range(5)]
    spec = oracle.make_agg_spec([t], STD_AGGS)
    out = oracle.hash_agg(spec, [kc], ac, r)
    assert_row_sets_equal(cols_rows(out), [tuple(x) for x in f["expected_rows"]], float_cols=[2])
    # the brute force agrees on the same input
    assert_row_sets_equal(cols_rows(brute.group_by([kc], STD_AGGS, ac)), [tuple(x) for x in f["expected_rows"]],
                          float_cols=[2])


@pytest.mark.parametrize("case", range(4))
def test_sum_long_wraps(case):
    c = FIX["sum_long"]["cases"][case]
    v = np.array(c["values"], np.int64)
    out = oracle.hash_agg(oracle.make_agg_spec([LONG], [(SUM, LONG, LONG)]), [Col(LONG, np.zeros(len(v), np.int64))],
                          [Col(LONG, v)], 1024)
    assert int(out[1].values[0]) == c["sum"]


def test_sum_decimal64_to_decimal128_fixture():
    f = FIX["sum_decimal64"]
    v = np.array(f["values"], np.int64)
    spec = oracle.make_agg_spec([LONG], [(SUM, DECIMAL64, DECIMAL128, f["in_scale"], f["out_scale"])])
    out = oracle.hash_agg(spec, [Col(LONG, np.zeros(3, np.int64))], [Col(DECIMAL64, v, scale=f["in_scale"])], 1024)
    got = rows_of([(c.type_id, c.values, c.nulls) for c in out])
    assert got == [(0, f["sum"])]
    out1 = oracle.hash_agg(spec, [Col(LONG, np.zeros(1, np.int64))], [Col(DECIMAL64, v[:1], scale=6)], 1024)
    assert rows_of([(c.type_id, c.values, c.nulls) for c in out1]) == [(0, f["sum_first_row"])]


def test_decimal_overflow_semantics():
    """SUM(decimal64 -> decimal64): a row-order partial sum leaving int64 makes the group NULL (or raises)."""
    v = np.array([2**62, 2**62, -2**62, 5, 7], np.int64)
    k = np.array([0, 0, 0, 1, 1], np.int64)
    aggs = [(SUM, DECIMAL64, DECIMAL64, 0, 0)]
    out = oracle.hash_agg(oracle.make_agg_spec([LONG], aggs, overflow_as_null=True), [Col(LONG, k)],
                          [Col(DECIMAL64, v)], 1024)
    assert sorted(cols_rows(out), key=lambda r: r[0]) == [(0, None), (1, 12)]
    with pytest.raises(OverflowError):
        oracle.hash_agg(oracle.make_agg_spec([LONG], aggs, overflow_as_null=False), [Col(LONG, k)],
                        [Col(DECIMAL64, v)], 1024)
    # The rescale at extraction is MulCheckedOverflow against the DECIMAL128 limits only; a DECIMAL64 output then
    # keeps static_cast<int64_t> of the result (sum_spark_decimal_aggregator.h:284-296, 580-589): 10^18 * 10^2 =
    # 10^20, whose low 64 bits are 10^20 - 5 * 2^64 = 7766279631452241920.
    out = oracle.hash_agg(oracle.make_agg_spec([LONG], [(SUM, DECIMAL64, DECIMAL64, 0, 2)]),
                          [Col(LONG, np.zeros(1, np.int64))], [Col(DECIMAL64, np.array([10**18], np.int64))], 16)
    assert cols_rows(out) == [(0, 7766279631452241920)]
    assert 10**20 - 5 * 2**64 == 7766279631452241920
    # past the DECIMAL128 limits the rescale overflows: 10^37 * 10^2 > 10^38 - 1
    big = brute.int_to_dec128([10**37])
    out = oracle.hash_agg(oracle.make_agg_spec([LONG], [(SUM, DECIMAL128, DECIMAL128, 0, 2)]),
                          [Col(LONG, np.zeros(1, np.int64))], [Col(DECIMAL128, big)], 16)
    assert cols_rows(out) == [(0, None)]


I64_MAX = 2**63 - 1
DEC38 = 10**38 - 1


@pytest.mark.parametrize("tin,tout,seq,expect", [
    # SUM(decimal64 -> decimal64) adds with __builtin_add_overflow in row order and the OVERFLOWED state sticks
    # (AddDecimalUseRowIndex / SumOp, sum_spark_decimal_aggregator.h:25-62, sum_aggregator.h:36-54): the same
    # values in another order may not overflow
    (DECIMAL64, DECIMAL64, [I64_MAX, 1, -2], None),
    (DECIMAL64, DECIMAL64, [I64_MAX, -2, 1], I64_MAX - 1),
    (DECIMAL64, DECIMAL64, [-2, I64_MAX, 1], I64_MAX - 1),
    (DECIMAL64, DECIMAL64, [-I64_MAX - 1, -1, 5], None),
    (DECIMAL64, DECIMAL64, [-I64_MAX - 1, 5, -1], -I64_MAX + 3),
    # SUM(decimal128 -> decimal128): AddCheckedOverflow against +-(10^38 - 1), sticky
    (DECIMAL128, DECIMAL128, [DEC38, 1, -2], None),
    (DECIMAL128, DECIMAL128, [DEC38, -2, 1], DEC38 - 1),
    (DECIMAL128, DECIMAL128, [-DEC38, -1, 3, 7], None),
    (DECIMAL128, DECIMAL128, [-DEC38, 3, -1, 7], -DEC38 + 9),
    # decimal64 inputs into a decimal128 sum cannot leave +-(10^38 - 1)
    (DECIMAL64, DECIMAL128, [I64_MAX, 1, -2], I64_MAX - 1),
])
def test_decimal_sum_sticky_overflow_in_row_order(tin, tout, seq, expect):
    vals = brute.int_to_dec128(seq) if tin == DECIMAL128 else np.array(seq, np.int64)
    aggs = [(SUM, tin, tout, 0, 0)]
    out = oracle.hash_agg(oracle.make_agg_spec([LONG], aggs, overflow_as_null=True),
                          [Col(LONG, np.zeros(len(seq), np.int64))], [Col(tin, vals)], 2)
    assert cols_rows(out) == [(0, expect)]
    if expect is None:
        with pytest.raises(OverflowError):
            oracle.hash_agg(oracle.make_agg_spec([LONG], aggs, overflow_as_null=False),
                            [Col(LONG, np.zeros(len(seq), np.int64))], [Col(tin, vals)], 2)
    # the sort-based brute force follows the same row-order rule
    b = brute.group_by([Col(LONG, np.zeros(len(seq), np.int64))], aggs, [Col(tin, vals)], overflow_as_null=True)
    assert cols_rows(b) == [(0, expect)]


def _random_case(seed, n, key_kind, with_nulls):
    rng = np.random.default_rng(seed)
    if key_kind == "long":
        keys = [Col(LONG, rng.integers(-50, 50, n, dtype=np.int64))]
    elif key_kind == "int":
        keys = [Col(INT, rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32) % 97)]
    elif key_kind == "composite":
        keys = [Col(INT, rng.integers(0, 7, n, dtype=np.int64).astype(np.int32)),
                Col(LONG, rng.integers(-2**63, 2**63 - 1, 13, dtype=np.int64)[rng.integers(0, 13, n)])]
    else:  # double key (serialize path, bit-pattern equality)
        keys = [Col(DOUBLE, rng.integers(0, 20, n).astype(np.float64) / 4)]
    if with_nulls:
        for c in keys:
            c.nulls = rng.random(n) < 0.1
    vals = rng.integers(-2**62, 2**62, n, dtype=np.int64)
    i32 = rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)
    dec = rng.integers(-10**12, 10**12, n, dtype=np.int64)
    dbl = rng.normal(100, 50, n)
    aggs = [(SUM, LONG, LONG), (COUNT_ALL, LONG, LONG), (COUNT_COLUMN, LONG, LONG), (MIN, LONG, LONG),
            (MAX, LONG, LONG), (MAX, INT, INT), (MIN, INT, INT), (SUM, DECIMAL64, DECIMAL128, 2, 2),
            (SUM, DECIMAL64, DECIMAL64, 2, 3), (SUM, DOUBLE, DOUBLE), (AVG, DOUBLE, DOUBLE), (AVG, LONG, DOUBLE),
            (SUM, INT, LONG)]
    nl = (lambda: rng.random(n) < 0.2) if with_nulls else (lambda: None)
    cols = [Col(LONG, vals, nl()), None, Col(LONG, vals, nl()), Col(LONG, vals, nl()), Col(LONG, vals, nl()),
            Col(INT, i32, nl()), Col(INT, i32, nl()), Col(DECIMAL64, dec, nl(), 2), Col(DECIMAL64, dec, nl(), 2),
            Col(DOUBLE, dbl, nl()), Col(DOUBLE, dbl, nl()), Col(LONG, rng.integers(-1000, 1000, n, dtype=np.int64), nl()),
            Col(INT, i32, nl())]
    return keys, aggs, cols


@pytest.mark.parametrize("key_kind", ["long", "int", "composite", "double"])
@pytest.mark.parametrize("with_nulls", [False, True])
def test_oracle_matches_brute_force(key_kind, with_nulls):
    keys, aggs, cols = _random_case(hash((key_kind, with_nulls)) % 2**32, 5000, key_kind, with_nulls)
    out = oracle.hash_agg(oracle.make_agg_spec([c.type_id for c in keys], aggs), keys, cols, 777)
    ref = brute.group_by(keys, aggs, cols)
    nk = len(keys)
    float_cols = [nk + i for i, a in enumerate(aggs) if a[2] == DOUBLE]
    assert_row_sets_equal(cols_rows(out), cols_rows(ref), float_cols=float_cols, rtol=1e-9)


def test_swiss_map_growth_and_null_slot():
    """BaseHashMap: degree 15 start, rehash once size exceeds ceil(0.9 * cap) before an insert, +2 degrees
    below 23 (base_hash_map.h:255-292, 528-547); the NULL key is a separate slot counted in the size."""
    m = oracle.SwissMap(15)
    assert m.capacity() == 2**15
    thr = int(np.ceil(2**15 * 0.9))
    for i in range(thr + 1):
        assert m.emplace(i * 7919, i)
    assert m.capacity() == 2**15
    assert m.emplace(-1, 10**6)
    assert m.capacity() == 2**17
    assert not m.emplace(7919 * 5, 0)
    assert m.emplace(None, 123) and not m.emplace(None, 5)
    assert m.size() == thr + 3
    vals = m.values_in_order()
    assert vals[-1] == 123 and sorted(vals[:-1].tolist()) == sorted(list(range(thr + 1)) + [10**6])


def test_partial_avg_output_definition():
    """The oracle's partial AVG column is the sum of the group's non-NULL inputs, 0 when there are none
    (average_flatim_aggregator.h:256-273), while the final AVG is NULL there."""
    rng = np.random.default_rng(3)
    n = 4000
    k = rng.integers(0, 40, n, dtype=np.int64)
    v = rng.integers(-1000, 1000, n, dtype=np.int64)
    nulls = rng.random(n) < 0.2
    nulls[k == 7] = True
    spec = oracle.make_agg_spec([LONG], [(AVG, LONG, DOUBLE, 0, 0, True), (AVG, LONG, DOUBLE)])
    outs = oracle.hash_agg(spec, [Col(LONG, k)], [Col(LONG, v, nulls), Col(LONG, v, nulls)], 512)
    keys, psum, mean = outs[0].values, outs[1], outs[2]
    for i, g in enumerate(keys.tolist()):
        sel = (k == g) & ~nulls
        assert not psum.nulls[i] and psum.values[i] == float(v[sel].sum())
        assert mean.nulls[i] == (sel.sum() == 0)


@pytest.mark.parametrize("tin,tout", [(DECIMAL64, DECIMAL64), (DECIMAL128, DECIMAL128), (DECIMAL64, DECIMAL128)])
def test_decimal_near_limit_oracle_equals_brute(tin, tout):
    """Mixed-sign values near the limits over many groups and batches: the oracle's Swiss-table aggregation and the
    sort-based brute force apply the same row-order checked add and give the same rows."""
    rng = np.random.default_rng(tin * 10 + tout)
    n, groups = 6000, 600
    k = Col(LONG, rng.integers(0, groups, n, dtype=np.int64))
    sign = np.where(rng.random(n) < 0.5, 1, -1)
    if tin == DECIMAL64:
        v = Col(DECIMAL64, np.where(rng.random(n) < 0.8, sign * rng.integers(2**60, 2**62, n, dtype=np.int64),
                                    rng.integers(-1000, 1000, n, dtype=np.int64)).astype(np.int64),
                rng.random(n) < 0.1)
    else:
        vals = [int(s) * int(x) * 10**19 for s, x in zip(sign, rng.integers(10**17, 4 * 10**18, n))]
        v = Col(DECIMAL128, brute.int_to_dec128(vals), rng.random(n) < 0.1)
    aggs = [(SUM, tin, tout, 0, 0), (SUM, tin, tout, 0, 2)]
    out = oracle.hash_agg(oracle.make_agg_spec([LONG], aggs), [k], [v, v], 700)
    got = cols_rows(out)
    want = cols_rows(brute.group_by([k], aggs, [v, v]))
    assert_row_sets_equal(got, want)
    if tout == DECIMAL64 or tin == DECIMAL128:
        assert 0 < sum(r[1] is None for r in got) < len(got)


def test_agg_stream_equals_one_operator_per_batch():
    """oracle.agg_stream (one operator per batch, batches shared by host threads) gives each batch the group count of
    its own hash_agg run, with one thread and with several; NULL keys and values at unaligned batch starts."""
    rng = np.random.default_rng(44)
    rows = rng.integers(0, 700, 60)
    rows[:2] = [0, 1]
    offsets = np.concatenate([[0], np.cumsum(rows)]).astype(np.int64)
    n = int(offsets[-1])
    key = Col(LONG, rng.integers(0, 90, n, dtype=np.int64), rng.random(n) < 0.05)
    val = Col(LONG, rng.integers(-10**6, 10**6, n, dtype=np.int64), rng.random(n) < 0.1)
    aggs = [(SUM, LONG, LONG), (COUNT_ALL, LONG, LONG)]
    spec = oracle.make_agg_spec([LONG], aggs)
    want = [len(oracle.hash_agg(spec, [key.slice(a, b)], [val.slice(a, b), None], max(1, b - a))[0].values)
            if b > a else 0 for a, b in zip(offsets[:-1], offsets[1:])]
    for threads in (1, 4):
        assert oracle.agg_stream(spec, key, [val, None], offsets, threads).tolist() == want


def test_partial_decimal_sum_output():
    """agg_partial_out on a decimal SUM: the sum column of SumSparkDecimalAggregator's partial state
    (sum_spark_decimal_aggregator.h:246-301 with OUTPUT_PARTIAL) -- the value itself for the reference's
    spark_sum_decimal64_normal row (aggregation_test.cpp:3680-3715: 999999999999.999999 from DECIMAL64(18, 6) into
    DECIMAL128(28, 6)), 0 (not NULL) for a group of NULL inputs, NULL for an overflowed group or rescale, and never
    an error."""
    f = FIX["sum_decimal64"]
    v1 = np.array(f["values"][:1], np.int64)
    spec = oracle.make_agg_spec([LONG], [(SUM, DECIMAL64, DECIMAL128, 6, 6, True)])
    out = oracle.hash_agg(spec, [Col(LONG, np.zeros(1, np.int64))], [Col(DECIMAL64, v1, scale=6)], 16)
    assert cols_rows(out) == [(0, f["sum_first_row"])]
    k = np.array([0, 0, 1, 1, 2], np.int64)
    v = np.array([I64_MAX, 1, 3, 4, 9], np.int64)
    nulls = np.array([False, False, True, True, False])
    for ovf_null in (False, True):
        out = oracle.hash_agg(oracle.make_agg_spec([LONG], [(SUM, DECIMAL64, DECIMAL64, 0, 0, True)],
                                                   overflow_as_null=ovf_null),
                              [Col(LONG, k)], [Col(DECIMAL64, v, nulls)], 16)
        assert sorted(cols_rows(out), key=lambda r: r[0]) == [(0, None), (1, 0), (2, 9)]
    # the rescale by 10^(out - in) overflowing the DECIMAL128 limits: NULL in the partial output, no error
    big = brute.int_to_dec128([10**37])
    out = oracle.hash_agg(oracle.make_agg_spec([LONG], [(SUM, DECIMAL128, DECIMAL128, 0, 2, True)]),
                          [Col(LONG, np.zeros(1, np.int64))], [Col(DECIMAL128, big)], 16)
    assert cols_rows(out) == [(0, None)]


def test_byte_keys_group_as_their_values():
    """BYTE (int8) key columns group by their values: the oracle's groups over BYTE keys, alone and beside a LONG
    key, are its groups over the same values stored as SHORT (pinned by the fixtures above), NULL keys included."""
    from oracle import BYTE
    from tests.helpers import rows_of
    rng = np.random.default_rng(18)
    n = 5000
    b = rng.integers(-128, 128, n).astype(np.int8)
    bn = rng.random(n) < 0.05
    l = rng.integers(-3, 3, n).astype(np.int64)
    v = rng.integers(-10**6, 10**6, n).astype(np.int64)
    cols = [Col(LONG, v)] * 5
    for keys_b, keys_s in (([Col(BYTE, b, bn)], [Col(SHORT, b.astype(np.int16), bn)]),
                           ([Col(BYTE, b, bn), Col(LONG, l)], [Col(SHORT, b.astype(np.int16), bn), Col(LONG, l)])):
        got = [oracle.hash_agg(oracle.make_agg_spec([c.type_id for c in k], STD_AGGS, True), k, cols, 1000)
               for k in (keys_b, keys_s)]
        rb, rs = (rows_of([(c.type_id, c.values, c.nulls) for c in outs]) for outs in got)
        assert got[0][0].values.dtype == np.int8
        assert sorted(rb, key=repr) == sorted(rs, key=repr)
        assert len(rb) > 250
