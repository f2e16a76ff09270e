"""The radix-partitioned group-by (OMNI_GPU_STRATEGY_RADIX, csrc/agg_radix.h) through the C ABI against the CPU
oracle: composite, DECIMAL128 and single keys with NULLs and the slot-sentinel key value, millions of groups, heavy
hitters, the overflow rounds (a capacity hint far below the true group count leaves pieces with more groups than
their LDS table), batches added after the groups were read, one add_batches call.  Bit-exact for integer, decimal,
COUNT, MIN and MAX results; double SUM / AVG within 1e-12 relative."""
import numpy as np
import pytest

import oracle
from oracle import (AVG, COUNT_ALL, COUNT_COLUMN, DECIMAL64, DECIMAL128, DOUBLE, INT, LONG, MAX, MIN, SHORT, SUM,
                    Col)
from omnioperator_amd import capi, workloads
from tests.helpers import assert_row_sets_equal, rows_of
from tests.test_gpu_agg import DOUBLE_RTOL, float_cols, run_gpu, run_oracle

pytestmark = pytest.mark.gpu

AGGS = [(SUM, LONG, LONG), (COUNT_ALL, LONG, LONG), (COUNT_COLUMN, LONG, LONG), (MIN, LONG, LONG), (MAX, LONG, LONG),
        (MAX, INT, INT), (MIN, INT, INT), (SUM, DECIMAL64, DECIMAL128, 2, 2), (SUM, DOUBLE, DOUBLE),
        (AVG, DOUBLE, DOUBLE), (AVG, LONG, DOUBLE), (SUM, INT, LONG)]


def agg_inputs(rng, n, with_nulls=True):
    nl = (lambda: rng.random(n) < 0.15) if with_nulls else (lambda: None)
    vals = rng.integers(-2**62, 2**62, n, dtype=np.int64)
    i32 = rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)
    dbl = rng.normal(100, 50, n)
    return [Col(LONG, vals, nl()), None, Col(LONG, vals, nl()), Col(LONG, vals, nl()), Col(LONG, vals, nl()),
            Col(INT, i32, nl()), Col(INT, i32, nl()), Col(DECIMAL64, rng.integers(-10**12, 10**12, n, dtype=np.int64),
                                                         nl(), 2),
            Col(DOUBLE, dbl, nl()), Col(DOUBLE, dbl, nl()), Col(LONG, rng.integers(-1000, 1000, n, dtype=np.int64),
                                                                nl()),
            Col(INT, i32, nl())]


def check(keys, aggs, cols, batch_rows, hint=0):
    got = run_gpu(keys, aggs, cols, batch_rows, capi.STRATEGY_RADIX, hint)
    want = run_oracle(keys, aggs, cols, batch_rows)
    assert_row_sets_equal(got, want, float_cols=float_cols(len(keys), aggs), rtol=DOUBLE_RTOL)
    return len(want)


def composite_keys(rng, n, distinct, nulls=True, zipf=None):
    if zipf:
        ids = np.minimum(rng.zipf(zipf, n) - 1, distinct - 1)
    else:
        ids = rng.integers(0, distinct, n)
    k1 = (ids % 997).astype(np.int32) - 400
    with np.errstate(over="ignore"):
        k2 = (ids.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)).view(np.int64)
    cols = [Col(INT, k1), Col(LONG, k2)]
    if nulls:
        cols[0].nulls = rng.random(n) < 0.03
        cols[1].nulls = rng.random(n) < 0.03
    return cols


@pytest.mark.parametrize("distinct", [50, 20_000, 600_000])
def test_composite_key_parity(distinct):
    rng = np.random.default_rng(distinct)
    n = 900_000
    keys = composite_keys(rng, n, distinct)
    groups = check(keys, AGGS, agg_inputs(rng, n), 65_536, hint=distinct)
    assert groups > min(distinct, 50)


def test_skewed_composite_key_with_heavy_hitters():
    rng = np.random.default_rng(7)
    n = 1_500_000
    keys = composite_keys(rng, n, 2_000_000, zipf=1.1)
    check(keys, AGGS, agg_inputs(rng, n), 1 << 20, hint=500_000)


def test_overflow_rounds_when_the_hint_is_far_too_low():
    """2M distinct keys planned as 1,000 groups: the 256 pieces hold thousands of groups each, far more than their
    LDS tables, so most rows go through the overflow rounds on lower hash bits."""
    rng = np.random.default_rng(11)
    n = 2_000_000
    keys = [Col(LONG, rng.permutation(n).astype(np.int64) * 7919 - 10**12), Col(SHORT, (np.arange(n) % 5).astype(np.int16))]
    aggs = [(SUM, LONG, LONG), (COUNT_ALL, LONG, LONG), (MAX, INT, INT)]
    cols = [Col(LONG, rng.integers(-10**9, 10**9, n, dtype=np.int64)), None,
            Col(INT, rng.integers(-100, 100, n, dtype=np.int64).astype(np.int32))]
    assert check(keys, aggs, cols, 1 << 19, hint=1000) == n


@pytest.mark.parametrize("key_type", [LONG, INT, SHORT])
def test_single_key_with_null_and_sentinel(key_type):
    rng = np.random.default_rng(30 + key_type)
    n = 400_000
    info = np.iinfo(oracle.np_dtype(key_type))
    k = rng.integers(info.min, info.max, n, dtype=np.int64) % 150_000
    if key_type == LONG:
        k[::97] = -2**63  # the value the hash table keeps as its empty-slot sentinel
    k = k.astype(oracle.np_dtype(key_type))
    keys = [Col(key_type, k, rng.random(n) < 0.02)]
    check(keys, AGGS, agg_inputs(rng, n), 50_000, hint=150_000)


def test_decimal128_key():
    rng = np.random.default_rng(5)
    n = 300_000
    ids = rng.integers(0, 80_000, n)
    k = np.stack([(ids * 2654435761).astype(np.uint64), (ids % 5 - 2).astype(np.int64).view(np.uint64)], 1)
    keys = [Col(DECIMAL128, k, rng.random(n) < 0.01)]
    check(keys, AGGS, agg_inputs(rng, n), 65_536, hint=80_000)


def test_batches_after_the_groups_were_read():
    """Rows added after num_groups go to the hash table, which first takes every group the radix pass made."""
    rng = np.random.default_rng(17)
    n = 400_000
    keys = composite_keys(rng, n, 100_000)
    cols = agg_inputs(rng, n)
    spec = capi.make_agg_spec([c.type_id for c in keys], AGGS, True, capi.STRATEGY_RADIX, 100_000)
    h = capi.HashAgg(spec)
    keep = []

    def add(a, b):
        kd = [capi.DeviceColumn(c.type_id, c.values[a:b], None if c.nulls is None else c.nulls[a:b]) for c in keys]
        ad = [None if c is None else capi.DeviceColumn(c.type_id, c.values[a:b], None if c.nulls is None else c.nulls[a:b])
              for c in cols]
        keep.append((kd, ad))
        h.add_batch([d.c() for d in kd], [None if d is None else d.c() for d in ad], b - a)

    add(0, 150_000)
    add(150_000, 250_000)
    first = h.num_groups()
    assert first > 0
    add(250_000, n)
    outs = h.output()
    h.close()
    types = [c.type_id for c in keys] + [a[2] for a in AGGS]
    got = rows_of([(t, v, nl) for t, (v, nl) in zip(types, outs)])
    want = run_oracle(keys, AGGS, cols, 100_000)
    assert_row_sets_equal(got, want, float_cols=float_cols(len(keys), AGGS), rtol=DOUBLE_RTOL)


def test_add_batches_call_and_wait():
    rng = np.random.default_rng(23)
    n = 300_000
    keys = composite_keys(rng, n, 60_000)
    cols = agg_inputs(rng, n)
    spec = capi.make_agg_spec([c.type_id for c in keys], AGGS, True, capi.STRATEGY_RADIX, 60_000)
    h = capi.HashAgg(spec)
    keep, batches = [], []
    for a, b in workloads.batches(n, 40_000):
        kd = [capi.DeviceColumn(c.type_id, c.values[a:b], None if c.nulls is None else c.nulls[a:b]) for c in keys]
        ad = [None if c is None else capi.DeviceColumn(c.type_id, c.values[a:b], None if c.nulls is None else c.nulls[a:b])
              for c in cols]
        keep.append((kd, ad))
        batches.append(([d.c() for d in kd], [None if d is None else d.c() for d in ad], b - a))
    h.add_batches(batches)
    h.wait()  # the batches are read here: the caller's columns may go afterwards
    del keep
    outs = h.output()
    h.close()
    types = [c.type_id for c in keys] + [a[2] for a in AGGS]
    got = rows_of([(t, v, nl) for t, (v, nl) in zip(types, outs)])
    want = run_oracle(keys, AGGS, cols, 40_000)
    assert_row_sets_equal(got, want, float_cols=float_cols(len(keys), AGGS), rtol=DOUBLE_RTOL)


def test_empty_input_and_rejected_aggregates():
    spec = capi.make_agg_spec([INT, LONG], [(SUM, LONG, LONG)], True, capi.STRATEGY_RADIX, 0)
    h = capi.HashAgg(spec)
    assert h.num_groups() == 0
    h.close()
    for aggs in ([(SUM, DECIMAL64, DECIMAL64, 2, 2)], [(MIN, DOUBLE, DOUBLE)], [(MAX, DOUBLE, DOUBLE)]):
        with pytest.raises(capi.GpuError):
            capi.HashAgg(capi.make_agg_spec([INT, LONG], aggs, True, capi.STRATEGY_RADIX, 0))
