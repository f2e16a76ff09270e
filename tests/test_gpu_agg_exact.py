"""Order-dependent aggregate results on the GPU against the CPU oracle, bit-exact:

- decimal SUM overflow: the reference adds in row order with a checked add and the first add that leaves the
  limits makes the group OVERFLOWED for good (AddDecimalUseRowIndex / SumOp, sum_spark_decimal_aggregator.h:25-62,
  sum_aggregator.h:12-54), then rescales against the DECIMAL128 limits and truncates a DECIMAL64 result
  (sum_spark_decimal_aggregator.h:246-301, 580-589);
- double MIN / MAX: std::min / std::max over the rows in order (max_aggregator.h:39-44, min_aggregator.h:40-44) --
  NaN inputs never replace the state, among equal inputs the first one stays (the sign of a zero result), and the
  state starts at +-DBL_MAX, so a group of only NaN or only -inf inputs gives -DBL_MAX for MAX.

Double results are compared by bit pattern (0.0 == -0.0 in Python)."""
import numpy as np
import pytest

import oracle
from oracle import COUNT_ALL, DECIMAL64, DECIMAL128, DOUBLE, LONG, MAX, MIN, SUM, Col, brute
from omnioperator_amd import capi, workloads
from tests.helpers import assert_row_sets_equal, rows_of
from tests.test_gpu_agg import run_gpu, run_oracle

pytestmark = pytest.mark.gpu

I64_MAX = 2**63 - 1
DEC38 = 10**38 - 1


def dec_col(t, vals):
    return brute.int_to_dec128(vals) if t == DECIMAL128 else np.array(vals, np.int64)


def run_gpu_batches(key_cols, aggs, agg_cols, batch_rows, overflow_as_null=True, strategy=capi.STRATEGY_AUTO):
    """run_gpu through one omni_gpu_agg_add_batches call (the staged path)."""
    spec = capi.make_agg_spec([c.type_id for c in key_cols], aggs, overflow_as_null, strategy, 0)
    h = capi.HashAgg(spec)
    keep, batches = [], []
    for a, b in workloads.batches(len(key_cols[0]), batch_rows):
        kd = [capi.DeviceColumn(c.type_id, c.values[a:b], None if c.nulls is None else c.nulls[a:b]) for c in key_cols]
        ad = [None if c is None else capi.DeviceColumn(c.type_id, c.values[a:b], None if c.nulls is None else c.nulls[a:b])
              for c in agg_cols]
        keep.append((kd, ad))
        batches.append(([d.c() for d in kd], [None if d is None else d.c() for d in ad], b - a))
    h.add_batches(batches)
    outs = h.output()
    types = [c.type_id for c in key_cols] + [a[2] for a in aggs]
    h.close()
    return rows_of([(t, v, nl) for t, (v, nl) in zip(types, outs)])


SEQS = [
    (DECIMAL64, DECIMAL64, [I64_MAX, 1, -2], None),
    (DECIMAL64, DECIMAL64, [I64_MAX, -2, 1], I64_MAX - 1),
    (DECIMAL64, DECIMAL64, [-2, I64_MAX, 1], I64_MAX - 1),
    (DECIMAL64, DECIMAL64, [-I64_MAX - 1, -1, 5], None),
    (DECIMAL64, DECIMAL64, [-I64_MAX - 1, 5, -1], -I64_MAX + 3),
    (DECIMAL128, DECIMAL128, [DEC38, 1, -2], None),
    (DECIMAL128, DECIMAL128, [DEC38, -2, 1], DEC38 - 1),
    (DECIMAL128, DECIMAL128, [-DEC38, -1, 3, 7], None),
    (DECIMAL128, DECIMAL128, [-DEC38, 3, -1, 7], -DEC38 + 9),
    (DECIMAL64, DECIMAL128, [I64_MAX, 1, -2], I64_MAX - 1),
]


@pytest.mark.parametrize("batch_rows", [1, 2, 1024])
@pytest.mark.parametrize("case", range(len(SEQS)))
def test_decimal_sequences_closed_form(case, batch_rows):
    """The row-order results that the reference's checked add fixes by hand, one group per case, with the rows in
    one launch or spread over several (the order across and within launches both count)."""
    tin, tout, seq, expect = SEQS[case]
    aggs = [(SUM, tin, tout, 0, 0)]
    k = [Col(LONG, np.zeros(len(seq), np.int64))]
    v = [Col(tin, dec_col(tin, seq))]
    assert run_gpu(k, aggs, v, batch_rows) == [(0, expect)]
    if expect is None:
        with pytest.raises(capi.GpuError) as e:
            run_gpu(k, aggs, v, batch_rows, overflow_as_null=False)
        assert e.value.code == -4


def test_decimal64_rescale_truncates():
    """SUM(decimal64 scale 0 -> decimal64 scale 2) of 10^18: MulCheckedOverflow against the DECIMAL128 limits passes,
    and static_cast<int64_t>(10^20) = 10^20 - 5 * 2^64."""
    got = run_gpu([Col(LONG, np.zeros(1, np.int64))], [(SUM, DECIMAL64, DECIMAL64, 0, 2)],
                  [Col(DECIMAL64, np.array([10**18], np.int64))], 16)
    assert got == [(0, 10**20 - 5 * 2**64)]
    got = run_gpu([Col(LONG, np.zeros(2, np.int64))], [(SUM, DECIMAL128, DECIMAL128, 0, 2)],
                  [Col(DECIMAL128, brute.int_to_dec128([10**37, 5]))], 16)
    assert got == [(0, None)]


def near_limit_case(seed, tin, n_groups, rows, nulls):
    """Rows of n_groups groups whose values are large steps of either sign (a group's running sum leaves the limits
    in some orders and not in others), a few small values, and NULL inputs."""
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, n_groups, rows, dtype=np.int64) * 1_000_003 - 17
    if tin == DECIMAL64:
        big = rng.integers(2**60, 2**62, rows, dtype=np.int64)
        small = rng.integers(-1000, 1000, rows, dtype=np.int64)
        sign = np.where(rng.random(rows) < 0.5, 1, -1)
        v = np.where(rng.random(rows) < 0.8, sign * big, small).astype(np.int64)
        col = Col(DECIMAL64, v)
    else:
        step = [int(x) * 10**19 for x in rng.integers(10**17, 4 * 10**18, rows)]
        sign = rng.random(rows) < 0.5
        small = rng.random(rows) < 0.2
        vals = [(int(s) % 1000 if sm else (x if sg else -x)) for x, sg, sm, s in
                zip(step, sign, small, rng.integers(0, 10**6, rows))]
        col = Col(DECIMAL128, brute.int_to_dec128(vals))
    if nulls:
        col.nulls = rng.random(rows) < 0.1
    return Col(LONG, keys), col


@pytest.mark.parametrize("tin,tout", [(DECIMAL64, DECIMAL64), (DECIMAL128, DECIMAL128), (DECIMAL64, DECIMAL128)])
@pytest.mark.parametrize("strategy", [capi.STRATEGY_AUTO, capi.STRATEGY_LDS, capi.STRATEGY_GLOBAL])
@pytest.mark.parametrize("groups,rows,batch", [(300, 3000, 256), (4000, 40_000, 5000)])
def test_decimal_near_limit_mixed_signs(tin, tout, strategy, groups, rows, batch):
    """Mixed-sign values near the limits over many groups and launches: every group's result (value, NULL from a
    row-order overflow) equals the oracle's; a few groups overflow in row order while their inputs' total fits, others
    pass their limits part-way in some order without overflowing in row order."""
    key, col = near_limit_case(hash((tin, tout, groups)) % 2**31, tin, groups, rows, nulls=True)
    aggs = [(SUM, tin, tout, 0, 0), (COUNT_ALL, LONG, LONG), (SUM, tin, tout, 0, 1)]
    cols = [col, None, col]
    got = run_gpu([key], aggs, cols, batch, strategy=strategy)
    want = run_oracle([key], aggs, cols, batch)
    assert_row_sets_equal(got, want)
    if tout == DECIMAL64 or tin == DECIMAL128:
        nul = sum(1 for r in want if r[1] is None)
        assert 0 < nul < len(want), "the case must hold groups with and without a row-order overflow"
    # the same rows through one staged launch
    assert_row_sets_equal(run_gpu_batches([key], aggs, cols, batch, strategy=strategy), want)


@pytest.mark.parametrize("tin,tout", [(DECIMAL64, DECIMAL64), (DECIMAL128, DECIMAL128)])
def test_decimal_overflow_raises_like_the_oracle(tin, tout):
    key, col = near_limit_case(77, tin, 30, 2000, nulls=False)
    aggs = [(SUM, tin, tout, 0, 0)]
    with pytest.raises(OverflowError):
        run_oracle([key], aggs, [col], 500, overflow_as_null=False)
    with pytest.raises(capi.GpuError) as e:
        run_gpu([key], aggs, [col], 500, overflow_as_null=False)
    assert e.value.code == -4
    # groups that never overflow raise nothing: only the small values of one group
    k = Col(LONG, np.zeros(100, np.int64))
    small = Col(tin, dec_col(tin, list(range(-50, 50))))
    assert run_gpu([k], aggs, [small], 7, overflow_as_null=False) == run_oracle([k], aggs, [small], 7, False)


def test_batched_stream_rejects_checked_decimal_sums():
    """The batched stream sums DECIMAL64 into DECIMAL128 only (its total can never overflow); the other decimal sums
    need the row-order check of the per-batch path and are refused with OMNI_GPU_ERR_INVALID."""
    for tin, tout in [(DECIMAL64, DECIMAL64), (DECIMAL128, DECIMAL128)]:
        spec = capi.make_agg_spec([LONG], [(SUM, tin, tout, 0, 0)])
        with pytest.raises(capi.GpuError) as e:
            capi.AggBatches(spec, 1024)
        assert e.value.code == -1


SPECIAL = np.array([np.nan, -0.0, 0.0, np.inf, -np.inf, 1.5, -1.5, np.finfo(np.float64).max,
                    -np.finfo(np.float64).max, 5e-324, -5e-324], np.float64)


def special_case(seed, rows, groups):
    """Double columns drawn from NaN, +-0, +-inf, +-DBL_MAX, +-denormal and small values, with groups of only NaN,
    only -inf, only zeros of both signs, and NULLs."""
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, groups, rows, dtype=np.int64)
    v = SPECIAL[rng.integers(0, len(SPECIAL), rows)]
    v = np.where(rng.random(rows) < 0.3, rng.normal(0, 3, rows).round(1), v)
    extra_k = np.repeat(np.array([groups, groups + 1, groups + 2, groups + 3], np.int64), 40)
    extra_v = np.concatenate([np.full(40, np.nan), np.full(40, -np.inf),
                              np.where(rng.random(40) < 0.5, -0.0, 0.0), np.full(40, np.inf)])
    order = rng.permutation(rows + len(extra_k))
    keys = np.concatenate([keys, extra_k])[order]
    v = np.concatenate([v, extra_v])[order]
    return Col(LONG, keys), Col(DOUBLE, v, rng.random(len(v)) < 0.05)


def bits_rows(rows, nkeys):
    """Rows with each double column replaced by its 64-bit pattern (None stays None)."""
    out = []
    for r in rows:
        out.append(tuple(x if (i < nkeys or x is None) else int(np.float64(x).view(np.int64)) for i, x in enumerate(r)))
    return sorted(out, key=lambda r: tuple((x is None, x or 0) for x in r))


@pytest.mark.parametrize("strategy", [capi.STRATEGY_AUTO, capi.STRATEGY_LDS, capi.STRATEGY_GLOBAL])
@pytest.mark.parametrize("batch", [97, 4096])
def test_double_min_max_special_values(strategy, batch):
    key, col = special_case(11 + strategy, 12_000, 60)
    aggs = [(MIN, DOUBLE, DOUBLE), (MAX, DOUBLE, DOUBLE), (COUNT_ALL, LONG, LONG)]
    cols = [col, col, None]
    got = run_gpu([key], aggs, cols, batch, strategy=strategy)
    want = run_oracle([key], aggs, cols, batch)
    assert bits_rows(got, 1) == bits_rows(want, 1)
    by_key = {r[0]: r for r in want}
    dmax = np.finfo(np.float64).max
    assert by_key[60][1] == dmax and by_key[60][2] == -dmax  # only NaN: the initial states stay
    assert by_key[61][2] == -dmax                            # only -inf: MAX stays lowest() = -DBL_MAX
    assert by_key[63][1] == dmax                             # only +inf: MIN stays DBL_MAX
    assert_row_sets_equal(run_gpu_batches([key], aggs, cols, batch, strategy=strategy), got)


def test_double_min_max_zero_sign_follows_row_order():
    """Among zeros the first one in row order decides the sign of a zero MIN / MAX, across launches too."""
    seqs = [[0.0, -0.0], [-0.0, 0.0], [1.0, -0.0, 0.0], [-3.0, 0.0, -0.0, -1.0], [-0.0, -0.0], [np.nan, -0.0, 0.0]]
    keys = np.concatenate([np.full(len(s), i, np.int64) for i, s in enumerate(seqs)])
    vals = np.concatenate([np.array(s, np.float64) for s in seqs])
    aggs = [(MIN, DOUBLE, DOUBLE), (MAX, DOUBLE, DOUBLE)]
    for batch in (1, 2, 64):
        got = run_gpu([Col(LONG, keys)], aggs, [Col(DOUBLE, vals), Col(DOUBLE, vals)], batch)
        want = run_oracle([Col(LONG, keys)], aggs, [Col(DOUBLE, vals), Col(DOUBLE, vals)], batch)
        assert bits_rows(got, 1) == bits_rows(want, 1)
    # std::min / std::max in row order, worked by hand
    expect = {0: (0.0, 0.0), 1: (-0.0, -0.0), 2: (-0.0, 1.0), 3: (-3.0, 0.0), 4: (-0.0, -0.0), 5: (-0.0, -0.0)}
    assert bits_rows(want, 1) == bits_rows([(k, a, b) for k, (a, b) in expect.items()], 1)


def test_double_min_max_batched_stream():
    """The batched stream (one independent group-by per batch) follows the same rules per batch."""
    key, col = special_case(5, 20_000, 300)
    offsets = np.array([0, 1, 700, 5000, 5001, 12_000, 20_160], np.int64)
    aggs = [(MIN, DOUBLE, DOUBLE), (MAX, DOUBLE, DOUBLE)]
    spec = capi.make_agg_spec([LONG], aggs)
    s = capi.AggBatches(spec, int(np.max(np.diff(offsets))))
    kd = capi.DeviceColumn(LONG, key.values, key.nulls)
    vd = capi.DeviceColumn(DOUBLE, col.values, col.nulls)
    for flags in (0, capi.AggBatches.FLAG_PER_BATCH):
        groups = s.run(kd.c(), [vd.c(), vd.c()], offsets, flags=flags)
        outs = s.outputs()
        ospec = oracle.make_agg_spec([LONG], aggs)
        for b in range(len(offsets) - 1):
            a0, a1, g = int(offsets[b]), int(offsets[b + 1]), int(groups[b])
            got = rows_of([(t, v[a0:a0 + g], nl[a0:a0 + g]) for t, (v, nl) in zip([LONG, DOUBLE, DOUBLE], outs)])
            ref = oracle.hash_agg(ospec, [key.slice(a0, a1)], [col.slice(a0, a1), col.slice(a0, a1)], a1 - a0)
            want = rows_of([(c.type_id, c.values, c.nulls) for c in ref])
            assert bits_rows(got, 1) == bits_rows(want, 1)
    s.close()
