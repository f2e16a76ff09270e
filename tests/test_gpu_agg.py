"""GPU hash aggregation (HIP kernels through the C ABI) against the CPU oracle: bit-exact for integer,
decimal, COUNT, MIN and MAX results; double SUM/AVG within 1e-12 relative (order of atomic additions)."""
import numpy as np
import pytest

import oracle
from oracle import brute
from oracle import (AVG, COUNT_ALL, COUNT_COLUMN, DECIMAL64, DECIMAL128, DOUBLE, INT, LONG, MAX, MIN, SHORT, SUM,
                    Col)
from omnioperator_amd import capi, workloads
from tests.helpers import assert_row_sets_equal, load_golden, rows_of
from tests.test_oracle_agg import I64_MAX, STD_AGGS, TYPES, _random_case

pytestmark = pytest.mark.gpu
FIX = load_golden("operator_fixtures.json")
DOUBLE_RTOL = 1e-12


def run_gpu(key_cols, aggs, agg_cols, batch_rows, strategy=capi.STRATEGY_AUTO, capacity_hint=0,
            overflow_as_null=True):
    spec = capi.make_agg_spec([c.type_id for c in key_cols], aggs, overflow_as_null, strategy, capacity_hint)
    h = capi.HashAgg(spec)
    n = len(key_cols[0])
    keep = []
    for a, b in workloads.batches(n, batch_rows):
        kd = [capi.DeviceColumn(c.type_id, c.values[a:b], None if c.nulls is None else c.nulls[a:b]) for c in key_cols]
        ad = [None if c is None else capi.DeviceColumn(c.type_id, c.values[a:b], None if c.nulls is None else c.nulls[a:b])
              for c in agg_cols]
        keep.append((kd, ad))
        h.add_batch([d.c() for d in kd], [None if d is None else d.c() for d in ad], b - a)
    outs = h.output()
    types = [c.type_id for c in key_cols] + [a[2] for a in aggs]
    h.close()
    return rows_of([(t, v, nl) for t, (v, nl) in zip(types, outs)])


def run_oracle(key_cols, aggs, agg_cols, batch_rows, overflow_as_null=True):
    spec = oracle.make_agg_spec([c.type_id for c in key_cols], aggs, overflow_as_null)
    outs = oracle.hash_agg(spec, key_cols, agg_cols, batch_rows)
    return rows_of([(c.type_id, c.values, c.nulls) for c in outs])


def float_cols(nkeys, aggs):
    return [nkeys + i for i, a in enumerate(aggs) if a[2] == DOUBLE]


def test_verify_correctness_fixture():
    f = FIX["agg_verify_correctness"]
    n = f["batches"] * f["rows_per_batch"]
    j = np.tile(np.arange(f["rows_per_batch"], dtype=np.int64) % f["cardinality"], f["batches"])
    got = run_gpu([Col(LONG, j), Col(LONG, j.copy())], STD_AGGS, [Col(LONG, np.ones(n, np.int64)) for _ in range(5)],
                  f["rows_per_batch"])
    e = f["expected"]
    want = [(k, k, e["sum"], e["avg"], e["count"], e["min"], e["max"]) for k in e["keys"]]
    assert_row_sets_equal(got, want, float_cols=[3], rtol=DOUBLE_RTOL)


@pytest.mark.parametrize("key_type", ["SHORT", "INT", "DATE32", "LONG", "TIMESTAMP", "DECIMAL64"])
def test_null_key_fixture(key_type):
    f = FIX["agg_null_keys"]
    t, r = TYPES[key_type], f["rows_per_batch"]
    k = np.tile(np.arange(r) % f["cardinality"], 2).astype(oracle.np_dtype(t))
    kn = np.zeros(2 * r, bool)
    for b in range(2):
        kn[[b * r + i for i in f["key_null_rows"][b]]] = True
    ac = []
    for a in range(5):
        m = np.zeros(2 * r, bool)
        for b in range(2):
            m[[b * r + i for i in f["agg_null_rows"][str(a + 1)][b]]] = True
        ac.append(Col(LONG, np.ones(2 * r, np.int64), m))
    got = run_gpu([Col(t, k, kn)], STD_AGGS, ac, r)
    assert_row_sets_equal(got, [tuple(x) for x in f["expected_rows"]], float_cols=[2], rtol=DOUBLE_RTOL)


@pytest.mark.parametrize("case", range(4))
def test_sum_long_wraps(case):
    c = FIX["sum_long"]["cases"][case]
    v = np.array(c["values"], np.int64)
    got = run_gpu([Col(LONG, np.zeros(len(v), np.int64))], [(SUM, LONG, LONG)], [Col(LONG, v)], 1024)
    assert got == [(0, c["sum"])]


def test_sum_decimal64_fixture():
    f = FIX["sum_decimal64"]
    got = run_gpu([Col(LONG, np.zeros(3, np.int64))], [(SUM, DECIMAL64, DECIMAL128, 6, 6)],
                  [Col(DECIMAL64, np.array(f["values"], np.int64), scale=6)], 1024)
    assert got == [(0, f["sum"])]


def test_decimal_overflow_null_and_error():
    v = np.array([2**62, 2**62, 5, 7, 10**18], np.int64)
    k = np.array([0, 0, 1, 1, 2], np.int64)
    aggs = [(SUM, DECIMAL64, DECIMAL64, 0, 2)]
    got = run_gpu([Col(LONG, k)], aggs, [Col(DECIMAL64, v)], 16)
    # group 2: 10^18 rescaled by 10^2 passes the DECIMAL128 check and keeps its low 64 bits (static_cast<int64_t>)
    assert sorted(got, key=lambda r: r[0]) == [(0, None), (1, 1200), (2, 10**20 - 5 * 2**64)]
    assert_row_sets_equal(got, run_oracle([Col(LONG, k)], aggs, [Col(DECIMAL64, v)], 16))
    with pytest.raises(capi.GpuError) as e:
        run_gpu([Col(LONG, k)], aggs, [Col(DECIMAL64, v)], 16, overflow_as_null=False)
    assert e.value.code == -4


@pytest.mark.parametrize("key_kind", ["long", "int", "composite", "double"])
@pytest.mark.parametrize("with_nulls", [False, True])
def test_random_parity(key_kind, with_nulls):
    keys, aggs, cols = _random_case(hash((key_kind, with_nulls, "gpu")) % 2**32, 20000, key_kind, with_nulls)
    got = run_gpu(keys, aggs, cols, 3000)
    want = run_oracle(keys, aggs, cols, 3000)
    assert_row_sets_equal(got, want, float_cols=float_cols(len(keys), aggs), rtol=DOUBLE_RTOL)


def test_growth_and_rehash():
    """A tiny capacity hint forces several grow-and-rehash rounds; results must not change."""
    rng = np.random.default_rng(5)
    n = 300_000
    k = rng.integers(0, 120_000, n, dtype=np.int64) * 1_000_003
    v = rng.integers(-10**9, 10**9, n, dtype=np.int64)
    aggs = [(SUM, LONG, LONG), (COUNT_ALL, LONG, LONG), (MIN, LONG, LONG)]
    cols = [Col(LONG, v), None, Col(LONG, v)]
    got = run_gpu([Col(LONG, k)], aggs, cols, 65536, capacity_hint=16)
    want = run_oracle([Col(LONG, k)], aggs, cols, 65536)
    assert_row_sets_equal(got, want)


def test_composite_growth():
    rng = np.random.default_rng(6)
    n = 200_000
    k1 = rng.integers(0, 300, n, dtype=np.int64).astype(np.int32)
    k2 = rng.integers(0, 500, n, dtype=np.int64) * 7
    aggs = [(COUNT_ALL, LONG, LONG), (MAX, INT, INT)]
    cols = [None, Col(INT, k1)]
    got = run_gpu([Col(INT, k1), Col(LONG, k2)], aggs, cols, 40000, capacity_hint=8)
    want = run_oracle([Col(INT, k1), Col(LONG, k2)], aggs, cols, 40000)
    assert_row_sets_equal(got, want)


def test_sentinel_and_extreme_keys():
    """INT64_MIN is the slot sentinel internally; as a real key it must still form its own group."""
    k = np.array([-2**63, 2**63 - 1, -2**63, 0, -1, 2**63 - 1], np.int64)
    n = np.array([False, False, False, True, False, False])
    aggs = [(COUNT_ALL, LONG, LONG), (SUM, LONG, LONG)]
    cols = [None, Col(LONG, np.arange(6, dtype=np.int64))]
    got = run_gpu([Col(LONG, k, n)], aggs, cols, 4)
    want = run_oracle([Col(LONG, k, n)], aggs, cols, 4)
    assert_row_sets_equal(got, want)
    assert len(got) == 4


def test_empty_input():
    spec = capi.make_agg_spec([LONG], [(SUM, LONG, LONG)])
    h = capi.HashAgg(spec)
    assert h.num_groups() == 0
    h.close()


def test_plain_agg_config_full_size():
    """1M rows in batches of 4096, int64 key over 1,000 values; SUM/COUNT/MIN/MAX bit-exact."""
    w = workloads.agg_plain()
    aggs = [(SUM, LONG, LONG), (COUNT_ALL, LONG, LONG), (MIN, LONG, LONG), (MAX, LONG, LONG)]
    cols = [Col(LONG, w.values), None, Col(LONG, w.values), Col(LONG, w.values)]
    got = run_gpu([Col(LONG, w.keys)], aggs, cols, w.batch_rows)
    want = run_oracle([Col(LONG, w.keys)], aggs, cols, w.batch_rows)
    assert len(got) == 1000
    assert_row_sets_equal(got, want)


def test_skewed_agg_reduced_size():
    """The skewed composite-key configuration at 2M rows (Zipf over 5e5 ranks): decimal64 -> decimal128 SUM,
    double SUM/AVG, COUNT, MAX(int32)."""
    w = workloads.agg_skewed(rows=2_000_000, k1_distinct=1000, k2_ranks=500_000, seed=0xC3)
    keys = [Col(INT, w.k1), Col(LONG, w.k2)]
    aggs = [(SUM, DECIMAL64, DECIMAL128, 2, 2), (SUM, DOUBLE, DOUBLE), (AVG, DOUBLE, DOUBLE), (COUNT_ALL, LONG, LONG),
            (MAX, INT, INT)]
    cols = [Col(DECIMAL64, w.v_dec, scale=2), Col(DOUBLE, w.v_dbl), Col(DOUBLE, w.v_dbl), None, Col(INT, w.v_i32)]
    got = run_gpu(keys, aggs, cols, 65536)
    want = run_oracle(keys, aggs, cols, 65536)
    assert_row_sets_equal(got, want, float_cols=float_cols(2, aggs), rtol=DOUBLE_RTOL)


@pytest.mark.parametrize("composite", [False, True])
def test_many_groups_chunked_ids(composite):
    """About 1M groups over 3M rows in 12 batches: waves take group ids in chunks once many groups exist, the
    table grows several times, and ids reserved but unused must not surface as groups."""
    rng = np.random.default_rng(21 + composite)
    n = 3_000_000
    k1 = rng.integers(0, 1_000_000, n, dtype=np.int64) * 2654435761 - 5
    k2 = rng.integers(0, 3, n, dtype=np.int64).astype(np.int32)
    v = rng.integers(-10**9, 10**9, n, dtype=np.int64)
    keys = [Col(LONG, k1, rng.random(n) < 0.001)] + ([Col(INT, k2)] if composite else [])
    aggs = [(SUM, LONG, LONG), (COUNT_ALL, LONG, LONG), (MIN, LONG, LONG), (MAX, LONG, LONG)]
    cols = [Col(LONG, v, rng.random(n) < 0.05), None, Col(LONG, v), Col(LONG, v)]
    got = run_gpu(keys, aggs, cols, 250_000, strategy=capi.STRATEGY_GLOBAL)
    want = run_oracle(keys, aggs, cols, 250_000)
    assert len(got) == len(want)
    assert_row_sets_equal(got, want)


@pytest.mark.parametrize("strategy", [capi.STRATEGY_LDS, capi.STRATEGY_GLOBAL, capi.STRATEGY_AUTO])
@pytest.mark.parametrize("groups", [40, 30_000])
def test_composite_key_strategy_parity(strategy, groups):
    """Composite (int32, int64) keys with NULLs in both columns through the LDS table of multi-word keys (hot
    keys aggregated in LDS, the rest spilled to the global table) and through the global table."""
    rng = np.random.default_rng(groups + strategy)
    n = 300_000
    r = np.minimum(rng.zipf(1.3, n), groups) - 1
    k1 = (r % 97).astype(np.int32)
    k2 = (r // 97).astype(np.int64) * 1_000_003 - 11
    keys = [Col(INT, k1, rng.random(n) < 0.01), Col(LONG, k2, rng.random(n) < 0.01)]
    v = rng.integers(-10**12, 10**12, n, dtype=np.int64)
    aggs = [(SUM, LONG, LONG), (COUNT_ALL, LONG, LONG), (MAX, LONG, LONG), (SUM, DECIMAL64, DECIMAL128, 0, 0),
            (AVG, DOUBLE, DOUBLE)]
    # doubles of one sign (N(100, 50)): sums without cancellation, so a relative tolerance applies
    cols = [Col(LONG, v, rng.random(n) < 0.05), None, Col(LONG, v), Col(DECIMAL64, v), Col(DOUBLE, rng.normal(100, 50, n))]
    got = run_gpu(keys, aggs, cols, 100_000, strategy=strategy)
    want = run_oracle(keys, aggs, cols, 100_000)
    assert_row_sets_equal(got, want, float_cols=float_cols(2, aggs), rtol=DOUBLE_RTOL)


@pytest.mark.parametrize("strategy", [capi.STRATEGY_LDS, capi.STRATEGY_GLOBAL])
@pytest.mark.parametrize("key_kind", ["long", "int", "double"])
@pytest.mark.parametrize("with_nulls", [False, True])
def test_strategy_parity(strategy, key_kind, with_nulls):
    """Every strategy gives the oracle's results (LDS-local partial tables merge into the global table)."""
    keys, aggs, cols = _random_case(hash((key_kind, with_nulls, "strategy")) % 2**32, 60000, key_kind, with_nulls)
    got = run_gpu(keys, aggs, cols, 20000, strategy=strategy)
    want = run_oracle(keys, aggs, cols, 20000)
    assert_row_sets_equal(got, want, float_cols=float_cols(len(keys), aggs), rtol=DOUBLE_RTOL)


def test_lds_strategy_spills_when_groups_exceed_lds():
    """More distinct keys than LDS slots: rows that find no LDS slot go to the global table."""
    rng = np.random.default_rng(9)
    n = 400_000
    k = rng.integers(0, 50_000, n, dtype=np.int64)
    v = rng.integers(-10**12, 10**12, n, dtype=np.int64)
    aggs = [(SUM, LONG, LONG), (COUNT_ALL, LONG, LONG), (MAX, LONG, LONG), (SUM, DECIMAL64, DECIMAL128, 0, 0)]
    cols = [Col(LONG, v), None, Col(LONG, v), Col(DECIMAL64, v)]
    got = run_gpu([Col(LONG, k)], aggs, cols, 100_000, strategy=capi.STRATEGY_LDS)
    want = run_oracle([Col(LONG, k)], aggs, cols, 100_000)
    assert_row_sets_equal(got, want)


def _uneven_spans(n, seed):
    """Cuts n rows into uneven batches, empty ones included."""
    rng = np.random.default_rng(seed)
    cuts = np.sort(rng.integers(0, n + 1, 40))
    cuts[5:8] = cuts[5]  # a run of empty batches
    edges = np.concatenate([[0], cuts, [n]])
    return [(int(a), int(b)) for a, b in zip(edges[:-1], edges[1:])]


def _dev_slice(c, a, b):
    """Device copy of rows [a, b) of column c; a slice without NULLs gets no bitmap."""
    nl = None if c.nulls is None or not c.nulls[a:b].any() else c.nulls[a:b]
    return capi.DeviceColumn(c.type_id, c.values[a:b], nl)


@pytest.mark.parametrize("key_kind", ["long", "int", "composite", "double"])
@pytest.mark.parametrize("with_nulls", [False, True])
def test_add_batches_parity(key_kind, with_nulls):
    """omni_gpu_agg_add_batches (one staging launch + one insert for 41 uneven batches, empty ones and batches
    with and without NULL bitmaps mixed) against the oracle: bit-exact, doubles within 1e-12."""
    keys, aggs, cols = _random_case(hash((key_kind, with_nulls, "batches")) % 2**32, 30000, key_kind, with_nulls)
    spec = capi.make_agg_spec([c.type_id for c in keys], aggs, True)
    h = capi.HashAgg(spec)
    keep, batches = [], []
    for a, b in _uneven_spans(len(keys[0]), 7):
        kd = [_dev_slice(c, a, b) for c in keys]
        ad = [None if c is None else _dev_slice(c, a, b) for c in cols]
        keep.append((kd, ad))
        batches.append(([d.c() for d in kd], [None if d is None else d.c() for d in ad], b - a))
    h.add_batches(batches)
    outs = h.output()
    types = [c.type_id for c in keys] + [x[2] for x in aggs]
    got = rows_of([(t, v, nl) for t, (v, nl) in zip(types, outs)])
    h.close()
    want = run_oracle(keys, aggs, cols, 4096)
    assert_row_sets_equal(got, want, float_cols=float_cols(len(keys), aggs), rtol=DOUBLE_RTOL)


def test_add_batches_shared_value_column():
    """The plain group-by shape: SUM, COUNT(*), MIN and MAX over one int64 value column passed for every
    aggregate (staged once), 100,000 rows in batches of 4096 added in two calls; equals the per-batch path."""
    rng = np.random.default_rng(5)
    n = 100_000
    k = rng.integers(0, 1000, n, dtype=np.int64)
    v = rng.integers(-10**6, 10**6 + 1, n, dtype=np.int64)
    aggs = [(SUM, LONG, LONG), (COUNT_ALL, LONG, LONG), (MIN, LONG, LONG), (MAX, LONG, LONG)]
    spec = capi.make_agg_spec([LONG], aggs, True)
    kd, vd = capi.DeviceColumn(LONG, k), capi.DeviceColumn(LONG, v)
    spans = workloads.batches(n, 4096)
    batches = [([kd.c(a)], [vd.c(a), None, vd.c(a), vd.c(a)], b - a) for a, b in spans]
    h = capi.HashAgg(spec)
    h.add_batches(batches[:10])
    h.add_batches(batches[10:])
    outs = h.output()
    got = rows_of([(t, v_, nl) for t, (v_, nl) in zip([LONG] + [a[2] for a in aggs], outs)])
    h.close()
    want = run_oracle([Col(LONG, k)], aggs, [Col(LONG, v), None, Col(LONG, v), Col(LONG, v)], 4096)
    assert len(got) == 1000
    assert_row_sets_equal(got, want)


@pytest.mark.parametrize("in_type", [LONG, DOUBLE])
def test_partial_avg_output(in_type):
    """agg_partial_out: AVG outputs its running double sum (0, not NULL, for a group of only NULL inputs) beside
    a COUNT_COLUMN of the same input -- the two columns of the reference's partial AVG state."""
    rng = np.random.default_rng(17)
    n = 50_000
    k = rng.integers(0, 700, n, dtype=np.int64)
    nulls = rng.random(n) < 0.1
    nulls[k == 5] = True  # group 5: every input NULL
    if in_type == DOUBLE:
        v = Col(DOUBLE, rng.normal(100, 50, n), nulls)
    else:
        v = Col(LONG, rng.integers(-10**6, 10**6, n, dtype=np.int64), nulls)
    aggs = [(AVG, in_type, DOUBLE, 0, 0, True), (COUNT_COLUMN, in_type, LONG), (AVG, in_type, DOUBLE)]
    got = run_gpu([Col(LONG, k)], aggs, [v, v, v], 4096)
    want = run_oracle([Col(LONG, k)], aggs, [v, v, v], 4096)
    assert_row_sets_equal(got, want, float_cols=[1, 3], rtol=DOUBLE_RTOL)
    row5 = [r for r in got if r[0] == 5][0]
    assert row5[1] == 0.0 and row5[2] == 0 and row5[3] is None


@pytest.mark.parametrize("tin,tout", [(DECIMAL64, DECIMAL64), (DECIMAL64, DECIMAL128), (DECIMAL128, DECIMAL128)])
def test_partial_decimal_sum_output(tin, tout):
    """agg_partial_out on a decimal SUM equals the oracle's partial sum column (0 for a group of NULL inputs, NULL
    for an overflowed group) and raises nothing even with overflow_as_null off; near-limit values of both signs
    make row-order overflows in some groups."""
    rng = np.random.default_rng(23)
    n = 40_000
    k = rng.integers(0, 400, n, dtype=np.int64)
    lim = I64_MAX if tout == DECIMAL64 else 10**38 - 1
    big = rng.random(n) < 0.02
    small = rng.integers(-10**6, 10**6, n, dtype=np.int64)
    if tin == DECIMAL64:
        vals = np.where(big, np.where(rng.random(n) < 0.5, I64_MAX // 2, -(I64_MAX // 2)), small).astype(np.int64)
    else:
        ints = [int(lim // 2) * (1 if rng.random() < 0.5 else -1) if b else int(x) for b, x in zip(big, small)]
        vals = brute.int_to_dec128(ints)
    nulls = rng.random(n) < 0.05
    nulls[k == 7] = True  # group 7: every input NULL
    col = Col(tin, vals, nulls)
    aggs = [(SUM, tin, tout, 0, 0, True), (COUNT_COLUMN, tin, LONG)]
    got = run_gpu([Col(LONG, k)], aggs, [col, col], 4096, overflow_as_null=False)
    want = run_oracle([Col(LONG, k)], aggs, [col, col], 4096)
    assert_row_sets_equal(got, want)
    assert [r for r in got if r[0] == 7][0][1:] == (0, 0)
    # some groups overflowed (DECIMAL64 inputs cannot leave a DECIMAL128 sum's limits)
    assert any(r[1] is None for r in got) == (tin == tout)


def run_gpu_rule(key_cols, aggs, agg_cols, batch_rows, strategy=capi.STRATEGY_AUTO):
    """run_gpu under the ARRAY hash-table rule; also returns the table decision."""
    spec = capi.make_agg_spec([c.type_id for c in key_cols], aggs, True, strategy, 0, capi.HASH_TABLE_ARRAY)
    h = capi.HashAgg(spec)
    keep = []
    for a, b in workloads.batches(len(key_cols[0]), batch_rows):
        kd = [capi.DeviceColumn(c.type_id, c.values[a:b], None if c.nulls is None else c.nulls[a:b]) for c in key_cols]
        ad = [None if c is None else capi.DeviceColumn(c.type_id, c.values[a:b], None if c.nulls is None else c.nulls[a:b])
              for c in agg_cols]
        keep.append((kd, ad))
        h.add_batch([d.c() for d in kd], [None if d is None else d.c() for d in ad], b - a)
    info = h.table_info()
    outs = h.output()
    types = [c.type_id for c in key_cols] + [a[2] for a in aggs]
    h.close()
    return rows_of([(t, v, nl) for t, (v, nl) in zip(types, outs)]), info


def expected_array_range(first, t):
    """VectorAnalyzer::CheckArrayMap over the first batch's non-NULL keys: (min - reserve, cells), saturated."""
    lim = {SHORT: (-2**15, 2**15 - 1), INT: (-2**31, 2**31 - 1), LONG: (-2**63, 2**63 - 1)}[t]
    mn, mx = int(first.min()), int(first.max())
    reserve = max(50, mx - mn)
    hi = mx + reserve if mx <= lim[1] - reserve else lim[1]
    lo = mn - reserve if mn >= lim[0] + reserve else lim[0]
    return lo, hi - lo + 1


@pytest.mark.parametrize("t", [SHORT, INT, LONG])
@pytest.mark.parametrize("strategy", [capi.STRATEGY_AUTO, capi.STRATEGY_GLOBAL])
def test_array_rule_parity(t, strategy):
    """ARRAY rule: keys of the first batch plan the direct-index map; later batches reach keys outside it, which go
    to the hash table.  The groups equal the oracle's whatever table found them."""
    rng = np.random.default_rng(31 + t + strategy)
    dt = oracle.np_dtype(t)
    n, batch = 60_000, 8192
    base = -2**15 + 10 if t == SHORT else -700
    k = (base + rng.integers(0, 1500, n)).astype(dt)
    far = rng.random(n) < 0.15
    far[:batch] = False
    k[far] = rng.integers(np.iinfo(dt).min // 2, np.iinfo(dt).max // 2, int(far.sum())).astype(dt)
    knull = rng.random(n) < 0.05
    v = Col(LONG, rng.integers(-10**6, 10**6, n, dtype=np.int64), rng.random(n) < 0.1)
    aggs = [(SUM, LONG, LONG), (COUNT_ALL, LONG, LONG), (MIN, LONG, LONG), (MAX, LONG, LONG)]
    got, info = run_gpu_rule([Col(t, k, knull)], aggs, [v, None, v, v], batch, strategy)
    want = run_oracle([Col(t, k, knull)], aggs, [v, None, v, v], batch)
    assert_row_sets_equal(got, want)
    first = k[:batch][~knull[:batch]]
    lo, cells = expected_array_range(first, t)
    assert info == (True, lo, cells)


def test_array_rule_declines():
    """No direct-index map when the first batch's keys are all NULL, span more than 2^21 - 2 values, or when the
    range widened by max(50, max - min) on each side plus the NULL slot exceeds 2^21 entries
    (vector_analyzer.cpp:46-77); results are unchanged."""
    aggs = [(SUM, LONG, LONG), (COUNT_ALL, LONG, LONG)]
    v = Col(LONG, np.arange(4000, dtype=np.int64))
    k = np.arange(4000, dtype=np.int64) % 37
    nulls = np.zeros(4000, bool)
    nulls[:2048] = True
    got, info = run_gpu_rule([Col(LONG, k, nulls)], aggs, [v, None], 2048)
    assert info == (False, 0, 0)
    assert_row_sets_equal(got, run_oracle([Col(LONG, k, nulls)], aggs, [v, None], 2048))
    k2 = np.array([0, 2**21 - 1] * 1000, np.int64)
    v2 = Col(LONG, v.values[:2000])
    got, info = run_gpu_rule([Col(LONG, k2)], aggs, [v2, None], 2000)
    assert info == (False, 0, 0)
    # span 2^21 - 2 passes the first check, but its widened range (3 * span + 1 cells + the NULL slot) does not
    k3 = np.array([0, 2**21 - 2] * 1000, np.int64)
    got, info = run_gpu_rule([Col(LONG, k3)], aggs, [v2, None], 2000)
    assert info == (False, 0, 0)
    assert_row_sets_equal(got, run_oracle([Col(LONG, k3)], aggs, [v2, None], 2000))
    # the largest span whose widened range fits: 3 * 699050 + 1 cells + the NULL slot = 2^21 exactly
    k4 = np.array([0, 699050] * 1000, np.int64)
    got, info = run_gpu_rule([Col(LONG, k4)], aggs, [v2, None], 2000)
    assert info == (True, -699050, 3 * 699050 + 1)
    assert_row_sets_equal(got, run_oracle([Col(LONG, k4)], aggs, [v2, None], 2000))
    k5 = np.array([0, 699051] * 1000, np.int64)
    got, info = run_gpu_rule([Col(LONG, k5)], aggs, [v2, None], 2000)
    assert info == (False, 0, 0)


@pytest.mark.gpu
@pytest.mark.parametrize("strategy", [capi.STRATEGY_AUTO, capi.STRATEGY_RADIX])
def test_byte_key_parity(strategy):
    """BYTE (int8) group-by keys, negative values and NULLs included, alone and in a composite key."""
    from oracle import BYTE
    rng = np.random.default_rng(180 + strategy)
    n = 60_000
    b = rng.integers(-128, 128, n).astype(np.int8)
    bn = rng.random(n) < 0.05
    i = rng.integers(-40, 40, n).astype(np.int32)
    v = rng.integers(-10**9, 10**9, n).astype(np.int64)
    cols = [Col(LONG, v)] * 5
    for keys in ([Col(BYTE, b, bn)], [Col(INT, i), Col(BYTE, b, bn)], [Col(BYTE, b), Col(LONG, v % 7)]):
        got = run_gpu(keys, STD_AGGS, cols, 7000, strategy=strategy)
        want = run_oracle(keys, STD_AGGS, cols, 7000)
        assert_row_sets_equal(got, want, float_cols=float_cols(len(keys), STD_AGGS), rtol=DOUBLE_RTOL)
