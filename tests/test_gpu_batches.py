"""Batched group-by over a stream of many uneven VectorBatches (each an independent group-by, as one
HashAggregationOperator AddInput + GetOutput per batch in core/src/operator/aggregation/group_aggregation.cpp)
against the CPU oracle run batch by batch.  Every launch form (two launches for the whole stream, one launch
per batch, hipGraph replay) and every per-batch route (LDS table first with exact HBM fallback, HBM table
directly) must give the same row sets."""
import numpy as np
import pytest

import oracle
from oracle import AVG, COUNT_ALL, DECIMAL64, DECIMAL128, DOUBLE, INT, LONG, MAX, MIN, SUM, Col
from omnioperator_amd import capi, workloads
from tests.helpers import assert_row_sets_equal, rows_of

pytestmark = pytest.mark.gpu

AGGS = [(SUM, LONG, LONG), (COUNT_ALL, LONG, LONG), (MIN, INT, INT), (AVG, DOUBLE, DOUBLE),
        (SUM, DECIMAL64, DECIMAL128, 2, 2), (MAX, LONG, LONG)]
TYPES = [LONG] + [a[2] for a in AGGS]
FLOAT_COLS = [1 + i for i, a in enumerate(AGGS) if a[2] == DOUBLE]


def make_stream(seed, nb=300, max_rows=5000):
    """nb batches with lognormal row counts (median 256, clipped to [1, max_rows]; the first three are the
    largest, a single row and an empty batch), per-batch distinct keys log-uniform in [1, 4000], 0.1% of keys
    equal to the table's empty-slot sentinel, 1% NULL keys, NULLs in the first aggregate's input."""
    rng = np.random.default_rng(seed)
    rows = np.clip(np.round(np.exp(np.log(256) + 1.2 * rng.standard_normal(nb))), 1, max_rows).astype(np.int64)
    rows[:3] = [max_rows, 1, 0]
    offsets = np.concatenate([[0], np.cumsum(rows)])
    n = int(offsets[-1])
    keys = np.empty(n, np.int64)
    for b in range(nb):
        d = max(1, int(np.exp(rng.uniform(0, np.log(4000)))))
        keys[offsets[b]:offsets[b + 1]] = rng.integers(0, d, rows[b]) * 7919 - 3
    keys[rng.random(n) < 0.001] = -2**63
    key_nulls = rng.random(n) < 0.01
    cols = [Col(LONG, rng.integers(-2**62, 2**62, n, dtype=np.int64), rng.random(n) < 0.05),
            None,
            Col(INT, rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)),
            Col(DOUBLE, rng.normal(100, 50, n)),
            Col(DECIMAL64, rng.integers(-10**15, 10**15, n, dtype=np.int64), None, 2),
            Col(LONG, rng.integers(-2**63, 2**63 - 1, n, dtype=np.int64))]
    return offsets, Col(LONG, keys, key_nulls), cols


def oracle_batches(offsets, key, cols, which):
    spec = oracle.make_agg_spec([LONG], AGGS)
    res = {}
    for b in which:
        a0, a1 = int(offsets[b]), int(offsets[b + 1])
        if a1 == a0:
            res[b] = []
            continue
        outs = oracle.hash_agg(spec, [key.slice(a0, a1)], [None if c is None else c.slice(a0, a1) for c in cols],
                               max(1, a1 - a0))
        res[b] = rows_of([(c.type_id, c.values, c.nulls) for c in outs])
    return res


def per_batch_rows(offsets, groups, outs):
    res = []
    for b in range(len(offsets) - 1):
        a0, g = int(offsets[b]), int(groups[b])
        res.append(rows_of([(t, v[a0:a0 + g], nl[a0:a0 + g]) for t, (v, nl) in zip(TYPES, outs)]))
    return res


def run_batches(offsets, key, cols, modes=None, flags=0, tune=None, repeat=1, parts=None):
    spec = capi.make_agg_spec([LONG], AGGS)
    s = capi.AggBatches(spec, int(np.max(np.diff(offsets))))
    if tune:
        s.tune(**tune)
    if parts:
        s.tune_partitions(**parts)
    kd = capi.DeviceColumn(LONG, key.values, key.nulls)
    ad = [None if c is None else capi.DeviceColumn(c.type_id, c.values, c.nulls) for c in cols]
    ac = [None if d is None else d.c() for d in ad]
    results = []
    for _ in range(repeat):
        groups = s.run(kd.c(), ac, offsets, modes, flags)
        results.append((groups, s.outputs()))
    stats = s.stats()
    stats["phases"] = s.phases()
    s.close()
    return results, stats


@pytest.mark.parametrize("form", ["two_launches", "sketch", "per_batch", "graph", "hbm_modes", "small_lds",
                                  "block512", "small_partitions", "partition_overflow", "fused", "fused_graph",
                                  "fused_block512", "fused_small", "fused_forced", "fused_overflow"])
def test_batches_match_oracle(form):
    """Every form of the batched group-by against the oracle, batch by batch.  small_partitions: 8 KiB phase-2
    tables (many partitions per batch); partition_overflow: partitions sized 16x above their table, so phase-2
    partitions overflow, the batch's partial phase-2 output is replaced by phase 3 (NULL bits included).  fused*:
    one workgroup per batch (flag bit 4; rows carry their row numbers here, as the aggregates read several columns
    with NULLs): fused_small with 8 KiB tables (many partitions per batch), fused_forced with every batch sent to
    the partitioned path, fused_overflow with partitions sized far above their table (phase 3 replaces them)."""
    offsets, key, cols = make_stream(1234)
    nb = len(offsets) - 1
    kwargs = {}
    if form == "per_batch":
        kwargs["flags"] = capi.AggBatches.FLAG_PER_BATCH
    elif form == "sketch":
        kwargs["flags"] = capi.AggBatches.FLAG_SKETCH
    elif form == "graph":
        kwargs["flags"] = capi.AggBatches.FLAG_GRAPH
        kwargs["repeat"] = 3  # captured on the first run, replayed on the next two
    elif form == "hbm_modes":
        kwargs["modes"] = np.where(np.arange(nb) % 2 == 0, 2, 0).astype(np.int32)
    elif form == "small_lds":
        kwargs["tune"] = {"block": 256, "lds_bytes": 8192, "max_probe": 4}
    elif form == "block512":
        kwargs["tune"] = {"block": 512, "lds_bytes": 65536, "max_probe": 64}
    elif form == "small_partitions":
        kwargs["tune"] = {"block": 256, "lds_bytes": 8192, "max_probe": 4}
        kwargs["parts"] = {"part_lds_bytes": 8192, "part_rows": 0}
    elif form == "partition_overflow":
        kwargs["tune"] = {"block": 256, "lds_bytes": 8192, "max_probe": 4}
        kwargs["parts"] = {"part_lds_bytes": 8192, "part_rows": 4096}
    elif form.startswith("fused"):
        kwargs["flags"] = capi.AggBatches.FLAG_FUSED
        if form == "fused_graph":
            kwargs["flags"] |= capi.AggBatches.FLAG_GRAPH
            kwargs["repeat"] = 3
        elif form == "fused_block512":
            kwargs["tune"] = {"block": 512, "lds_bytes": 32768, "max_probe": 32}
        elif form == "fused_small":
            kwargs["parts"] = {"part_lds_bytes": 8192, "part_rows": 0}
        elif form == "fused_forced":
            kwargs["modes"] = np.full(nb, 2, np.int32)
        elif form == "fused_overflow":
            kwargs["parts"] = {"part_lds_bytes": 8192, "part_rows": 4096}
    results, stats = run_batches(offsets, key, cols, **kwargs)
    want = oracle_batches(offsets, key, cols, range(nb))
    for groups, outs in results:
        got = per_batch_rows(offsets, groups, outs)
        for b in range(nb):
            assert_row_sets_equal(got[b], want[b], float_cols=FLOAT_COLS, rtol=1e-12)
    assert stats["lds_batches"] + stats["hbm_batches"] == nb
    p1, p2, p3 = stats["phases"]
    assert p1 + p2 + p3 == nb and p1 == stats["lds_batches"]
    if form == "small_lds":
        assert stats["hbm_batches"] > 0  # 8 KiB tables cannot hold the larger batches
    if form in ("small_lds", "small_partitions", "two_launches", "sketch"):
        assert p3 == 0  # partitions sized to half a table do not overflow on this stream
    if form == "small_partitions":
        assert p2 > 0
    if form == "partition_overflow":
        assert p3 > 0
    if form == "per_batch":
        assert p2 == 0
    if form == "hbm_modes":
        assert stats["hbm_batches"] >= (nb + 1) // 2
    if form == "block512":
        assert stats["lds_batches"] > 0
    if form in ("fused", "fused_graph", "fused_block512", "fused_small"):
        assert p1 > 0 and p2 > 0 and p3 == 0  # direct and partitioned batches, no overflow
    if form == "fused_forced":
        assert p2 > 0 and p3 == 0
    if form == "fused_overflow":
        assert p3 > 0


@pytest.mark.parametrize("graph", [False, True])
def test_advisor_form_matches_oracle(graph):
    """The learned form (flag bit 3: the stream advisor sketches every batch and picks its first phase on the
    device) gives every batch the oracle's rows, eager and under hipGraph replay, and does route batches both
    ways; the advisor's features equal omni_gpu_batch_features."""
    from omnioperator_amd import batch_advisor
    offsets, key, cols = make_stream(4321)
    nb = len(offsets) - 1
    net = batch_advisor.network()
    spec = capi.make_agg_spec([LONG], AGGS)
    s = capi.AggBatches(spec, int(np.max(np.diff(offsets))))
    s.set_advisor(net)
    kd = capi.DeviceColumn(LONG, key.values, key.nulls)
    ad = [None if c is None else capi.DeviceColumn(c.type_id, c.values, c.nulls) for c in cols]
    ac = [None if d is None else d.c() for d in ad]
    flags = capi.AggBatches.FLAG_ADVISOR | (capi.AggBatches.FLAG_GRAPH if graph else 0)
    want = oracle_batches(offsets, key, cols, range(nb))
    for _ in range(3 if graph else 1):
        groups = s.run(kd.c(), ac, offsets, None, flags)
        got = per_batch_rows(offsets, groups, s.outputs())
        for b in range(nb):
            assert_row_sets_equal(got[b], want[b], float_cols=FLOAT_COLS, rtol=1e-12)
    ph = s.phase_of(nb)
    assert (ph == 1).any() and (ph == 2).any()
    # batches the network predicts above half the LDS table start in the partitioned passes
    X = capi.batch_features(kd.c(), offsets)
    y = net.forward(X)
    limit = 0.5 * s.info()["lds_slots"]
    skip = np.exp2(y[:, 0] * 32.0) - 1.0 > 1.05 * limit  # clear of the float rounding at the limit
    assert np.all(ph[skip] >= 2)
    s.close()
    net.close()


@pytest.mark.parametrize("form", ["fused", "fused_graph", "fused_advisor", "fused_advisor_graph", "fused_small"])
def test_fused_carried_values_match_oracle(form):
    """One workgroup per batch with the rows carrying their value word (SUM(int64) and COUNT(*) over NULL-free
    columns, the batched stream's aggregates) against the oracle, batch by batch; with the stream advisor's
    predicted group counts choosing direct or partitioned and the partition count (fused_advisor*)."""
    from omnioperator_amd import batch_advisor
    rng = np.random.default_rng(99)
    nb = 400
    rows = np.clip(np.round(np.exp(np.log(1024) + 1.4 * rng.standard_normal(nb))), 0, 20000).astype(np.int64)
    rows[:3] = [20000, 1, 0]
    offsets = np.concatenate([[0], np.cumsum(rows)])
    n = int(offsets[-1])
    keys = np.empty(n, np.int64)
    for b in range(nb):
        d = max(1, int(np.exp(rng.uniform(0, np.log(30000)))))
        keys[offsets[b]:offsets[b + 1]] = rng.integers(0, d, rows[b]) * 2654435761 + b
    keys[rng.random(n) < 0.0005] = -2**63
    vals = rng.integers(-2**62, 2**62, n, dtype=np.int64)
    aggs = [(SUM, LONG, LONG), (COUNT_ALL, LONG, LONG)]
    s = capi.AggBatches(capi.make_agg_spec([LONG], aggs), int(rows.max()))
    net = None
    flags = capi.AggBatches.FLAG_FUSED
    if "graph" in form:
        flags |= capi.AggBatches.FLAG_GRAPH
    if "advisor" in form:
        net = batch_advisor.network()
        s.set_advisor(net)
        flags |= capi.AggBatches.FLAG_ADVISOR
    if form == "fused_small":
        s.tune_partitions(part_lds_bytes=8192, part_rows=0)
    kd, vd = capi.DeviceColumn(LONG, keys), capi.DeviceColumn(LONG, vals)
    spec_o = oracle.make_agg_spec([LONG], aggs)
    want = []
    for b in range(nb):
        a0, a1 = int(offsets[b]), int(offsets[b + 1])
        ref = oracle.hash_agg(spec_o, [Col(LONG, keys[a0:a1])], [Col(LONG, vals[a0:a1]), None], max(1, a1 - a0))
        want.append(rows_of([(c.type_id, c.values, c.nulls) for c in ref]))
    for _ in range(3 if "graph" in form else 1):
        groups = s.run(kd.c(), [vd.c(), None], offsets, None, flags)
        outs = s.outputs()
        for b in range(nb):
            a0, g = int(offsets[b]), int(groups[b])
            got = rows_of([(LONG, v[a0:a0 + g], nl[a0:a0 + g]) for v, nl in outs])
            assert_row_sets_equal(got, want[b])
    p1, p2, p3 = s.phases()
    assert p1 > 0 and p2 > 0 and p3 == 0
    s.close()
    if net is not None:
        net.close()


def test_graph_replays_match_eager():
    """hipGraph replays of the steady state, with eager runs between them, give the eager result each time."""
    offsets, key, cols = make_stream(77, nb=120)
    spec = capi.make_agg_spec([LONG], AGGS)
    s = capi.AggBatches(spec, int(np.max(np.diff(offsets))))
    kd = capi.DeviceColumn(LONG, key.values, key.nulls)
    ad = [None if c is None else capi.DeviceColumn(c.type_id, c.values, c.nulls) for c in cols]
    ac = [None if d is None else d.c() for d in ad]
    eager_groups = s.run(kd.c(), ac, offsets, None, 0)
    eager = per_batch_rows(offsets, eager_groups, s.outputs())
    for _ in range(4):
        g = s.run(kd.c(), ac, offsets, None, capi.AggBatches.FLAG_GRAPH)
        assert np.array_equal(g, eager_groups)
        cur = per_batch_rows(offsets, g, s.outputs())
        for b in range(len(offsets) - 1):
            assert_row_sets_equal(cur[b], eager[b], float_cols=FLOAT_COLS, rtol=1e-12)
        s.run(kd.c(), ac, offsets, None, 0)
    s.close()


def test_batches_decimal_overflow_null_and_error():
    """SUM(decimal64 scale 0 -> decimal128 scale 20): a group whose rescaled sum leaves +-(10^38 - 1) is NULL with
    overflowAsNull and an error without it (MulCheckedOverflow, sum_spark_decimal_aggregator.h:284-296), in each of
    two batches; the values equal the oracle's."""
    v = np.array([2**62, 2**62, 5, 7, 10**18] * 2, np.int64)
    k = np.array([0, 0, 1, 1, 2] * 2, np.int64)
    offsets = np.array([0, 5, 10], np.int64)
    aggs = [(SUM, DECIMAL64, DECIMAL128, 0, 20)]
    kd = capi.DeviceColumn(LONG, k)
    vd = capi.DeviceColumn(DECIMAL64, v)
    s = capi.AggBatches(capi.make_agg_spec([LONG], aggs, overflow_as_null=True), 16)
    groups = s.run(kd.c(), [vd.c()], offsets)
    outs = s.outputs()
    s.close()
    want = rows_of([(c.type_id, c.values, c.nulls) for c in oracle.hash_agg(
        oracle.make_agg_spec([LONG], aggs), [Col(LONG, k[:5])], [Col(DECIMAL64, v[:5])], 5)])
    assert sorted(want, key=lambda r: r[0]) == [(0, None), (1, 12 * 10**20), (2, None)]
    for b in range(2):
        a0, g = int(offsets[b]), int(groups[b])
        got = rows_of([(LONG, outs[0][0][a0:a0 + g], outs[0][1][a0:a0 + g]),
                       (DECIMAL128, outs[1][0][a0:a0 + g], outs[1][1][a0:a0 + g])])
        assert sorted(got, key=lambda r: r[0]) == sorted(want, key=lambda r: r[0])
    s = capi.AggBatches(capi.make_agg_spec([LONG], aggs, overflow_as_null=False), 16)
    with pytest.raises(capi.GpuError) as e:
        s.run(kd.c(), [vd.c()], offsets)
    assert e.value.code == -4
    s.close()


@pytest.mark.parametrize("form", ["phases", "fused_learned"])
def test_batched_stream_config_full_size(form):
    """20,000 batches (lognormal rows, median 2048, clipped to [64, 65536]), per-batch distinct keys
    log-uniform in [8, 1e6]; SUM(int64) + COUNT, in the phased form and in the one-workgroup-per-batch form with
    the stream advisor (the benchmarked learned + batched form).  Size-independent checks on every batch (COUNT
    sums to the batch's rows, SUM equals the batch's wrapping total, group count equals the batch's distinct keys)
    and bit-exact parity with the oracle on a sample of batches."""
    from omnioperator_amd import batch_advisor
    w = workloads.batched_stream()
    offsets = np.concatenate([[0], np.cumsum(w.row_counts)])
    aggs = [(SUM, LONG, LONG), (COUNT_ALL, LONG, LONG)]
    spec = capi.make_agg_spec([LONG], aggs)
    s = capi.AggBatches(spec, 65536)
    kd = capi.DeviceColumn(LONG, w.keys)
    vd = capi.DeviceColumn(LONG, w.values)
    flags = capi.AggBatches.FLAG_GRAPH
    net = None
    if form == "fused_learned":
        net = batch_advisor.network()
        s.set_advisor(net)
        flags |= capi.AggBatches.FLAG_FUSED | capi.AggBatches.FLAG_ADVISOR
    for _ in range(2):  # captured, then replayed
        groups = s.run(kd.c(), [vd.c(), None], offsets, None, flags)
    outs = s.outputs()
    s.close()
    if net is not None:
        net.close()
    kv, sv, cv = outs[0][0], outs[1][0], outs[2][0]
    nb = len(w.row_counts)
    owner = np.repeat(np.arange(nb), w.row_counts)
    valid = np.arange(len(w.keys)) - offsets[owner] < groups[owner]
    cnt = np.bincount(owner[valid], minlength=nb)
    assert np.array_equal(cnt, groups)
    count_tot = np.zeros(nb, np.int64)
    np.add.at(count_tot, owner[valid], cv[valid])
    assert np.array_equal(count_tot, w.row_counts)
    with np.errstate(over="ignore"):
        sum_tot = np.zeros(nb, np.int64)
        np.add.at(sum_tot, owner[valid], sv[valid])
        want = np.zeros(nb, np.int64)
        np.add.at(want, owner, w.values)
    assert np.array_equal(sum_tot, want)
    order = np.lexsort((w.keys, owner))
    sk, so = w.keys[order], owner[order]
    distinct = np.bincount(so[np.concatenate([[True], (sk[1:] != sk[:-1]) | (so[1:] != so[:-1])])], minlength=nb)
    assert np.array_equal(distinct, groups)
    rng = np.random.default_rng(3)
    sample = rng.choice(nb, 40, replace=False)
    spec_o = oracle.make_agg_spec([LONG], aggs)
    for b in sample:
        a0, a1 = int(offsets[b]), int(offsets[b + 1])
        ref = oracle.hash_agg(spec_o, [Col(LONG, w.keys[a0:a1])], [Col(LONG, w.values[a0:a1]), None], a1 - a0)
        g = int(groups[b])
        got = rows_of([(LONG, kv[a0:a0 + g], None), (LONG, sv[a0:a0 + g], None), (LONG, cv[a0:a0 + g], None)])
        assert_row_sets_equal(got, rows_of([(c.type_id, c.values, c.nulls) for c in ref]))
