"""Batched stream of independent group-bys (one persistent launch over many uneven batches) against the CPU
oracle run batch by batch.  Every launch form (persistent, one launch per batch, hipGraph replay) and every
per-batch strategy (LDS first with exact spill, scratch table directly) must give the same row sets."""
import numpy as np
import pytest

import oracle
from oracle import AVG, COUNT_ALL, DECIMAL64, DECIMAL128, DOUBLE, INT, LONG, MAX, MIN, SUM, Col
from omnioperator_amd import capi, workloads
from tests.helpers import assert_row_sets_equal, rows_of

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason="the first persistent-queue stream kernel (agg_stream) took a GPU memory fault on its "
                            "first run; its indexing is clean in a concurrent CPU build under AddressSanitizer and no "
                            "cause is shown, so that kernel and its entry points (omni_gpu_agg_stream_*) were "
                            "removed; the batched path is rebuilt as agg_batch_* and its persistent queue-draining "
                            "kernel as batch_fused (tests/test_gpu_batches.py)"),
]

AGGS = [(SUM, LONG, LONG), (COUNT_ALL, LONG, LONG), (MIN, INT, INT), (AVG, DOUBLE, DOUBLE),
        (SUM, DECIMAL64, DECIMAL128, 2, 2), (MAX, LONG, LONG)]


def make_stream(seed, nb=300, max_rows=5000):
    rng = np.random.default_rng(seed)
    rows = np.clip(np.round(np.exp(np.log(256) + 1.2 * rng.standard_normal(nb))), 1, max_rows).astype(np.int64)
    rows[:3] = [max_rows, 1, 0]  # edge batches: the largest, a single row, an empty one
    offsets = np.concatenate([[0], np.cumsum(rows)])
    n = int(offsets[-1])
    keys = np.empty(n, np.int64)
    for b in range(nb):
        d = max(1, int(np.exp(rng.uniform(0, np.log(4000)))))
        keys[offsets[b]:offsets[b + 1]] = rng.integers(0, d, rows[b]) * 7919 - 3
    keys[rng.random(n) < 0.001] = -2**63  # the slot sentinel as a real key
    key_nulls = rng.random(n) < 0.01
    cols = [Col(LONG, rng.integers(-2**62, 2**62, n, dtype=np.int64), rng.random(n) < 0.05),
            None,
            Col(INT, rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)),
            Col(DOUBLE, rng.normal(100, 50, n)),
            Col(DECIMAL64, rng.integers(-10**15, 10**15, n, dtype=np.int64), None, 2),
            Col(LONG, rng.integers(-2**63, 2**63 - 1, n, dtype=np.int64))]
    return offsets, Col(LONG, keys, key_nulls), cols


def run_stream(offsets, key, cols, modes=None, flags=0, tune=None, repeat=1):
    spec = capi.make_agg_spec([LONG], AGGS)
    s = capi.AggStream(spec, int(np.max(np.diff(offsets))) if len(offsets) > 1 else 1)
    if tune:
        s.tune(**tune)
    kd = capi.DeviceColumn(LONG, key.values, key.nulls)
    ad = [None if c is None else capi.DeviceColumn(c.type_id, c.values, c.nulls) for c in cols]
    for _ in range(repeat):
        groups = s.run(kd.c(), [None if d is None else d.c() for d in ad], offsets, modes, flags)
    outs = s.outputs()
    stats = s.stats()
    s.close()
    types = [LONG] + [a[2] for a in AGGS]
    per_batch = []
    for b in range(len(offsets) - 1):
        a0, g = int(offsets[b]), int(groups[b])
        per_batch.append(rows_of([(t, v[a0:a0 + g], nl[a0:a0 + g]) for t, (v, nl) in zip(types, outs)]))
    return per_batch, stats


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


FLOAT_COLS = [1 + i for i, a in enumerate(AGGS) if a[2] == DOUBLE]


@pytest.mark.parametrize("form", ["persistent", "per_batch", "graph", "scratch_modes", "small_lds", "block512"])
def test_stream_matches_oracle(form):
    offsets, key, cols = make_stream(1234)
    nb = len(offsets) - 1
    kwargs = {}
    if form == "per_batch":
        kwargs["flags"] = capi.AggStream.FLAG_PER_BATCH
    elif form == "graph":
        kwargs["flags"] = capi.AggStream.FLAG_GRAPH
        kwargs["repeat"] = 3  # capture once, replay: every replay must give the same result
    elif form == "scratch_modes":
        kwargs["modes"] = np.where(np.arange(nb) % 2 == 0, 2, 0).astype(np.int32)
    elif form == "small_lds":
        kwargs["tune"] = {"block": 256, "lds_bytes": 8192, "max_probe": 4}
    elif form == "block512":
        kwargs["tune"] = {"block": 512, "lds_bytes": 65536, "max_probe": 64}
    got, stats = run_stream(offsets, key, cols, **kwargs)
    want = oracle_batches(offsets, key, cols, range(nb))
    for b in range(nb):
        assert_row_sets_equal(got[b], want[b], float_cols=FLOAT_COLS, rtol=1e-12)
    if form == "small_lds":
        assert stats["scratch_batches"] > 0  # tiny LDS tables must have spilled
    if form == "scratch_modes":
        assert stats["scratch_batches"] >= nb // 2


def test_graph_replays_match_eager():
    """hipGraph replay of the steady state, with eager work between replays, gives the eager result."""
    offsets, key, cols = make_stream(77, nb=120)
    spec = capi.make_agg_spec([LONG], AGGS)
    s = capi.AggStream(spec, int(np.max(np.diff(offsets))))
    kd = capi.DeviceColumn(LONG, key.values, key.nulls)
    ad = [None if c is None else capi.DeviceColumn(c.type_id, c.values, c.nulls) for c in cols]
    ac = [None if d is None else d.c() for d in ad]
    eager_groups = s.run(kd.c(), ac, offsets, None, 0)
    eager = s.outputs()
    for _ in range(4):
        g = s.run(kd.c(), ac, offsets, None, capi.AggStream.FLAG_GRAPH)
        assert np.array_equal(g, eager_groups)
        cur = s.outputs()
        for b in range(len(offsets) - 1):
            a0, n = int(offsets[b]), int(g[b])
            types = [LONG] + [a[2] for a in AGGS]
            x = rows_of([(t, v[a0:a0 + n], nl[a0:a0 + n]) for t, (v, nl) in zip(types, cur)])
            y = rows_of([(t, v[a0:a0 + n], nl[a0:a0 + n]) for t, (v, nl) in zip(types, eager)])
            assert_row_sets_equal(x, y, float_cols=FLOAT_COLS, rtol=1e-12)
        s.run(kd.c(), ac, offsets, None, 0)  # eager work between replays
    s.close()


def test_batched_stream_config_full_size():
    """20,000 batches (lognormal rows, median 2048, clipped to [64, 65536]), per-batch distinct keys
    log-uniform in [8, 1e6]; SUM(int64) + COUNT.  Size-independent checks on every batch (COUNT sums to the
    batch's rows, SUM equals the batch's wrapping total, group count equals the batch's distinct keys) and
    bit-exact parity with the oracle on a sample of batches."""
    w = workloads.batched_stream()
    offsets = np.concatenate([[0], np.cumsum(w.row_counts)])
    aggs = [(SUM, LONG, LONG), (COUNT_ALL, LONG, LONG)]
    spec = capi.make_agg_spec([LONG], aggs)
    s = capi.AggStream(spec, 65536)
    kd = capi.DeviceColumn(LONG, w.keys)
    vd = capi.DeviceColumn(LONG, w.values)
    groups = s.run(kd.c(), [vd.c(), None], offsets, None, capi.AggStream.FLAG_GRAPH)
    outs = s.outputs()
    s.close()
    kv, sv, cv = outs[0][0], outs[1][0], outs[2][0]
    nb = len(w.row_counts)
    owner = np.repeat(np.arange(nb), w.row_counts)
    valid = np.arange(len(w.keys)) - offsets[owner] < groups[owner]
    cnt = np.bincount(owner[valid], weights=None, minlength=nb)
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
