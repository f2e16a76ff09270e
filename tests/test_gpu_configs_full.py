"""The skewed group-by and the learned task set at their configured sizes, on the GPU.

- Skewed group-by: 200M rows, key (int32 uniform over 1e4, int64 Zipf s=1.1 over 5e7 ranks), SUM(decimal64 ->
  decimal128), SUM and AVG(double), COUNT(*), MAX(int32), under the radix-partitioned and the hash-table
  strategies.  Every group is checked against an exact sort-based group-by of the same rows on the host
  (tests/bulk_checks.py, itself pinned to the oracle): the group set, COUNT, MAX and the decimal sums exactly,
  double SUM/AVG within 1e-12 relative.  The first 4M rows are also run through the oracle, bit-exact.
- Learned task set: the 512 seeded group-by and join tasks; a subset of 29 spanning its full row and distinct
  ranges (every 20th task by row count, plus the largest group-by, the largest join and the group-by with the most
  distinct keys) runs in the configuration the trained advisor picks.  Group-bys are checked in full by
  keyed checksums (exact COUNT and SUM of every group w.h.p., distinct group keys), joins by their exact matches
  per probe row and a payload checksum; tasks of at most 1M rows are also compared with the oracle, bit-exact.
"""
import time

import numpy as np
import pytest

import oracle
from oracle import AVG, COUNT_ALL, DECIMAL64, DECIMAL128, DOUBLE, INT, LONG, MAX, SUM, Col
from omnioperator_amd import capi, learned, workloads
from tests import bulk_checks as bc
from tests.test_gpu_agg import float_cols, run_gpu, run_oracle
from tests.helpers import assert_row_sets_equal, rows_of

pytestmark = pytest.mark.gpu

SKEWED_AGGS = [(SUM, DECIMAL64, DECIMAL128, 2, 2), (SUM, DOUBLE, DOUBLE), (AVG, DOUBLE, DOUBLE),
               (COUNT_ALL, LONG, LONG), (MAX, INT, INT)]


def _log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def skewed():
    t0 = time.perf_counter()
    w = workloads.agg_skewed(keep_ranks=True)
    packed = (w.k1.astype(np.int64) << 26) | w.ranks
    ukeys, counts, red = bc.sorted_groupby(packed, {"dec": (w.v_dec, "sum"), "dbl": (w.v_dbl, "sum"),
                                                    "max": (w.v_i32, "max")})
    del packed
    _log(f"skewed workload and host reference: {len(ukeys)} groups, {time.perf_counter() - t0:.1f} s")
    return w, ukeys, counts, red


@pytest.mark.parametrize("strategy", [capi.STRATEGY_RADIX, capi.STRATEGY_AUTO])
def test_skewed_agg_config_full_size(skewed, strategy):
    w, ukeys, counts, red = skewed
    n = len(w.k1)
    _log(f"skewed full size: {n} rows, strategy {strategy}")
    cols = [capi.DeviceColumn(INT, w.k1), capi.DeviceColumn(LONG, w.k2), capi.DeviceColumn(DECIMAL64, w.v_dec),
            capi.DeviceColumn(DOUBLE, w.v_dbl), capi.DeviceColumn(INT, w.v_i32)]
    k1, k2, vd, vf, vi = cols
    h = capi.HashAgg(capi.make_agg_spec([INT, LONG], SKEWED_AGGS, strategy=strategy))
    for a, b in workloads.batches(n, 1 << 20):
        h.add_batch([k1.c(a), k2.c(a)], [vd.c(a), vf.c(a), vf.c(a), None, vi.c(a)], b - a)
    g = h.num_groups()
    out = h.output()
    h.close()
    del cols
    assert g == len(ukeys) == 78_979_901
    for v, nl in out:
        assert not nl.any()
    packed = (out[0][0].astype(np.int64) << 26) | bc.unsplitmix64(out[1][0]).view(np.int64)
    o = np.argsort(packed)
    assert np.array_equal(packed[o], ukeys), "group keys differ"
    del packed
    dec = out[2][0][o]
    assert np.all(dec[:, 1] == 0) and np.array_equal(dec[:, 0], red["dec"]), "decimal SUM differs"
    assert int(dec[:, 0].sum()) == int(w.v_dec.sum())
    assert np.allclose(out[3][0][o], red["dbl"], rtol=1e-12, atol=0), "double SUM beyond 1e-12"
    assert np.allclose(out[4][0][o], red["dbl"] / counts, rtol=1e-12, atol=0), "double AVG beyond 1e-12"
    assert np.array_equal(out[5][0][o], counts) and int(counts.sum()) == n, "COUNT differs"
    assert np.array_equal(out[6][0][o], red["max"]), "MAX(int32) differs"


def test_skewed_agg_first_4m_rows_bit_exact_with_oracle(skewed):
    w = skewed[0]
    m = 4_000_000
    keys = [Col(INT, w.k1[:m]), Col(LONG, w.k2[:m])]
    cols = [Col(DECIMAL64, w.v_dec[:m], scale=2), Col(DOUBLE, w.v_dbl[:m]), Col(DOUBLE, w.v_dbl[:m]), None,
            Col(INT, w.v_i32[:m])]
    want = run_oracle(keys, SKEWED_AGGS, cols, 1 << 20)
    for strategy in (capi.STRATEGY_RADIX, capi.STRATEGY_AUTO):
        got = run_gpu(keys, SKEWED_AGGS, cols, 1 << 20, strategy=strategy)
        assert_row_sets_equal(got, want, float_cols=float_cols(2, SKEWED_AGGS), rtol=1e-12)


def c5_subset():
    tasks = learned.sample_tasks(512, seed=0xC5, min_rows=1e5, max_rows=2e8, join_share=0.5)
    order = np.argsort([t.rows for t in tasks], kind="stable")
    sel = [int(i) for i in order[::20]]
    aggs = [i for i, t in enumerate(tasks) if t.op == "agg"]
    joins = [i for i, t in enumerate(tasks) if t.op == "join"]
    for i in (max(aggs, key=lambda i: tasks[i].rows), max(joins, key=lambda i: tasks[i].rows),
              max(aggs, key=lambda i: tasks[i].distinct)):
        if i not in sel:
            sel.append(i)
    return [(i, tasks[i]) for i in sel]


ORACLE_ROWS = 1_000_000


def _check_agg_task(data, adv):
    task = data.task
    cap, strategy = adv.plan(data)
    g, _, out = learned.run_agg(data, strategy, cap, output=True)
    keys_in = [k.numpy() for k in data.keys]
    values = data.values.numpy()
    nk = len(data.types)
    for v, nl in out:
        assert not nl.any()
    bc.check_sum_count_groupby(keys_in, values, [o[0] for o in out[:nk]], out[nk][0], out[nk + 1][0])
    if task.rows <= ORACLE_ROWS:
        ocols = [Col(t, k) for t, k in zip(data.types, keys_in)]
        want = run_oracle(ocols, learned.AGGS, [Col(LONG, values), None], learned.BATCH_ROWS)
        got = rows_of([(t, v, nl) for t, (v, nl) in zip(data.types + [LONG, LONG], out)])
        assert_row_sets_equal(got, want)
    return f"{g} groups, strategy {strategy}, capacity {cap}"


def _check_join_task(data, adv):
    task = data.task
    _, cls = adv.plan(data)
    m, _, outs = learned.run_join(data, cls)
    rows, pay = outs[0].numpy(m), outs[1][0][0].numpy(m)
    pk, bk, bv = data.keys[0].numpy(), data.build_keys[0].numpy(), data.build_values.numpy()
    pk, bk = pk.astype(np.int64), bk.astype(np.int64)
    assert bc.check_inner_join(pk, bk, bv, rows, pay) == m
    if task.rows <= ORACLE_ROWS:
        kt = data.types[0]
        r = oracle.hash_join(Col(kt, data.build_keys[0].numpy()), [Col(LONG, bv)], Col(kt, data.keys[0].numpy()),
                             learned.BATCH_ROWS, learned.BATCH_ROWS)
        assert sorted(zip(rows.tolist(), pay.tolist())) == \
            sorted(zip(r.probe_rows.tolist(), r.payloads[0].values.tolist()))
    return f"{m} matches, class {cls}"


C5_SUBSET = c5_subset()


def test_learned_task_subset_spans_the_set():
    rows = [t.rows for _, t in C5_SUBSET]
    assert len(C5_SUBSET) == 29 and min(rows) < 2e5 and max(rows) > 1.9e8
    assert max(t.distinct for _, t in C5_SUBSET if t.op == "agg") > 5e7
    assert min(t.distinct for _, t in C5_SUBSET) < 100
    assert {t.kind for _, t in C5_SUBSET if t.op == "agg"} == {0, 1, 2}
    assert {t.kind for _, t in C5_SUBSET if t.op == "join"} == {0, 1}


@pytest.fixture(scope="module")
def advisor():
    adv = learned.Advisor()
    yield adv
    adv.close()


@pytest.mark.parametrize("pos", range(len(C5_SUBSET)))
def test_learned_task_full_size(advisor, pos):
    i, task = C5_SUBSET[pos]
    _log(f"task {i}: {task.op}, kind {task.kind}, {task.rows} rows, distinct {int(task.distinct)}, s {task.zipf_s:.2f}")
    t0 = time.perf_counter()
    data = learned.TaskData(task)
    what = _check_agg_task(data, advisor) if task.op == "agg" else _check_join_task(data, advisor)
    _log(f"task {i}: {what}, {time.perf_counter() - t0:.1f} s")
