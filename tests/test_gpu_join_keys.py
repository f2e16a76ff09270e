"""GPU joins on several key columns and on a DECIMAL128 key (omni_gpu_join_spec.num_keys / key_types, the key
translation of join_keys.h) through the C ABI, against the CPU oracle (pinned by tests/test_oracle_join_keys.py) and
the reference's fixtures: every key variant (int32 / int64 / int128 packing, Decimal128, serialised), every join
type, NULLs in each key column, ComputeMultiColKey's sign-extension collisions, per-batch and batched probes, the
partitioned probe.  Results compare bit-exact as row sets."""
import numpy as np
import pytest

import oracle
from oracle import BYTE, DATE32, INT, LONG, SHORT, Col
from omnioperator_amd import capi, workloads
from tests.helpers import load_golden
from tests.test_gpu_join_types import gpu_set, keeps_build, oracle_outer_set, oracle_set, outer_set
from tests.test_oracle_join_keys import (ALL_TYPES, FIX, KEY_SETS, _arg, _norm, _sort_key, fixture_col,
                                         random_key_cols)

pytestmark = pytest.mark.gpu


def _dev(c: Col, a: int, b: int):
    return capi.DeviceColumn(c.type_id, c.values[a:b], None if c.nulls is None else c.nulls[a:b])


def gpu_multi(bkey, bpay, pkey, build_batch, probe_batch, jt, bs, batched=False, partitions=None, info=None):
    """As tests.test_gpu_join_types.gpu_typed, with bkey / pkey lists of key columns."""
    types = [c.type_id for c in bkey]
    j = capi.HashJoin(types if len(types) > 1 else types[0], [p.type_id for p in bpay])
    j.set_type(jt, bs)
    if partitions is not None:
        j.tune(1, partitions)
    keep = []
    nb = len(bkey[0])
    for a, b in workloads.batches(nb, build_batch):
        kd = [_dev(c, a, b) for c in bkey]
        pd = [_dev(p, a, b) for p in bpay]
        keep.append((kd, pd))
        j.add_build([k.c() for k in kd], [d.c() for d in pd], b - a)
    j.finish()
    del keep
    npr = len(pkey[0])
    cols = [(a, b, [_dev(c, a, b) for c in pkey]) for a, b in workloads.batches(npr, probe_batch)]
    if jt == oracle.JOIN_EXISTENCE:
        flags = np.concatenate([j.probe_exists([k.c() for k in c], b - a) for a, b, c in cols])
        j.close()
        return None, None, flags
    if info is not None:
        info.update(j.info())
    cap = max(1, npr * max(1, j.info()["max_dup"]))
    rows, pays = [], [[] for _ in bpay]
    if batched:
        out_rows, po, oarr = j._outs(cap, True)
        flat = [k.c() for _, _, c in cols for k in c]
        m = j.probe_batches_raw(flat, [b - a for a, b, _ in cols], out_rows, oarr, cap)
        rows.append(out_rows.numpy(m))
        for i, (v, nbm) in enumerate(po):
            pays[i].append((v.numpy(m), capi.unpack_nulls(nbm.numpy(), m)))
    else:
        for a, b, c in cols:
            r, p = j.probe([k.c() for k in c], b - a, row_base=a, cap=cap)
            rows.append(r)
            for i, x in enumerate(p):
                pays[i].append(x)
    outer = None
    if keeps_build(jt, bs):
        orow, op = j.outer_output()
        outer = (orow, [list(zip(v.tolist(), n.tolist())) for v, n in op])
    j.close()
    rows = np.concatenate(rows) if rows else np.zeros(0, np.int64)
    plist = [list(zip(np.concatenate([x[0] for x in pl]).tolist(), np.concatenate([x[1] for x in pl]).tolist()))
             if pl else [] for pl in pays]
    return (rows, plist), outer, None


def check(bkey, pays, pkey, jt, bs, bb, pbs, batched=False, partitions=None):
    want = oracle.hash_join_typed(_arg(bkey), pays, _arg(pkey), bb, pbs, jt, bs)
    info = {}
    got, outer, flags = gpu_multi(bkey, pays, pkey, bb, pbs, jt, bs, batched, partitions, info)
    if jt == oracle.JOIN_EXISTENCE:
        assert flags.tolist() == want.exists.tolist()
        return
    assert not info["is_array"]  # the reference never builds an array table for these keys
    with_pay = jt not in (oracle.JOIN_LEFT_SEMI, oracle.JOIN_LEFT_ANTI)
    assert gpu_set(*got, with_pay) == oracle_set(want, with_pay)
    if outer is not None:
        assert outer_set(*outer) == oracle_outer_set(want)


def payloads(rng, n):
    return [Col(LONG, np.arange(n, dtype=np.int64)),
            Col(SHORT, rng.integers(-2**15, 2**15, n).astype(np.int16), rng.random(n) < 0.1)]


@pytest.mark.parametrize("case", [k for k in FIX if not k.startswith("_")])
def test_gpu_multi_key_fixture(case):
    f = FIX[case]
    types, kc = f["types"], f["key_cols"]
    bcols = [fixture_col(t, v) for t, v in zip(types, f["build"])]
    pcols = [fixture_col(t, v) for t, v in zip(types, f["probe"])]
    nb, npr = len(f["build"][0]), len(f["probe"][0])
    got, outer, _ = gpu_multi([bcols[c] for c in kc], [Col(LONG, np.arange(nb, dtype=np.int64))],
                              [pcols[c] for c in kc], nb, npr, f["join_type"], f["build_side"])
    b_rows = [[_norm(f["build"][c][r]) for c in range(len(types))] for r in range(nb)]
    p_rows = [[_norm(f["probe"][c][r]) for c in range(len(types))] for r in range(npr)]
    none = [None] * len(types)
    rows = [tuple(p_rows[p] + (none if n else b_rows[b])) for p, (b, n) in zip(got[0].tolist(), got[1][0])]
    want = [tuple(_norm(v) for v in r) for r in f["expected_rows"]]
    assert sorted(rows, key=_sort_key) == sorted(want, key=_sort_key)
    if "expected_outer" in f:
        orows = [tuple(none + b_rows[b]) for b in outer[0].tolist()]
        want_outer = [tuple(_norm(v) for v in r) for r in f["expected_outer"]]
        assert sorted(orows, key=_sort_key) == sorted(want_outer, key=_sort_key)


@pytest.mark.parametrize("name", list(KEY_SETS))
def test_gpu_multi_key_inner(name):
    types = KEY_SETS[name]
    rng = np.random.default_rng(50 + list(KEY_SETS).index(name))
    bkey = random_key_cols(types, 6000, rng, pool=2000)
    pkey = random_key_cols(types, 9000, rng, pool=2600)
    check(bkey, payloads(rng, 6000), pkey, oracle.JOIN_INNER, 0, 1500, 2048)


@pytest.mark.parametrize("jt,bs", ALL_TYPES)
@pytest.mark.parametrize("name", ["short_short", "short_int_long", "date32_double", "decimal128"])
def test_gpu_multi_key_join_types(jt, bs, name):
    types = KEY_SETS[name]
    rng = np.random.default_rng(70 + 10 * jt + bs)
    bkey = random_key_cols(types, 3000, rng, pool=1200)
    pkey = random_key_cols(types, 5000, rng, pool=1500)
    check(bkey, payloads(rng, 3000), pkey, jt, bs, 700, 1000)


@pytest.mark.parametrize("name", ["int_int", "long_long", "decimal128_int"])
@pytest.mark.parametrize("jt", [oracle.JOIN_INNER, oracle.JOIN_FULL])
def test_gpu_multi_key_batched_probe(name, jt):
    types = KEY_SETS[name]
    rng = np.random.default_rng(90 + jt)
    bkey = random_key_cols(types, 4000, rng, pool=1500)
    pkey = random_key_cols(types, 7000, rng, pool=2000)
    check(bkey, payloads(rng, 4000), pkey, jt, 0, 1000, 333, batched=True)


@pytest.mark.parametrize("name", ["short_int", "long_long", "date32_double"])
def test_gpu_multi_key_partitioned_probe(name):
    types = KEY_SETS[name]
    rng = np.random.default_rng(110)
    bkey = random_key_cols(types, 50000, rng, pool=30000)
    pkey = random_key_cols(types, 60000, rng, pool=40000)
    check(bkey, payloads(rng, 50000), pkey, oracle.JOIN_INNER, 0, 8192, 30000, partitions=16)


def test_gpu_sign_extension_collisions():
    def cols(t0, a, t1, b):
        return [Col(t0, np.array(a, oracle.np_dtype(t0))), Col(t1, np.array(b, oracle.np_dtype(t1)))]
    cases = [(cols(INT, [5, 5], SHORT, [-1, 1]), cols(INT, [7, 7], SHORT, [-1, 1]), [(0, 0)]),
             (cols(SHORT, [5], SHORT, [-1]), cols(SHORT, [9], SHORT, [-1]), [(0, 0)]),
             (cols(LONG, [3], LONG, [-1]), cols(LONG, [4], LONG, [-1]), [(0, 0)]),
             (cols(DATE32, [5], SHORT, [-1]), cols(DATE32, [7], SHORT, [-1]), []),
             (cols(SHORT, [3, 3], BYTE, [-1, 4]), cols(SHORT, [8, 8], BYTE, [-1, 4]), [(0, 0)]),
             (cols(BYTE, [3, -7], INT, [-1, 9]), cols(BYTE, [5, -7], INT, [-1, 9]), [(0, 0), (1, 1)])]
    for bk, pk, want in cases:
        n = len(bk[0])
        got, _, _ = gpu_multi(bk, [Col(LONG, np.arange(n, dtype=np.int64))], pk, n, n, oracle.JOIN_INNER, 0)
        assert sorted(zip(got[0].tolist(), [v for v, _ in got[1][0]])) == want


def test_gpu_multi_key_null_and_empty():
    a = Col(INT, np.array([1, 1, 1], np.int32), np.array([False, True, False]))
    b = Col(INT, np.array([2, 2, 2], np.int32), np.array([False, False, True]))
    pay = [Col(LONG, np.arange(3, dtype=np.int64))]
    check([a, b], pay, [a, b], oracle.JOIN_FULL, 0, 3, 3)
    check([a, b], pay, [a, b], oracle.JOIN_INNER, 0, 2, 2)
    # empty build side: every probe row unmatched; empty probe batches translate to nothing
    e = [Col(INT, np.zeros(0, np.int32)), Col(INT, np.zeros(0, np.int32))]
    check(e, [Col(LONG, np.zeros(0, np.int64))], [a, b], oracle.JOIN_LEFT, 0, 4, 2)


def test_gpu_multi_key_rejects_mismatched_columns():
    j = capi.HashJoin([INT, LONG], [LONG])
    k0 = capi.DeviceColumn(INT, np.arange(4, dtype=np.int32))
    k1 = capi.DeviceColumn(INT, np.arange(4, dtype=np.int32))  # second column should be LONG
    p = capi.DeviceColumn(LONG, np.arange(4, dtype=np.int64))
    with pytest.raises(capi.GpuError):
        j.add_build([k0.c(), k1.c()], [p.c()], 4)
    j.close()
