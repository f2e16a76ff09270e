"""GPU inner hash join (HIP kernels through the C ABI) against the CPU oracle and the reference fixtures.
Join matches are bit-exact as (probe row, payload) row sets."""
import numpy as np
import pytest

import oracle
from oracle import BOOLEAN, DOUBLE, INT, LONG, SHORT, Col
from omnioperator_amd import capi, workloads
from tests.helpers import load_golden

pytestmark = pytest.mark.gpu
FIX = load_golden("operator_fixtures.json")


def gpu_join(bk: Col, bpay, pk: Col, build_batch, probe_batch, strategy=0, batched=False, probe_group=None,
             partitions=-1):
    j = capi.HashJoin(bk.type_id, [p.type_id for p in bpay], strategy)
    if probe_group is not None or partitions != -1:
        j.tune(probe_group or 1, partitions)
    keep = []  # build batches stay alive until finish (the copies are asynchronous)
    for a, b in workloads.batches(len(bk), build_batch):
        kd = capi.DeviceColumn(bk.type_id, bk.values[a:b], None if bk.nulls is None else bk.nulls[a:b])
        pd = [capi.DeviceColumn(p.type_id, p.values[a:b], None if p.nulls is None else p.nulls[a:b]) for p in bpay]
        keep.append((kd, pd))
        j.add_build(kd.c(), [d.c() for d in pd], b - a)
    j.finish()
    del keep
    info = j.info()
    rows_out, pays_out = [], [[] for _ in bpay]
    if batched:
        cols = [capi.DeviceColumn(pk.type_id, pk.values[a:b], None if pk.nulls is None else pk.nulls[a:b])
                for a, b in workloads.batches(len(pk), probe_batch)]
        cap = len(pk) * max(1, info["max_dup"])
        out_rows, pays, oarr = j._outs(cap, True)
        m = j.probe_batches_raw([c.c() for c in cols], [c.rows for c in cols], out_rows, oarr, cap)
        rows_out = [out_rows.numpy(m)]
        for i, (v, nb) in enumerate(pays):
            pays_out[i].append((v.numpy(m), capi.unpack_nulls(nb.numpy(), m)))
    else:
        for a, b in workloads.batches(len(pk), probe_batch):
            kd = capi.DeviceColumn(pk.type_id, pk.values[a:b], None if pk.nulls is None else pk.nulls[a:b])
            r, p = j.probe(kd.c(), b - a, row_base=a)
            rows_out.append(r)
            for i, x in enumerate(p):
                pays_out[i].append(x)
    j.close()
    rows = np.concatenate(rows_out) if rows_out else np.zeros(0, np.int64)
    pays = [(np.concatenate([x[0] for x in pl]) if pl else np.zeros(0),
             np.concatenate([x[1] for x in pl]) if pl else np.zeros(0, bool)) for pl in pays_out]
    return rows, pays, info


def as_set(rows, pays):
    cols = [rows.tolist()] + [[None if n else int(v) for v, n in zip(pv, pn)] for pv, pn in pays]
    return sorted(zip(*cols), key=lambda r: tuple((x is None, x or 0) for x in r))


def oracle_set(bk, bpay, pk, bb, pb):
    r = oracle.hash_join(bk, bpay, pk, bb, pb)
    return as_set(r.probe_rows, [(c.values, c.nulls) for c in r.payloads]), r.is_array


def test_join_simple_fixture():
    f = FIX["join_simple"]
    bk = Col(LONG, np.array(f["build_keys"], np.int64))
    bp = Col(LONG, np.array(f["build_payload"], np.int64))
    pk = Col(LONG, np.array(f["probe_keys"], np.int64))
    rows, pays, info = gpu_join(bk, [bp], pk, 1024, 1024)
    pp = np.array(f["probe_payload"], np.int64)
    got = sorted(zip(pp[rows].tolist(), pays[0][0].tolist()))
    want = sorted(zip(pp[np.array(f["expected_probe_rows"])].tolist(), f["expected_build_payload"]))
    assert got == want
    assert info["max_dup"] == 3


@pytest.mark.parametrize("strategy", [0, 1])
def test_join_array_table_fixture(strategy):
    f = FIX["join_array_table"]
    bk = Col(SHORT, np.array(f["build_keys"], np.int16))
    bp = Col(SHORT, np.array(f["build_payload"], np.int16))
    pk = np.array(f["probe_keys"], np.int16)
    pp = np.array(f["probe_payload"], np.int16)
    rows, pays, info = gpu_join(bk, [bp], Col(SHORT, pk), 1024, 1024, strategy=strategy)
    assert info["is_array"] == (strategy == 0)
    got = sorted(zip(pk[rows].tolist(), pp[rows].tolist(), pk[rows].tolist(), pays[0][0].tolist()))
    assert got == sorted(tuple(r) for r in f["expected_rows"])


@pytest.mark.parametrize("key_type", [DOUBLE, oracle.DECIMAL64, BOOLEAN])
def test_array_table_only_for_integer_key_types(key_type):
    """JoinHashTableVariants::BuildHashTable tries the array table for INT / DATE32 / SHORT / LONG / TIMESTAMP keys
    only (join_hash_table_variants.cpp:700-736): dense DOUBLE, DECIMAL64 and BOOLEAN keys build a hash table, on the
    GPU as in the oracle, with the same matches."""
    rng = np.random.default_rng(7 + key_type)
    if key_type == BOOLEAN:
        bkv, pkv = rng.integers(0, 2, 500).astype(np.uint8), rng.integers(0, 2, 2000).astype(np.uint8)
    else:
        ints = np.arange(1, 3001, dtype=np.int64)
        pint = rng.integers(0, 3500, 20_000, dtype=np.int64)
        bkv = ints.view(np.float64) if key_type == DOUBLE else ints
        pkv = pint.view(np.float64) if key_type == DOUBLE else pint
    bk = Col(key_type, bkv)
    pk = Col(key_type, pkv)
    pays = [Col(INT, np.arange(len(bkv), dtype=np.int32))]
    rows, gp, info = gpu_join(bk, pays, pk, 1024, 4096)
    want, is_array = oracle_set(bk, pays, pk, 1024, 4096)
    assert not is_array and info["is_array"] == 0
    assert as_set(rows, gp) == want


def test_destroy_releases_device_memory():
    """create / build / probe / destroy in a loop leaves the library's live device memory where it started: the
    partitioned build (tables of more than 16,384 build rows), its overflow list and the outer join's visited
    marks are all released."""
    rng = np.random.default_rng(5)
    bk = Col(LONG, rng.permutation(np.arange(40_000, dtype=np.int64)) * 7919)
    pk = Col(LONG, bk.values[rng.integers(0, 40_000, 50_000)])
    pays = [Col(INT, np.arange(40_000, dtype=np.int32))]
    kd = capi.DeviceColumn(LONG, bk.values)
    pd = capi.DeviceColumn(INT, pays[0].values)
    qd = capi.DeviceColumn(LONG, pk.values)
    used = []
    for it in range(4):
        j = capi.HashJoin(LONG, [INT], 0)
        if it % 2:
            j.set_type(capi.JOIN_FULL, capi.BUILD_RIGHT)
        j.add_build(kd.c(), [pd.c()], len(bk))
        j.finish()
        assert j.info()["is_array"] == 0
        j.probe(qd.c(), len(pk))
        j.close()
        used.append(capi.memory_stats()[0])
    assert used[1] == used[0] and used[2] == used[0] and used[3] == used[0], used


@pytest.mark.parametrize("dense", [False, True])
@pytest.mark.parametrize("nulls", [False, True])
@pytest.mark.parametrize("batched", [False, True])
def test_random_parity(dense, nulls, batched):
    rng = np.random.default_rng(100 + dense + 2 * nulls + 4 * batched)
    nb, npr = 50_000, 300_000
    if dense:
        bkv = rng.integers(0, 30_000, nb, dtype=np.int64)
        pkv = rng.integers(-1000, 32_000, npr, dtype=np.int64)
    else:
        pool = rng.integers(-2**63, 2**63 - 1, 40_000, dtype=np.int64)
        bkv = pool[rng.integers(0, 40_000, nb)]
        pkv = np.where(rng.random(npr) < 0.8, pool[rng.integers(0, 40_000, npr)],
                       rng.integers(-2**63, 2**63 - 1, npr, dtype=np.int64))
    bk = Col(LONG, bkv, rng.random(nb) < 0.03 if nulls else None)
    pk = Col(LONG, pkv, rng.random(npr) < 0.03 if nulls else None)
    pays = [Col(INT, rng.integers(-2**31, 2**31, nb, dtype=np.int64).astype(np.int32), rng.random(nb) < 0.1 if nulls else None),
            Col(LONG, rng.integers(-2**63, 2**63 - 1, nb, dtype=np.int64))]
    rows, gp, info = gpu_join(bk, pays, pk, 8192, 65536, batched=batched)
    want, is_array = oracle_set(bk, pays, pk, 8192, 65536)
    assert info["is_array"] == is_array
    assert as_set(rows, gp) == want


@pytest.mark.parametrize("probe_group", [1, 2, 4, 16])
@pytest.mark.parametrize("unique", [False, True])
def test_probe_width_parity(probe_group, unique):
    """Every probe width gives the oracle's row set: hash table with duplicate keys (chains), unique keys
    with an inline payload, and the array table; NULL probe keys and a ragged last batch included."""
    rng = np.random.default_rng(7 + probe_group)
    nb, npr = 20_000, 100_003
    if unique:
        bkv = rng.permutation(np.arange(-2**40, -2**40 + 3 * nb, 3, dtype=np.int64))[:nb]
    else:
        bkv = rng.integers(-2**62, -2**62 + 15_000, nb, dtype=np.int64)
    pkv = np.where(rng.random(npr) < 0.7, bkv[rng.integers(0, nb, npr)], rng.integers(-2**63, 2**63 - 1, npr, dtype=np.int64))
    bk = Col(LONG, bkv)
    pk = Col(LONG, pkv, rng.random(npr) < 0.02)
    pays = [Col(INT, rng.integers(-2**31, 2**31, nb, dtype=np.int64).astype(np.int32))]
    for strategy in (0, 1):
        rows, gp, _ = gpu_join(bk, pays, pk, 4096, 65536, strategy=strategy, batched=True, probe_group=probe_group)
        want, _ = oracle_set(bk, pays, pk, 4096, 65536)
        assert as_set(rows, gp) == want


def test_probe_width_rejects_other_values():
    j = capi.HashJoin(LONG, [])
    with pytest.raises(capi.GpuError):
        j.tune(3)
    with pytest.raises(capi.GpuError):
        j.tune(1, 12)
    j.close()


@pytest.mark.parametrize("partitions", [8, 16, 64, 1024])
@pytest.mark.parametrize("unique", [False, True])
@pytest.mark.parametrize("wide", [False, True])
def test_partitioned_probe_parity(partitions, unique, wide):
    """Radix-partitioned probe against the oracle.  The table has 131,072 slots: 8 partitions (regions of 16,384
    slots) take the per-XCD persistent L2 probe, 16 / 64 / 1024 (regions of 8,192 / 2,048 / 128 slots) the LDS
    window probe.  Duplicate-key chains and inline payloads (one int32) or gathered payloads (int64 + int16 with
    NULLs), NULL probe keys, the slot sentinel as a real key on both sides, ragged batches."""
    rng = np.random.default_rng(11 + partitions + unique + 5 * wide)
    nb, npr = 30_000, 250_001
    if unique:
        bkv = rng.permutation(np.arange(-2**50, -2**50 + 5 * nb, 5, dtype=np.int64))[:nb]
    else:
        bkv = rng.integers(2**40, 2**40 + 20_000, nb, dtype=np.int64)
    bkv[7] = -2**63
    pkv = np.where(rng.random(npr) < 0.75, bkv[rng.integers(0, nb, npr)],
                   rng.integers(-2**63, 2**63 - 1, npr, dtype=np.int64))
    pkv[[0, 99, npr - 1]] = -2**63
    bk = Col(LONG, bkv)
    pk = Col(LONG, pkv, rng.random(npr) < 0.02)
    if wide:
        pays = [Col(LONG, rng.integers(-2**63, 2**63 - 1, nb, dtype=np.int64)),
                Col(SHORT, rng.integers(-2**15, 2**15, nb).astype(np.int16), rng.random(nb) < 0.1)]
    else:
        pays = [Col(INT, rng.integers(-2**31, 2**31, nb, dtype=np.int64).astype(np.int32))]
    rows, gp, info = gpu_join(bk, pays, pk, 8192, 65536, strategy=1, batched=True, partitions=partitions)
    want, _ = oracle_set(bk, pays, pk, 8192, 65536)
    assert as_set(rows, gp) == want
    assert not info["is_array"] and info["table_slots"] == 131072


_FIB = 0x9E3779B97F4A7C15
_FIB_INV = pow(_FIB, -1, 1 << 64)


def keys_with_homes(homes, bits, rng):
    """int64 keys whose home slot in a table of 2**bits slots is `homes`: the home slot is the top `bits` bits of
    (k ^ k >> 29) * 0x9E3779B97F4A7C15 mod 2**64, a bijection of k, so a chosen product is inverted (odd-constant
    inverse, then the inverse of the xor-shift)."""
    homes = np.asarray(homes, np.uint64)
    low = rng.integers(0, 1 << (64 - bits), len(homes), dtype=np.uint64)
    x = (homes << np.uint64(64 - bits)) | low
    m = x * np.uint64(_FIB_INV)
    k = m ^ (m >> np.uint64(29)) ^ (m >> np.uint64(58))
    y = (k ^ (k >> np.uint64(29))) * np.uint64(_FIB)  # forward map, checked
    assert np.array_equal(y >> np.uint64(64 - bits), homes)
    return k.view(np.int64)


@pytest.mark.parametrize("kind", ["inline", "gathered", "dups"])
def test_window_probe_fallbacks(kind):
    """The LDS window probe's two global-walk cases: a region holding more keys than its LDS table takes (4,000
    keys homed in region 0), and a region whose probe runs fill the slots staged past its end (400 keys homed
    in the last slot of region 1).  10,000 keys: a 65,536-slot table in 8 regions of 8,192 slots."""
    rng = np.random.default_rng(5)
    bits = 16
    homes = np.concatenate([rng.integers(0, 8192, 4000), np.full(400, 16383), rng.integers(0, 1 << bits, 5600)])
    bkv = keys_with_homes(homes, bits, rng)
    assert len(np.unique(bkv)) == len(bkv)
    miss = keys_with_homes(np.concatenate([rng.integers(0, 8192, 500), np.full(50, 16383)]), bits, rng)
    if kind == "dups":
        bkv = np.concatenate([bkv, bkv[rng.integers(0, len(bkv), 3000)]])
    pkv = np.concatenate([bkv, miss, rng.integers(-2**63, 2**63 - 1, 2000, dtype=np.int64)])
    pkv = pkv[rng.permutation(len(pkv))]
    nb = len(bkv)
    if kind == "inline":
        pays = [Col(INT, rng.integers(-2**31, 2**31, nb, dtype=np.int64).astype(np.int32))]
    else:
        pays = [Col(LONG, rng.integers(-2**63, 2**63 - 1, nb, dtype=np.int64))]
    bk, pk = Col(LONG, bkv), Col(LONG, pkv)
    rows, gp, info = gpu_join(bk, pays, pk, 4096, 5000, strategy=1, batched=True, partitions=8)
    want, _ = oracle_set(bk, pays, pk, 4096, 5000)
    assert info["table_slots"] == 1 << bits
    assert as_set(rows, gp) == want


def test_sentinel_key_inline_payload():
    """Unique keys with a <= 4-byte payload keep the payload in the slot; the sentinel key's row lives
    outside the table and must still return its own payload."""
    bk = Col(LONG, np.array([-2**63, 3, 9, 2**63 - 1], np.int64))
    pk = Col(LONG, np.array([9, -2**63, 4, -2**63, 3], np.int64))
    pay = [Col(INT, np.array([101, 102, 103, 104], np.int32))]
    for batched in (False, True):
        rows, gp, info = gpu_join(bk, pay, pk, 4, 4, strategy=1, batched=batched)
        want, _ = oracle_set(bk, pay, pk, 4, 4)
        assert as_set(rows, gp) == want
        assert info["max_dup"] == 1


@pytest.mark.parametrize("key_type", [INT, SHORT])
def test_narrow_keys_and_sentinel(key_type):
    dt = oracle.np_dtype(key_type)
    info_ = np.iinfo(dt)
    bkv = np.array([info_.min, info_.max, 0, -1, info_.min, 5], dt)
    pkv = np.array([info_.min, 5, 7, info_.max, -1, -1], dt)
    bk, pk = Col(key_type, bkv), Col(key_type, pkv)
    pay = [Col(INT, np.arange(6, dtype=np.int32))]
    for strategy in (0, 1):
        rows, gp, _ = gpu_join(bk, pay, pk, 4, 3, strategy=strategy)
        want, _ = oracle_set(bk, pay, pk, 4, 3)
        assert as_set(rows, gp) == want


def test_int64_sentinel_key():
    """INT64_MIN is the internal empty-slot marker; as a real key it must still join."""
    bk = Col(LONG, np.array([-2**63, 3, -2**63, 2**63 - 1], np.int64))
    pk = Col(LONG, np.array([-2**63, 2**63 - 1, 4, -2**63], np.int64))
    pay = [Col(LONG, np.array([10, 11, 12, 13], np.int64))]
    rows, gp, info = gpu_join(bk, pay, pk, 2, 2, strategy=1)
    want, _ = oracle_set(bk, pay, pk, 2, 2)
    assert as_set(rows, gp) == want
    assert info["max_dup"] == 2


def test_output_full_reports_needed_size():
    bk = Col(LONG, np.array([1, 1, 1, 2], np.int64))
    pk = Col(LONG, np.array([1, 2, 1], np.int64))
    j = capi.HashJoin(LONG, [LONG])
    kd = capi.DeviceColumn(LONG, bk.values)
    pd = capi.DeviceColumn(LONG, np.arange(4, dtype=np.int64))
    j.add_build(kd.c(), [pd.c()], 4)
    j.finish()
    pcol = capi.DeviceColumn(LONG, pk.values)
    with pytest.raises(capi.GpuError) as e:
        j.probe(pcol.c(), 3, cap=3)
    assert e.value.code == -5
    r, p = j.probe(pcol.c(), 3, cap=7)
    assert len(r) == 7
    j.close()


def test_probe_before_finish_is_an_error():
    j = capi.HashJoin(LONG, [])
    pcol = capi.DeviceColumn(LONG, np.array([1], np.int64))
    with pytest.raises(capi.GpuError) as e:
        j.probe(pcol.c(), 1, cap=1)
    assert e.value.code == -6
    j.close()


@pytest.mark.parametrize("dense", [False, True])
@pytest.mark.parametrize("partitions", [-1, 0])
def test_inner_join_config_full_size(dense, partitions):
    """Build 2M unique int64 keys + int32 payload, probe 64M keys (80% hits) in batches of 65,536, all
    batches in one launch.  Size-independent checks: exact match count, every hit probe row exactly once,
    every payload equal to its build row's payload; plus bit-exact parity with the oracle on a 1M-row slice."""
    w = workloads.join_inner(dense=dense)
    bk = Col(LONG, w.build_keys)
    j = capi.HashJoin(LONG, [INT])
    j.tune(1, partitions)
    keep = []
    for a, b in workloads.batches(len(w.build_keys), w.batch_rows):
        kd, pd = capi.DeviceColumn(LONG, w.build_keys[a:b]), capi.DeviceColumn(INT, w.build_payload[a:b])
        keep.append((kd, pd))
        j.add_build(kd.c(), [pd.c()], b - a)
    j.finish()
    info = j.info()
    assert info["is_array"] == dense and info["max_dup"] == 1
    probe = capi.DeviceColumn(LONG, w.probe_keys)
    spans = workloads.batches(len(w.probe_keys), w.batch_rows)
    cols = [probe.c(a) for a, _ in spans]
    cap = len(w.probe_keys)
    out_rows, pays, oarr = j._outs(cap, False)
    m = j.probe_batches_raw(cols, [b - a for a, b in spans], out_rows, oarr, cap)
    assert m == w.hits
    # automatic mode partitions the 128 MiB table into 1024 regions of 8192 slots (LDS window probe)
    assert j.last_partitions() == (1 if dense or partitions == 0 else 1024)
    rows = out_rows.numpy(m)
    pay = pays[0][0].numpy(m)
    order = np.argsort(w.build_keys, kind="stable")
    pos = order[np.searchsorted(w.build_keys[order], w.probe_keys[rows])]
    assert np.array_equal(w.build_keys[pos], w.probe_keys[rows])
    assert np.array_equal(pay, w.build_payload[pos])
    hit_rows = np.sort(rows)
    assert np.all(np.diff(hit_rows) > 0)  # each probe row at most once (unique build keys)
    j.close()
    # bit-exact against the oracle on a slice
    sl = slice(0, 1_000_000)
    r = oracle.hash_join(bk, [Col(INT, w.build_payload)], Col(LONG, w.probe_keys[sl]), w.batch_rows, w.batch_rows)
    sel = rows < 1_000_000
    assert sorted(zip(rows[sel].tolist(), pay[sel].tolist())) == sorted(zip(r.probe_rows.tolist(),
                                                                             r.payloads[0].values.tolist()))


@pytest.mark.parametrize("strategy", [0, 1])
def test_double_keys_join_on_bit_patterns(strategy):
    """DOUBLE keys join on their 64-bit pattern, as the reference's int64 table compares them: 0.0 and -0.0 differ
    (-0.0's pattern is the slot sentinel), NaNs match only the same pattern."""
    rng = np.random.default_rng(21)
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1.5, -1.5], np.float64)
    other_nan = np.array([0x7FF8000000000001], np.int64).view(np.float64)
    bkv = np.concatenate([special, rng.normal(0, 1e3, 3000).round(2)])
    bkv = np.concatenate([bkv, bkv[rng.integers(0, len(bkv), 500)]])  # duplicates
    pkv = np.concatenate([special, other_nan, bkv[rng.integers(0, len(bkv), 20000)],
                          rng.normal(0, 1e3, 5000).round(2)])
    bk = Col(DOUBLE, bkv, rng.random(len(bkv)) < 0.02)
    pk = Col(DOUBLE, pkv, rng.random(len(pkv)) < 0.02)
    pays = [Col(LONG, np.arange(len(bkv), dtype=np.int64))]
    rows, gp, _ = gpu_join(bk, pays, pk, 1024, 4096, strategy=strategy)
    want, _ = oracle_set(bk, pays, pk, 1024, 4096)
    assert as_set(rows, gp) == want
    # probe row 1 is -0.0: it never meets a +0.0 build row (payload = build row id)
    pos0 = {i for i, v in enumerate(bkv) if v == 0 and not np.signbit(v)}
    assert not ({b for p, b in zip(rows.tolist(), gp[0][0].tolist()) if p == 1} & pos0)


def test_boolean_keys():
    rng = np.random.default_rng(22)
    bkv = (rng.random(300) < 0.5).astype(np.uint8)
    pkv = (rng.random(1000) < 0.5).astype(np.uint8)
    bk = Col(BOOLEAN, bkv, rng.random(300) < 0.1)
    pk = Col(BOOLEAN, pkv, rng.random(1000) < 0.1)
    pays = [Col(INT, np.arange(300, dtype=np.int32))]
    for strategy in (0, 1):
        rows, gp, _ = gpu_join(bk, pays, pk, 128, 256, strategy=strategy)
        want, _ = oracle_set(bk, pays, pk, 128, 256)
        assert as_set(rows, gp) == want


@pytest.mark.gpu
def test_gather_rows_matches_numpy():
    """omni_gpu_join_gather_rows: output row k of each column is the input row rows[k], value and NULL bit, for
    1-, 2-, 4-, 8- and 16-byte columns (DECIMAL128), a row count that ends inside a wave and a NULL-free column."""
    rng = np.random.default_rng(31)
    src_rows, n = 5000, 1000 + 37
    specs = [(BOOLEAN, np.uint8), (SHORT, np.int16), (INT, np.int32), (LONG, np.int64), (oracle.DECIMAL128, None)]
    cols, host = [], []
    for i, (t, dt) in enumerate(specs):
        vals = (rng.integers(0, 2**62, (src_rows, 2), dtype=np.int64) if dt is None
                else rng.integers(0, 100, src_rows).astype(dt))
        nulls = rng.random(src_rows) < (0.2 if i != 3 else 0.0)
        cols.append(capi.DeviceColumn(t, vals, nulls))
        host.append((vals, nulls))
    rows = rng.integers(0, src_rows, n, dtype=np.int64)
    drows = capi.DeviceArray.from_numpy(rows)
    outs, keep = (capi.OutColumn * len(specs))(), []
    for i, (t, _) in enumerate(specs):
        v = capi.DeviceArray(n * capi.WIDTH[t])
        b = capi.DeviceArray(4 * ((n + 31) // 32))
        keep += [v, b]
        outs[i].type_id, outs[i].values, outs[i].nulls = t, v.ptr, b.ptr
    carr = (capi.Column * len(specs))(*[c.c() for c in cols])
    capi.check(capi.lib().omni_gpu_join_gather_rows(carr, len(specs), drows.ptr, n, outs))
    for i, (t, dt) in enumerate(specs):
        vals, nulls = host[i]
        got = keep[2 * i].numpy().view(np.int64).reshape(n, 2) if dt is None else keep[2 * i].numpy().view(dt)[:n]
        bits = np.unpackbits(keep[2 * i + 1].numpy(), bitorder="little")[:n].astype(bool)
        np.testing.assert_array_equal(bits, nulls[rows])
        np.testing.assert_array_equal(got, vals[rows])


def _probe_partitions_oracle(cols, mask):
    """PopulateProbeHashes + GetRawHashPartition from the oracle's KAT-pinned HashUtil primitives
    (hash_value_long / _double / _decimal128, raw_hash_partition): h = 0, then h = 31 h + HashValue(v) for each
    non-NULL key column, wrapping in 64 bits."""
    n = len(cols[0][1])
    out = np.zeros(n, np.uint32)
    for r in range(n):
        h = 0
        for t, vals, nulls in cols:
            if nulls[r]:
                continue
            if t == BOOLEAN:
                hv = 1231 if vals[r] else 1237
            elif t == oracle.DECIMAL64:
                hv = int(vals[r])
            elif t == DOUBLE:
                hv = oracle.hash_value_double(float(vals[r]))
            elif t == oracle.DECIMAL128:
                hv = oracle.hash_value_decimal128(int(vals[r][0]), int(vals[r][1]))
            else:
                hv = oracle.hash_value_i64(int(vals[r]))
            h = (31 * h + hv) & (2**64 - 1)
        out[r] = oracle.raw_hash_partition(h - 2**64 if h >= 2**63 else h, mask)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("types", [(LONG,), (INT, LONG), (SHORT, DOUBLE, BOOLEAN), (oracle.DECIMAL64,),
                                   (oracle.DECIMAL128, INT)])
def test_join_partition_ids_match_oracle(types):
    """omni_gpu_join_partition_ids equals the oracle's composition of the reference's hash primitives, NULLs
    skipped, for 1 to 3 key columns (DOUBLE with NaN payloads and -0.0)."""
    rng = np.random.default_rng(41 + len(types))
    n = 3000
    cols, dcols = [], []
    for t in types:
        if t == DOUBLE:
            v = rng.normal(0, 1e6, n)
            v[::17] = np.nan
            v[5::19] = -0.0
            v[7::23] = np.frombuffer(np.array([0x7FF0000000000001], np.uint64).tobytes(), np.float64)[0]
        elif t == BOOLEAN:
            v = rng.integers(0, 2, n).astype(np.uint8)
        elif t == SHORT:
            v = rng.integers(-2**15, 2**15, n).astype(np.int16)
        elif t == INT:
            v = rng.integers(-2**31, 2**31, n).astype(np.int32)
        elif t == oracle.DECIMAL128:
            v = rng.integers(-2**62, 2**62, (n, 2), dtype=np.int64)
        else:
            v = rng.integers(-2**62, 2**62, n, dtype=np.int64)
        nulls = rng.random(n) < 0.1
        cols.append((t, v, nulls))
        dcols.append(capi.DeviceColumn(t, v, nulls))
    out = capi.DeviceArray(4 * n)
    arr = (capi.Column * len(types))(*[c.c() for c in dcols])
    capi.check(capi.lib().omni_gpu_join_partition_ids(arr, len(types), n, 7, out.ptr))
    got = out.numpy().view(np.uint32)[:n]
    np.testing.assert_array_equal(got, _probe_partitions_oracle(cols, 7))
