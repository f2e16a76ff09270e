"""GPU join types other than inner (omni_gpu_join_set_type / probe / probe_exists / outer_output through the C
ABI) against the CPU oracle and the reference's fixtures.  Output rows compare bit-exact as row sets."""
import numpy as np
import pytest

import oracle
from oracle import INT, LONG, SHORT, Col
from omnioperator_amd import capi, workloads
from tests.helpers import load_golden
from tests.test_oracle_join_types import ALL_TYPES, _sort_key, key_col

pytestmark = pytest.mark.gpu
FIX = load_golden("operator_fixtures.json")["join_types"]


def keeps_build(jt, bs):
    return oracle.join_mode(jt, bs)[1]


def gpu_typed(bk, bpay, pk, build_batch, probe_batch, jt, bs, strategy=0, batched=False, repeat=1, partitions=None,
              info=None):
    """GPU output: (probe rows, payload lists), outer (build rows, payload lists) or None, existence flags or None.
    repeat > 1 probes the same batches again (the visited marks accumulate, the outer pass is unchanged)."""
    j = capi.HashJoin(bk.type_id, [p.type_id for p in bpay], strategy)
    j.set_type(jt, bs)
    if partitions is not None:
        j.tune(1, partitions)
    keep = []
    for a, b in workloads.batches(len(bk), build_batch):
        kd = capi.DeviceColumn(bk.type_id, bk.values[a:b], None if bk.nulls is None else bk.nulls[a:b])
        pd = [capi.DeviceColumn(p.type_id, p.values[a:b], None if p.nulls is None else p.nulls[a:b]) for p in bpay]
        keep.append((kd, pd))
        j.add_build(kd.c(), [d.c() for d in pd], b - a)
    j.finish()
    del keep
    cols = [(a, b, capi.DeviceColumn(pk.type_id, pk.values[a:b], None if pk.nulls is None else pk.nulls[a:b]))
            for a, b in workloads.batches(len(pk), probe_batch)]
    if jt == oracle.JOIN_EXISTENCE:
        flags = np.concatenate([j.probe_exists(c.c(), b - a) for a, b, c in cols]) if cols else np.zeros(0)
        j.close()
        return None, None, flags
    cap = max(1, len(pk) * max(1, j.info()["max_dup"]))
    rows, pays = [], [[] for _ in bpay]
    for _ in range(repeat):
        rows, pays = [], [[] for _ in bpay]
        if batched:
            out_rows, po, oarr = j._outs(cap, True)
            m = j.probe_batches_raw([c.c() for _, _, c in cols], [b - a for a, b, _ in cols], out_rows, oarr, cap)
            rows.append(out_rows.numpy(m))
            for i, (v, nb) in enumerate(po):
                pays[i].append((v.numpy(m), capi.unpack_nulls(nb.numpy(), m)))
        else:
            for a, b, c in cols:
                r, p = j.probe(c.c(), b - a, row_base=a, cap=cap)
                rows.append(r)
                for i, x in enumerate(p):
                    pays[i].append(x)
    if info is not None:
        info["last_partitions"] = j.last_partitions()
    outer = None
    if keeps_build(jt, bs):
        orow, op = j.outer_output()
        outer = (orow, [list(zip(v.tolist(), n.tolist())) for v, n in op])
    j.close()
    rows = np.concatenate(rows) if rows else np.zeros(0, np.int64)
    plist = [list(zip(np.concatenate([x[0] for x in pl]).tolist(), np.concatenate([x[1] for x in pl]).tolist()))
             if pl else [] for pl in pays]
    return (rows, plist), outer, None


def val(v, n):
    return None if n else v


def gpu_set(rows, plist, with_pay=True):
    if not with_pay:
        return sorted(rows.tolist())
    return sorted(zip(rows.tolist(), *[[val(v, n) for v, n in pl] for pl in plist]), key=_sort_key)


def oracle_set(res, with_pay=True):
    if not with_pay:
        return sorted(res.probe_rows.tolist())
    cols = [[None if n else v for v, n in zip(c.values.tolist(), c.nulls.tolist())] for c in res.payloads]
    return sorted(zip(res.probe_rows.tolist(), *cols), key=_sort_key)


def outer_set(rows, plist):
    return sorted(zip(rows.tolist(), *[[val(v, n) for v, n in pl] for pl in plist]), key=_sort_key)


def oracle_outer_set(res):
    cols = [[None if n else v for v, n in zip(c.values.tolist(), c.nulls.tolist())] for c in res.outer_payloads]
    return sorted(zip(res.outer_build_rows.tolist(), *cols), key=_sort_key)


def make_inputs(rng, dense, kind, nb=3000, npr=20000):
    hi = 6000 if dense else 2**62
    if kind == "inline":  # unique keys, one int32 payload without NULLs: the payload lives in the slot
        bkv = rng.choice(hi, nb, replace=False).astype(np.int64) if dense else rng.integers(0, hi, nb, dtype=np.int64)
        pays = [Col(INT, rng.integers(-2**31, 2**31, nb, dtype=np.int64).astype(np.int32))]
        bnull = None
    else:  # duplicate keys, NULL build keys, gathered payloads with NULLs
        pool = rng.integers(0, hi, 1500, dtype=np.int64)
        bkv = pool[rng.integers(0, len(pool), nb)]
        pays = [Col(LONG, rng.integers(-2**63, 2**63 - 1, nb, dtype=np.int64)),
                Col(SHORT, rng.integers(-2**15, 2**15, nb).astype(np.int16), rng.random(nb) < 0.1)]
        bnull = rng.random(nb) < 0.05
    pkv = np.where(rng.random(npr) < 0.6, bkv[rng.integers(0, nb, npr)], rng.integers(0, hi, npr, dtype=np.int64))
    return Col(LONG, bkv, bnull), pays, Col(LONG, pkv, rng.random(npr) < 0.03)


def check(bk, pays, pk, jt, bs, bb, pbs, strategy=0, batched=False, repeat=1, partitions=None, info=None):
    want = oracle.hash_join_typed(bk, pays, pk, bb, pbs, jt, bs)
    got, outer, flags = gpu_typed(bk, pays, pk, bb, pbs, jt, bs, strategy, batched, repeat, partitions, info)
    if jt == oracle.JOIN_EXISTENCE:
        assert flags.tolist() == want.exists.tolist()
        return
    with_pay = jt not in (oracle.JOIN_LEFT_SEMI, oracle.JOIN_LEFT_ANTI)
    assert gpu_set(*got, with_pay) == oracle_set(want, with_pay)
    if keeps_build(jt, bs):
        assert outer_set(*outer) == oracle_outer_set(want)


@pytest.mark.parametrize("jt,bs", ALL_TYPES)
@pytest.mark.parametrize("dense", [False, True])
@pytest.mark.parametrize("kind", ["inline", "dups"])
def test_join_types_parity(jt, bs, dense, kind):
    rng = np.random.default_rng(1000 + 10 * jt + bs + 2 * dense + (kind == "dups"))
    bk, pays, pk = make_inputs(rng, dense, kind)
    check(bk, pays, pk, jt, bs, 1024, 6000, strategy=0, batched=(jt + bs) % 2 == 0)


@pytest.mark.parametrize("jt,bs", [t for t in ALL_TYPES if t[0] != oracle.JOIN_EXISTENCE])
@pytest.mark.parametrize("partitions", [16, 64, 8])
@pytest.mark.parametrize("dups", [False, True])
def test_join_types_partitioned(jt, bs, partitions, dups):
    """The radix-partitioned window probe with the join-type rules, and the separate launch for NULL probe keys.
    The hash table has 131,072 slots: 16 and 64 partitions (regions of 8,192 and 2,048 slots) take the window
    probe; 8 (regions of 16,384 slots) falls back to the direct probe (the inner join takes the L2 probe)."""
    rng = np.random.default_rng(50 + 10 * jt + bs + partitions + dups)
    nb, npr = 30_000, 250_001
    if dups:
        bkv = rng.integers(2**40, 2**40 + 20_000, nb, dtype=np.int64)
    else:
        bkv = rng.permutation(np.arange(-2**50, -2**50 + 5 * nb, 5, dtype=np.int64))[:nb]
    bkv[7] = -2**63
    pkv = np.where(rng.random(npr) < 0.7, bkv[rng.integers(0, nb, npr)],
                   rng.integers(-2**63, 2**63 - 1, npr, dtype=np.int64))
    pkv[[0, 99, npr - 1]] = -2**63
    bk = Col(LONG, bkv, rng.random(nb) < 0.01)
    pk = Col(LONG, pkv, rng.random(npr) < 0.02)
    pays = [Col(INT, rng.integers(-2**31, 2**31, nb, dtype=np.int64).astype(np.int32))]
    info = {}
    check(bk, pays, pk, jt, bs, 8192, 65536, strategy=1, batched=True, partitions=partitions, info=info)
    typed = oracle.join_mode(jt, bs) != (0, False)  # the plain inner join keeps the per-XCD L2 probe at 8
    assert info["last_partitions"] == (1 if typed and partitions == 8 else partitions)


@pytest.mark.parametrize("jt,bs", [(oracle.JOIN_FULL, 0), (oracle.JOIN_LEFT, oracle.BUILD_LEFT),
                                   (oracle.JOIN_RIGHT, oracle.BUILD_RIGHT)])
def test_outer_output_after_repeated_probes(jt, bs):
    """The visited marks accumulate over every probe call; probing the same rows twice changes nothing."""
    rng = np.random.default_rng(7)
    bk, pays, pk = make_inputs(rng, False, "dups")
    check(bk, pays, pk, jt, bs, 1024, 6000, strategy=1, repeat=2)


@pytest.mark.parametrize("jt", [oracle.JOIN_LEFT, oracle.JOIN_FULL, oracle.JOIN_LEFT_SEMI, oracle.JOIN_LEFT_ANTI,
                                oracle.JOIN_EXISTENCE, oracle.JOIN_RIGHT])
def test_empty_build_side(jt):
    """No build rows: outer and anti joins emit every probe row, semi none, existence all false."""
    rng = np.random.default_rng(3)
    bk = Col(LONG, np.zeros(0, np.int64))
    pays = [Col(LONG, np.zeros(0, np.int64))]
    pk = Col(LONG, rng.integers(0, 100, 5000, dtype=np.int64), rng.random(5000) < 0.1)
    check(bk, pays, pk, jt, oracle.BUILD_UNKNOWN, 1024, 2000)


@pytest.mark.parametrize("case", [k for k in FIX if not k.startswith("_") and k != "right" and k != "left_semi"])
def test_join_type_fixture_on_gpu(case):
    """The reference fixtures with integer columns, run through the GPU path (payload = the build row's other
    column; the key column's NULLs as in the source)."""
    f = FIX[case]
    kc = f["key_col"]
    bcols, pcols = f["build"], f["probe"]
    nb = len(bcols[0])
    other = 1 - kc
    bk = key_col(bcols[kc])
    pay = [key_col(bcols[other])]
    pk = key_col(pcols[kc])
    want = oracle.hash_join_typed(bk, pay, pk, nb, len(pcols[0]), f["join_type"], f["build_side"])
    got, outer, flags = gpu_typed(bk, pay, pk, nb, len(pcols[0]), f["join_type"], f["build_side"])
    if "expected_exists" in f:
        assert flags.astype(bool).tolist() == f["expected_exists"]
        return
    with_pay = f["join_type"] not in (oracle.JOIN_LEFT_SEMI, oracle.JOIN_LEFT_ANTI)
    assert gpu_set(*got, with_pay) == oracle_set(want, with_pay)
    if keeps_build(f["join_type"], f["build_side"]):
        assert outer_set(*outer) == oracle_outer_set(want)


def test_type_errors():
    j = capi.HashJoin(LONG, [LONG])
    with pytest.raises(capi.GpuError):
        j.set_type(7)
    with pytest.raises(capi.GpuError):
        j.set_type(oracle.JOIN_LEFT, 3)
    j.set_type(oracle.JOIN_EXISTENCE)
    kd = capi.DeviceColumn(LONG, np.array([1, 2], np.int64))
    pd = capi.DeviceColumn(LONG, np.array([5, 6], np.int64))
    j.add_build(kd.c(), [pd.c()], 2)
    j.finish()
    with pytest.raises(capi.GpuError) as e:  # an existence join is probed with probe_exists
        j.probe(kd.c(), 2)
    assert e.value.code == -6
    with pytest.raises(capi.GpuError) as e:  # no outer pass without a build-preserving type
        j.outer_output()
    assert e.value.code == -6
    with pytest.raises(capi.GpuError) as e:  # the type is fixed once the build finished
        j.set_type(oracle.JOIN_INNER)
    assert e.value.code == -6
    assert j.probe_exists(kd.c(), 2).tolist() == [1, 1]
    j.close()
