"""Pins the CPU oracle's inner hash join: reference fixtures, the array-table decision rule, and the
sort-based brute force (oracle.brute.inner_join) on seeded inputs with duplicates and NULLs."""
import numpy as np
import pytest

import oracle
from oracle import brute
from oracle import INT, LONG, SHORT, Col
from tests.helpers import load_golden

FIX = load_golden("operator_fixtures.json")


def pairs(res):
    o = np.lexsort((res.build_rows, res.probe_rows))
    return res.probe_rows[o], res.build_rows[o]


def test_join_simple_fixture():
    f = FIX["join_simple"]
    bk = Col(LONG, np.array(f["build_keys"], np.int64))
    bp = Col(LONG, np.array(f["build_payload"], np.int64))
    pk = Col(LONG, np.array(f["probe_keys"], np.int64))
    res = oracle.hash_join(bk, [bp], pk, 1024, 1024)
    probe_payload = np.array(f["probe_payload"], np.int64)
    got = sorted(zip(probe_payload[res.probe_rows].tolist(), res.payloads[0].values.tolist()))
    want = sorted(zip(probe_payload[np.array(f["expected_probe_rows"])].tolist(), f["expected_build_payload"]))
    assert got == want
    assert sorted(res.probe_rows.tolist()) == sorted(f["expected_probe_rows"])


def test_join_array_table_fixture():
    f = FIX["join_array_table"]
    bk = Col(SHORT, np.array(f["build_keys"], np.int16))
    bp = Col(SHORT, np.array(f["build_payload"], np.int16))
    pk = np.array(f["probe_keys"], np.int16)
    pp = np.array(f["probe_payload"], np.int16)
    res = oracle.hash_join(bk, [bp], Col(SHORT, pk), 1024, 1024)
    assert res.is_array
    got = sorted(zip(pk[res.probe_rows].tolist(), pp[res.probe_rows].tolist(), pk[res.probe_rows].tolist(),
                     res.payloads[0].values.tolist()))
    assert got == sorted(tuple(r) for r in f["expected_rows"])


def test_array_table_rule():
    """TryToBuildArrayTable (join_hash_table_variants.cpp:748-808): range <= 8 * 2^ceil(log2(n/0.75))."""
    n = 1000
    deg_len = 2 ** int(np.ceil(np.log2(n / 0.75)))  # 2048
    keys = np.arange(n, dtype=np.int64)
    keys[-1] = 8 * deg_len  # range == 8 * len -> array table
    assert oracle.hash_join(Col(LONG, keys), [], Col(LONG, keys[:1]), 4096, 4096).is_array
    keys[-1] = 8 * deg_len + 1  # one past -> hash table
    assert not oracle.hash_join(Col(LONG, keys), [], Col(LONG, keys[:1]), 4096, 4096).is_array


@pytest.mark.parametrize("dense", [False, True])
@pytest.mark.parametrize("nulls", [False, True])
def test_oracle_join_matches_brute_force(dense, nulls):
    rng = np.random.default_rng(11 + dense + 2 * nulls)
    nb, npr = 3000, 20000
    if dense:
        bkv = rng.integers(0, 1500, nb, dtype=np.int64)          # duplicates, dense range
        pkv = rng.integers(-100, 1700, npr, dtype=np.int64)
    else:
        pool = rng.integers(-2**63, 2**63 - 1, 2000, dtype=np.int64)
        bkv = pool[rng.integers(0, 2000, nb)]
        pkv = np.where(rng.random(npr) < 0.7, pool[rng.integers(0, 2000, npr)],
                       rng.integers(-2**63, 2**63 - 1, npr, dtype=np.int64))
    bk = Col(LONG, bkv, rng.random(nb) < 0.05 if nulls else None)
    pk = Col(LONG, pkv, rng.random(npr) < 0.05 if nulls else None)
    res = oracle.hash_join(bk, [Col(INT, np.arange(nb, dtype=np.int32))], pk, 1000, 4096)
    assert res.is_array == dense
    gp, gb = pairs(res)
    wp, wb = brute.inner_join(bk, pk)
    assert np.array_equal(gp, wp) and np.array_equal(gb, wb)
    assert np.array_equal(res.payloads[0].values[np.lexsort((res.build_rows, res.probe_rows))], wb.astype(np.int32))
