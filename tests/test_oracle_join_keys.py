"""Pins the CPU oracle's joins on several key columns and on a DECIMAL128 key: the reference's own fixtures
(join_test.cpp, transcribed in tests/golden/operator_fixtures.json), the sign-extension collisions that
ComputeMultiColKey (join_hash_table_variants.cpp:105-146) makes, worked out by hand, and the brute-force restatement
(oracle.brute.join_keys + typed_join) on seeded inputs of every key variant with NULLs in each key column."""
import numpy as np
import pytest

import oracle
from oracle import (BOOLEAN, BYTE, DATE32, DECIMAL128, DECIMAL64, DOUBLE, INT, LONG, SHORT, TIMESTAMP, Col, brute)
from tests.helpers import load_golden

FIX = load_golden("operator_fixtures.json")["join_multi_key"]
CASES = [k for k in FIX if not k.startswith("_")]


def fixture_col(type_id, values):
    nulls = np.array([v is None for v in values])
    if type_id == DECIMAL128:
        vals = np.array([[0, 0] if v is None else v for v in values], np.uint64).reshape(-1, 2)
    else:
        vals = np.array([0 if v is None else v for v in values], oracle.np_dtype(type_id))
    return Col(type_id, vals, nulls if nulls.any() else None)


def _norm(v):
    return tuple(v) if isinstance(v, list) else v


def _sort_key(row):
    return tuple((0, ()) if v is None else (1, v if isinstance(v, tuple) else (v,)) for v in row)


def run_fixture(f):
    types, kc = f["types"], f["key_cols"]
    bcols = [fixture_col(t, v) for t, v in zip(types, f["build"])]
    pcols = [fixture_col(t, v) for t, v in zip(types, f["probe"])]
    nb, npr = len(f["build"][0]), len(f["probe"][0])
    bkey = [bcols[c] for c in kc]
    pkey = [pcols[c] for c in kc]
    if len(kc) == 1:
        bkey, pkey = bkey[0], pkey[0]
    row_id = Col(LONG, np.arange(nb, dtype=np.int64))
    res = oracle.hash_join_typed(bkey, [row_id], pkey, nb, npr, f["join_type"], f["build_side"])
    b_rows = [[_norm(f["build"][c][r]) for c in range(len(types))] for r in range(nb)]
    p_rows = [[_norm(f["probe"][c][r]) for c in range(len(types))] for r in range(npr)]
    none = [None] * len(types)
    rows = [tuple(p_rows[p] + (none if b < 0 else b_rows[b]))
            for p, b in zip(res.probe_rows.tolist(), res.build_rows.tolist())]
    outer = [tuple(none + b_rows[b]) for b in res.outer_build_rows.tolist()]
    return rows, outer


@pytest.mark.parametrize("case", CASES)
def test_multi_key_fixture(case):
    f = FIX[case]
    rows, outer = run_fixture(f)
    want = [tuple(_norm(v) for v in r) for r in f["expected_rows"]]
    assert sorted(rows, key=_sort_key) == sorted(want, key=_sort_key)
    if "expected_outer" in f:
        want_outer = [tuple(_norm(v) for v in r) for r in f["expected_outer"]]
        assert sorted(outer, key=_sort_key) == sorted(want_outer, key=_sort_key)


def _pairs(bkey, pkey, jt=oracle.JOIN_INNER):
    nb = len(bkey[0]) if isinstance(bkey, list) else len(bkey)
    res = oracle.hash_join_typed(bkey, [Col(LONG, np.arange(nb, dtype=np.int64))], pkey, 64, 64, jt)
    return sorted(zip(res.probe_rows.tolist(), res.build_rows.tolist()))


def test_sign_extension_collisions():
    """ComputeMultiColKey ORs each value sign-extended to the key's width into the shifted key: a negative later
    column sets every earlier bit, so (5, -1) and (7, -1) as (INT, SHORT) are the same int64 key (all ones) and
    join; as (SHORT, SHORT) -- an int32 key -- (5, -1) and (9, -1) collide the same way; (LONG, LONG) -- an int128
    key -- (3, -1) and (4, -1) too.  A non-negative later column keeps the earlier one: (5, 1) != (7, 1)."""
    def cols(t0, a, t1, b):
        return [Col(t0, np.array(a, oracle.np_dtype(t0))), Col(t1, np.array(b, oracle.np_dtype(t1)))]
    assert _pairs(cols(INT, [5, 5], SHORT, [-1, 1]), cols(INT, [7, 7], SHORT, [-1, 1])) == [(0, 0)]
    assert _pairs(cols(SHORT, [5], SHORT, [-1]), cols(SHORT, [9], SHORT, [-1])) == [(0, 0)]
    assert _pairs(cols(LONG, [3], LONG, [-1]), cols(LONG, [4], LONG, [-1])) == [(0, 0)]
    # the serialised variant compares the stored values: no such collision for (DATE32, SHORT)
    assert _pairs(cols(DATE32, [5], SHORT, [-1]), cols(DATE32, [7], SHORT, [-1])) == []
    # BYTE packs in 8 bits (GetTypeLength, hash_builder.cpp:96-119), sign-extended like the others: (SHORT, BYTE)
    # (3, -1) and (8, -1) are the same int32 key; (INT, INT, BYTE) -- 72 bits, an int128 key -- (1, 2, -1) and
    # (6, 7, -1) too; with a non-negative BYTE the SHORT survives
    assert _pairs(cols(SHORT, [3, 3], BYTE, [-1, 4]), cols(SHORT, [8, 8], BYTE, [-1, 4])) == [(0, 0)]
    three = [Col(INT, np.array([1], np.int32)), Col(INT, np.array([2], np.int32)), Col(BYTE, np.array([-1], np.int8))]
    other = [Col(INT, np.array([6], np.int32)), Col(INT, np.array([7], np.int32)), Col(BYTE, np.array([-1], np.int8))]
    assert _pairs(three, other) == [(0, 0)]
    # brute force agrees on each of them
    assert brute.join_keys(cols(INT, [5], SHORT, [-1])) == brute.join_keys(cols(INT, [7], SHORT, [-1]))
    assert brute.join_keys(cols(DATE32, [5], SHORT, [-1])) != brute.join_keys(cols(DATE32, [7], SHORT, [-1]))


KEY_SETS = {
    "short_short": [SHORT, SHORT],                 # int32 key
    "int_int": [INT, INT],                         # int64 key
    "short_int": [SHORT, INT],                     # int64 key, 48 bits
    "long_long": [LONG, LONG],                     # int128 key
    "short_int_long": [SHORT, INT, LONG],          # int128 key, 112 bits
    "dec64_int_short": [DECIMAL64, INT, SHORT],    # int128 key, 112 bits
    "long_long_int": [LONG, LONG, INT],            # 160 bits: serialised
    "date32_double": [DATE32, DOUBLE],             # serialised
    "int_timestamp": [INT, TIMESTAMP],             # serialised
    "boolean_long": [BOOLEAN, LONG],               # serialised
    "decimal128_int": [DECIMAL128, INT],           # serialised
    "decimal128": [DECIMAL128],                    # one Decimal128 key
    "four_shorts": [SHORT, SHORT, SHORT, SHORT],   # int64 key, 64 bits
    "byte": [BYTE],                                # one int8 key
    "short_byte": [SHORT, BYTE],                   # int32 key, 24 bits
    "byte_int": [BYTE, INT],                       # int64 key, 40 bits
    "long_int_short_byte": [LONG, INT, SHORT, BYTE],  # int128 key, 120 bits
    "byte_double": [BYTE, DOUBLE],                 # serialised
}


def random_key_cols(types, n, rng, pool=None, null_rate=0.04):
    """n rows of key columns drawn from a pool of distinct tuples (so that keys repeat), values small and negative
    too, a NULL in each column with probability null_rate."""
    cols = []
    m = pool if pool is not None else max(4, n // 3)
    pick = rng.integers(0, m, n)
    for c, t in enumerate(types):
        prng = np.random.default_rng(1000 + 17 * c + t)  # the same pool for both sides
        if t == DECIMAL128:
            lo = prng.integers(0, 2**63, m, dtype=np.int64).astype(np.uint64)
            hi = prng.integers(-3, 3, m, dtype=np.int64).astype(np.uint64)
            v = np.stack([lo, hi], 1)[pick]
        elif t == DOUBLE:
            pool_v = np.round(prng.standard_normal(m), 2)
            pool_v[:2] = [0.0, -0.0]  # stored values: the two zeros are different keys
            v = pool_v[pick]
        elif t == BOOLEAN:
            v = prng.integers(0, 2, m).astype(np.uint8)[pick]
        else:
            info = np.iinfo(oracle.np_dtype(t))
            lo, hi = max(info.min, -50), min(info.max, 50)
            v = prng.integers(lo, hi, m).astype(oracle.np_dtype(t))[pick]
        nulls = rng.random(n) < null_rate
        cols.append(Col(t, v, nulls if nulls.any() else None))
    return cols


def _arg(cols):
    return cols[0] if len(cols) == 1 else cols


@pytest.mark.parametrize("name", list(KEY_SETS))
def test_multi_key_inner_matches_brute_force(name):
    types = KEY_SETS[name]
    rng = np.random.default_rng(list(KEY_SETS).index(name))
    bcols = random_key_cols(types, 2000, rng, pool=700)
    pcols = random_key_cols(types, 3000, rng, pool=900)
    got = _pairs(_arg(bcols), _arg(pcols))
    want, _ = brute.typed_join(brute.join_keys(bcols), brute.join_keys(pcols), oracle.JOIN_INNER)
    assert got == want
    assert len(want) > 500  # the pools overlap, so most probe rows match


ALL_TYPES = [(oracle.JOIN_INNER, 0), (oracle.JOIN_LEFT, 0), (oracle.JOIN_LEFT, oracle.BUILD_LEFT),
             (oracle.JOIN_RIGHT, 0), (oracle.JOIN_RIGHT, oracle.BUILD_RIGHT), (oracle.JOIN_FULL, 0),
             (oracle.JOIN_LEFT_SEMI, 0), (oracle.JOIN_LEFT_ANTI, 0), (oracle.JOIN_EXISTENCE, 0)]


@pytest.mark.parametrize("jt,bs", ALL_TYPES)
@pytest.mark.parametrize("name", ["short_int", "long_long", "date32_double", "decimal128"])
def test_multi_key_join_types_match_brute_force(jt, bs, name):
    types = KEY_SETS[name]
    rng = np.random.default_rng(7 + 10 * jt + bs)
    bcols = random_key_cols(types, 900, rng, pool=400)
    pcols = random_key_cols(types, 1100, rng, pool=600)
    nb = len(bcols[0])
    res = oracle.hash_join_typed(_arg(bcols), [Col(LONG, np.arange(nb, dtype=np.int64))], _arg(pcols), 300, 250,
                                 jt, bs)
    want, want_outer = brute.typed_join(brute.join_keys(bcols), brute.join_keys(pcols), jt, bs)
    if jt == oracle.JOIN_EXISTENCE:
        assert res.exists.astype(bool).tolist() == want
        return
    if jt in (oracle.JOIN_LEFT_SEMI, oracle.JOIN_LEFT_ANTI):
        assert sorted(res.probe_rows.tolist()) == want
        return
    assert sorted(zip(res.probe_rows.tolist(), res.build_rows.tolist())) == want
    if want_outer is None:
        assert len(res.outer_build_rows) == 0
    else:
        assert res.outer_build_rows.tolist() == want_outer


def test_null_in_any_key_column_matches_nothing():
    a = Col(INT, np.array([1, 1, 1], np.int32), np.array([False, True, False]))
    b = Col(INT, np.array([2, 2, 2], np.int32), np.array([False, False, True]))
    # build rows 1 and 2 each have one NULL key column; probe the same three tuples
    assert _pairs([a, b], [a, b]) == [(0, 0)]
    res = oracle.hash_join_typed([a, b], [Col(LONG, np.arange(3, dtype=np.int64))], [a, b], 3, 3, oracle.JOIN_FULL)
    assert sorted(zip(res.probe_rows.tolist(), res.build_rows.tolist())) == [(0, 0), (1, -1), (2, -1)]
    assert res.outer_build_rows.tolist() == [1, 2]
