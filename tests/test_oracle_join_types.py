"""Pins the CPU oracle's join types other than inner (left / right / full outer, left semi, left anti, existence,
and the outer pass over unvisited build rows): the reference's own fixtures (join_test.cpp, transcribed in
tests/golden/operator_fixtures.json) and a dictionary-based brute force of each type's definition on seeded inputs
with duplicate keys and NULLs on both sides."""
import numpy as np
import pytest

import oracle
from oracle import LONG, Col
from tests.helpers import load_golden

FIX = load_golden("operator_fixtures.json")["join_types"]
CASES = [k for k in FIX if not k.startswith("_")]


def _sort_key(row):
    return tuple((0, "") if v is None else (1, v) if isinstance(v, int) else (2, v) for v in row)


def key_col(values):
    keys = np.array([0 if v is None else v for v in values], np.int64)
    nulls = np.array([v is None for v in values])
    return Col(LONG, keys, nulls if nulls.any() else None)


def run_fixture(f):
    """Oracle output of one fixture in the fixture's row representation."""
    kc = f["key_col"]
    bcols, pcols = f["build"], f["probe"]
    nb, npr = len(bcols[0]), len(pcols[0])
    row_id = Col(LONG, np.arange(nb, dtype=np.int64))  # payload = build row, to read its columns back
    res = oracle.hash_join_typed(key_col(bcols[kc]), [row_id], key_col(pcols[kc]), nb, npr, f["join_type"],
                                 f["build_side"])
    if res.exists is not None:
        return res, None, None
    if f["join_type"] in (oracle.JOIN_LEFT_SEMI, oracle.JOIN_LEFT_ANTI):
        rows = [[pcols[0][p], pcols[1][p]] for p in res.probe_rows.tolist()]
        return res, rows, None
    rows = []
    for p, b in zip(res.probe_rows.tolist(), res.build_rows.tolist()):
        rows.append([pcols[0][p], pcols[1][p]] + ([None, None] if b < 0 else [bcols[0][b], bcols[1][b]]))
    outer = [[None, None, bcols[0][b], bcols[1][b]] for b in res.outer_build_rows.tolist()]
    return res, rows, outer


@pytest.mark.parametrize("case", CASES)
def test_join_type_fixture(case):
    f = FIX[case]
    res, rows, outer = run_fixture(f)
    if "expected_exists" in f:
        assert res.exists.astype(bool).tolist() == f["expected_exists"]
        return
    assert sorted(rows, key=_sort_key) == sorted(f["expected_rows"], key=_sort_key)
    if outer is not None:
        assert sorted(outer, key=_sort_key) == sorted(f["expected_outer"], key=_sort_key)


def brute(bk, bnull, pk, pnull, join_type, build_side):
    """Each type from its definition (one equality key; NULL equals nothing)."""
    index = {}
    for b, (k, n) in enumerate(zip(bk.tolist(), bnull.tolist())):
        if not n:
            index.setdefault(k, []).append(b)
    matches = [[] if n else index.get(k, []) for k, n in zip(pk.tolist(), pnull.tolist())]
    same = (join_type, build_side) in ((oracle.JOIN_LEFT, oracle.BUILD_LEFT), (oracle.JOIN_RIGHT, oracle.BUILD_RIGHT))
    keep_probe = join_type == oracle.JOIN_FULL or (join_type in (oracle.JOIN_LEFT, oracle.JOIN_RIGHT) and not same)
    keep_build = join_type == oracle.JOIN_FULL or same
    if join_type == oracle.JOIN_EXISTENCE:
        return [bool(m) for m in matches], None
    if join_type == oracle.JOIN_LEFT_SEMI:
        return sorted(p for p, m in enumerate(matches) if m), None
    if join_type == oracle.JOIN_LEFT_ANTI:
        return sorted(p for p, m in enumerate(matches) if not m), None
    pairs = [(p, b) for p, m in enumerate(matches) for b in m]
    if keep_probe:
        pairs += [(p, -1) for p, m in enumerate(matches) if not m]
    outer = None
    if keep_build:
        seen = {b for m in matches for b in m}
        outer = [b for b in range(len(bk)) if b not in seen]
    return sorted(pairs), outer


ALL_TYPES = [(oracle.JOIN_INNER, 0), (oracle.JOIN_LEFT, 0), (oracle.JOIN_LEFT, oracle.BUILD_LEFT),
             (oracle.JOIN_LEFT, oracle.BUILD_RIGHT), (oracle.JOIN_RIGHT, 0), (oracle.JOIN_RIGHT, oracle.BUILD_RIGHT),
             (oracle.JOIN_FULL, 0), (oracle.JOIN_LEFT_SEMI, 0), (oracle.JOIN_LEFT_ANTI, 0), (oracle.JOIN_EXISTENCE, 0)]


@pytest.mark.parametrize("jt,bs", ALL_TYPES)
@pytest.mark.parametrize("dense", [False, True])
def test_join_types_match_brute_force(jt, bs, dense):
    rng = np.random.default_rng(100 + 10 * jt + bs + dense)
    nb, npr = 3000, 5000
    hi = 4000 if dense else 2**62
    keys = rng.integers(0, hi, 2500, dtype=np.int64)
    bk = keys[rng.integers(0, len(keys), nb)]
    pk = np.where(rng.random(npr) < 0.6, bk[rng.integers(0, nb, npr)], rng.integers(0, hi, npr, dtype=np.int64))
    bnull, pnull = rng.random(nb) < 0.05, rng.random(npr) < 0.05
    res = oracle.hash_join_typed(Col(LONG, bk, bnull), [Col(LONG, np.arange(nb, dtype=np.int64))],
                                 Col(LONG, pk, pnull), 700, 900, jt, bs)
    want, want_outer = brute(bk, bnull, pk, pnull, jt, bs)
    if jt == oracle.JOIN_EXISTENCE:
        assert res.exists.astype(bool).tolist() == want
        return
    if jt in (oracle.JOIN_LEFT_SEMI, oracle.JOIN_LEFT_ANTI):
        assert sorted(res.probe_rows.tolist()) == want
        return
    assert sorted(zip(res.probe_rows.tolist(), res.build_rows.tolist())) == want
    # gathered payload = build row id, NULL exactly where the build row is missing
    assert np.array_equal(res.payloads[0].nulls, res.build_rows < 0)
    assert np.array_equal(res.payloads[0].values[res.build_rows >= 0], res.build_rows[res.build_rows >= 0])
    if want_outer is None:
        assert len(res.outer_build_rows) == 0
    else:
        assert res.outer_build_rows.tolist() == want_outer


@pytest.mark.parametrize("jt,bs", ALL_TYPES)
def test_empty_build_side(jt, bs):
    rng = np.random.default_rng(9)
    pk, pnull = rng.integers(0, 50, 300, dtype=np.int64), rng.random(300) < 0.1
    empty = np.zeros(0, np.int64)
    res = oracle.hash_join_typed(Col(LONG, empty), [Col(LONG, empty)], Col(LONG, pk, pnull), 64, 100, jt, bs)
    want, want_outer = brute(empty, np.zeros(0, bool), pk, pnull, jt, bs)
    if jt == oracle.JOIN_EXISTENCE:
        assert res.exists.astype(bool).tolist() == want
    elif jt in (oracle.JOIN_LEFT_SEMI, oracle.JOIN_LEFT_ANTI):
        assert sorted(res.probe_rows.tolist()) == want
    else:
        assert sorted(zip(res.probe_rows.tolist(), res.build_rows.tolist())) == want
        assert res.outer_build_rows.tolist() == (want_outer or [])
