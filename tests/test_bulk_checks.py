"""The full-size checkers (tests/bulk_checks.py) pinned on the CPU: each accepts the oracle's result and rejects
the same result with a planted mistake, so a full-size GPU test that passes them has the oracle's result."""
import numpy as np
import pytest

import oracle
from oracle import AVG, COUNT_ALL, DECIMAL64, DECIMAL128, DOUBLE, INT, LONG, MAX, SUM, Col
from omnioperator_amd import workloads
from tests import bulk_checks as bc


def test_unsplitmix64_inverts_the_workload_bijection():
    rng = np.random.default_rng(5)
    x = np.r_[np.array([0, 1, 2**63 - 1, 2**64 - 1], np.uint64), rng.integers(0, 2**63, 10000).astype(np.uint64)]
    y = workloads.splitmix64(x)
    assert np.array_equal(bc.splitmix64(x), y)
    assert np.array_equal(bc.unsplitmix64(y), x)


def _oracle_sum_count(keys, values):
    spec = oracle.make_agg_spec([k.type_id for k in keys], [(SUM, LONG, LONG), (COUNT_ALL, LONG, LONG)])
    out = oracle.hash_agg(spec, keys, [Col(LONG, values), None], 4096)
    return [c.values for c in out[:len(keys)]], out[len(keys)].values, out[len(keys) + 1].values


@pytest.mark.parametrize("composite", [False, True])
def test_sum_count_checksum_accepts_oracle_and_rejects_planted_mistakes(composite):
    rng = np.random.default_rng(11 + composite)
    n = 50_000
    k64 = rng.integers(0, 3000, n).astype(np.int64) * 7919 - 10**6
    keys = [Col(LONG, k64)]
    if composite:
        keys.append(Col(INT, rng.integers(0, 3, n).astype(np.int32)))
    v = rng.integers(-10**6, 10**6, n).astype(np.int64)
    ok, s, c = _oracle_sum_count(keys, v)
    kin = [k.values for k in keys]
    bc.check_sum_count_groupby(kin, v, ok, s, c)

    def rejects(ok2, s2, c2):
        with pytest.raises(AssertionError):
            bc.check_sum_count_groupby(kin, v, ok2, s2, c2)

    # one count moved between two groups (counts still add up)
    c2 = c.copy()
    c2[0] += 1
    c2[1] -= 1
    rejects(ok, s, c2)
    # one sum off by one
    s2 = s.copy()
    s2[5] += 1
    rejects(ok, s2, c)
    # the sums of two groups swapped
    s2 = s.copy()
    s2[[3, 4]] = s2[[4, 3]]
    assert not np.array_equal(s2, s)
    rejects(ok, s2, c)
    # a group dropped and its rows counted into another
    rejects([k[1:] for k in ok], np.r_[s[1:2] + s[0], s[2:]], np.r_[c[1:2] + c[0], c[2:]])
    # a group split into two output rows with the same key (both checksums still hold)
    i = int(np.argmax(c))
    c2, s2 = np.r_[c, 1], np.r_[s, 5]
    c2[i] -= 1
    s2[i] -= 5
    rejects([np.r_[k, k[i:i + 1]] for k in ok], s2, c2)
    # one key changed
    ok2 = [k.copy() for k in ok]
    ok2[0][7] += 1
    rejects(ok2, s, c)


@pytest.mark.parametrize("dup_build", [False, True])
def test_inner_join_check_accepts_oracle_and_rejects_planted_mistakes(dup_build):
    rng = np.random.default_rng(3 + dup_build)
    bk = workloads.splitmix64(np.arange(5000, dtype=np.uint64)).view(np.int64)
    if dup_build:  # every third build key repeated once or twice
        bk = np.r_[bk, bk[::3], bk[::9]]
    bv = rng.integers(-10**6, 10**6, len(bk)).astype(np.int64)
    pk = np.where(rng.random(40_000) < 0.8, bk[rng.integers(0, len(bk), 40_000)], rng.integers(0, 2**62, 40_000))
    r = oracle.hash_join(Col(LONG, bk), [Col(LONG, bv)], Col(LONG, pk), 4096, 4096)
    rows, pay = r.probe_rows, r.payloads[0].values
    assert bc.check_inner_join(pk, bk, bv, rows, pay) == len(rows)
    assert bc.check_inner_join(pk, bk, bv, rows[::-1], pay[::-1]) == len(rows)  # order does not matter
    dup_row = int(np.flatnonzero(np.bincount(rows) > 1)[0]) if dup_build else None
    bad = [(rows[1:], pay[1:]),                          # a match dropped
           (np.r_[rows, rows[:1]], np.r_[pay, pay[:1]]),  # a match emitted twice
           (rows, np.r_[pay[:1] + 1, pay[1:]])]          # a wrong payload
    if dup_build:  # two payloads of one probe row made equal: counts hold, the payload multiset does not
        i, j = np.flatnonzero(rows == dup_row)[:2]
        if pay[i] != pay[j]:
            p2 = pay.copy()
            p2[j] = p2[i]
            bad.append((rows, p2))
    for bad_rows, bad_pay in bad:
        with pytest.raises(AssertionError):
            bc.check_inner_join(pk, bk, bv, bad_rows, bad_pay)


def test_sorted_groupby_equals_oracle_on_skewed_rows():
    """The sort-based reference of the full-size skewed group-by test equals the oracle on 300,000 rows of the
    same distribution (decimal64 SUM, double SUM/AVG, COUNT, MAX(int32))."""
    w = workloads.agg_skewed(rows=300_000, k1_distinct=100, k2_ranks=50_000, seed=0xC3, keep_ranks=True)
    packed = (w.k1.astype(np.int64) << 26) | w.ranks
    ukeys, counts, red = bc.sorted_groupby(packed, {"dec": (w.v_dec, "sum"), "dbl": (w.v_dbl, "sum"),
                                                    "max": (w.v_i32, "max")})
    aggs = [(SUM, DECIMAL64, DECIMAL128, 2, 2), (SUM, DOUBLE, DOUBLE), (AVG, DOUBLE, DOUBLE), (COUNT_ALL, LONG, LONG),
            (MAX, INT, INT)]
    out = oracle.hash_agg(oracle.make_agg_spec([INT, LONG], aggs),
                          [Col(INT, w.k1), Col(LONG, w.k2)],
                          [Col(DECIMAL64, w.v_dec, scale=2), Col(DOUBLE, w.v_dbl), Col(DOUBLE, w.v_dbl), None,
                           Col(INT, w.v_i32)], 65536)
    opacked = (out[0].values.astype(np.int64) << 26) | bc.unsplitmix64(out[1].values).view(np.int64)
    o = np.argsort(opacked)
    assert np.array_equal(opacked[o], ukeys)
    dec = out[2].values[o]
    assert np.all(dec[:, 1] == 0) and np.array_equal(dec[:, 0], red["dec"])
    assert np.allclose(out[3].values[o], red["dbl"], rtol=1e-12, atol=0)
    assert np.allclose(out[4].values[o], red["dbl"] / counts, rtol=1e-12, atol=0)
    assert np.array_equal(out[5].values[o], counts)
    assert np.array_equal(out[6].values[o], red["max"])
