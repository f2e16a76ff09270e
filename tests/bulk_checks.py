"""Checks of full-size GPU results on the host, for workloads too large for the oracle (hundreds of millions of
rows).  They are test infrastructure, pinned against the oracle and against planted mistakes by
tests/test_bulk_checks.py.

- sorted_groupby: an exact sort-based group-by (argsort of a packed key, segment reductions).
- check_sum_count_groupby: verifies a GPU SUM(value) / COUNT(*) group-by in O(rows) without sorting the input:
  the output keys are distinct, and for keyed pseudo-random hashes h, sum over rows of h(key) equals sum over
  groups of h(group key) * count, and sum over rows of h(key) * value equals sum over groups of
  h(group key) * sum (mod 2^64).  A wrong count or sum of any group, a missing group or a group that does not
  exist changes either side by a pseudo-random multiple, so a mistake passes with probability about 2^-64 per
  seed.
- check_inner_join: an inner join's matches per probe row, exactly (sorted probe and build keys), and the payloads
  of each probe row by a keyed checksum of (row, payload) pairs; duplicate build keys allowed.
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

U64 = np.uint64
C1 = 0xBF58476D1CE4E5B9
C2 = 0x94D049BB133111EB
GOLDEN = 0x9E3779B97F4A7C15


def splitmix64(x) -> np.ndarray:
    z = np.asarray(x).astype(U64) + U64(GOLDEN)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> U64(30))) * U64(C1)
        z = (z ^ (z >> U64(27))) * U64(C2)
    return z ^ (z >> U64(31))


def _unxorshift(y: np.ndarray, s: int) -> np.ndarray:
    z = y.copy()
    for _ in range(64 // s + 1):
        z = y ^ (z >> U64(s))
    return z


def unsplitmix64(y) -> np.ndarray:
    """Inverse of splitmix64 (each step is a bijection: xorshifts and odd multipliers mod 2^64)."""
    z = np.asarray(y).astype(U64)
    with np.errstate(over="ignore"):
        z = _unxorshift(z, 31) * U64(pow(C2, -1, 1 << 64))
        z = _unxorshift(z, 27) * U64(pow(C1, -1, 1 << 64))
        z = _unxorshift(z, 30)
        return z - U64(GOLDEN)


def key_hash(cols: Sequence[np.ndarray], seed: int) -> np.ndarray:
    """Keyed pseudo-random 64-bit hash of composite keys (one array per key column)."""
    h = np.full(len(cols[0]), U64(seed), U64)
    with np.errstate(over="ignore"):
        for c in cols:
            h = splitmix64(h ^ np.asarray(c).astype(np.int64).view(U64))
    return h


def _wrap_sum(x: np.ndarray) -> int:
    return int(np.sum(x, dtype=U64))


def sorted_groupby(packed: np.ndarray, columns: dict):
    """Exact group-by of rows keyed by the int64 array `packed`.  columns: name -> (values, reduction) with
    reduction in {"sum", "max", "min"}; sums of integers are exact int64, sums of doubles run in the sorted order
    (any order within a group, as the reference's own order is not the GPU's either).
    Returns (sorted distinct keys, counts, {name: per-group result})."""
    order = np.argsort(packed)
    sp = packed[order]
    starts = np.flatnonzero(np.r_[True, sp[1:] != sp[:-1]]) if len(sp) else np.zeros(0, np.int64)
    counts = np.diff(np.r_[starts, len(sp)])
    out = {}
    for name, (vals, red) in columns.items():
        v = vals[order]
        ufunc = {"sum": np.add, "max": np.maximum, "min": np.minimum}[red]
        out[name] = ufunc.reduceat(v, starts) if len(starts) else v[:0]
        del v
    return sp[starts], counts, out


def check_sum_count_groupby(in_keys: Sequence[np.ndarray], values: np.ndarray, out_keys: Sequence[np.ndarray],
                            out_sum: np.ndarray, out_count: np.ndarray, seeds=(0x1234567,)) -> None:
    """Asserts that (out_keys, out_sum, out_count) is the SUM(values), COUNT(*) GROUP BY in_keys (no NULLs)."""
    rows = len(values)
    g = len(out_count)
    assert all(len(k) == g for k in out_keys) and len(out_sum) == g
    assert g <= rows and (g > 0) == (rows > 0)
    assert int(out_count.min(initial=1)) >= 1, "a group with count 0"
    assert int(out_count.sum()) == rows, "counts do not add up to the rows"
    # distinct output keys: sort a 64-bit mix of the key, and compare the full keys wherever two mixes are equal
    mix = key_hash(out_keys, 0x5DEECE66D)
    order = np.argsort(mix)
    sm = mix[order]
    dup = np.flatnonzero(sm[1:] == sm[:-1])
    for i in dup:
        a, b = order[i], order[i + 1]
        assert any(k[a] != k[b] for k in out_keys), f"group key emitted twice at output rows {a} and {b}"
    v = np.asarray(values).astype(np.int64).view(U64)
    s = np.asarray(out_sum).astype(np.int64).view(U64)
    c = np.asarray(out_count).astype(np.int64).view(U64)
    with np.errstate(over="ignore"):
        for seed in seeds:
            hi = key_hash(in_keys, seed)
            ho = key_hash(out_keys, seed)
            assert _wrap_sum(hi) == _wrap_sum(ho * c), f"COUNT checksum differs (seed {seed:#x})"
            assert _wrap_sum(hi * v) == _wrap_sum(ho * s), f"SUM checksum differs (seed {seed:#x})"


def check_inner_join(probe_keys: np.ndarray, build_keys: np.ndarray, build_values: np.ndarray,
                     out_rows: np.ndarray, out_pay: np.ndarray, seed: int = 0x2545F491) -> int:
    """Asserts that the pairs (out_rows[i], out_pay[i]) are the inner join of probe_keys with (build_keys,
    build_values), duplicate build keys included: each probe row appears exactly as often as its key occurs among
    the build keys (exact, through sorted build and probe keys), and the multiset of payloads of every probe row is
    its matching build rows' values (keyed checksum of (row, payload) pairs).  Returns the match count."""
    n = len(probe_keys)
    out_rows = np.asarray(out_rows, np.int64)
    assert len(out_pay) == len(out_rows)
    assert len(out_rows) == 0 or (int(out_rows.min()) >= 0 and int(out_rows.max()) < n), "row id out of range"
    bo = np.argsort(build_keys)
    sb = build_keys[bo]
    po = np.argsort(probe_keys)
    sp = probe_keys[po]
    lo = np.searchsorted(sb, sp, "left")
    hi = np.searchsorted(sb, sp, "right")
    del sp
    mult = np.empty(n, np.int64)
    mult[po] = hi - lo
    got = np.bincount(out_rows, minlength=n)
    assert np.array_equal(got, mult), f"matches per probe row differ ({len(out_rows)} matches, want {int(mult.sum())})"
    with np.errstate(over="ignore"):
        hv = key_hash([build_values[bo]], seed)
        cs = np.r_[np.zeros(1, U64), np.cumsum(hv, dtype=U64)]
        per_row = np.empty(n, U64)
        per_row[po] = cs[hi] - cs[lo]
        hr = key_hash([np.arange(n, dtype=np.int64)], seed ^ 0xA5A5)
        want = _wrap_sum(hr * per_row)
        have = _wrap_sum(hr[out_rows] * key_hash([out_pay], seed))
    assert have == want, "payload checksum differs"
    return len(out_rows)


def zero_nulls(mask: Optional[np.ndarray]) -> bool:
    return mask is None or not bool(np.any(mask))
