"""Brute-force, hash-free restatement of the path's results (test infrastructure only).

Group-by is done by sorting (numpy lexsort + reduceat) and the inner join by sort + searchsorted, so
neither shares any code or data structure with the Swiss-table oracle (`oracle/omni_ref_cpu.cpp`) or
with the GPU kernels.  It encodes only the result semantics the reference pins:
  - SUM(int) wraps mod 2^64 (core/test/operator/aggregation_test.cpp:4386-4422);
  - SUM(decimal) is NULL (or raises) once a row-order partial sum leaves the representable range
    (sum_spark_decimal_aggregator.h:15-66, base_operations.h:23-35);
  - SUM/AVG/MIN/MAX over only NULLs give NULL, COUNT gives 0; a NULL key forms its own group
    (sum_flatim_aggregator.h:243-257, max_aggregator.cpp:34-55, column_marshaller.h:225-234);
  - the inner join skips NULL keys on both sides (join_hash_table_variants.cpp:254-283,
    lookup_join.cpp:973-976);
  - several join key columns compare as the reference's key variant does (hash_builder.cpp:44-119): SHORT / INT /
    LONG / DECIMAL64 columns of at most 128 bits in total as one integer, key = (key * 2^bits OR (value mod 2^W))
    mod 2^W with W the key width 32, 64 or 128 (ComputeMultiColKey, join_hash_table_variants.cpp:105-146, where the
    value is sign-extended into the key), other combinations and a single DECIMAL128 key by the tuple of their
    stored values; a NULL in any key column matches nothing (:340-400).
"""
from __future__ import annotations

from typing import List, Optional, Sequence

import numpy as np

from . import (AVG, BUILD_LEFT, BUILD_RIGHT, BYTE, COUNT_ALL, COUNT_COLUMN, DECIMAL128, DECIMAL64, DOUBLE, INT,
               JOIN_EXISTENCE, JOIN_FULL, JOIN_LEFT, JOIN_LEFT_ANTI, JOIN_LEFT_SEMI, JOIN_RIGHT, LONG, MAX, MIN, SHORT,
               SUM, Col, np_dtype)

DEC128_MAX = 10**38 - 1
I64_MIN, I64_MAX = -(2**63), 2**63 - 1


def dec128_to_int(v: np.ndarray) -> List[int]:
    """(n, 2) uint64 [low, high] -> Python ints."""
    lo = v[:, 0].astype(object)
    hi = v[:, 1].astype(np.int64).astype(object)
    return [int(h) * (1 << 64) + int(l) for l, h in zip(lo, hi)]


def int_to_dec128(xs: Sequence[int]) -> np.ndarray:
    out = np.zeros((len(xs), 2), np.uint64)
    for i, x in enumerate(xs):
        x &= (1 << 128) - 1
        out[i, 0] = x & ((1 << 64) - 1)
        out[i, 1] = x >> 64
    return out


def _key_order(key_cols: Sequence[Col]):
    """Stable lexicographic order over (null flag, value) per key column; returns (order, starts)."""
    n = len(key_cols[0])
    sort_keys = []
    for c in reversed(key_cols):
        nulls = np.zeros(n, bool) if c.nulls is None else c.nulls
        if c.type_id == DECIMAL128:
            sort_keys += [np.where(nulls, 0, c.values[:, 0]), np.where(nulls, 0, c.values[:, 1].astype(np.int64))]
        elif c.type_id == DOUBLE:  # group by bit pattern (the reference serialises the 8 bytes)
            sort_keys.append(np.where(nulls, 0, c.values.view(np.int64)))
        else:
            sort_keys.append(np.where(nulls, 0, c.values.astype(np.int64)))
        sort_keys.append(nulls)
    order = np.lexsort(sort_keys) if sort_keys else np.arange(n)
    # boundaries: a new group starts where any key component changes
    change = np.zeros(n, bool)
    if n:
        change[0] = True
    for sk in sort_keys:
        s = sk[order]
        change[1:] |= s[1:] != s[:-1]
    starts = np.flatnonzero(change)
    return order, starts


def group_by(key_cols: Sequence[Col], aggs: Sequence[tuple], agg_cols: Sequence[Optional[Col]],
             overflow_as_null: bool = True) -> List[Col]:
    """aggs: tuples (func, in_type, out_type[, in_scale, out_scale]).  Returns output columns
    (keys, then aggregates), one row per group, in sorted key order."""
    n = len(key_cols[0])
    order, starts = _key_order(key_cols)
    g = len(starts)
    ends = np.append(starts[1:], n)
    outs: List[Col] = []
    for c in key_cols:
        vals = c.values[order][starts] if g else c.values[:0]
        nl = (np.zeros(n, bool) if c.nulls is None else c.nulls)[order][starts] if g else np.zeros(0, bool)
        vals = vals.copy()
        vals[nl] = 0
        outs.append(Col(c.type_id, vals, nl, c.scale))
    for spec, c in zip(aggs, agg_cols):
        func, tin, tout = spec[0], spec[1], spec[2]
        in_scale, out_scale = (spec[3], spec[4]) if len(spec) > 3 else (0, 0)
        if func == COUNT_ALL:
            outs.append(Col(tout, (ends - starts).astype(np.int64), np.zeros(g, bool)))
            continue
        valid = ~(np.zeros(n, bool) if c.nulls is None else c.nulls)[order]
        cnt = np.add.reduceat(valid.astype(np.int64), starts) if g else np.zeros(0, np.int64)
        if func == COUNT_COLUMN:
            outs.append(Col(tout, cnt, np.zeros(g, bool)))
            continue
        if tin == DECIMAL128:
            vals_py = dec128_to_int(c.values[order])
        elif tin in (DECIMAL64,) and func == SUM:
            vals_py = [int(x) for x in c.values[order]]
        else:
            vals_py = None
        if func == SUM and tin in (DECIMAL64, DECIMAL128):
            res, nul = [], []
            scale = 10 ** (out_scale - in_scale)
            for s, e in zip(starts, ends):
                acc, seen, ovf = 0, False, False
                for i in range(s, e):
                    if not valid[i] or ovf:
                        continue
                    acc += vals_py[i]
                    seen = True
                    lim_hi, lim_lo = (DEC128_MAX, -DEC128_MAX) if tout == DECIMAL128 else (I64_MAX, I64_MIN)
                    if acc > lim_hi or acc < lim_lo:
                        ovf = True
                if seen and not ovf:
                    # the rescale is checked against the DECIMAL128 limits whatever the output type; a DECIMAL64
                    # output then keeps the low 64 bits of the result (two's complement)
                    acc *= scale
                    if acc > DEC128_MAX or acc < -DEC128_MAX:
                        ovf = True
                    elif tout != DECIMAL128:
                        acc = (acc + 2**63) % 2**64 - 2**63
                if ovf and not overflow_as_null:
                    raise OverflowError("decimal overflow")
                res.append(0 if (ovf or not seen) else acc)
                nul.append(ovf or not seen)
            vals = int_to_dec128(res) if tout == DECIMAL128 else np.array(res, dtype=np.int64)
            outs.append(Col(tout, vals, np.array(nul, bool), out_scale))
            continue
        v = c.values[order]
        if func == SUM and tin == DOUBLE:
            sums = np.array([float(np.sum(v[s:e][valid[s:e]])) if cnt[k] else 0.0
                             for k, (s, e) in enumerate(zip(starts, ends))], dtype=np.float64)
            outs.append(Col(tout, sums, cnt == 0))
        elif func == SUM:  # integer: wraps mod 2^64
            vv = np.where(valid, v.astype(np.int64), 0)
            with np.errstate(over="ignore"):
                sums = np.add.reduceat(vv, starts) if g else np.zeros(0, np.int64)
            sums = np.where(cnt == 0, 0, sums)
            outs.append(Col(tout, sums.astype(np.int64), cnt == 0))
        elif func == AVG:
            vv = np.where(valid, v.astype(np.float64), 0.0)
            sums = np.array([float(np.sum(vv[s:e])) for s, e in zip(starts, ends)], dtype=np.float64)
            with np.errstate(invalid="ignore", divide="ignore"):
                avg = np.where(cnt == 0, 0.0, sums / np.maximum(cnt, 1))
            outs.append(Col(tout, avg, cnt == 0))
        elif func in (MIN, MAX):
            if tin == DECIMAL128:
                res = []
                for s, e in zip(starts, ends):
                    xs = [vals_py[i] for i in range(s, e) if valid[i]]
                    res.append((min(xs) if func == MIN else max(xs)) if xs else 0)
                outs.append(Col(tout, int_to_dec128(res), cnt == 0))
            else:
                dt = np_dtype(tin)
                info = np.finfo(dt) if np.issubdtype(dt, np.floating) else np.iinfo(dt)
                fill = info.max if func == MIN else info.min
                vv = np.where(valid, v, fill).astype(dt)
                red = (np.minimum if func == MIN else np.maximum).reduceat(vv, starts) if g else vv[:0]
                red = np.where(cnt == 0, 0, red).astype(np_dtype(tout))
                outs.append(Col(tout, red, cnt == 0))
        else:
            raise ValueError(f"unsupported aggregate {func}")
    return outs


def inner_join(build_key: Col, probe_key: Col):
    """All (probe_row, build_row) pairs with equal non-NULL keys, sorted by (probe_row, build_row)."""
    bk = build_key.values.astype(np.int64)
    bvalid = np.ones(len(bk), bool) if build_key.nulls is None else ~build_key.nulls
    bidx = np.flatnonzero(bvalid)
    order = bidx[np.argsort(bk[bidx], kind="stable")]
    sk = bk[order]
    pk = probe_key.values.astype(np.int64)
    pvalid = np.ones(len(pk), bool) if probe_key.nulls is None else ~probe_key.nulls
    lo = np.searchsorted(sk, pk, side="left")
    hi = np.searchsorted(sk, pk, side="right")
    cnt = np.where(pvalid, hi - lo, 0)
    probe_rows = np.repeat(np.arange(len(pk), dtype=np.int64), cnt)
    starts = np.repeat(lo, cnt)
    offs = np.arange(len(probe_rows)) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    build_rows = order[starts + offs].astype(np.int64)
    o = np.lexsort((build_rows, probe_rows))
    return probe_rows[o], build_rows[o]


_KEY_BITS = {BYTE: 8, SHORT: 16, INT: 32, LONG: 64, DECIMAL64: 64}


def join_keys(cols: Sequence[Col]) -> List[Optional[object]]:
    """The value a row's key columns compare by (None: a NULL key column), for one or several key columns."""
    n = len(cols[0])
    null = np.zeros(n, bool)
    for c in cols:
        if c.nulls is not None:
            null |= c.nulls
    bits = [_KEY_BITS.get(c.type_id) for c in cols]
    total = sum(b for b in bits if b is not None)
    width = None
    if len(cols) > 1 and None not in bits:
        width = 32 if total <= 32 else 64 if total <= 64 else 128 if total <= 128 else None
    if width is not None:
        mod = 1 << width
        vals = [c.values.astype(np.int64).tolist() for c in cols]
        out = []
        for r in range(n):
            if null[r]:
                out.append(None)
                continue
            k = 0
            for b, v in zip(bits, vals):
                k = ((k << b) | (v[r] % mod)) % mod  # the value sign-extended to `width` bits, then OR-ed in
            out.append(k)
        return out
    stored = []
    for c in cols:
        if c.type_id == DECIMAL128:
            stored.append([tuple(x) for x in c.values.astype(np.uint64).tolist()])
        elif c.type_id == DOUBLE:
            stored.append(c.values.view(np.int64).tolist())
        else:
            stored.append(c.values.tolist())
    return [None if null[r] else tuple(s[r] for s in stored) for r in range(n)]


def typed_join(build_keys: Sequence[Optional[object]], probe_keys: Sequence[Optional[object]], join_type: int,
               build_side: int = 0):
    """Each join type from its definition over hashable keys (None matches nothing).  Returns (pairs, outer):
    sorted (probe, build) pairs with build -1 for a kept unmatched probe row and the unmatched build rows when the
    type keeps them (else None); semi / anti give the sorted probe rows, existence one flag per probe row."""
    index = {}
    for b, k in enumerate(build_keys):
        if k is not None:
            index.setdefault(k, []).append(b)
    matches = [[] if k is None else index.get(k, []) for k in probe_keys]
    same = (join_type, build_side) in ((JOIN_LEFT, BUILD_LEFT), (JOIN_RIGHT, BUILD_RIGHT))
    keep_probe = join_type == JOIN_FULL or (join_type in (JOIN_LEFT, JOIN_RIGHT) and not same)
    keep_build = join_type == JOIN_FULL or same
    if join_type == JOIN_EXISTENCE:
        return [bool(m) for m in matches], None
    if join_type == JOIN_LEFT_SEMI:
        return sorted(p for p, m in enumerate(matches) if m), None
    if join_type == JOIN_LEFT_ANTI:
        return sorted(p for p, m in enumerate(matches) if not m), None
    pairs = [(p, b) for p, m in enumerate(matches) for b in m]
    if keep_probe:
        pairs += [(p, -1) for p, m in enumerate(matches) if not m]
    outer = None
    if keep_build:
        seen = {b for m in matches for b in m}
        outer = [b for b in range(len(build_keys)) if b not in seen]
    return sorted(pairs), outer
