"""Shared test utilities: fixture loading and row-set comparison (VecBatchMatchIgnoreOrder semantics)."""
from __future__ import annotations

import json
import math
import os
from typing import List, Optional, Sequence, Tuple

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

DEC128 = 7
DOUBLE = 3


def load_golden(name: str):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def to_py(type_id: int, values: np.ndarray, nulls: Optional[np.ndarray]) -> List:
    n = values.shape[0]
    nulls = np.zeros(n, bool) if nulls is None else nulls
    if type_id == DEC128:
        lo = values[:, 0].astype(object)
        hi = values[:, 1].astype(np.int64).astype(object)
        vals = [int(h) * (1 << 64) + int(l) for l, h in zip(lo, hi)]
    elif type_id == DOUBLE:
        vals = [float(x) for x in values]
    else:
        vals = [int(x) for x in values]
    return [None if nl else v for v, nl in zip(vals, nulls)]


def rows_of(cols: Sequence[Tuple[int, np.ndarray, Optional[np.ndarray]]]) -> List[tuple]:
    """cols: (type_id, values, null mask) -> list of row tuples."""
    pys = [to_py(t, v, nl) for t, v, nl in cols]
    return list(zip(*pys)) if pys else []


def _sort_key(row, exact_idx):
    return tuple((0, 0) if row[i] is None else (1, row[i]) for i in exact_idx)


def assert_row_sets_equal(got: List[tuple], want: List[tuple], float_cols: Sequence[int] = (), rtol: float = 1e-12):
    """Exact equality of the row multisets; columns in float_cols are matched within rtol after
    aligning rows on the exact columns."""
    assert len(got) == len(want), f"row count {len(got)} != {len(want)}"
    if not got:
        return
    exact = [i for i in range(len(want[0])) if i not in set(float_cols)]
    g = sorted(got, key=lambda r: _sort_key(r, exact))
    w = sorted(want, key=lambda r: _sort_key(r, exact))
    for a, b in zip(g, w):
        for i in exact:
            assert a[i] == b[i], f"row mismatch at column {i}: got {a} want {b}"
        for i in float_cols:
            if a[i] is None or b[i] is None:
                assert a[i] is None and b[i] is None, f"null mismatch at column {i}: got {a} want {b}"
            else:
                assert math.isclose(a[i], b[i], rel_tol=rtol, abs_tol=0.0) or a[i] == b[i], \
                    f"column {i}: {a[i]!r} vs {b[i]!r} (rtol {rtol})"
