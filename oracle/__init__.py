"""CPU oracle for the hash-aggregation / hash-join hot path (test infrastructure only).

`oracle/omni_ref_cpu.cpp` restates the reference operators (Swiss-table BaseHashMap, CRC32C and
xxHash64-style hashing, HashAggregationOperator, HashBuilderOperator + LookupJoinOperator) in C++;
this module loads it (`oracle/lib/liboracle.so`, built by `make -C oracle`) through ctypes and
offers numpy-level entry points.  `oracle.brute` holds an independent sort-based restatement used
to pin this one.

Only tests, `__graft_entry__.smoke()` and bench.py's CPU-baseline leg use this package.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "liboracle.so")

# DataTypeId values (core/src/type/data_type.h:40-65) and FunctionType values
# (core/src/operator/util/function_type.h:11-27).
INT, LONG, DOUBLE, BOOLEAN, SHORT, DECIMAL64, DECIMAL128, DATE32, TIMESTAMP = 1, 2, 3, 4, 5, 6, 7, 8, 12
BYTE = 18
SUM, COUNT_COLUMN, COUNT_ALL, AVG, MAX, MIN = 0, 1, 2, 3, 5, 6
MAX_KEYS, MAX_AGGS = 4, 16

NP_DTYPE = {
    INT: np.int32, DATE32: np.int32, LONG: np.int64, TIMESTAMP: np.int64, DECIMAL64: np.int64,
    DOUBLE: np.float64, SHORT: np.int16, BOOLEAN: np.uint8, BYTE: np.int8,
}


def np_dtype(type_id: int):
    """numpy storage type of a column; DECIMAL128 is stored as (n, 2) uint64 [low, high]."""
    return np.uint64 if type_id == DECIMAL128 else NP_DTYPE[type_id]


@dataclass
class Col:
    """A flat column: values + optional boolean NULL mask (True = NULL)."""
    type_id: int
    values: np.ndarray
    nulls: Optional[np.ndarray] = None
    scale: int = 0

    def __len__(self):
        return int(self.values.shape[0])

    def slice(self, a: int, b: int) -> "Col":
        return Col(self.type_id, self.values[a:b], None if self.nulls is None else self.nulls[a:b], self.scale)


def pack_nulls(mask: Optional[np.ndarray]) -> Optional[np.ndarray]:
    """Boolean mask -> OmniVec LSB-first bitmap (bit = 1 means NULL)."""
    if mask is None:
        return None
    return np.packbits(np.asarray(mask, dtype=bool), bitorder="little")


def unpack_nulls(bitmap: np.ndarray, n: int) -> np.ndarray:
    return np.unpackbits(bitmap, count=n, bitorder="little").astype(bool)


class AggSpec(C.Structure):
    """Layout of omni_gpu_agg_spec (include/omni_gpu.h)."""
    _fields_ = [
        ("num_keys", C.c_int32), ("key_types", C.c_int32 * MAX_KEYS),
        ("num_aggs", C.c_int32), ("agg_funcs", C.c_int32 * MAX_AGGS),
        ("agg_in_types", C.c_int32 * MAX_AGGS), ("agg_out_types", C.c_int32 * MAX_AGGS),
        ("agg_in_scales", C.c_int32 * MAX_AGGS), ("agg_out_scales", C.c_int32 * MAX_AGGS),
        ("overflow_as_null", C.c_int32), ("strategy", C.c_int32), ("capacity_hint", C.c_int64),
        ("agg_partial_out", C.c_int32 * MAX_AGGS), ("hash_table_rule", C.c_int32),
    ]


def make_agg_spec(key_types: Sequence[int], aggs: Sequence[tuple], overflow_as_null: bool = True,
                  strategy: int = 0, capacity_hint: int = 0) -> AggSpec:
    """aggs: tuples (func, in_type, out_type[, in_scale, out_scale[, partial_out]])."""
    s = AggSpec()
    s.num_keys = len(key_types)
    for i, t in enumerate(key_types):
        s.key_types[i] = t
    s.num_aggs = len(aggs)
    for i, a in enumerate(aggs):
        f, ti, to = a[0], a[1], a[2]
        s.agg_funcs[i], s.agg_in_types[i], s.agg_out_types[i] = f, ti, to
        if len(a) > 3:
            s.agg_in_scales[i], s.agg_out_scales[i] = a[3], a[4]
        if len(a) > 5:
            s.agg_partial_out[i] = 1 if a[5] else 0
    s.overflow_as_null = 1 if overflow_as_null else 0
    s.strategy = strategy
    s.capacity_hint = capacity_hint
    return s


class _OColumn(C.Structure):
    _fields_ = [("type_id", C.c_int32), ("scale", C.c_int32), ("values", C.c_void_p), ("nulls", C.c_void_p)]


class _OutColumn(C.Structure):
    _fields_ = [("type_id", C.c_int32), ("reserved", C.c_int32), ("values", C.c_void_p), ("nulls", C.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        u64, i64, i32, u32, vp = C.c_uint64, C.c_int64, C.c_int32, C.c_uint32, C.c_void_p
        sig = {
            "oracle_hash_uint64": (u64, [u64]),
            "oracle_crc32c_step_software": (u32, [u32, u64]),
            "oracle_crc_hash_int": (u64, [i64, i32]),
            "oracle_crc_hash_decimal128": (u64, [u64, u64]),
            "oracle_crc32c_extend": (u32, [u32, vp, C.c_size_t]),
            "oracle_join_hash_int": (u64, [i64]),
            "oracle_hash_value_i64": (i64, [i64]),
            "oracle_hash_value_double": (i64, [C.c_double]),
            "oracle_hash_value_decimal128": (i64, [i64, i64]),
            "oracle_xxhash64": (i64, [vp, i32]),
            "oracle_raw_hash_partition": (u32, [i64, u32]),
            "oracle_raw_hash_position": (u32, [i64, u64]),
            "oracle_serialize_row": (i32, [vp, i32, i64, vp, i32]),
            "oracle_map_create": (vp, [i32]),
            "oracle_map_emplace": (i32, [vp, i64, i32, u32]),
            "oracle_map_size": (i64, [vp]),
            "oracle_map_capacity": (i64, [vp]),
            "oracle_map_values_in_order": (i64, [vp, vp]),
            "oracle_map_destroy": (None, [vp]),
            "oracle_agg_create": (vp, [vp]),
            "oracle_agg_add_batch": (i32, [vp, vp, vp, i64]),
            "oracle_agg_num_groups": (i64, [vp]),
            "oracle_agg_get_output": (i32, [vp, vp]),
            "oracle_agg_destroy": (None, [vp]),
            "oracle_agg_stream": (i64, [vp, vp, vp, vp, i32, i32, vp]),
            "oracle_join_create": (vp, [i32]),
            "oracle_join_add_batch": (i32, [vp, vp, vp, i32, i64]),
            "oracle_join_create_multi": (vp, [vp, i32]),
            "oracle_join_translate": (None, [vp, vp, i32, i64, i32, vp, vp]),
            "oracle_join_finish": (i32, [vp]),
            "oracle_join_is_array": (i32, [vp]),
            "oracle_join_probe": (i64, [vp, vp, i64, i64, vp, vp, i64]),
            "oracle_join_gather": (None, [vp, i32, vp, i64, vp, vp]),
            "oracle_join_probe_typed": (i64, [vp, vp, i64, i64, i32, i32, vp, vp, i64]),
            "oracle_join_exists": (None, [vp, vp, i64, vp]),
            "oracle_join_outer": (i64, [vp, vp, i64]),
            "oracle_join_destroy": (None, [vp]),
            "oracle_mm3_hash_int": (C.c_uint32, [C.c_uint32, C.c_uint32]),
            "oracle_mm3_hash_long": (C.c_uint32, [C.c_uint64, C.c_uint32]),
            "oracle_partition_ids": (None, [vp, i32, i64, i32, vp]),
            "oracle_last_error": (C.c_char_p, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _lib = L
    return _lib


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


def _ocols(cols: Sequence[Optional[Col]], n: int):
    arr = (_OColumn * n)()
    keep = []
    for i, c in enumerate(cols):
        if c is None:
            continue
        v = np.ascontiguousarray(c.values)
        nb = pack_nulls(c.nulls)
        keep += [v, nb]
        arr[i].type_id, arr[i].scale, arr[i].values, arr[i].nulls = c.type_id, c.scale, _ptr(v), _ptr(nb)
    return arr, keep


# ---- hashes -------------------------------------------------------------------------------------------
def hash_uint64(x: int) -> int:
    return lib().oracle_hash_uint64(x & (2**64 - 1))


def crc32c_step_software(crc: int, x: int) -> int:
    return lib().oracle_crc32c_step_software(crc, x & (2**64 - 1))


def crc_hash_int(v: int, type_id: int = LONG) -> int:
    return lib().oracle_crc_hash_int(v, type_id)


def crc_hash_decimal128(lo: int, hi: int) -> int:
    return lib().oracle_crc_hash_decimal128(lo & (2**64 - 1), hi & (2**64 - 1))


def crc32c_extend(crc: int, data: bytes) -> int:
    b = np.frombuffer(data, dtype=np.uint8).copy()
    return lib().oracle_crc32c_extend(crc, _ptr(b), len(data))


def join_hash_int(v: int) -> int:
    return lib().oracle_join_hash_int(v)


def hash_value_i64(v: int) -> int:
    return lib().oracle_hash_value_i64(v)


def hash_value_double(v: float) -> int:
    return lib().oracle_hash_value_double(v)


def hash_value_decimal128(low: int, high: int) -> int:
    return lib().oracle_hash_value_decimal128(low, high)


def xxhash64(data: bytes) -> int:
    b = np.frombuffer(data, dtype=np.uint8).copy() if data else np.zeros(1, np.uint8)
    return lib().oracle_xxhash64(_ptr(b), len(data))


def raw_hash_partition(h: int, mask: int) -> int:
    return lib().oracle_raw_hash_partition(h, mask)


def raw_hash_position(h: int, mask: int) -> int:
    return lib().oracle_raw_hash_position(h, mask)


def serialize_row(cols: Sequence[Col], row: int) -> bytes:
    arr, keep = _ocols(cols, len(cols))
    out = np.zeros(256, np.uint8)
    n = lib().oracle_serialize_row(arr, len(cols), row, _ptr(out), 256)
    return bytes(out[:n])


class SwissMap:
    """BaseHashMap<int64, uint32> with the aggregation hasher (CRC32C), for hash-map behaviour tests."""

    def __init__(self, degree: int = 15):
        self.h = lib().oracle_map_create(degree)

    def emplace(self, key: Optional[int], value: int) -> bool:
        return bool(lib().oracle_map_emplace(self.h, 0 if key is None else key, 1 if key is None else 0, value))

    def size(self) -> int:
        return lib().oracle_map_size(self.h)

    def capacity(self) -> int:
        return lib().oracle_map_capacity(self.h)

    def values_in_order(self) -> np.ndarray:
        out = np.zeros(self.size(), np.uint32)
        n = lib().oracle_map_values_in_order(self.h, _ptr(out))
        return out[:n]

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_map_destroy(self.h)
            self.h = None


# ---- aggregation ------------------------------------------------------------------------------------------
def type_width(type_id: int) -> int:
    return {SHORT: 2, INT: 4, DATE32: 4, LONG: 8, TIMESTAMP: 8, DECIMAL64: 8, DOUBLE: 8, DECIMAL128: 16,
            BOOLEAN: 1, BYTE: 1}[type_id]


def alloc_out(type_id: int, n: int) -> Col:
    shape = (n, 2) if type_id == DECIMAL128 else (n,)
    return Col(type_id, np.zeros(shape, dtype=np_dtype(type_id)), np.zeros(n, dtype=bool))


def hash_agg(spec: AggSpec, key_cols: Sequence[Col], agg_cols: Sequence[Optional[Col]], batch_rows: int) -> List[Col]:
    """Runs the restated HashAggregationOperator over the input cut into VectorBatches of batch_rows
    rows; returns the output columns (keys, then one per aggregate) in the reference's slot order."""
    L = lib()
    h = L.oracle_agg_create(C.byref(spec))
    try:
        n = len(key_cols[0])
        for a in range(0, max(n, 1) if n else 0, batch_rows):
            b = min(n, a + batch_rows)
            karr, k1 = _ocols([c.slice(a, b) for c in key_cols], MAX_KEYS)
            aarr, k2 = _ocols([None if c is None else c.slice(a, b) for c in agg_cols], MAX_AGGS)
            L.oracle_agg_add_batch(h, karr, aarr, b - a)
        g = L.oracle_agg_num_groups(h)
        out_types = [spec.key_types[i] for i in range(spec.num_keys)] + \
                    [spec.agg_out_types[i] for i in range(spec.num_aggs)]
        outs = [alloc_out(t, g) for t in out_types]
        bitmaps = [np.zeros((g + 7) // 8 + 1, np.uint8) for _ in outs]
        oarr = (_OutColumn * len(outs))()
        for i, (o, bm) in enumerate(zip(outs, bitmaps)):
            oarr[i].type_id, oarr[i].values, oarr[i].nulls = o.type_id, _ptr(o.values), _ptr(bm)
        rc = L.oracle_agg_get_output(h, oarr)
        if rc != 0:
            raise OverflowError(L.oracle_last_error().decode())
        for o, bm in zip(outs, bitmaps):
            o.nulls = unpack_nulls(bm, g)
        return outs
    finally:
        L.oracle_agg_destroy(h)


def agg_stream(spec: AggSpec, key_col: Col, agg_cols: Sequence[Optional[Col]], offsets: np.ndarray,
               threads: int = 1) -> np.ndarray:
    """One restated HashAggregationOperator per batch (rows [offsets[b], offsets[b + 1])): AddInput, then the
    output of every group; batches spread over `threads` host threads.  Returns the group count of each batch."""
    L = lib()
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    nb = len(offsets) - 1
    karr, k1 = _ocols([key_col], MAX_KEYS)
    aarr, k2 = _ocols(list(agg_cols), MAX_AGGS)
    groups = np.zeros(max(1, nb), np.int64)
    total = L.oracle_agg_stream(C.byref(spec), karr, aarr, _ptr(offsets), nb, threads, _ptr(groups))
    if total < 0:
        raise OverflowError(L.oracle_last_error().decode())
    return groups[:nb]


# ---- join ---------------------------------------------------------------------------------------------------
@dataclass
class JoinResult:
    probe_rows: np.ndarray
    build_rows: np.ndarray
    payloads: List[Col]
    is_array: bool


def _is_wide(key) -> bool:
    return isinstance(key, (list, tuple)) or key.type_id == DECIMAL128


def _join_create(L, build_key):
    """A restated join for a single fixed-width key, or for several key columns / one DECIMAL128 key (their keys
    translated into int64 join keys by the oracle's key codec)."""
    if not _is_wide(build_key):
        return L.oracle_join_create(build_key.type_id)
    cols = list(build_key) if isinstance(build_key, (list, tuple)) else [build_key]
    types = np.array([c.type_id for c in cols], np.int32)
    return L.oracle_join_create_multi(_ptr(types), len(cols))


def _key_slice(L, h, key, a: int, b: int, insert: bool):
    """The key of rows [a, b) as one oracle column array: the column itself, or the translated int64 join key."""
    if not _is_wide(key):
        return _ocols([key.slice(a, b)], 1)
    cols = list(key) if isinstance(key, (list, tuple)) else [key]
    karr, keep = _ocols([c.slice(a, b) for c in cols], len(cols))
    out = np.zeros(max(1, b - a), np.int64)
    bm = np.zeros((b - a + 7) // 8 + 1, np.uint8)
    L.oracle_join_translate(h, karr, len(cols), b - a, 1 if insert else 0, _ptr(out), _ptr(bm))
    return _ocols([Col(LONG, out[:b - a], unpack_nulls(bm, b - a))], 1)


def _key_len(key) -> int:
    return len(key[0]) if isinstance(key, (list, tuple)) else len(key)


def hash_join(build_key, build_payloads: Sequence[Col], probe_key, build_batch: int,
              probe_batch: int) -> JoinResult:
    """Restated HashBuilderOperator + LookupJoinOperator (inner join).  build_key / probe_key: one key column, or a
    list of key columns."""
    L = lib()
    h = _join_create(L, build_key)
    try:
        nb = _key_len(build_key)
        for a in range(0, nb, build_batch):
            b = min(nb, a + build_batch)
            karr, k1 = _key_slice(L, h, build_key, a, b, True)
            parr, k2 = _ocols([p.slice(a, b) for p in build_payloads], max(1, len(build_payloads)))
            L.oracle_join_add_batch(h, karr, parr, len(build_payloads), b - a)
        L.oracle_join_finish(h)
        is_array = bool(L.oracle_join_is_array(h))
        np_rows, nb_rows = [], []
        npr = _key_len(probe_key)
        for a in range(0, npr, probe_batch):
            b = min(npr, a + probe_batch)
            karr, k1 = _key_slice(L, h, probe_key, a, b, False)
            m = L.oracle_join_probe(h, karr, b - a, a, None, None, 0)
            pr = np.zeros(max(m, 1), np.int64)
            br = np.zeros(max(m, 1), np.int64)
            L.oracle_join_probe(h, karr, b - a, a, _ptr(pr), _ptr(br), m)
            np_rows.append(pr[:m])
            nb_rows.append(br[:m])
        probe_rows = np.concatenate(np_rows) if np_rows else np.zeros(0, np.int64)
        build_rows = np.concatenate(nb_rows) if nb_rows else np.zeros(0, np.int64)
        pays = []
        m = len(build_rows)
        for p, bp in enumerate(build_payloads):
            o = alloc_out(bp.type_id, m)
            bm = np.zeros((m + 7) // 8 + 1, np.uint8)
            L.oracle_join_gather(h, p, _ptr(build_rows), m, _ptr(o.values), _ptr(bm))
            o.nulls = unpack_nulls(bm, m)
            pays.append(o)
        return JoinResult(probe_rows, build_rows, pays, is_array)
    finally:
        L.oracle_join_destroy(h)


# JoinType / BuildSide values (core/src/plannode/planNode.h:406-414, 28-32).
JOIN_INNER, JOIN_LEFT, JOIN_RIGHT, JOIN_FULL, JOIN_LEFT_SEMI, JOIN_LEFT_ANTI, JOIN_EXISTENCE = range(7)
BUILD_UNKNOWN, BUILD_LEFT, BUILD_RIGHT = 0, 1, 2


def join_mode(join_type: int, build_side: int):
    """(probe emission mode, marks visited build rows) of a join type, restated from LookupJoinOperator::ProcessProbe
    (core/src/operator/join/lookup_join.cpp:344-367: LEFT / RIGHT with the build on the same side take the
    same-side loop, otherwise the opposite-side loop) and HashBuilderOperatorFactory (hash_builder.cpp:36-40:
    visited flags for FULL and the same-side outer joins).  Modes: 0 each match, 1 each match or one row without
    a build row, 2 semi, 3 anti; None for EXISTENCE."""
    same = (join_type == JOIN_LEFT and build_side == BUILD_LEFT) or \
        (join_type == JOIN_RIGHT and build_side == BUILD_RIGHT)
    if join_type in (JOIN_LEFT, JOIN_RIGHT):
        return (0, True) if same else (1, False)
    return {JOIN_INNER: (0, False), JOIN_FULL: (1, True), JOIN_LEFT_SEMI: (2, False),
            JOIN_LEFT_ANTI: (3, False), JOIN_EXISTENCE: (None, False)}[join_type]


@dataclass
class TypedJoinResult:
    probe_rows: np.ndarray            # probe row of each output row
    build_rows: np.ndarray            # its build row, -1 when the build columns are NULL
    payloads: List[Col]               # build payloads of each output row
    outer_build_rows: np.ndarray      # build rows the outer pass emits (build-preserving types), else empty
    outer_payloads: List[Col]
    exists: Optional[np.ndarray]      # EXISTENCE: one flag per probe row


def hash_join_typed(build_key, build_payloads: Sequence[Col], probe_key, build_batch: int,
                    probe_batch: int, join_type: int, build_side: int = BUILD_UNKNOWN) -> TypedJoinResult:
    """Restated HashBuilderOperator + LookupJoinOperator (+ LookupOuterJoinOperator) for any join type; keys as in
    hash_join."""
    mode, mark = join_mode(join_type, build_side)
    L = lib()
    h = _join_create(L, build_key)
    try:
        nb = _key_len(build_key)
        for a in list(range(0, nb, build_batch)) or [0]:  # an empty build side still declares its payload types
            b = min(nb, a + build_batch)
            karr, k1 = _key_slice(L, h, build_key, a, b, True)
            parr, k2 = _ocols([p.slice(a, b) for p in build_payloads], max(1, len(build_payloads)))
            L.oracle_join_add_batch(h, karr, parr, len(build_payloads), b - a)
        L.oracle_join_finish(h)
        npr = _key_len(probe_key)
        if mode is None:
            ex = np.zeros(max(1, npr), np.uint8)
            for a in range(0, npr, probe_batch):
                b = min(npr, a + probe_batch)
                karr, k1 = _key_slice(L, h, probe_key, a, b, False)
                L.oracle_join_exists(h, karr, b - a, _ptr(ex[a:b]))
            return TypedJoinResult(np.zeros(0, np.int64), np.zeros(0, np.int64), [], np.zeros(0, np.int64), [],
                                   ex[:npr])
        np_rows, nb_rows = [], []
        for a in range(0, npr, probe_batch):
            b = min(npr, a + probe_batch)
            karr, k1 = _key_slice(L, h, probe_key, a, b, False)
            m = L.oracle_join_probe_typed(h, karr, b - a, a, mode, int(mark), None, None, 0)
            pr = np.zeros(max(m, 1), np.int64)
            br = np.zeros(max(m, 1), np.int64)
            # the visited marks of the counting call are the same as the second call's
            L.oracle_join_probe_typed(h, karr, b - a, a, mode, int(mark), _ptr(pr), _ptr(br), m)
            np_rows.append(pr[:m])
            nb_rows.append(br[:m])
        probe_rows = np.concatenate(np_rows) if np_rows else np.zeros(0, np.int64)
        build_rows = np.concatenate(nb_rows) if nb_rows else np.zeros(0, np.int64)

        def gather(rows):
            out = []
            for p, bp in enumerate(build_payloads):
                o = alloc_out(bp.type_id, len(rows))
                bm = np.zeros((len(rows) + 7) // 8 + 1, np.uint8)
                L.oracle_join_gather(h, p, _ptr(np.ascontiguousarray(rows)), len(rows), _ptr(o.values), _ptr(bm))
                o.nulls = unpack_nulls(bm, len(rows))
                out.append(o)
            return out

        pays = gather(build_rows) if mode in (0, 1) else []
        outer = np.zeros(0, np.int64)
        if mark:
            m = L.oracle_join_outer(h, None, 0)
            outer = np.zeros(max(1, m), np.int64)
            L.oracle_join_outer(h, _ptr(outer), m)
            outer = outer[:m]
        return TypedJoinResult(probe_rows, build_rows, pays, outer, gather(outer) if mark else [], None)
    finally:
        L.oracle_join_destroy(h)


def mm3_hash_int(v: int, seed: int = 42) -> int:
    """Murmur3 HashInt (core/src/operator/util/mm3_util.h) of a 32-bit value, as a signed int32."""
    return int(np.uint32(lib().oracle_mm3_hash_int(v & 0xFFFFFFFF, seed & 0xFFFFFFFF)).view(np.int32))


def mm3_hash_long(v: int, seed: int = 42) -> int:
    """Murmur3 HashLong of a 64-bit value, as a signed int32."""
    return int(np.uint32(lib().oracle_mm3_hash_long(v & 0xFFFFFFFFFFFFFFFF, seed & 0xFFFFFFFF)).view(np.int32))


def partition_ids(cols: Sequence[Col], partitions: int) -> np.ndarray:
    """Restated HashUtil::ComputePartitionIds (core/src/operator/hash_util.cpp:40-82): one int32 per row."""
    n = len(cols[0]) if cols else 0
    out = np.zeros(max(1, n), np.int32)
    arr, keep = _ocols(list(cols), max(1, len(cols)))
    lib().oracle_partition_ids(arr, len(cols), n, partitions, _ptr(out))
    return out[:n]
