"""ctypes binding of the C ABI in include/omni_gpu.h (libomni_gpu.so).

This is a thin Python face of the native library for tests and benchmarks; every computation runs in
the HIP kernels behind the C ABI.  There is no CPU fallback: if the library is missing or no GPU is
present, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libomni_gpu.so")

# DataTypeId / FunctionType values shared with the reference (core/src/type/data_type.h:40-65,
# core/src/operator/util/function_type.h:11-27).
INT, LONG, DOUBLE, BOOLEAN, SHORT, DECIMAL64, DECIMAL128, DATE32, TIMESTAMP = 1, 2, 3, 4, 5, 6, 7, 8, 12
BYTE = 18
SUM, COUNT_COLUMN, COUNT_ALL, AVG, MAX, MIN = 0, 1, 2, 3, 5, 6
STRATEGY_AUTO, STRATEGY_LDS, STRATEGY_GLOBAL, STRATEGY_RADIX = 0, 1, 2, 3
MAX_KEYS, MAX_AGGS = 4, 16
HASH_TABLE_NORMAL, HASH_TABLE_ARRAY = 0, 1  # AggHashTableRule (core/src/util/policy.h:51)
# JoinType / BuildSide values shared with the reference (core/src/plannode/planNode.h:406-414, 28-32).
JOIN_INNER, JOIN_LEFT, JOIN_RIGHT, JOIN_FULL, JOIN_LEFT_SEMI, JOIN_LEFT_ANTI, JOIN_EXISTENCE = range(7)
BUILD_UNKNOWN, BUILD_LEFT, BUILD_RIGHT = 0, 1, 2

ERRORS = {-1: "INVALID", -2: "HIP", -3: "NOMEM", -4: "OVERFLOW", -5: "OUTPUT_FULL", -6: "STATE"}

WIDTH = {SHORT: 2, INT: 4, DATE32: 4, LONG: 8, TIMESTAMP: 8, DECIMAL64: 8, DOUBLE: 8, DECIMAL128: 16, BOOLEAN: 1, BYTE: 1}
NP = {INT: np.int32, DATE32: np.int32, LONG: np.int64, TIMESTAMP: np.int64, DECIMAL64: np.int64,
      DOUBLE: np.float64, SHORT: np.int16, BOOLEAN: np.uint8, BYTE: np.int8, DECIMAL128: np.uint64}


class GpuError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"omni_gpu error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class Column(C.Structure):
    _fields_ = [("type_id", C.c_int32), ("reserved", C.c_int32), ("values", C.c_void_p), ("nulls", C.c_void_p)]


class OutColumn(C.Structure):
    _fields_ = [("type_id", C.c_int32), ("reserved", C.c_int32), ("values", C.c_void_p), ("nulls", C.c_void_p)]


class AggSpec(C.Structure):
    _fields_ = [
        ("num_keys", C.c_int32), ("key_types", C.c_int32 * MAX_KEYS),
        ("num_aggs", C.c_int32), ("agg_funcs", C.c_int32 * MAX_AGGS),
        ("agg_in_types", C.c_int32 * MAX_AGGS), ("agg_out_types", C.c_int32 * MAX_AGGS),
        ("agg_in_scales", C.c_int32 * MAX_AGGS), ("agg_out_scales", C.c_int32 * MAX_AGGS),
        ("overflow_as_null", C.c_int32), ("strategy", C.c_int32), ("capacity_hint", C.c_int64),
        ("agg_partial_out", C.c_int32 * MAX_AGGS), ("hash_table_rule", C.c_int32),
    ]


class JoinSpec(C.Structure):
    _fields_ = [("key_type", C.c_int32), ("num_payloads", C.c_int32), ("payload_types", C.c_int32 * MAX_AGGS),
                ("strategy", C.c_int32), ("num_keys", C.c_int32), ("key_types", C.c_int32 * MAX_KEYS)]


class Sketch(C.Structure):
    _fields_ = [("features", C.c_float * 16), ("rows", C.c_int64), ("sampled", C.c_int64),
                ("distinct", C.c_int64), ("f1", C.c_int64), ("f2", C.c_int64), ("max_count", C.c_int64),
                ("heavy", C.c_int64), ("nulls", C.c_int64), ("key_min", C.c_int64), ("key_max", C.c_int64)]


# Every exported symbol of include/omni_gpu.h with its ctypes signature.
_I, _I32, _I64, _VP, _SZ, _D = C.c_int, C.c_int32, C.c_int64, C.c_void_p, C.c_size_t, C.c_double
_P = C.POINTER
SIGNATURES = {
    "omni_gpu_init": (_I, [_I]),
    "omni_gpu_last_error": (C.c_char_p, []),
    "omni_gpu_malloc": (_I, [_P(_VP), _SZ]),
    "omni_gpu_free": (_I, [_VP]),
    "omni_gpu_memcpy_h2d": (_I, [_VP, _VP, _SZ]),
    "omni_gpu_memcpy_d2h": (_I, [_VP, _VP, _SZ]),
    "omni_gpu_memcpy_d2d": (_I, [_VP, _VP, _SZ]),
    "omni_gpu_memset": (_I, [_VP, _I, _SZ]),
    "omni_gpu_host_alloc": (_I, [_P(_VP), _SZ]),
    "omni_gpu_host_free": (_I, [_VP]),
    "omni_gpu_device_synchronize": (_I, []),
    "omni_gpu_memory_stats": (_I, [_P(_I64), _P(_I64)]),
    "omni_gpu_profile_enable": (_I, [_I]),
    "omni_gpu_profile_reset": (_I, []),
    "omni_gpu_profile_get": (_I, [C.c_char_p, _P(_I64), _P(_D)]),
    "omni_gpu_agg_create": (_I, [_P(AggSpec), _P(_VP)]),
    "omni_gpu_agg_add_batch": (_I, [_VP, _P(Column), _P(Column), _I64]),
    "omni_gpu_agg_add_batches": (_I, [_VP, _P(Column), _P(Column), _P(_I64), _I32]),
    "omni_gpu_agg_wait": (_I, [_VP]),
    "omni_gpu_agg_num_groups": (_I, [_VP, _P(_I64)]),
    "omni_gpu_agg_get_output": (_I, [_VP, _I64, _I64, _P(OutColumn)]),
    "omni_gpu_agg_destroy": (_I, [_VP]),
    "omni_gpu_agg_batches_create": (_I, [_P(AggSpec), _I64, _P(_VP)]),
    "omni_gpu_agg_batches_tune": (_I, [_VP, _I32, _I32, _I32]),
    "omni_gpu_agg_batches_info": (_I, [_VP, _P(_I32), _P(_I64)]),
    "omni_gpu_agg_batches_run": (_I, [_VP, _P(Column), _P(Column), _P(_I64), _I32, _VP, _I32, _P(OutColumn),
                                      _P(_I64)]),
    "omni_gpu_agg_batches_stats": (_I, [_VP, _P(_I64), _P(_I64)]),
    "omni_gpu_agg_batches_phases": (_I, [_VP, _P(_I64), _P(_I64), _P(_I64)]),
    "omni_gpu_agg_batches_tune_partitions": (_I, [_VP, _I32, _I32]),
    "omni_gpu_agg_batches_destroy": (_I, [_VP]),
    "omni_gpu_agg_batches_set_advisor": (_I, [_VP, _VP]),
    "omni_gpu_agg_batches_phase_of": (_I, [_VP, _VP]),
    "omni_gpu_batch_features": (_I, [_P(Column), _VP, _I32, _VP]),
    "omni_gpu_join_build_create": (_I, [_P(JoinSpec), _P(_VP)]),
    "omni_gpu_join_build_add": (_I, [_VP, _P(Column), _P(Column), _I64]),
    "omni_gpu_join_build_finish": (_I, [_VP]),
    "omni_gpu_join_build_info": (_I, [_VP, _P(_I32), _P(_I64), _P(_I64)]),
    "omni_gpu_join_tune": (_I, [_VP, _I32, _I32]),
    "omni_gpu_join_probe_info": (_I, [_VP, _P(_I32)]),
    "omni_gpu_join_probe": (_I, [_VP, _P(Column), _I64, _I64, _VP, _P(OutColumn), _I64, _P(_I64)]),
    "omni_gpu_join_probe_batches": (_I, [_VP, _P(Column), _P(_I64), _I32, _VP, _P(OutColumn), _I64, _P(_I64)]),
    "omni_gpu_agg_table_info": (_I, [_VP, _P(_I32), _P(_I64), _P(_I64)]),
    "omni_gpu_join_set_type": (_I, [_VP, _I32, _I32]),
    "omni_gpu_join_probe_exists": (_I, [_VP, _P(Column), _I64, _VP]),
    "omni_gpu_join_gather_rows": (_I, [_P(Column), _I32, _VP, _I64, _P(OutColumn)]),
    "omni_gpu_join_partition_ids": (_I, [_P(Column), _I32, _I64, C.c_uint32, _VP]),
    "omni_gpu_join_outer_output": (_I, [_VP, _VP, _P(OutColumn), _I64, _P(_I64)]),
    "omni_gpu_join_destroy": (_I, [_VP]),
    "omni_gpu_partition_ids": (_I, [_P(Column), _I32, _I64, _I32, _VP]),
    "omni_gpu_sketch_keys": (_I, [_P(Column), _I32, _I64, _I32, _P(Sketch)]),
    "omni_gpu_sketch_keys_noalloc": (_I, [_P(Column), _I32, _I64, _I32, _VP, _VP]),
    "omni_gpu_advise": (_I, [_VP, _P(Column), _I32, _I64, _I32, _P(C.c_float)]),
    "omni_gpu_mlp_create": (_I, [C.c_uint64, _P(_VP)]),
    "omni_gpu_mlp_num_params": (_I, []),
    "omni_gpu_mlp_set_params": (_I, [_VP, _P(C.c_float)]),
    "omni_gpu_mlp_get_params": (_I, [_VP, _P(C.c_float)]),
    "omni_gpu_mlp_forward": (_I, [_VP, _P(C.c_float), _I32, _P(C.c_float)]),
    "omni_gpu_mlp_gradient": (_I, [_VP, _P(C.c_float), _P(C.c_float), _P(_I32), _I32, _P(C.c_float), _P(C.c_float)]),
    "omni_gpu_mlp_train": (_I, [_VP, _P(C.c_float), _P(C.c_float), _P(_I32), _I32, _I32, _I32, C.c_float, C.c_uint64,
                                _P(C.c_float)]),
    "omni_gpu_mlp_destroy": (_I, [_VP]),
    "omni_gpu_generate_keys": (_I, [_I32, _I64, _D, _D, C.c_uint64, _VP, _VP, _VP]),
}

_lib = None
_inited = False


def lib():
    """Loads libomni_gpu.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m omnioperator_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _lib = L
    return _lib


def check(rc: int):
    if rc != 0:
        raise GpuError(rc, lib().omni_gpu_last_error().decode(errors="replace"))


def init(device: int = 0):
    global _inited
    if not _inited:
        check(lib().omni_gpu_init(device))
        _inited = True


class DeviceArray:
    """Owning device buffer holding a numpy-shaped array."""

    def __init__(self, nbytes: int, dtype=np.uint8, shape=None):
        init()
        self.nbytes = int(nbytes)
        self.dtype = np.dtype(dtype)
        self.shape = shape if shape is not None else (self.nbytes // max(1, self.dtype.itemsize),)
        p = C.c_void_p()
        check(lib().omni_gpu_malloc(C.byref(p), max(1, self.nbytes)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes, a.dtype, a.shape)
        if a.nbytes:
            check(lib().omni_gpu_memcpy_h2d(d.ptr, a.ctypes.data, a.nbytes))
        return d

    @classmethod
    def zeros(cls, shape, dtype) -> "DeviceArray":
        shape = (shape,) if isinstance(shape, int) else tuple(shape)
        nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        d = cls(nbytes, dtype, shape)
        check(lib().omni_gpu_memset(d.ptr, 0, max(1, nbytes)))
        return d

    def numpy(self, count: Optional[int] = None) -> np.ndarray:
        shape = self.shape if count is None else (count,) + tuple(self.shape[1:])
        out = np.empty(shape, self.dtype)
        if out.nbytes:
            check(lib().omni_gpu_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def free(self):
        if getattr(self, "ptr", None):
            lib().omni_gpu_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def pack_nulls(mask: Optional[np.ndarray]) -> Optional[np.ndarray]:
    if mask is None:
        return None
    return np.packbits(np.asarray(mask, dtype=bool), bitorder="little")


def unpack_nulls(bitmap: np.ndarray, n: int) -> np.ndarray:
    return np.unpackbits(np.asarray(bitmap, dtype=np.uint8), count=n, bitorder="little").astype(bool)


class DeviceColumn:
    """A column resident in device memory (values + optional LSB-first null bitmap)."""

    def __init__(self, type_id: int, values: np.ndarray, nulls: Optional[np.ndarray] = None):
        self.type_id = type_id
        self.rows = int(values.shape[0])
        self.values = DeviceArray.from_numpy(values)
        bm = pack_nulls(nulls)
        self.nulls = DeviceArray.from_numpy(bm) if bm is not None and nulls.any() else None

    def c(self, offset_rows: int = 0) -> Column:
        col = Column()
        col.type_id = self.type_id
        col.values = self.values.ptr + offset_rows * WIDTH[self.type_id]
        col.nulls = self.nulls.ptr if self.nulls is not None else None
        if offset_rows and self.nulls is not None:
            raise ValueError("sliced columns with NULLs need an own bitmap")
        return col


def make_agg_spec(key_types: Sequence[int], aggs: Sequence[tuple], overflow_as_null: bool = True,
                  strategy: int = STRATEGY_AUTO, capacity_hint: int = 0, hash_table_rule: int = 0) -> AggSpec:
    """aggs: tuples (func, in_type, out_type[, in_scale, out_scale[, partial_out]])."""
    s = AggSpec()
    s.num_keys = len(key_types)
    for i, t in enumerate(key_types):
        s.key_types[i] = t
    s.num_aggs = len(aggs)
    for i, a in enumerate(aggs):
        s.agg_funcs[i], s.agg_in_types[i], s.agg_out_types[i] = a[0], a[1], a[2]
        if len(a) > 3:
            s.agg_in_scales[i], s.agg_out_scales[i] = a[3], a[4]
        if len(a) > 5:
            s.agg_partial_out[i] = 1 if a[5] else 0
    s.overflow_as_null = 1 if overflow_as_null else 0
    s.strategy = strategy
    s.capacity_hint = capacity_hint
    s.hash_table_rule = hash_table_rule
    return s


class HashAgg:
    """GPU group-by aggregation handle (omni_gpu_agg_*)."""

    def __init__(self, spec: AggSpec):
        init()
        self.spec = spec
        h = C.c_void_p()
        check(lib().omni_gpu_agg_create(C.byref(spec), C.byref(h)))
        self.h = h.value

    def add_batch(self, keys: Sequence[Column], aggs: Sequence[Optional[Column]], rows: int):
        karr = (Column * MAX_KEYS)(*keys)
        aarr = (Column * MAX_AGGS)()
        for i, a in enumerate(aggs):
            if a is not None:
                aarr[i] = a
        check(lib().omni_gpu_agg_add_batch(self.h, karr, aarr, rows))

    def batch_arrays(self, batches):
        """ctypes arrays for add_batches_raw from [(keys, aggs, rows)] (aggs entries may be None)."""
        return HashAgg.batch_arrays_for(self.spec, batches)

    @staticmethod
    def batch_arrays_for(spec, batches):
        nk, na = spec.num_keys, spec.num_aggs
        karr = (Column * max(1, len(batches) * nk))()
        aarr = (Column * max(1, len(batches) * na))()
        rarr = (C.c_int64 * max(1, len(batches)))()
        for b, (keys, aggs, rows) in enumerate(batches):
            for c, k in enumerate(keys):
                karr[b * nk + c] = k
            for i, a in enumerate(aggs):
                if a is not None:
                    aarr[b * na + i] = a
            rarr[b] = rows
        return karr, aarr, rarr, len(batches)

    def add_batches(self, batches):
        """Adds [(keys, aggs, rows)] with one omni_gpu_agg_add_batches call."""
        self.add_batches_raw(*self.batch_arrays(batches))

    def add_batches_raw(self, karr, aarr, rarr, n: int):
        check(lib().omni_gpu_agg_add_batches(self.h, karr, aarr, rarr, n))

    def table_info(self):
        """(direct-index map in use, its first key, its cells) -- the ARRAY rule's decision."""
        a, lo, n = C.c_int32(), C.c_int64(), C.c_int64()
        check(lib().omni_gpu_agg_table_info(self.h, C.byref(a), C.byref(lo), C.byref(n)))
        return bool(a.value), lo.value, n.value

    def wait(self):
        """omni_gpu_agg_wait: every added batch consumed (the caller's columns may be released afterwards)."""
        check(lib().omni_gpu_agg_wait(self.h))

    def num_groups(self) -> int:
        g = C.c_int64()
        check(lib().omni_gpu_agg_num_groups(self.h, C.byref(g)))
        return g.value

    def output(self, start: int = 0, count: Optional[int] = None):
        """Returns [(values ndarray, null mask)] for keys then aggregates."""
        if count is None:
            count = self.num_groups() - start
        types = [self.spec.key_types[i] for i in range(self.spec.num_keys)] + \
                [self.spec.agg_out_types[i] for i in range(self.spec.num_aggs)]
        bufs = []
        oarr = (OutColumn * len(types))()
        for i, t in enumerate(types):
            shape = (count, 2) if t == DECIMAL128 else (count,)
            v = DeviceArray.zeros(shape, NP[t])
            nb = DeviceArray.zeros(max(1, (count + 7) // 8), np.uint8)
            bufs.append((v, nb))
            oarr[i].type_id, oarr[i].values, oarr[i].nulls = t, v.ptr, nb.ptr
        check(lib().omni_gpu_agg_get_output(self.h, start, count, oarr))
        return [(v.numpy(), unpack_nulls(nb.numpy(), count)) for v, nb in bufs]

    def alloc_output(self, count: int):
        """Device output buffers for `count` groups (keys, then aggregates)."""
        types = [self.spec.key_types[i] for i in range(self.spec.num_keys)] + \
                [self.spec.agg_out_types[i] for i in range(self.spec.num_aggs)]
        bufs, oarr = [], (OutColumn * len(types))()
        for i, t in enumerate(types):
            shape = (count, 2) if t == DECIMAL128 else (count,)
            v = DeviceArray(max(1, count) * WIDTH[t], NP[t], shape)
            nb = DeviceArray(max(8, (count + 63) // 64 * 8), np.uint8)
            bufs.append((v, nb))
            oarr[i].type_id, oarr[i].values, oarr[i].nulls = t, v.ptr, nb.ptr
        return (bufs, oarr, count)

    def output_into(self, bufs, count: int, start: int = 0):
        """Writes groups [start, start + count) into buffers from alloc_output (device-resident)."""
        assert count <= bufs[2]
        check(lib().omni_gpu_agg_get_output(self.h, start, count, bufs[1]))

    def close(self):
        if getattr(self, "h", None):
            lib().omni_gpu_agg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class AggBatches:
    """Batched stream of independent group-bys (omni_gpu_agg_batches_*)."""

    FLAG_PER_BATCH, FLAG_GRAPH, FLAG_SKETCH, FLAG_ADVISOR, FLAG_FUSED = 1, 2, 4, 8, 16

    def __init__(self, spec: AggSpec, max_batch_rows: int):
        init()
        self.spec = spec
        h = C.c_void_p()
        check(lib().omni_gpu_agg_batches_create(C.byref(spec), max_batch_rows, C.byref(h)))
        self.h = h.value
        self._outs = None

    def tune(self, block: int = 256, lds_bytes: int = 32768, max_probe: int = 32):
        check(lib().omni_gpu_agg_batches_tune(self.h, block, lds_bytes, max_probe))

    def tune_partitions(self, part_lds_bytes: int = 65536, part_rows: int = 0):
        check(lib().omni_gpu_agg_batches_tune_partitions(self.h, part_lds_bytes, part_rows))

    def info(self):
        s, p = C.c_int32(), C.c_int64()
        check(lib().omni_gpu_agg_batches_info(self.h, C.byref(s), C.byref(p)))
        return {"lds_slots": s.value, "pool_bytes": p.value}

    def alloc_outputs(self, total_rows: int):
        """Device output columns (key, then one per aggregate) for a stream of total_rows rows."""
        types = [self.spec.key_types[0]] + [self.spec.agg_out_types[i] for i in range(self.spec.num_aggs)]
        bufs, oarr = [], (OutColumn * len(types))()
        for i, t in enumerate(types):
            shape = (total_rows, 2) if t == DECIMAL128 else (total_rows,)
            v = DeviceArray(max(1, total_rows) * WIDTH[t], NP[t], shape)
            nb = DeviceArray(max(4, ((total_rows + 31) // 32) * 4), np.uint8)
            bufs.append((v, nb))
            oarr[i].type_id, oarr[i].values, oarr[i].nulls = t, v.ptr, nb.ptr
        self._outs = (types, bufs, oarr, total_rows)
        return self._outs

    def run(self, keys: Column, aggs: Sequence[Optional[Column]], offsets: np.ndarray, modes=None, flags: int = 0):
        """Runs every batch; returns groups per batch (numpy int64).  Outputs stay on the device (outputs())."""
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        nb = len(offsets) - 1
        total = int(offsets[-1])
        if self._outs is None or self._outs[3] != total:
            self.alloc_outputs(total)
        aarr = (Column * MAX_AGGS)()
        for i, a in enumerate(aggs):
            if a is not None:
                aarr[i] = a
        groups = np.zeros(nb, np.int64)
        mp = None
        if modes is not None:
            modes = np.ascontiguousarray(modes, dtype=np.int32)
            mp = modes.ctypes.data
        karr = (Column * 1)(keys)
        check(lib().omni_gpu_agg_batches_run(self.h, karr, aarr, offsets.ctypes.data_as(C.POINTER(C.c_int64)), nb, mp,
                                             flags, self._outs[2], groups.ctypes.data_as(C.POINTER(C.c_int64))))
        return groups

    def outputs(self):
        """[(values, null mask)] of the whole output buffer (rows past each batch's group count are junk)."""
        types, bufs, _, total = self._outs
        return [(v.numpy(), unpack_nulls(nb.numpy(), total)) for v, nb in bufs]

    def set_advisor(self, mlp: Optional["Mlp"]):
        """Attaches the advisor network used under FLAG_ADVISOR (None detaches it)."""
        check(lib().omni_gpu_agg_batches_set_advisor(self.h, mlp.h if mlp is not None else None))
        self._advisor = mlp  # keeps the network alive while attached

    def phase_of(self, nb: int) -> np.ndarray:
        """The phase (1, 2, 3) that finished each batch of the last run."""
        out = np.zeros(max(1, nb), np.int32)
        check(lib().omni_gpu_agg_batches_phase_of(self.h, out.ctypes.data))
        return out[:nb]

    def stats(self):
        a, b = C.c_int64(), C.c_int64()
        check(lib().omni_gpu_agg_batches_stats(self.h, C.byref(a), C.byref(b)))
        return {"lds_batches": a.value, "hbm_batches": b.value}

    def phases(self):
        """Batches of the last run finished by phase 1 / 2 / 3."""
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().omni_gpu_agg_batches_phases(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def close(self):
        if getattr(self, "h", None):
            lib().omni_gpu_agg_batches_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _key_arg(key):
    """A key Column, a sequence of them (several key columns) or a prebuilt ctypes array of them, as the pointer
    the C ABI takes."""
    if isinstance(key, C.Array):
        return key
    if isinstance(key, (list, tuple)):
        return (Column * len(key))(*key)
    return C.byref(key)


class HashJoin:
    """GPU hash join handle (omni_gpu_join_*)."""

    def __init__(self, key_type, payload_types: Sequence[int] = (), strategy: int = 0):
        """key_type: one key type, or a sequence of them (several key columns)."""
        init()
        s = JoinSpec()
        types = list(key_type) if isinstance(key_type, (list, tuple)) else [key_type]
        s.key_type = types[0]
        s.num_keys = len(types)
        for i, t in enumerate(types):
            s.key_types[i] = t
        s.num_payloads = len(payload_types)
        for i, t in enumerate(payload_types):
            s.payload_types[i] = t
        s.strategy = strategy
        self.spec = s
        self.payload_types = list(payload_types)
        self.build_rows = 0
        h = C.c_void_p()
        check(lib().omni_gpu_join_build_create(C.byref(s), C.byref(h)))
        self.h = h.value

    def add_build(self, key: Column, payloads, rows: int):
        """payloads: sequence of Columns or a prebuilt ctypes array of them."""
        parr = payloads if isinstance(payloads, C.Array) else (Column * max(1, len(payloads)))(*payloads)
        check(lib().omni_gpu_join_build_add(self.h, _key_arg(key), parr, rows))
        self.build_rows += rows

    def set_type(self, join_type: int, build_side: int = 0):
        """Join type (OMNI_GPU_JOIN_*) and build side (OMNI_GPU_BUILD_*), before finish."""
        check(lib().omni_gpu_join_set_type(self.h, join_type, build_side))

    def probe_exists(self, key: Column, rows: int) -> np.ndarray:
        """Existence flags (uint8, one per probe row) of one batch."""
        out = DeviceArray(max(1, rows), np.uint8, (rows,))
        check(lib().omni_gpu_join_probe_exists(self.h, _key_arg(key), rows, out.ptr))
        return out.numpy(rows)

    def outer_output_raw(self, out_rows: "DeviceArray", oarr, cap: int) -> int:
        """Device-resident outer pass (omni_gpu_join_outer_output); returns the row count."""
        m = C.c_int64()
        check(lib().omni_gpu_join_outer_output(self.h, out_rows.ptr, oarr, cap, C.byref(m)))
        return m.value

    def outer_output(self, cap: Optional[int] = None, with_nulls: bool = True):
        """Build rows no probe matched: (build row ids, [(payload values, null mask)])."""
        if cap is None:
            cap = self.build_rows
        out_rows, pays, oarr = self._outs(cap, with_nulls)
        m = C.c_int64()
        check(lib().omni_gpu_join_outer_output(self.h, out_rows.ptr, oarr, cap, C.byref(m)))
        n = m.value
        res = [(v.numpy(n), unpack_nulls(nb.numpy(), n) if nb is not None else np.zeros(n, bool)) for v, nb in pays]
        return out_rows.numpy(n), res

    def finish(self):
        check(lib().omni_gpu_join_build_finish(self.h))

    def tune(self, probe_group: int = 1, partitions: int = -1):
        check(lib().omni_gpu_join_tune(self.h, probe_group, partitions))

    def last_partitions(self) -> int:
        p = C.c_int32()
        check(lib().omni_gpu_join_probe_info(self.h, C.byref(p)))
        return p.value

    def info(self):
        a, d, s = C.c_int32(), C.c_int64(), C.c_int64()
        check(lib().omni_gpu_join_build_info(self.h, C.byref(a), C.byref(d), C.byref(s)))
        return {"is_array": bool(a.value), "max_dup": d.value, "table_slots": s.value}

    def _outs(self, cap: int, with_nulls: bool):
        rows = DeviceArray(max(1, cap) * 8, np.int64, (cap,))
        pays, oarr = [], (OutColumn * max(1, len(self.payload_types)))()
        for i, t in enumerate(self.payload_types):
            shape = (cap, 2) if t == DECIMAL128 else (cap,)
            v = DeviceArray(max(1, cap) * WIDTH[t], NP[t], shape)
            nb = DeviceArray(max(4, ((cap + 31) // 32) * 4), np.uint8) if with_nulls else None
            pays.append((v, nb))
            oarr[i].type_id, oarr[i].values = t, v.ptr
            oarr[i].nulls = nb.ptr if nb is not None else None
        return rows, pays, oarr

    def probe(self, key: Column, rows: int, row_base: int = 0, cap: Optional[int] = None, with_nulls: bool = True):
        """Probes one batch; returns (probe_rows, [(payload values, null mask)])."""
        if cap is None:
            cap = rows * max(1, self.info()["max_dup"])
        out_rows, pays, oarr = self._outs(cap, with_nulls)
        m = C.c_int64()
        check(lib().omni_gpu_join_probe(self.h, _key_arg(key), rows, row_base, out_rows.ptr, oarr, cap, C.byref(m)))
        n = m.value
        res = [(v.numpy(n), unpack_nulls(nb.numpy(), n) if nb is not None else np.zeros(n, bool)) for v, nb in pays]
        return out_rows.numpy(n), res

    def probe_batches_raw(self, keys: Sequence[Column], rows: Sequence[int], out_rows: DeviceArray,
                          oarr, cap: int) -> int:
        """Device-resident batched probe (one launch); returns the match count.  keys / rows may be prebuilt
        ctypes arrays (probe_arrays) so that repeated calls do not rebuild them."""
        karr = keys if isinstance(keys, C.Array) else (Column * len(keys))(*keys)
        rarr = rows if isinstance(rows, C.Array) else (C.c_int64 * len(rows))(*rows)
        m = C.c_int64()
        check(lib().omni_gpu_join_probe_batches(self.h, karr, rarr, len(rarr), out_rows.ptr, oarr, cap, C.byref(m)))
        return m.value

    @staticmethod
    def probe_arrays(keys: Sequence[Column], rows: Sequence[int]):
        """ctypes arrays of probe batch columns and row counts for probe_batches_raw."""
        return (Column * len(keys))(*keys), (C.c_int64 * len(rows))(*rows)

    def close(self):
        if getattr(self, "h", None):
            lib().omni_gpu_join_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def partition_ids_raw(cols: Sequence[Column], rows: int, partitions: int, out: "DeviceArray"):
    """Shuffle partition ids of device columns into a device int32 array (omni_gpu_partition_ids)."""
    init()
    arr = cols if isinstance(cols, C.Array) else (Column * len(cols))(*cols)
    check(lib().omni_gpu_partition_ids(arr, len(arr), rows, partitions, out.ptr))


def partition_ids(cols: Sequence[Column], rows: int, partitions: int) -> np.ndarray:
    out = DeviceArray(max(1, rows) * 4, np.int32, (rows,))
    partition_ids_raw(cols, rows, partitions, out)
    return out.numpy(rows)


def sketch_keys(keys: Sequence[Column], rows: int, is_join: bool = False) -> Sketch:
    """Sample sketch of one or two device key columns (omni_gpu_sketch_keys)."""
    init()
    karr = (Column * len(keys))(*keys)
    out = Sketch()
    check(lib().omni_gpu_sketch_keys(karr, len(keys), rows, 1 if is_join else 0, C.byref(out)))
    return out


def batch_features(key: Column, offsets: np.ndarray) -> np.ndarray:
    """The advisor's 16 features of each batch of a stream (omni_gpu_batch_features): float32 [nb, 16]."""
    init()
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    nb = len(offsets) - 1
    out = np.zeros((max(1, nb), 16), np.float32)
    check(lib().omni_gpu_batch_features(C.byref(key), offsets.ctypes.data, nb, out.ctypes.data))
    return out[:nb]


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Mlp:
    """The advisor's 16-64-64-4 fp32 MLP on the GPU (omni_gpu_mlp_*)."""

    def __init__(self, seed: int = 0):
        init()
        h = C.c_void_p()
        check(lib().omni_gpu_mlp_create(seed, C.byref(h)))
        self.h = h.value
        self.n_params = lib().omni_gpu_mlp_num_params()

    def params(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        check(lib().omni_gpu_mlp_get_params(self.h, _fptr(p)))
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.shape == (self.n_params,)
        check(lib().omni_gpu_mlp_set_params(self.h, _fptr(p)))

    def advise(self, keys: Sequence[Column], rows: int, is_join: bool = False) -> np.ndarray:
        """Sketch of the device key columns + inference in one call; returns the 4 outputs."""
        karr = (Column * len(keys))(*keys)
        y = np.zeros(4, np.float32)
        check(lib().omni_gpu_advise(self.h, karr, len(keys), rows, 1 if is_join else 0, _fptr(y)))
        return y

    def forward(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((len(X), 4), np.float32)
        check(lib().omni_gpu_mlp_forward(self.h, _fptr(X), len(X), _fptr(Y)))
        return Y

    def gradient(self, X: np.ndarray, t: np.ndarray, c: np.ndarray):
        X = np.ascontiguousarray(X, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        c = np.ascontiguousarray(c, dtype=np.int32)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_float()
        check(lib().omni_gpu_mlp_gradient(self.h, _fptr(X), _fptr(t), c.ctypes.data_as(C.POINTER(C.c_int32)), len(X),
                                          _fptr(g), C.byref(loss)))
        return g, loss.value

    def train(self, X, t, c, batch: int = 128, epochs: int = 50, lr: float = 3e-3, seed: int = 0) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        c = np.ascontiguousarray(c, dtype=np.int32)
        losses = np.zeros(epochs, np.float32)
        check(lib().omni_gpu_mlp_train(self.h, _fptr(X), _fptr(t), c.ctypes.data_as(C.POINTER(C.c_int32)), len(X),
                                       batch, epochs, lr, seed, _fptr(losses)))
        return losses

    def close(self):
        if getattr(self, "h", None):
            lib().omni_gpu_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


GEN_SEQUENTIAL = 8   # row i gets rank i + 1 (unique keys: a join build side)
GEN_MISSES = 16      # about 20% of rows get keys outside the rank range (join probe misses)


def generate_keys(kind: int, rows: int, distinct: float, zipf_s: float, seed: int, with_values: bool = True):
    """Device key columns from the seeded generator: kind & 7 = 0 int64, 1 int32, 2 (int32, int64), optionally
    | GEN_SEQUENTIAL / GEN_MISSES; returns ([key DeviceArrays...], value DeviceArray or None)."""
    init()
    t = kind & 7
    k64 = DeviceArray(max(1, rows) * 8, np.int64, (rows,)) if t != 1 else None
    k32 = DeviceArray(max(1, rows) * 4, np.int32, (rows,)) if t != 0 else None
    v = DeviceArray(max(1, rows) * 8, np.int64, (rows,)) if with_values else None
    check(lib().omni_gpu_generate_keys(kind, rows, float(distinct), float(zipf_s), seed, k64.ptr if k64 else None,
                                       k32.ptr if k32 else None, v.ptr if v else None))
    keys = [k64] if t == 0 else ([k32] if t == 1 else [k32, k64])
    return keys, v


def memory_stats():
    """(bytes in live buffers of open handles, bytes of freed blocks kept for reuse)."""
    used, cached = C.c_int64(), C.c_int64()
    check(lib().omni_gpu_memory_stats(C.byref(used), C.byref(cached)))
    return used.value, cached.value


def profile_enable(on: bool = True):
    check(lib().omni_gpu_profile_enable(1 if on else 0))


def profile_reset():
    check(lib().omni_gpu_profile_reset())


def profile_get(prefix: str):
    n, ms = C.c_int64(), C.c_double()
    check(lib().omni_gpu_profile_get(prefix.encode(), C.byref(n), C.byref(ms)))
    return n.value, ms.value


def synchronize():
    check(lib().omni_gpu_device_synchronize())
