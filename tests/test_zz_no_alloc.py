"""omni_gpu_sketch_keys_noalloc, the sketch without device allocation, against omni_gpu_sketch_keys through the C ABI:
the same statistics and the same feature bits, on consecutive calls that share the library's fixed scratch array
(the second smaller than the first, so stale table slots of the first call would show), with the library's device
memory unchanged by the calls."""
import ctypes as C

import numpy as np
import pytest

from omnioperator_amd import capi

pytestmark = pytest.mark.gpu

FIELDS = ("rows", "sampled", "distinct", "f1", "f2", "max_count", "heavy", "nulls", "key_min", "key_max")


def as_tuple(s):
    return tuple(getattr(s, f) for f in FIELDS) + tuple(int(np.float32(x).view(np.int32)) for x in s.features)


def sketch_noalloc(keys, rows, is_join, out):
    karr = (capi.Column * len(keys))(*keys)
    capi.check(capi.lib().omni_gpu_sketch_keys_noalloc(karr, len(keys), rows, 1 if is_join else 0, None, out.ptr))
    capi.synchronize()
    host = capi.Sketch()
    capi.check(capi.lib().omni_gpu_memcpy_d2h(C.addressof(host), out.ptr, C.sizeof(capi.Sketch)))
    return host


def cases():
    rng = np.random.default_rng(0x5C)
    n = 2_000_000
    ranks = np.minimum(rng.zipf(1.2, n), 300_000).astype(np.int64)
    k64 = ranks * 2654435761 % (2**40) - 2**39
    k32 = (ranks % 977).astype(np.int32)
    big = [capi.DeviceColumn(capi.INT, k32, rng.random(n) < 0.02), capi.DeviceColumn(capi.LONG, k64)]
    small = [capi.DeviceColumn(capi.LONG, rng.integers(-50, 50, 3000, dtype=np.int64), rng.random(3000) < 0.1)]
    return [(big, n, False), (small, 3000, True), (small[:1], 0, False)]


def test_noalloc_sketch_equals_sketch_keys_on_consecutive_calls():
    capi.init(0)
    out = capi.DeviceArray(C.sizeof(capi.Sketch), np.uint8)
    runs = cases()
    want = [as_tuple(capi.sketch_keys([c.c() for c in cols], rows, join)) for cols, rows, join in runs]
    before = capi.memory_stats()
    got = [as_tuple(sketch_noalloc([c.c() for c in cols], rows, join, out)) for cols, rows, join in runs]
    assert capi.memory_stats() == before
    assert got[0][1] == 20_000 and got[1][1] == 3000 and got[2][1] == 0  # sample sizes: 1%, every row, none
    assert got == want
    # and again in the opposite order: a small call, then the large one, then the small one
    for i in (1, 0, 1):
        cols, rows, join = runs[i]
        assert as_tuple(sketch_noalloc([c.c() for c in cols], rows, join, out)) == want[i]


def test_noalloc_sketch_rejects_bad_arguments():
    capi.init(0)
    out = capi.DeviceArray(C.sizeof(capi.Sketch), np.uint8)
    col = capi.DeviceColumn(capi.LONG, np.arange(10, dtype=np.int64))
    karr = (capi.Column * 3)(col.c(), col.c(), col.c())
    assert capi.lib().omni_gpu_sketch_keys_noalloc(karr, 3, 10, 0, None, out.ptr) == -1  # at most two key columns
    assert capi.lib().omni_gpu_sketch_keys_noalloc(karr, 1, 10, 0, None, None) == -1     # no output
    assert capi.lib().omni_gpu_sketch_keys_noalloc(karr, 1, -1, 0, None, out.ptr) == -1  # negative rows
