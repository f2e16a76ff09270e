"""GPU shuffle partitioner (omni_gpu_partition_ids through the C ABI) against the CPU oracle: partition ids equal
row for row."""
import numpy as np
import pytest

import oracle
from oracle import BOOLEAN, DATE32, DECIMAL64, DECIMAL128, DOUBLE, INT, LONG, SHORT, TIMESTAMP, Col
from omnioperator_amd import capi

pytestmark = pytest.mark.gpu


def gpu_ids(cols, partitions):
    n = len(cols[0])
    dev = [capi.DeviceColumn(c.type_id, c.values, c.nulls) for c in cols]
    return capi.partition_ids([d.c() for d in dev], n, partitions)


def random_col(rng, t, n, nulls):
    if t in (LONG, TIMESTAMP, DECIMAL64):
        v = rng.integers(-2**63, 2**63 - 1, n, dtype=np.int64)
    elif t in (INT, DATE32):
        v = rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)
    elif t == DOUBLE:
        v = rng.normal(0, 1e9, n)
        special = [0.0, -0.0, np.inf, -np.inf, np.nan][:n]
        v[:len(special)] = special
    else:
        v = (rng.random(n) < 0.5).astype(np.uint8)
    return Col(t, v, (rng.random(n) < 0.1) if nulls else None)


@pytest.mark.parametrize("t", [INT, DATE32, LONG, TIMESTAMP, DECIMAL64, DOUBLE, BOOLEAN])
@pytest.mark.parametrize("nulls", [False, True])
def test_single_column_parity(t, nulls):
    rng = np.random.default_rng(t * 10 + nulls)
    cols = [random_col(rng, t, 100_003, nulls)]
    for p in (1, 7, 200, 2**31 - 1):
        assert np.array_equal(gpu_ids(cols, p), oracle.partition_ids(cols, p))


@pytest.mark.parametrize("n", [1, 255, 1023, 1024, 1025, 300_001])
def test_multi_column_parity(n):
    rng = np.random.default_rng(n)
    cols = [random_col(rng, t, n, True) for t in (LONG, INT, BOOLEAN, DOUBLE)]
    assert np.array_equal(gpu_ids(cols, 200), oracle.partition_ids(cols, 200))


def test_all_null_row_keeps_the_seed():
    """A row whose columns are all NULL hashes to the seed 42: Pmod(42, p)."""
    cols = [Col(LONG, np.arange(10, dtype=np.int64), np.ones(10, bool)), Col(INT, np.arange(10, dtype=np.int32),
                                                                             np.ones(10, bool))]
    assert gpu_ids(cols, 5).tolist() == [42 % 5] * 10


def test_rejected_types_and_arguments():
    for t, dt in ((SHORT, np.int16), (DECIMAL128, np.uint64)):
        d = capi.DeviceColumn(t, np.zeros(8 if t == SHORT else (8, 2), dt))
        with pytest.raises(capi.GpuError):
            capi.partition_ids([d.c()], 8, 4)
    d = capi.DeviceColumn(LONG, np.zeros(8, np.int64))
    with pytest.raises(capi.GpuError):
        capi.partition_ids([d.c()], 8, 0)
    assert len(capi.partition_ids([d.c()], 0, 4)) == 0


def test_full_size_properties():
    """64M int64 keys into 200 partitions: every id in range, the partition sizes near uniform, and a slice equal
    to the oracle."""
    n = 64_000_000
    keys = np.random.default_rng(0xC2).integers(-2**63, 2**63 - 1, n, dtype=np.int64)
    ids = gpu_ids([Col(LONG, keys)], 200)
    assert ids.min() >= 0 and ids.max() < 200
    counts = np.bincount(ids, minlength=200)
    assert abs(counts - n / 200).max() < 6 * np.sqrt(n / 200)
    assert np.array_equal(ids[:1_000_000], oracle.partition_ids([Col(LONG, keys[:1_000_000])], 200))
