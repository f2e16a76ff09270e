"""Pins the oracle's hash functions against the reference's own known-answer tests
(core/test/operator/hash_util_test.cpp) and against closed forms."""
import struct

import numpy as np
import pytest

import oracle
from tests.helpers import load_golden

KAT = load_golden("hash_util_kat.json")


def s64(x):
    return (x + 2**63) % 2**64 - 2**63


@pytest.mark.parametrize("vals,hashes,parts,poss", [
    ("values", "hashes", "partitions_mask7", "positions_mask7"),
    ("negative_values", "negative_hashes", "negative_partitions_mask7", "negative_positions_mask7"),
])
def test_hash_value_long_kat(vals, hashes, parts, poss):
    k = KAT["hash_value_long"]
    for v, h, p, q in zip(k[vals], k[hashes], k[parts], k[poss]):
        assert oracle.hash_value_i64(v) == h
        assert oracle.raw_hash_partition(h, 7) == p
        assert oracle.raw_hash_position(h, 7) == q


@pytest.mark.parametrize("vals,hashes,parts,poss", [
    ("values", "hashes", "partitions_mask7", "positions_mask7"),
    ("negative_values", "negative_hashes", "negative_partitions_mask7", "negative_positions_mask7"),
])
def test_hash_value_double_kat(vals, hashes, parts, poss):
    k = KAT["hash_value_double"]
    for v, h, p, q in zip(k[vals], k[hashes], k[parts], k[poss]):
        assert oracle.hash_value_double(v) == h
        assert oracle.raw_hash_partition(h, 7) == p
        assert oracle.raw_hash_position(h, 7) == q


def test_hash_value_decimal128_kat():
    k = KAT["hash_value_decimal128"]
    for (lo, hi), h in zip(k["values"], k["hashes"]):
        assert oracle.hash_value_decimal128(lo, hi) == h


def test_xxhash_varchar_kat():
    """HashUtil::HashValue(int8*, len) = XxHash64Hash(0, ...): the serialize-mode group key hash."""
    k = KAT["hash_value_varchar"]
    for s, h in zip(k["values"], k["hashes"]):
        assert oracle.xxhash64(s.encode()) == h


def test_crc_identity_and_decimal128_kat():
    k = KAT["crc_numeric"]
    for v in k["identity_values"]:
        assert oracle.join_hash_int(v) == v % 2**64
    for (hi, lo), h in zip(k["decimal128_high_low"], k["decimal128_hashes"]):
        assert oracle.crc_hash_decimal128(lo, hi) == h


def test_crc_varchar_kat():
    k = KAT["crc_varchar"]
    for s, h in zip(k["values"], k["hashes"]):
        assert oracle.crc32c_extend(0, s.encode()) == h


def test_crc_step_hardware_matches_bitwise_definition():
    """The SSE4.2 step used by HashUint64 equals the bitwise reflected-polynomial definition."""
    rng = np.random.default_rng(7)
    for v in list(rng.integers(-2**63, 2**63 - 1, 200, dtype=np.int64)) + [0, -1, 1, 2**63 - 1]:
        assert oracle.hash_uint64(int(v)) == oracle.crc32c_step_software(0xFFFFFFFF, int(v) % 2**64)


def test_group_hash_zero_extends_narrow_ints():
    """CRC32HasherForInt copies the key into a zeroed uint64 (crc_hasher.h:44-54)."""
    assert oracle.crc_hash_int(-1, oracle.INT) == oracle.hash_uint64(0xFFFFFFFF)
    assert oracle.crc_hash_int(-1, oracle.SHORT) == oracle.hash_uint64(0xFFFF)
    assert oracle.crc_hash_int(-1, oracle.LONG) == oracle.hash_uint64(2**64 - 1)


def test_serialize_format():
    """FixedLenTypeSerializer: [size][value truncated to size bytes]; NULL -> [0]; the size is the
    smallest power of two that keeps the sign (vector_marshaller.cpp:155-182)."""
    cols = [oracle.Col(oracle.INT, np.array([0, 127, 128, -129, 2**31 - 1], np.int32)),
            oracle.Col(oracle.LONG, np.array([5, -1, 2**40, 0, -2**63], np.int64),
                       np.array([False, False, False, True, False]))]
    assert oracle.serialize_row(cols, 0) == bytes([1, 0, 1, 5])
    assert oracle.serialize_row(cols, 1) == bytes([1, 127, 1, 0xFF])
    assert oracle.serialize_row(cols, 2) == bytes([2]) + struct.pack("<h", 128) + bytes([8]) + struct.pack("<q", 2**40)
    assert oracle.serialize_row(cols, 3) == bytes([2]) + struct.pack("<h", -129) + bytes([0])
    assert oracle.serialize_row(cols, 4) == bytes([4]) + struct.pack("<i", 2**31 - 1) + bytes([8]) + struct.pack("<q", -2**63)
