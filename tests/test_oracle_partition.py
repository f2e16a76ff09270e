"""Pins the oracle's shuffle partitioner (HashUtil::ComputePartitionIds): the reference's Murmur3 known answers
(tests/golden/mm3_kat.json, from core/test/codegen/function_test.cpp), equality with the textbook byte-wise
Murmur3_x86_32 over the value's little-endian bytes, and the per-row fold over columns with NULLs and Pmod."""
import struct

import numpy as np
import pytest

import oracle
from oracle import BOOLEAN, DATE32, DECIMAL64, DOUBLE, INT, LONG, Col
from tests.helpers import load_golden

KAT = load_golden("mm3_kat.json")
M32 = 0xFFFFFFFF


def murmur3_x86_32(data: bytes, seed: int) -> int:
    """Textbook MurmurHash3_x86_32 (Austin Appleby), 4-byte little-endian blocks, tail, finalisation."""
    def rotl(x, r):
        return ((x << r) | (x >> (32 - r))) & M32
    c1, c2 = 0xCC9E2D51, 0x1B873593
    h = seed & M32
    n = len(data) // 4
    for i in range(n):
        k = int.from_bytes(data[4 * i:4 * i + 4], "little")
        k = (rotl((k * c1) & M32, 15) * c2) & M32
        h = (rotl(h ^ k, 13) * 5 + 0xE6546B64) & M32
    tail = data[4 * n:]
    k = 0
    for i, b in enumerate(tail):
        k |= b << (8 * i)
    if tail:
        k = (rotl((k * c1) & M32, 15) * c2) & M32
        h ^= k
    h ^= len(data)
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & M32
    h ^= h >> 13
    h = (h * 0xC2B2AE35) & M32
    h ^= h >> 16
    return h


def s32(x):
    return x - (1 << 32) if x >= 1 << 31 else x


def test_reference_known_answers():
    for v, want in KAT["int32"]:
        assert oracle.mm3_hash_int(v) == want
    for v, want in KAT["int64"]:
        assert oracle.mm3_hash_long(v) == want
    for v, want in KAT["double"]:
        assert oracle.mm3_hash_long(struct.unpack("<q", struct.pack("<d", v))[0]) == want


def test_matches_textbook_murmur3():
    rng = np.random.default_rng(1)
    for v in [0, 1, -1, 2**31 - 1, -2**31] + rng.integers(-2**31, 2**31, 200).tolist():
        for seed in (42, 0, 123456789):
            assert oracle.mm3_hash_int(v, seed) == s32(murmur3_x86_32(struct.pack("<i", v), seed))
    for v in [0, 1, -1, 2**63 - 1, -2**63] + rng.integers(-2**63, 2**63 - 1, 200, dtype=np.int64).tolist():
        assert oracle.mm3_hash_long(v) == s32(murmur3_x86_32(struct.pack("<q", v), 42))


def brute_partition_ids(cols, partitions):
    """Fold of the textbook hash over each row's non-NULL values, seed 42, then a floor modulo (equal to Pmod
    for a positive partition count)."""
    n = len(cols[0])
    out = []
    for r in range(n):
        h = 42
        for c in cols:
            if c.nulls is not None and c.nulls[r]:
                continue
            if c.type_id == BOOLEAN:
                data = struct.pack("<i", 1 if c.values[r] else 0)
            elif c.type_id in (INT, DATE32):
                data = struct.pack("<i", int(c.values[r]))
            elif c.type_id == DOUBLE:
                data = struct.pack("<d", float(c.values[r]))
            else:
                data = struct.pack("<q", int(c.values[r]))
            h = murmur3_x86_32(data, h)
        out.append(s32(h) % partitions)
    return out


@pytest.mark.parametrize("partitions", [1, 7, 200, 2**31 - 1])
def test_partition_ids_match_brute_force(partitions):
    rng = np.random.default_rng(partitions % 1000)
    n = 2000
    cols = [Col(LONG, rng.integers(-2**63, 2**63 - 1, n, dtype=np.int64), rng.random(n) < 0.1),
            Col(INT, rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32), rng.random(n) < 0.1),
            Col(DOUBLE, np.concatenate([[0.0, -0.0, np.inf, -np.inf, np.nan], rng.normal(0, 1e6, n - 5)])),
            Col(BOOLEAN, (rng.random(n) < 0.5).astype(np.uint8), rng.random(n) < 0.1),
            Col(DECIMAL64, rng.integers(-10**15, 10**15, n, dtype=np.int64))]
    got = oracle.partition_ids(cols, partitions)
    assert got.tolist() == brute_partition_ids(cols, partitions)
    assert got.min() >= 0 and got.max() < partitions
