"""Seeded synthetic inputs of the benchmark configurations (host numpy arrays).

Every generator is deterministic for a given seed, so the GPU path and the CPU oracle see identical data.
  agg_plain      1M rows, int64 key uniform over 1,000 values, int64 value U[-1e6, 1e6]
  join_inner     build 2M unique int64 keys + int32 payload, probe 64M keys (80% hits), batches of 65,536
  agg_skewed     (int32 U[0,1e4), int64 Zipf(s) rank over 5e7) keys, decimal64 / double / int32 values
  batched_stream uneven batches (lognormal row counts) of independent group-bys
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """Vectorised splitmix64 finaliser: a bijection on 64-bit words (distinct inputs give distinct keys)."""
    z = np.asarray(x, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


@dataclass
class AggPlain:
    keys: np.ndarray     # int64
    values: np.ndarray   # int64
    batch_rows: int


def agg_plain(rows: int = 1_000_000, distinct: int = 1000, batch_rows: int = 4096, seed: int = 0xC1) -> AggPlain:
    rng = np.random.Generator(np.random.PCG64(seed))
    keys = rng.integers(0, distinct, rows, dtype=np.int64)
    values = rng.integers(-1_000_000, 1_000_001, rows, dtype=np.int64)
    return AggPlain(keys, values, batch_rows)


@dataclass
class JoinInner:
    build_keys: np.ndarray     # int64, unique
    build_payload: np.ndarray  # int32
    probe_keys: np.ndarray     # int64
    batch_rows: int
    hits: int                  # probe rows whose key is a build key


def join_inner(build_rows: int = 2_000_000, probe_rows: int = 64_000_000, hit_rate: float = 0.8,
               batch_rows: int = 65536, dense: bool = False, seed: int = 0xC2) -> JoinInner:
    """dense=False: build keys are distinct splitmix64 words (the reference would use its hash table);
    dense=True: build keys are a permutation of [0, build_rows) (the reference would use its array table).
    Misses come from a key set disjoint from the build keys."""
    rng = np.random.Generator(np.random.PCG64(seed))
    if dense:
        build = rng.permutation(build_rows).astype(np.int64)
    else:
        build = splitmix64(np.arange(build_rows, dtype=np.uint64)).view(np.int64)
    payload = rng.integers(0, 2**31, build_rows, dtype=np.int64).astype(np.int32)
    hit = rng.random(probe_rows) < hit_rate
    idx = rng.integers(0, build_rows, probe_rows, dtype=np.int64)
    if dense:
        miss = build_rows + rng.integers(0, 1 << 40, probe_rows, dtype=np.int64)
    else:
        miss = splitmix64(np.uint64(build_rows) + rng.integers(0, 1 << 40, probe_rows, dtype=np.int64).astype(np.uint64)).view(np.int64)
    probe = np.where(hit, build[idx], miss)
    return JoinInner(build, payload, probe, batch_rows, int(hit.sum()))


def zipf_ranks(rng: np.random.Generator, n: int, s: float, max_rank: int) -> np.ndarray:
    """Zipf(s) ranks in [1, max_rank]: numpy's unbounded Zipf sampler with rejection above max_rank
    (s > 1); s <= 1 falls back to inverse-CDF sampling over the explicit finite distribution."""
    if s > 1.0:
        out = np.empty(n, np.int64)
        filled = 0
        while filled < n:
            draw = rng.zipf(s, size=max(1024, int((n - filled) * 1.3)))
            draw = draw[draw <= max_rank]
            take = min(n - filled, draw.size)
            out[filled:filled + take] = draw[:take]
            filled += take
        return out
    w = 1.0 / np.arange(1, max_rank + 1, dtype=np.float64) ** s
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    return np.searchsorted(cdf, rng.random(n)).astype(np.int64) + 1


@dataclass
class AggSkewed:
    k1: np.ndarray       # int32
    k2: np.ndarray       # int64
    v_dec: np.ndarray    # int64 (decimal64, scale 2)
    v_dbl: np.ndarray    # float64
    v_i32: np.ndarray    # int32
    ranks: Optional[np.ndarray] = None  # int64 Zipf rank of k2 (k2 = splitmix64(rank)), when asked for


def agg_skewed(rows: int = 200_000_000, k1_distinct: int = 10_000, k2_ranks: int = 50_000_000, s: float = 1.1,
               seed: int = 0xC3, keep_ranks: bool = False) -> AggSkewed:
    rng = np.random.Generator(np.random.PCG64(seed))
    k1 = rng.integers(0, k1_distinct, rows, dtype=np.int64).astype(np.int32)
    ranks = zipf_ranks(rng, rows, s, k2_ranks)
    k2 = splitmix64(ranks.astype(np.uint64)).view(np.int64)  # fixed bijection of the rank
    v_dec = rng.integers(0, 1_000_000_001, rows, dtype=np.int64)
    v_dbl = rng.normal(100.0, 50.0, rows)
    v_i32 = rng.integers(-2**31, 2**31, rows, dtype=np.int64).astype(np.int32)
    return AggSkewed(k1, k2, v_dec, v_dbl, v_i32, ranks if keep_ranks else None)


@dataclass
class BatchedStream:
    row_counts: np.ndarray   # int64 per batch
    distinct: np.ndarray     # int64 per batch
    keys: np.ndarray         # int64, all batches concatenated
    values: np.ndarray       # int64


def batched_stream(batches: int = 20_000, median_rows: int = 2048, sigma: float = 1.4, min_rows: int = 64,
                   max_rows: int = 65536, min_distinct: int = 8, max_distinct: int = 1_000_000,
                   seed: int = 0xC4) -> BatchedStream:
    rng = np.random.Generator(np.random.PCG64(seed))
    rows = np.clip(np.round(np.exp(np.log(median_rows) + sigma * rng.standard_normal(batches))), min_rows,
                   max_rows).astype(np.int64)
    distinct = np.floor(np.exp(rng.uniform(np.log(min_distinct), np.log(max_distinct), batches))).astype(np.int64)
    total = int(rows.sum())
    keys = np.empty(total, np.int64)
    off = 0
    for b in range(batches):
        n, d = int(rows[b]), int(distinct[b])
        dict_ids = rng.integers(0, d, n, dtype=np.int64)
        keys[off:off + n] = splitmix64((np.uint64(b) << np.uint64(32)) + dict_ids.astype(np.uint64)).view(np.int64)
        off += n
    values = rng.integers(-1_000_000, 1_000_001, total, dtype=np.int64)
    return BatchedStream(rows, distinct, keys, values)


def batches(n: int, batch_rows: int) -> List[tuple]:
    """[(start, stop)] cutting n rows into VectorBatches of batch_rows rows (the last one shorter)."""
    return [(a, min(n, a + batch_rows)) for a in range(0, n, batch_rows)]
