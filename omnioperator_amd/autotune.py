"""Autotuner.  For the batched group-by it picks threads per workgroup, LDS table bytes, the LDS probe limit and
the phase-2 partition table bytes (for the one-workgroup-per-batch form: threads and its table bytes), per (key
type, aggregate signature, cardinality bucket); for the hash join the probe width (lookups in flight per lane of
the direct probe) and the radix partition count of the probe (the window probe's region size), per (key types,
build-rows bucket, probe-to-build ratio bucket).  Both time candidate runs of the workload itself; winners are
cached in a JSON file (OMNI_GPU_AUTOTUNE_CACHE, default ~/.cache/omnioperator_amd/autotune.json) so later runs of
the same kind skip the sweep.

Tuning never changes results: every candidate is an exact configuration of the same kernels (a batch that
does not fit the LDS table is aggregated exactly in HBM; every probe width and partition count finds the same
matches).
"""
from __future__ import annotations

import json
import math
import os
import time
from dataclasses import asdict, dataclass
from typing import Dict, Iterable, List, Optional

import numpy as np

CACHE = os.environ.get("OMNI_GPU_AUTOTUNE_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "omnioperator_amd", "autotune.json"))


@dataclass(frozen=True)
class StreamConfig:
    block: int = 256
    lds_bytes: int = 32768
    max_probe: int = 32
    part_lds_bytes: int = 65536  # phase-2 partition table


DEFAULT = StreamConfig()

CANDIDATES: List[StreamConfig] = [
    StreamConfig(b, l, p, q)
    for b in (256, 512)
    for l in (16384, 32768, 65536)
    for p in (8, 32)
    for q in (16384, 32768, 65536)
]


# the one-workgroup-per-batch form (flag bit 4) reads only the block size and the table bytes
FUSED_CANDIDATES: List[StreamConfig] = [
    StreamConfig(b, 32768, 32, q) for b in (256, 512) for q in (24576, 32768, 40960, 49152, 65536)
]


def apply(stream, cfg: StreamConfig):
    """Sets every knob of cfg on a capi.AggBatches stream."""
    stream.tune(block=cfg.block, lds_bytes=cfg.lds_bytes, max_probe=cfg.max_probe)
    stream.tune_partitions(part_lds_bytes=cfg.part_lds_bytes, part_rows=0)


def cardinality_bucket(mean_distinct: float) -> int:
    """log4 bucket of the mean per-batch distinct count."""
    return int(math.floor(math.log(max(1.0, mean_distinct), 4)))


def key_of(key_type: int, aggs: Iterable[tuple], bucket: int) -> str:
    sig = ",".join(f"{a[0]}:{a[1]}:{a[2]}" for a in aggs)
    return f"k{key_type}|{sig}|c{bucket}"


def _load() -> Dict[str, dict]:
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(d: Dict[str, dict]):
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        tmp = CACHE + ".tmp"
        with open(tmp, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, CACHE)
    except OSError:
        pass


def lookup(key: str) -> Optional[StreamConfig]:
    e = _load().get(key)
    return StreamConfig(**e["config"]) if e else None


def tune_stream(stream, run, key: str, candidates: List[StreamConfig] = CANDIDATES, repeats: int = 2,
                use_cache: bool = True) -> StreamConfig:
    """stream: capi.AggBatches; run(): runs the stream once (synchronizing).  Returns the fastest config and
    leaves the stream tuned to it."""
    if use_cache:
        hit = lookup(key)
        if hit is not None:
            apply(stream, hit)
            return hit
    best, best_t, results = None, float("inf"), []
    for cfg in candidates:
        apply(stream, cfg)
        run()  # warm: first launch of this configuration
        ts = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            run()
            ts.append(time.perf_counter() - t0)
        t = float(np.median(ts))
        results.append((asdict(cfg), t))
        if t < best_t:
            best, best_t = cfg, t
    apply(stream, best)
    d = _load()
    d[key] = {"config": asdict(best), "seconds": best_t, "swept": results}
    _store(d)
    return best


@dataclass(frozen=True)
class JoinConfig:
    probe_group: int = 1   # lookups in flight per lane of the direct probe (omni_gpu_join_tune)
    partitions: int = -1   # radix partitions of the probe: -1 the library's rule, 0 none, else P


# the direct probe at every width, the library's rule, and the partition counts whose regions fit the LDS window
# of tables from 2^20 to 2^26 slots
JOIN_CANDIDATES: List[JoinConfig] = (
    [JoinConfig(g, 0) for g in (1, 2, 4, 8, 16)] + [JoinConfig(1, -1)] +
    [JoinConfig(1, p) for p in (256, 512, 1024, 2048, 4096)]
)


def join_key_of(key_types: Iterable[int], build_rows: int, probe_rows: int) -> str:
    """Cache key of a join: its key types, build rows (power-of-two bucket) and probe / build ratio bucket."""
    b = int(math.log2(max(1, build_rows)))
    r = int(math.log2(max(1, probe_rows // max(1, build_rows)) + 1))
    return "join|" + ",".join(str(t) for t in key_types) + f"|b{b}|r{r}"


def tune_join(run, key: str, candidates: List[JoinConfig] = JOIN_CANDIDATES, repeats: int = 2,
              use_cache: bool = True) -> JoinConfig:
    """run(cfg): one build + probe of the workload with the probe tuned to cfg (synchronizing); raising for a
    candidate the table cannot take (e.g. a partition count with regions smaller than 64 slots) skips it.  Returns
    the fastest configuration."""
    if use_cache:
        e = _load().get(key)
        if e:
            return JoinConfig(**e["config"])
    best, best_t, results = None, float("inf"), []
    for cfg in candidates:
        try:
            run(cfg)  # warm
            ts = []
            for _ in range(repeats):
                t0 = time.perf_counter()
                run(cfg)
                ts.append(time.perf_counter() - t0)
        except Exception as e:  # a candidate the library rejects for this table
            results.append((asdict(cfg), str(e)))
            continue
        t = float(np.median(ts))
        results.append((asdict(cfg), t))
        if t < best_t:
            best, best_t = cfg, t
    if best is None:
        raise RuntimeError("no join candidate ran")
    d = _load()
    d[key] = {"config": asdict(best), "seconds": best_t, "swept": results}
    _store(d)
    return best

