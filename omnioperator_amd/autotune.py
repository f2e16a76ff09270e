"""Autotuner for the batched group-by: picks threads per workgroup, LDS table bytes, the LDS probe limit and the
phase-2 partition table bytes (for the one-workgroup-per-batch form: threads and its table bytes) by
timing candidate runs on the stream itself, per (key type, aggregate signature, cardinality bucket); winners
are cached in a JSON file (OMNI_GPU_AUTOTUNE_CACHE, default ~/.cache/omnioperator_amd/autotune.json) so later
runs of the same kind skip the sweep.

Tuning never changes results: every candidate is an exact configuration of the same kernels (a batch that
does not fit the LDS table is aggregated exactly in HBM).
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
