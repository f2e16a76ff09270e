"""The advisor of the batched group-by stream: the learned half of its learned + batched form.

Each batch of a stream is sketched on the device (an evenly strided sample of at most 512 rows, the 16 features of
the column sketch), the advisor network (16-64-64-4 fp32, the library's own HIP kernels) predicts log2 of the
batch's group count, and a batch whose predicted groups exceed half the per-batch LDS table's slots skips that
table and starts in the hash-partitioned passes (omni_gpu_agg_batches_run, flag bit 3).  A batch the network sends
the wrong way still gets exact results: phase 1 abandons a batch that does not fit and phase 3 catches a partition
that overflows.

Training data are runs of the stream itself over seeded streams of randomized shape (rows lognormal with median in
[256, 8192] and sigma in [0.6, 1.6], clipped to [64, 65536]; distinct keys per batch log-uniform in [8, 1e6]; half
of the batches draw their keys by a Zipf law with s in [0, 1.2]): the per-batch features, the exact group count
and the phase that finished the batch when every batch starts in the LDS table.

    python -m omnioperator_amd.batch_advisor [--streams 12] [--batches 2000] [--epochs 300] [--report PATH]
"""
from __future__ import annotations

import argparse
import math
import os
import time
from typing import Optional, Tuple

import numpy as np

from . import capi, learned

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PATH = os.path.join(HERE, "models", "batch_advisor_mlp.npz")
SLOT_SHARE = 0.5   # a batch predicted to need more than this share of the phase-1 table's slots skips it


def splitmix64(x: np.ndarray) -> np.ndarray:
    z = x + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def training_stream(seed: int, batches: int):
    """(offsets, keys, values) of a seeded stream of randomized shape (see the module docstring)."""
    rng = np.random.default_rng(seed)
    median = math.exp(rng.uniform(math.log(256), math.log(8192)))
    sigma = rng.uniform(0.6, 1.6)
    rows = np.clip(np.round(np.exp(math.log(median) + sigma * rng.standard_normal(batches))), 64, 65536)
    rows = rows.astype(np.int64)
    distinct = np.floor(np.exp(rng.uniform(math.log(8), math.log(1e6), batches))).astype(np.int64)
    skew = np.where(rng.random(batches) < 0.5, rng.uniform(0.0, 1.2, batches), 0.0)
    offsets = np.concatenate([[0], np.cumsum(rows)]).astype(np.int64)
    keys = np.empty(int(offsets[-1]), np.int64)
    with np.errstate(over="ignore"):
        for b in range(batches):
            n, d, s = int(rows[b]), int(distinct[b]), float(skew[b])
            u = rng.random(n)
            if s < 1e-3:
                ids = np.floor(u * d).astype(np.int64)
            elif abs(s - 1.0) < 1e-3:
                ids = np.floor(np.exp(u * math.log(d + 1.0))).astype(np.int64) - 1
            else:
                e = 1.0 - s
                ids = np.floor((u * ((d + 1.0) ** e - 1.0) + 1.0) ** (1.0 / e)).astype(np.int64) - 1
            ids = np.clip(ids, 0, d - 1)
            keys[offsets[b]:offsets[b + 1]] = splitmix64((np.uint64(seed & 0xFFFF) << np.uint64(48)) +
                                                         (np.uint64(b) << np.uint64(24)) +
                                                         ids.astype(np.uint64)).view(np.int64)
    values = rng.integers(-1_000_000, 1_000_001, len(keys), dtype=np.int64)
    return offsets, keys, values


def label_stream(offsets: np.ndarray, keys: np.ndarray, values: np.ndarray, tiles=None):
    """Runs the stream with every batch starting in the LDS table; returns (features, exact groups, phases)."""
    kd, vd = capi.DeviceColumn(capi.LONG, keys), capi.DeviceColumn(capi.LONG, values)
    X = capi.batch_features(kd.c(), offsets)
    aggs = [(capi.SUM, capi.LONG, capi.LONG), (capi.COUNT_ALL, capi.LONG, capi.LONG)]
    s = capi.AggBatches(capi.make_agg_spec([capi.LONG], aggs), 65536)
    if tiles is not None:
        from . import autotune
        autotune.apply(s, tiles)
    groups = s.run(kd.c(), [vd.c(), None], offsets)
    phases = s.phase_of(len(offsets) - 1)
    s.close()
    return X, groups, phases


def collect(streams: int, batches: int, seed: int, log=None):
    Xs, gs, ps = [], [], []
    for i in range(streams):
        t0 = time.time()
        offsets, keys, values = training_stream(seed + i, batches)
        X, g, ph = label_stream(offsets, keys, values)
        Xs.append(X)
        gs.append(g)
        ps.append(ph)
        if log:
            log(f"stream {i}: {len(offsets) - 1} batches, {offsets[-1]} rows, {int(g.sum())} groups, phase-1 share "
                f"{float(np.mean(ph == 1)):.2f} ({time.time() - t0:.1f} s)")
    X = np.concatenate(Xs)
    groups = np.concatenate(gs)
    phases = np.concatenate(ps)
    t = (np.log2(groups + 1.0) / learned.T_SCALE).astype(np.float32)
    c = np.where(phases == 1, 0, 2).astype(np.int32)
    return X.astype(np.float32), t, c, groups


def load(path: str = MODEL_PATH) -> Optional[np.ndarray]:
    return learned.load(path)


def network(params: Optional[np.ndarray] = None) -> capi.Mlp:
    """An MLP handle holding the stream advisor's trained parameters (from MODEL_PATH unless given)."""
    p = params if params is not None else load()
    if p is None:
        raise FileNotFoundError(f"no trained stream advisor at {MODEL_PATH}")
    m = capi.Mlp(0)
    m.set_params(p)
    return m


def main(argv=None):
    p = argparse.ArgumentParser()
    p.add_argument("--streams", type=int, default=12)
    p.add_argument("--batches", type=int, default=2000)
    p.add_argument("--epochs", type=int, default=300)
    p.add_argument("--seed", type=int, default=0xBA7C4)
    p.add_argument("--out", default=MODEL_PATH)
    p.add_argument("--report", default=None)
    args = p.parse_args(argv)
    capi.init(0)
    t0 = time.time()
    X, t, c, groups = collect(args.streams, args.batches, args.seed, log=lambda m: print(m, flush=True))
    t1 = time.time()
    n = len(X)
    rng = np.random.default_rng(args.seed)
    perm = rng.permutation(n)
    held, fit = perm[: n // 10], perm[n // 10:]
    params, losses = learned.train(X[fit], t[fit], c[fit], seed=args.seed & 0xFFFF, epochs=args.epochs, batch=256)
    t2 = time.time()
    m = network(params)
    y = m.forward(X[held])
    m.close()
    pred = np.exp2(y[:, 0] * learned.T_SCALE) - 1.0
    err = np.abs(np.log2(pred + 1) - np.log2(groups[held] + 1))
    acc = float(np.mean(np.where(np.argmax(y[:, 1:], axis=1) == 0, 0, 2) == c[held]))
    learned.save(params, args.out, losses=losses, held_median_log2_error=float(np.median(err)),
                 held_phase_accuracy=acc)
    lines = [
        f"stream advisor: {args.streams} streams x {args.batches} batches = {n} batches ({len(fit)} fitted, "
        f"{len(held)} held out); collection {t1 - t0:.1f} s, training {t2 - t1:.1f} s ({args.epochs} epochs, "
        "batch 256, Adam lr 2e-3, on the GPU)",
        f"loss: first epoch {losses[0]:.4f}, last epoch {losses[-1]:.4f}",
        f"held-out |log2(predicted groups + 1) - log2(groups + 1)|: median {np.median(err):.3f}, 90th percentile "
        f"{np.percentile(err, 90):.3f}",
        f"held-out phase-1 outcome accuracy {acc:.3f} (share finished by phase 1: {float(np.mean(c == 0)):.3f})",
    ]
    text = "\n".join(lines)
    print(text, flush=True)
    if args.report:
        with open(args.report, "w") as f:
            f.write(text + "\n")


if __name__ == "__main__":
    main()
