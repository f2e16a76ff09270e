"""Trains the learned strategy/capacity advisor on the GPU and writes omnioperator_amd/models/advisor_mlp.npz.

    python -m omnioperator_amd.train_advisor [--runs 4096] [--epochs 400] [--out PATH] [--report PATH]

Each run generates a seeded key distribution on the device (rows log-uniform in [1e4, 4e6], distinct keys
log-uniform in [10, 1e8] capped at the rows, Zipf s uniform in [0, 1.5], int32 / int64 / (int32, int64) keys),
sketches it, and runs SUM + COUNT(*) GROUP BY key with every strategy to record the true group count and the
fastest strategy.  The MLP is then trained by the library's own HIP kernels (Adam, mini-batches of 128); 10% of the
runs are held out to report the prediction error.
"""
from __future__ import annotations

import argparse
import math
import time

import numpy as np

from . import capi, learned


def main(argv=None):
    p = argparse.ArgumentParser()
    p.add_argument("--runs", type=int, default=4096)
    p.add_argument("--epochs", type=int, default=400)
    p.add_argument("--seed", type=int, default=2024)
    p.add_argument("--out", default=learned.MODEL_PATH)
    p.add_argument("--report", default=None, help="write a text summary of the training here")
    args = p.parse_args(argv)
    capi.init(0)
    t0 = time.time()
    tasks = learned.sample_tasks(args.runs, seed=args.seed, min_rows=1e4, max_rows=4e6)
    X, t, c, groups = learned.collect(tasks, repeats=1)
    t1 = time.time()
    n = len(tasks)
    rng = np.random.default_rng(args.seed)
    perm = rng.permutation(n)
    held = perm[: n // 10]
    fit = perm[n // 10:]
    params, losses = learned.train(X[fit], t[fit], c[fit], seed=args.seed, epochs=args.epochs)
    t2 = time.time()
    adv = learned.Advisor(params)
    pred, cls = adv.predict(X[held])
    adv.close()
    err = np.abs(np.log2(pred + 1) - np.log2(groups[held] + 1))
    acc = float(np.mean(cls == c[held]))
    base = float(np.mean(c[held] == np.bincount(c[fit], minlength=3).argmax()))
    learned.save(params, args.out, losses=losses, held_median_log2_error=float(np.median(err)),
                 held_strategy_accuracy=acc)
    lines = [
        f"runs: {n} ({len(fit)} fitted, {len(held)} held out); collection {t1 - t0:.1f} s, training {t2 - t1:.1f} s "
        f"({args.epochs} epochs, batch 128, Adam lr 2e-3, on the GPU)",
        f"loss: first epoch {losses[0]:.4f}, last epoch {losses[-1]:.4f}",
        f"held-out |log2(predicted groups + 1) - log2(groups + 1)|: median {np.median(err):.3f}, "
        f"90th percentile {np.percentile(err, 90):.3f}",
        f"held-out strategy accuracy {acc:.3f} (majority class {base:.3f}); "
        f"class counts lds/global/radix = {np.bincount(c, minlength=3).tolist()}",
    ]
    text = "\n".join(lines)
    print(text, flush=True)
    if args.report:
        with open(args.report, "w") as f:
            f.write(text + "\n")


if __name__ == "__main__":
    main()
