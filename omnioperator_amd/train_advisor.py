"""Trains the learned strategy/capacity advisor on the GPU and writes omnioperator_amd/models/advisor_mlp.npz.

    python -m omnioperator_amd.train_advisor [--runs 4096] [--epochs 400] [--out PATH] [--report PATH]

Each run generates a seeded key distribution on the device (rows log-uniform in [1e4, 1.6e7], distinct keys
log-uniform in [10, 1e8] capped at the rows, Zipf s uniform in [0, 1.5], int32 / int64 / (int32, int64) keys),
sketches it, and runs either SUM + COUNT(*) GROUP BY key (LDS-local, global-atomic and radix-partitioned) or, for half of
the runs, an inner join of unique build keys with the rows as probe side (direct and partitioned probes), to
record the true group count (build rows for a join) and the fastest strategy.  The MLP is then trained by the library's own HIP kernels (Adam, mini-batches of 128); 10% of the
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
    p.add_argument("--join-share", type=float, default=0.5, help="share of join tasks among the runs")
    p.add_argument("--max-rows", type=float, default=1.6e7, help="largest rows of a run (log-uniform from 1e4)")
    p.add_argument("--scale-runs", type=int, default=512,
                   help="additional runs with rows log-uniform in [1e5, 2e8], the evaluation's scale")
    args = p.parse_args(argv)
    capi.init(0)
    t0 = time.time()
    tasks = learned.sample_tasks(args.runs, seed=args.seed, min_rows=1e4, max_rows=args.max_rows,
                                 join_share=args.join_share)
    if args.scale_runs:
        # runs at the evaluation's row scale (another seed than any evaluation set): the fastest strategy of a
        # large table is not the fastest of a small one, and the features alone cannot extrapolate that
        tasks += learned.sample_tasks(args.scale_runs, seed=args.seed + 1, min_rows=1e5, max_rows=2e8,
                                      join_share=args.join_share)
    X, t, c, groups = learned.collect(tasks, repeats=1, log=lambda m: print(m, flush=True))
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
    agg_held = X[held, 14] < 0.5
    err = np.abs(np.log2(pred + 1) - np.log2(groups[held] + 1))[agg_held]  # capacity is predicted for group-bys
    acc = float(np.mean(cls == c[held]))
    base = float(np.mean(c[held] == np.bincount(c[fit], minlength=3).argmax()))
    learned.save(params, args.out, losses=losses, held_median_log2_error=float(np.median(err)),
                 held_strategy_accuracy=acc)
    lines = [
        f"runs: {n} ({len(fit)} fitted, {len(held)} held out; {args.runs} with rows log-uniform [1e4, "
        f"{args.max_rows:.2g}] and {args.scale_runs} in [1e5, 2e8], {args.join_share:.0%} joins); "
        f"collection {t1 - t0:.1f} s, training {t2 - t1:.1f} s "
        f"({args.epochs} epochs, batch 128, Adam lr 2e-3, on the GPU)",
        f"loss: first epoch {losses[0]:.4f}, last epoch {losses[-1]:.4f}",
        f"held-out group-bys |log2(predicted groups + 1) - log2(groups + 1)|: median {np.median(err):.3f}, "
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
