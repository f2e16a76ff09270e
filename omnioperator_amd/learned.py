"""Learned strategy / capacity advisor for the GPU group-by.

A 1% sample sketch of the key columns (omni_gpu_sketch_keys) feeds a 16-64-64-4 fp32 MLP (omni_gpu_mlp_*, trained
and run by the library's own HIP kernels).  Output 0 predicts log2 of the group count, which sizes the hash table
once instead of growing it x4 several times; outputs 1..3 pick the strategy (LDS-local, global-atomic,
radix-partitioned).  The plain method is the fallback and the judge: a capacity that is too small only costs a
grow-and-rehash, and every strategy gives the same groups, so results never depend on the prediction.

Training data come from runs of the aggregation itself over seeded key distributions generated on the device
(omni_gpu_generate_keys): for each run the sketch features, the true group count and the fastest strategy.
"""
from __future__ import annotations

import math
import os
import time
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import capi

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PATH = os.path.join(HERE, "models", "advisor_mlp.npz")
STRATEGIES = (capi.STRATEGY_LDS, capi.STRATEGY_GLOBAL)   # classes 0 and 1; class 2 (radix) has no agg kernel yet
KIND_TYPES = {0: [capi.LONG], 1: [capi.INT], 2: [capi.INT, capi.LONG]}
T_SCALE = 32.0          # the regression target is log2(groups + 1) / T_SCALE
CAPACITY_MARGIN = 1.25  # capacity hint = predicted groups x margin


@dataclass
class Task:
    kind: int          # 0 int64 key, 1 int32 key, 2 (int32, int64) key
    rows: int
    distinct: float    # ranks of the Zipf law (an upper bound of the group count)
    zipf_s: float
    seed: int


def sample_tasks(n: int, seed: int, min_rows: float = 1e5, max_rows: float = 2e8, min_distinct: float = 10,
                 max_distinct: float = 1e8) -> List[Task]:
    """Tasks of the advisor's workload: rows and distinct counts log-uniform, Zipf s uniform in [0, 1.5],
    key width int32, int64 or (int32, int64); the distinct count is capped at the rows."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        rows = int(math.exp(rng.uniform(math.log(min_rows), math.log(max_rows))))
        d = min(float(rows), math.exp(rng.uniform(math.log(min_distinct), math.log(max_distinct))))
        out.append(Task(int(rng.integers(0, 3)), rows, max(1.0, math.floor(d)), float(rng.uniform(0.0, 1.5)),
                        int(rng.integers(1, 2**62))))
    return out


def _col(type_id: int, arr: capi.DeviceArray, offset: int = 0) -> capi.Column:
    c = capi.Column()
    c.type_id = type_id
    c.values = arr.ptr + offset * capi.WIDTH[type_id]
    c.nulls = None
    return c


class TaskData:
    """Device columns of one task (keys + an int64 value column)."""

    def __init__(self, task: Task):
        self.task = task
        self.keys, self.values = capi.generate_keys(task.kind, task.rows, task.distinct, task.zipf_s, task.seed)
        self.types = KIND_TYPES[task.kind]

    def key_cols(self, offset: int = 0):
        return [_col(t, k, offset) for t, k in zip(self.types, self.keys)]

    def sketch(self) -> capi.Sketch:
        return capi.sketch_keys(self.key_cols(), self.task.rows)


AGGS = [(capi.SUM, capi.LONG, capi.LONG), (capi.COUNT_ALL, capi.LONG, capi.LONG)]


def run_agg(data: TaskData, strategy: int, capacity_hint: int, batch_rows: int = 1 << 22, output: bool = False):
    """Runs SUM(value), COUNT(*) GROUP BY key over the task in VectorBatches of batch_rows rows; returns
    (groups, seconds[, outputs])."""
    spec = capi.make_agg_spec(data.types, AGGS, True, strategy, int(capacity_hint))
    capi.synchronize()
    t0 = time.perf_counter()
    h = capi.HashAgg(spec)
    n = data.task.rows
    for a in range(0, n, batch_rows):
        h.add_batch(data.key_cols(a), [_col(capi.LONG, data.values, a), None], min(batch_rows, n - a))
    g = h.num_groups()
    capi.synchronize()
    dt = time.perf_counter() - t0
    outs = h.output() if output else None
    h.close()
    return (g, dt, outs) if output else (g, dt)


def label_task(data: TaskData, repeats: int = 2) -> Tuple[int, int, List[float]]:
    """(true groups, fastest strategy class, seconds per class).  A first run measures the group count; each
    strategy is then timed with that exact capacity, so that the label reflects the kernel choice and not the
    growth schedule."""
    groups, _ = run_agg(data, capi.STRATEGY_GLOBAL, 0)
    times = []
    for s in STRATEGIES:
        best = float("inf")
        for _ in range(repeats):
            _, dt = run_agg(data, s, groups)
            best = min(best, dt)
        times.append(best)
    return groups, int(np.argmin(times)), times


def collect(tasks: Sequence[Task], repeats: int = 2):
    """Features, regression targets and class labels of runs over `tasks` (the training set)."""
    X = np.zeros((len(tasks), 16), np.float32)
    t = np.zeros(len(tasks), np.float32)
    c = np.zeros(len(tasks), np.int32)
    groups = np.zeros(len(tasks), np.int64)
    for i, task in enumerate(tasks):
        data = TaskData(task)
        X[i] = np.frombuffer(bytes(data.sketch().features), np.float32)
        g, cls, _ = label_task(data, repeats)
        groups[i] = g
        t[i] = math.log2(g + 1) / T_SCALE
        c[i] = cls
    return X, t, c, groups


def train(X, t, c, seed: int = 0, epochs: int = 300, batch: int = 128, lr: float = 2e-3):
    """Trains a fresh MLP on the GPU; returns (params, per-epoch losses)."""
    m = capi.Mlp(seed)
    losses = m.train(X, t, c, batch=batch, epochs=epochs, lr=lr, seed=seed)
    p = m.params()
    m.close()
    return p, losses


def save(params: np.ndarray, path: str = MODEL_PATH, **meta):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, params=params.astype(np.float32), **{k: np.asarray(v) for k, v in meta.items()})


def load(path: str = MODEL_PATH) -> Optional[np.ndarray]:
    if not os.path.exists(path):
        return None
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32)


class Advisor:
    """Predicts (capacity hint, strategy) for a group-by from its key columns."""

    def __init__(self, params: Optional[np.ndarray] = None):
        if params is None:
            params = load()
        if params is None:
            raise RuntimeError(f"no trained advisor at {MODEL_PATH}")
        self.mlp = capi.Mlp(0)
        self.mlp.set_params(params)

    def predict(self, features: np.ndarray):
        y = self.mlp.forward(np.asarray(features, np.float32).reshape(-1, 16))
        groups = np.exp2(np.clip(y[:, 0], 0.0, 1.0) * T_SCALE) - 1.0
        cls = np.argmax(y[:, 1:3], axis=1)  # the radix class has no aggregation kernel yet
        return groups, cls

    def plan(self, data: TaskData) -> Tuple[int, int]:
        y = self.mlp.advise(data.key_cols(), data.task.rows)
        groups = [2.0 ** (min(max(float(y[0]), 0.0), 1.0) * T_SCALE) - 1.0]
        cls = [int(np.argmax(y[1:3]))]  # the radix class has no aggregation kernel yet
        strategy = STRATEGIES[int(cls[0])]
        cap = int(min(float(data.task.rows), groups[0] * CAPACITY_MARGIN)) + 16
        return cap, strategy

    def close(self):
        self.mlp.close()
