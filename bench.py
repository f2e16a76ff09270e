#!/usr/bin/env python3
"""Benchmark of the MI355X hash-aggregation / hash-join hot path.

Default workload (the configuration the headline metric, "learned+batched vs plain GPU agg/join", is quoted on):
the batched stream of 20,000 independent group-bys -- VectorBatches of lognormal row counts (median 2048, clipped
to [64, 65536], 102,826,994 rows), int64 keys with a per-batch distinct count log-uniform in [8, 1e6], SUM(int64) +
COUNT(*) -- in its learned + batched form: a per-batch sketch through the stream advisor network picks each batch's
first phase, the autotuner picks the tiles, and the whole run is one hipGraph launch.  One step = one run over all
20,000 batches with every group written to device-resident output columns.  Inputs are resident in HBM before
timing starts.  The same process times the plain form (one launch per batch and phase) and the fixed-strategy
batched forms for the speed-ups in `config`.

Prints one JSON line with rows/s, the run's roofline figure (algorithmic bytes / the graph launch's duration from
HIP events on its stream) and a CPU baseline: the oracle (the C++ restatement of the reference operator, one
operator per batch) over the same stream on all host cores and on one core.  --workload selects the other
configurations (join, agg, skewed, learned, partition), each a line of the same form.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "rows/s and % HBM roofline, learned+batched vs plain GPU agg/join"


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


# --join-type: (JoinType, BuildSide) values of the reference (core/src/plannode/planNode.h:406-414, 28-32)
JOIN_TYPES = {"inner": (0, 0), "left": (1, 0), "right_build_right": (2, 2), "full": (3, 0), "semi": (4, 0),
              "anti": (5, 0)}
JOIN_NAMES = {"inner": "inner", "left": "left outer", "right_build_right": "right outer (build right)",
              "full": "full outer", "semi": "left semi", "anti": "left anti"}


def bench_join(args):
    from omnioperator_amd import capi, workloads

    capi.init(0)
    w = workloads.join_inner(build_rows=args.build_rows, probe_rows=args.probe_rows, dense=args.dense)
    nb, npr = len(w.build_keys), len(w.probe_keys)
    # device-resident inputs, cut into VectorBatches; the key as one int64 column, or as the key columns of
    # --join-keys (each a function of the int64 key, so the same rows match)
    ktypes, bkcols, pkcols = join_key_columns(capi, w, args.join_keys)
    bkey = [capi.DeviceColumn(t, v) for t, v in zip(ktypes, bkcols)]
    bpay = capi.DeviceColumn(capi.INT, w.build_payload)
    probe = [capi.DeviceColumn(t, v) for t, v in zip(ktypes, pkcols)]
    bspans = workloads.batches(nb, w.batch_rows)
    pspans = workloads.batches(npr, w.batch_rows)
    pcols, prows = capi.HashJoin.probe_arrays([k.c(a) for a, _ in pspans for k in probe], [b - a for a, b in pspans])
    jt, bs = JOIN_TYPES[args.join_type]
    cap = npr
    out_rows = capi.DeviceArray(cap * 8, np.int64, (cap,))
    out_pay = capi.DeviceArray(cap * 4, np.int32, (cap,))
    # outer joins write NULL payloads for unmatched probe rows: their payload output carries a bitmap
    with_nulls = args.join_type in ("left", "full")
    out_nulls = capi.DeviceArray(4 * ((cap + 31) // 32), np.uint8) if with_nulls else None
    oarr = (capi.OutColumn * 1)()
    oarr[0].type_id, oarr[0].values = capi.INT, out_pay.ptr
    oarr[0].nulls = out_nulls.ptr if with_nulls else None
    keeps_build = args.join_type in ("right_build_right", "full")
    outer_rows = capi.DeviceArray(nb * 8, np.int64, (nb,)) if keeps_build else None
    outer_pay = capi.DeviceArray(nb * 4, np.int32, (nb,)) if keeps_build else None
    oouter = (capi.OutColumn * 1)()
    if keeps_build:
        oouter[0].type_id, oouter[0].values, oouter[0].nulls = capi.INT, outer_pay.ptr, None

    # per-batch column descriptors prepared once (the C++ operator path has no such host cost)
    bcols = [((capi.Column * len(bkey))(*[k.c(a) for k in bkey]), (capi.Column * 1)(bpay.c(a)), b - a)
             for a, b in bspans]

    tuned = {"probe_group": args.probe_group, "partitions": args.partitions}

    def step():
        j = capi.HashJoin(ktypes if len(ktypes) > 1 else ktypes[0], [capi.INT])
        j.tune(tuned["probe_group"], tuned["partitions"])
        if jt != capi.JOIN_INNER:
            j.set_type(jt, bs)
        for kc, pc, n in bcols:
            j.add_build(kc, pc, n)
        j.finish()
        m = j.probe_batches_raw(pcols, prows, out_rows, oarr, cap)
        if keeps_build:  # the outer pass over unmatched build rows belongs to the job
            outer[0] = j.outer_output_raw(outer_rows, oouter, nb)
        parts[0] = j.last_partitions()
        is_array[0] = j.info()["is_array"]
        j.close()
        return m

    parts, is_array, outer = [0], [False], [0]

    if args.join_autotune:
        # probe width and partition count from the autotuner (cached per key types and size buckets)
        from omnioperator_amd import autotune

        def run(cfg):
            tuned.update(probe_group=cfg.probe_group, partitions=cfg.partitions)
            step()
            capi.synchronize()

        best = autotune.tune_join(run, autotune.join_key_of(ktypes, nb, npr))
        tuned.update(probe_group=best.probe_group, partitions=best.partitions)

    for _ in range(args.warmup):
        m = step()
    capi.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        m = step()
    capi.synchronize()
    t1 = time.perf_counter()
    # stage durations from HIP events in separate steps: the events around each stage's launches hold the
    # stream for a few microseconds each, so the timed steps above run without them
    prof_steps = max(1, min(args.steps, 5))
    capi.profile_reset()
    capi.profile_enable(True)
    for _ in range(prof_steps):
        step()
    capi.synchronize()
    capi.profile_enable(False)
    launches, probe_ms = capi.profile_get("join_probe")
    _, build_ms = capi.profile_get("join_build")
    _, tbuild_ms = capi.profile_get("join_translate_build")
    _, tprobe_ms = capi.profile_get("join_translate_probe")
    _, outer_ms = capi.profile_get("join_outer")
    expected = {"inner": w.hits, "left": npr, "right_build_right": w.hits, "full": npr, "semi": w.hits,
                "anti": npr - w.hits}[args.join_type]
    assert m == expected, f"output rows {m} != expected {expected}"
    if keeps_build:  # unique build keys: the build rows no probe row hit
        assert outer[0] == nb - len(np.unique(w.probe_keys[np.isin(w.probe_keys, w.build_keys)])), "outer rows"
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    rows = nb + npr
    value = rows / (ms_per_step / 1e3)
    # the probe stage of one step: the probe kernel, or the partition chain (histogram, scan, scatter,
    # persistent probe) when the table is radix-partitioned
    avg_probe_ms = probe_ms / prof_steps
    # 8 B per probe key; per output row its int64 probe row id, plus the int32 payload (and its NULL bit) for
    # the types that output build columns
    out_row_bytes = 8 if args.join_type in ("semi", "anti") else 12
    probe_bytes = npr * 8 + m * out_row_bytes + (m + 7) // 8 * with_nulls
    achieved = probe_bytes / (avg_probe_ms / 1e3) / 1e9
    # the profiled configuration is the default one: inner join, sparse keys, automatic partitioning
    # (1024 partitions)
    profiled = (not args.dense and parts[0] == 1024 and (nb, npr) == (2_000_000, 64_000_000)
                and args.join_type == "inner")
    traffic = profiled_traffic("probe stage (chain)") if profiled else None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "join_probe" if parts[0] <= 1 else
                          f"join_probe_* (radix-partitioned, {parts[0]} partitions: hist+scan+scatter+probe)",
                "kernel_ms": round(avg_probe_ms, 4), "algorithmic_bytes": probe_bytes}
    extra = {"build_kernel_ms": round(build_ms / prof_steps, 4), "matches": m, "probe_group": tuned["probe_group"],
             "partitions_setting": tuned["partitions"], "autotuned": bool(args.join_autotune),
             "probe_partitions": parts[0], "join_type": args.join_type}
    if keeps_build:
        extra["outer_rows"] = outer[0]
        extra["outer_kernel_ms"] = round(outer_ms / prof_steps, 4)
    if parts[0] > 1:
        extra["probe_stage_ms"] = {k: round(capi.profile_get("join_probe_" + k)[1] / prof_steps, 4)
                                   for k in ("hist", "scan", "scatter", "window", "part")}
    if args.join_keys != "single":
        # key translation: per probe row its key columns read, the int64 join key and its NULL bit written, and
        # for a dictionary mode one dictionary slot (state word + key words) read
        kbytes = sum(capi.WIDTH[t] for t in ktypes)
        dict_bytes = 0 if args.join_keys == "int_int" else 4 + kbytes_words(ktypes) * 8
        tbytes = npr * (kbytes + 8 + dict_bytes) + npr // 8
        tms = tprobe_ms / prof_steps
        extra["key_translation"] = {
            "key_types": [TYPE_NAMES[t] for t in ktypes],
            "build_translate_ms": round(tbuild_ms / prof_steps, 4), "probe_translate_ms": round(tms, 4),
            "probe_translate_algorithmic_bytes": tbytes,
            "probe_translate_gbs": round(tbytes / (tms / 1e3) / 1e9, 1) if tms > 0 else None}
    cpu = cpu_baseline_join(w, args, jt, bs, ktypes, bkcols, pkcols) if not args.no_cpu else None
    name = JOIN_NAMES[args.join_type]
    keys_desc = {"single": "int64 keys", "int_int": "(INT, INT) keys", "long_long": "(LONG, LONG) keys",
                 "decimal128": "DECIMAL128 keys"}[args.join_keys]
    cfg = {"workload": f"{name} hash join: build 2M unique {keys_desc} + int32 payload, probe 64M {keys_desc} "
                       "(80% hit), VectorBatches of 65536" if (nb, npr) == (2_000_000, 64_000_000) else
                       f"{name} hash join on {keys_desc}: build {nb} + probe {npr} rows",
           "build_rows": nb, "probe_rows": npr, "batch_rows": w.batch_rows,
           "keys": "dense permutation" if args.dense else "sparse 64-bit (splitmix64)",
           "output": "probe rowid int64" if args.join_type in ("semi", "anti") else
                     "probe rowid int64 + build payload int32" + (" (NULL when unmatched)" if with_nulls else "")}
    cfg.update(extra)
    return {"value": value, "ms_per_step": ms_per_step, "config": cfg, "roofline": roofline,
            "cpu_baseline": cpu, "dtype": "int64"}


def timed_steps(step, args, capi):
    """Warm-up, then args.steps timed steps, then args.steps more with the per-launch HIP events on (the events
    hold the stream for a few microseconds each, so the timed steps run without them); returns the last step's
    result and the timed milliseconds per step."""
    for _ in range(args.warmup):
        step()
    capi.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = step()
    capi.synchronize()
    t1 = time.perf_counter()
    capi.profile_reset()
    capi.profile_enable(True)
    for _ in range(args.steps):
        step()
    capi.synchronize()
    capi.profile_enable(False)
    return r, (t1 - t0) * 1e3 / args.steps


def kernel_roofline(capi, prefix: str, algorithmic_bytes: float, launches_per_step: float, args):
    n, ms = capi.profile_get(prefix)
    avg = ms / max(1, n)
    per_launch = algorithmic_bytes / max(1.0, launches_per_step)
    achieved = per_launch / (avg / 1e3) / 1e9 if avg > 0 else 0.0
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": prefix,
            "kernel_ms": round(avg, 4), "algorithmic_bytes": int(per_launch)}


def bench_partition(args):
    """Shuffle partition ids (HashUtil::ComputePartitionIds) of the C2 probe side's 64M int64 keys into 200
    partitions (Spark's default shuffle partition count), one launch per step."""
    from omnioperator_amd import capi

    capi.init(0)
    n = args.probe_rows
    keys = np.random.default_rng(0xC2).integers(-2**63, 2**63 - 1, n, dtype=np.int64)
    col = capi.DeviceColumn(capi.LONG, keys)
    cols = (capi.Column * 1)(col.c())
    out = capi.DeviceArray(n * 4, np.int32, (n,))

    def step():
        capi.partition_ids_raw(cols, n, args.partitions_out, out)

    _, ms = timed_steps(step, args, capi)
    ids = out.numpy()
    assert ids.min() >= 0 and ids.max() < args.partitions_out
    nbytes = n * (8 + 4)  # the key in, the int32 id out
    roof = kernel_roofline(capi, "partition_ids", nbytes, 1, args)
    if (n, args.partitions_out) == (64_000_000, 200):  # the profiled configuration
        roof["traffic"] = profiled_traffic("partition_ids", "partition_ids.md")
    cpu = None
    if not args.no_cpu:
        import oracle
        sample = min(n, args.cpu_probe_rows)
        t0 = time.perf_counter()
        oracle.partition_ids([oracle.Col(oracle.LONG, keys[:sample])], args.partitions_out)
        t1 = time.perf_counter()
        cpu = {"value": sample / (t1 - t0), "unit": "rows/s", "cores": 1, "kind": "port",
               "sample": f"first {sample} keys; host CPU {cpu_model()}"}
    cfg = {"workload": f"shuffle partition ids: {n} int64 keys (uniform 64-bit) into {args.partitions_out} "
                       "partitions, Murmur3 seed 42 + pmod",
           "rows": n, "partitions": args.partitions_out, "output": "partition id int32 per row"}
    return {"value": n / (ms / 1e3), "ms_per_step": ms, "config": cfg, "roofline": roof, "cpu_baseline": cpu,
            "dtype": "int64"}


def bench_agg_plain(args):
    """1M rows, int64 key over 1,000 values, VectorBatches of 4096; SUM/COUNT/MIN/MAX."""
    from omnioperator_amd import capi, workloads
    from omnioperator_amd.capi import COUNT_ALL, LONG, MAX, MIN, SUM
    capi.init(0)
    w = workloads.agg_plain()
    n = len(w.keys)
    kd, vd = capi.DeviceColumn(LONG, w.keys), capi.DeviceColumn(LONG, w.values)
    aggs = [(SUM, LONG, LONG), (COUNT_ALL, LONG, LONG), (MIN, LONG, LONG), (MAX, LONG, LONG)]
    spans = workloads.batches(n, w.batch_rows)
    cols = [([kd.c(a)], [vd.c(a), None, vd.c(a), vd.c(a)], b - a) for a, b in spans]

    rule = capi.HASH_TABLE_ARRAY if args.agg_rule == "array" else capi.HASH_TABLE_NORMAL
    spec = capi.make_agg_spec([LONG], aggs, hash_table_rule=rule)
    per_batch = args.agg_form == "per_batch"
    arrays = capi.HashAgg.batch_arrays_for(spec, cols)

    outs = {}

    def step():
        h = capi.HashAgg(spec)
        if per_batch:
            for kc, ac, m in cols:
                h.add_batch(kc, ac, m)
        else:  # all VectorBatches in one omni_gpu_agg_add_batches call: one staging launch + one insert
            h.add_batches_raw(*arrays)
        g = h.num_groups()
        if "bufs" not in outs:  # device-resident output, as the other workloads write theirs
            outs["bufs"] = h.alloc_output(g)
        h.output_into(outs["bufs"], g)
        h.close()
        return g

    groups, ms = timed_steps(step, args, capi)
    alg = n * 16 + groups * 40
    roof = kernel_roofline(capi, "agg_insert", n * 16, len(spans) if per_batch else 1, args)
    cpu = None
    if not args.no_cpu:
        import oracle
        spec = oracle.make_agg_spec([LONG], aggs)
        t0 = time.perf_counter()
        oracle.hash_agg(spec, [oracle.Col(LONG, w.keys)], [oracle.Col(LONG, w.values), None, oracle.Col(LONG, w.values),
                                                            oracle.Col(LONG, w.values)], w.batch_rows)
        t1 = time.perf_counter()
        cpu = {"value": n / (t1 - t0), "unit": "rows/s", "cores": 1, "kind": "port",
               "sample": f"the full workload ({n} rows); host CPU {cpu_model()}"}
    cfg = {"workload": "group-by: 1M rows, int64 key over 1,000 values, int64 value, VectorBatches of 4096; "
                       "SUM/COUNT(*)/MIN/MAX", "rows": n, "groups": groups, "batch_rows": w.batch_rows,
           "form": "one omni_gpu_agg_add_batch per VectorBatch" if per_batch else
                   "all VectorBatches in one omni_gpu_agg_add_batches call",
           "hash_table_rule": args.agg_rule,
           "output": "every group written to device-resident output columns",
           "algorithmic_bytes_per_step": alg}
    return {"value": n / (ms / 1e3), "ms_per_step": ms, "config": cfg, "roofline": roof, "cpu_baseline": cpu,
            "dtype": "int64"}


def bench_agg_skewed(args):
    """200M rows, key (int32 U[1e4], int64 Zipf 1.1 over 5e7); SUM(dec64->dec128), SUM/AVG(double), COUNT, MAX(int32)."""
    from omnioperator_amd import capi, workloads
    from omnioperator_amd.capi import AVG, COUNT_ALL, DECIMAL64, DECIMAL128, DOUBLE, INT, LONG, MAX, SUM
    capi.init(0)
    w = workloads.agg_skewed(rows=args.skewed_rows)
    n = len(w.k1)
    k1, k2 = capi.DeviceColumn(INT, w.k1), capi.DeviceColumn(LONG, w.k2)
    vd, vf, vi = capi.DeviceColumn(DECIMAL64, w.v_dec), capi.DeviceColumn(DOUBLE, w.v_dbl), capi.DeviceColumn(INT, w.v_i32)
    aggs = [(SUM, DECIMAL64, DECIMAL128, 2, 2), (SUM, DOUBLE, DOUBLE), (AVG, DOUBLE, DOUBLE), (COUNT_ALL, LONG, LONG),
            (MAX, INT, INT)]
    batch = 1 << 20
    spans = workloads.batches(n, batch)
    cols = [([k1.c(a), k2.c(a)], [vd.c(a), vf.c(a), vf.c(a), None, vi.c(a)], b - a) for a, b in spans]
    outs = {}

    strategy = {"radix": capi.STRATEGY_RADIX, "auto": capi.STRATEGY_AUTO, "global": capi.STRATEGY_GLOBAL,
                "lds": capi.STRATEGY_LDS}[args.agg_strategy]

    def step():
        h = capi.HashAgg(capi.make_agg_spec([INT, LONG], aggs, strategy=strategy))
        for kc, ac, m in cols:
            h.add_batch(kc, ac, m)
        g = h.num_groups()
        if "bufs" not in outs:
            outs["bufs"] = h.alloc_output(g)
        h.output_into(outs["bufs"], g)
        h.close()
        return g

    groups, ms = timed_steps(step, args, capi)
    in_bytes = n * (4 + 8 + 8 + 8 + 4)
    out_bytes = groups * (4 + 8 + 16 + 8 + 8 + 8 + 4)
    if args.agg_strategy == "radix":
        # the aggregation proper is the radix pipeline: its input rows read once and each group written once (as
        # many bytes as its output row) are the algorithmic bytes; the output pass is timed apart
        stages = {k: capi.profile_get("agg_radix_" + k) for k in ("hist", "scan", "scatter", "pieces", "final",
                                                                   "gather")}
        pipe_ms = sum(v[1] for v in stages.values()) / args.steps
        state_bytes = out_bytes
        achieved = (in_bytes + state_bytes) / (pipe_ms / 1e3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": profiled_traffic("radix pipeline (all kernels)", "agg_radix_c3.md"),
                "kernel": "radix pipeline: agg_radix_hist + scan + scatter (two levels) + agg_radix_final",
                "kernel_ms": round(pipe_ms, 4), "algorithmic_bytes": in_bytes + state_bytes,
                "stage_ms": {k: round(v[1] / args.steps, 4) for k, v in stages.items()},
                "output_ms": round((capi.profile_get("agg_output")[1] + capi.profile_get("agg_compact")[1]) /
                                   args.steps, 4)}
    else:
        roof = kernel_roofline(capi, "agg_insert_multi", in_bytes, len(spans), args)
    cpu = None
    if not args.no_cpu:
        import oracle
        sample = min(n, args.cpu_skewed_rows)
        spec = oracle.make_agg_spec([INT, LONG], aggs)
        t0 = time.perf_counter()
        oracle.hash_agg(spec, [oracle.Col(INT, w.k1[:sample]), oracle.Col(LONG, w.k2[:sample])],
                        [oracle.Col(DECIMAL64, w.v_dec[:sample], None, 2), oracle.Col(DOUBLE, w.v_dbl[:sample]),
                         oracle.Col(DOUBLE, w.v_dbl[:sample]), None, oracle.Col(INT, w.v_i32[:sample])], batch)
        t1 = time.perf_counter()
        cpu = {"value": sample / (t1 - t0), "unit": "rows/s", "cores": 1, "kind": "port",
               "sample": f"first {sample} rows of the workload; host CPU {cpu_model()}"}
    cfg = {"workload": f"skewed group-by: {n} rows, key (int32 U[0,1e4), int64 Zipf(1.1) over 5e7), VectorBatches of "
                       f"{batch}; SUM(decimal64->decimal128), SUM/AVG(double), COUNT(*), MAX(int32); output of "
                       "every group included", "rows": n, "groups": groups,
           "algorithmic_bytes_per_step": in_bytes + out_bytes, "strategy": args.agg_strategy}
    return {"value": n / (ms / 1e3), "ms_per_step": ms, "config": cfg, "roofline": roof, "cpu_baseline": cpu,
            "dtype": "mixed: int32/int64 keys, decimal64->decimal128, f64, int32"}


def bench_batched(args):
    """The headline configuration: 20,000 independent group-bys (lognormal rows, median 2048, clipped to
    [64, 65536]; per-batch distinct keys log-uniform in [8, 1e6]), SUM(int64) + COUNT(*), one run over the stream.
    Default form "learned": the learned + batched form -- the stream advisor (a per-batch sketch through the MLP,
    on the device) predicts each batch's group count, which decides direct LDS aggregation or hash partitions
    (and their number) inside the one-workgroup-per-batch kernel; tiles from the autotuner; the whole run one
    hipGraph launch.  It is compared in the same process with the plain form (one launch per batch and phase, the
    reference's one operator per VectorBatch), with the same kernel without the advisor (its own distinct
    estimate) and with the earlier phased batched forms (every batch LDS-first, every batch partitioned)."""
    from omnioperator_amd import autotune, batch_advisor, capi, workloads
    capi.init(0)
    w = workloads.batched_stream()
    offsets = np.concatenate([[0], np.cumsum(w.row_counts)]).astype(np.int64)
    nb = len(w.row_counts)
    n = int(offsets[-1])
    kd, vd = capi.DeviceColumn(capi.LONG, w.keys), capi.DeviceColumn(capi.LONG, w.values)
    aggs = [(capi.SUM, capi.LONG, capi.LONG), (capi.COUNT_ALL, capi.LONG, capi.LONG)]
    spec = capi.make_agg_spec([capi.LONG], aggs)
    s = capi.AggBatches(spec, 65536)
    F = capi.AggBatches
    forms = {"learned": F.FLAG_GRAPH | F.FLAG_ADVISOR | F.FLAG_FUSED, "fused": F.FLAG_GRAPH | F.FLAG_FUSED,
             "learned_phases": F.FLAG_GRAPH | F.FLAG_ADVISOR, "graph": F.FLAG_GRAPH, "two_launches": 0,
             "per_batch": F.FLAG_PER_BATCH}
    flags = forms[args.batched_form]
    net = None
    if flags & F.FLAG_ADVISOR:
        net = batch_advisor.network()
        s.set_advisor(net)
    all_part = np.full(nb, 2, np.int32)

    def run(fl, modes=None):
        return s.run(kd.c(), [vd.c(), None], offsets, modes, fl)

    def step():
        return int(run(flags).sum())

    def tune(fl, form):
        key = autotune.key_of(capi.LONG, aggs, autotune.cardinality_bucket(float(np.mean(w.distinct)))) + f"|{form}"
        cands = autotune.FUSED_CANDIDATES if fl & F.FLAG_FUSED else autotune.CANDIDATES
        return autotune.tune_stream(s, lambda: (run(fl), capi.synchronize()), key, cands, use_cache=False)

    tuned = None
    if args.tiles:  # fixed tiles, e.g. for a profile of one configuration
        b, l, p_, q = (int(x) for x in args.tiles.split(","))
        tuned = autotune.StreamConfig(b, l, p_, q)
        autotune.apply(s, tuned)
    elif args.batched_form != "per_batch":  # the plain one-launch-per-batch form keeps the default tiles
        tuned = tune(flags, args.batched_form)

    groups, ms = timed_steps(step, args, capi)
    prof = "agg_batches_graph" if flags & F.FLAG_GRAPH else \
        ("agg_batches_per_batch" if flags & F.FLAG_PER_BATCH else "agg_batches")
    n_launch, gpu_ms = capi.profile_get(prof)
    run_ms = gpu_ms / max(1, n_launch)
    ph = s.phases()
    assert groups == 41_697_991, f"groups {groups}"

    compare = {}
    if not args.no_compare and args.batched_form == "learned":
        def timed(fl, modes=None, steps=args.steps):
            run(fl, modes)
            capi.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                g = int(run(fl, modes).sum())
            capi.synchronize()
            assert g == groups
            return (time.perf_counter() - t0) * 1e3 / steps
        compare["fused_without_advisor_ms"] = timed(F.FLAG_GRAPH | F.FLAG_FUSED)
        phased_tiles = tune(F.FLAG_GRAPH, "graph")
        compare["phased_lds_first_ms"] = timed(F.FLAG_GRAPH)
        compare["phased_partitioned_ms"] = timed(F.FLAG_GRAPH, all_part)
        compare["phased_learned_ms"] = timed(F.FLAG_GRAPH | F.FLAG_ADVISOR)
        best_fixed = min(compare["fused_without_advisor_ms"], compare["phased_lds_first_ms"],
                         compare["phased_partitioned_ms"])
        autotune.apply(s, autotune.DEFAULT)
        compare["plain_one_launch_per_batch_ms"] = timed(F.FLAG_PER_BATCH, steps=1)
        if tuned is not None:
            autotune.apply(s, tuned)
        compare["speedup_vs_plain"] = compare["plain_one_launch_per_batch_ms"] / ms
        compare["speedup_vs_best_fixed_batched"] = best_fixed / ms
        compare["speedup_vs_fused_without_advisor"] = compare["fused_without_advisor_ms"] / ms
        compare = {k: round(v, 4) for k, v in compare.items()}
        compare["phased_tiles"] = {"block": phased_tiles.block, "lds_bytes": phased_tiles.lds_bytes,
                                   "max_probe": phased_tiles.max_probe, "part_lds_bytes": phased_tiles.part_lds_bytes}
    s.close()
    if net is not None:
        net.close()
    alg = n * 16 + groups * 24
    achieved = alg / (run_ms / 1e3) / 1e9
    fused = bool(flags & F.FLAG_FUSED)
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": profiled_traffic("stream (all kernels)", "batched_groupby_c4.md")
            if args.batched_form == "learned" and not args.tiles else None,
            "kernel": ("the stream's launch sequence (one hipGraph launch: advisor sketch + MLP, one workgroup per batch "
                       "doing its whole group-by in LDS, region scan, exact fallback)" if fused else
                       "the stream's launch sequence (one hipGraph launch: advisor sketch + MLP, per-batch LDS table, "
                       "scan, split, partition tables, exact fallback)") if flags & F.FLAG_GRAPH else prof,
            "kernel_ms": round(run_ms, 4), "algorithmic_bytes": alg}
    cpu = cpu_baseline_stream(w, offsets, spec) if not args.no_cpu else None
    cfg = {"workload": f"batched group-by: {nb} independent VectorBatches, {n} rows (lognormal rows, median 2048, "
                       "[64, 65536]), per-batch distinct keys log-uniform in [8, 1e6]; SUM(int64) + COUNT(*)",
           "form": {"learned": "learned + batched (stream advisor predicting each batch's groups, one workgroup per "
                               "batch, autotuned tiles, hipGraph)",
                    "fused": "batched, one workgroup per batch, own distinct estimate (hipGraph)",
                    "learned_phases": "learned + batched, phased kernels (hipGraph)",
                    "graph": "batched, phased, every batch LDS-first (hipGraph)", "two_launches": "batched, no graph",
                    "per_batch": "plain: one launch per batch and phase"}[args.batched_form],
           "rows": n, "batches": nb, "groups": groups,
           "batches_by_phase": ({"direct": ph[0], "partitioned": ph[1], "hbm_table": ph[2]} if fused else
                                {"lds_table": ph[0], "lds_partitions": ph[1], "hbm_table": ph[2]}),
           "tiles": None if tuned is None else {"block": tuned.block, "lds_bytes": tuned.lds_bytes,
                                                "max_probe": tuned.max_probe, "part_lds_bytes": tuned.part_lds_bytes},
           "algorithmic_bytes_per_step": alg, "bytes_per_row_in": 16, "bytes_per_group_out": 24}
    cfg.update(compare)
    return {"value": n / (ms / 1e3), "ms_per_step": ms, "config": cfg, "roofline": roof, "cpu_baseline": cpu,
            "dtype": "int64"}


def cpu_baseline_stream(w, offsets, spec):
    """The oracle (the C++ restatement of HashAggregationOperator, one operator per batch) over the same stream:
    all host cores (P threads taking batches from a shared counter) on the full stream, and one core on the first
    2,000 batches; each a warm-up run, then the best of 3."""
    import oracle
    key, val = oracle.Col(oracle.LONG, w.keys), oracle.Col(oracle.LONG, w.values)
    ospec = oracle.make_agg_spec([oracle.LONG], [(oracle.SUM, oracle.LONG, oracle.LONG),
                                                 (oracle.COUNT_ALL, oracle.LONG, oracle.LONG)])

    def best(offs, threads):
        times = []
        for i in range(4):
            t0 = time.perf_counter()
            oracle.agg_stream(ospec, key, [val, None], offs, threads)
            if i:
                times.append(time.perf_counter() - t0)
        return min(times)

    P = max(1, min(16, os.cpu_count() or 1))
    t_all = best(offsets, P)
    sub = offsets[:2001]
    t_one = best(sub, 1)
    return {"value": float(offsets[-1]) / t_all, "unit": "rows/s", "cores": P, "kind": "port",
            "sample": f"the full stream ({len(offsets) - 1} batches, {offsets[-1]} rows) on {P} threads, one "
                      f"operator per batch, warm-up + best of 3; host CPU {cpu_model()}",
            "single_core": {"value": float(sub[-1]) / t_one, "unit": "rows/s", "cores": 1,
                            "sample": f"first 2000 batches ({sub[-1]} rows), warm-up + best of 3"}}


def bench_operator(args, mode: str):
    """C1 (mode "agg") or C2 (mode "join") end to end through the GPU operators' Operator surface: host VectorBatches
    in, host VectorBatches out, every host<->device copy inside the timed region (omni_operator_bench, a C++ driver
    of GpuHashAggregationOperator / GpuHashBuilderOperator + GpuLookupJoinOperator)."""
    import subprocess
    exe = os.path.join(ROOT, "omnioperator_amd", "lib", "omni_operator_bench")
    if not os.path.exists(exe):
        raise SystemExit("omni_operator_bench is not built (python -c 'import __graft_entry__ as g; g.build()')")
    r = subprocess.run([exe, mode, str(args.steps), str(args.warmup)], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=1200)
    if r.returncode != 0:
        raise SystemExit(f"omni_operator_bench {mode} failed ({r.returncode}): {r.stderr.strip()} {r.stdout.strip()}")
    d = json.loads(r.stdout.strip().splitlines()[-1])
    ms = d["ms_per_step"]
    if mode == "agg":
        alg = d["rows"] * 16
        work = "group-by through GpuHashAggregationOperator: 1M rows, int64 key over 1,000 values, int64 value, " \
               "VectorBatches of 4096; SUM/COUNT(*)/MIN/MAX; host batches in and out, copies included"
        cfg_extra = {"groups": d["groups"], "groups_ok": d["groups_ok"]}
    else:
        alg = d["probe_rows"] * 8 + int(d["matches"]) * 12
        work = "inner join through GpuHashBuilderOperator + GpuLookupJoinOperator: build 2M unique int64 keys + int32 " \
               "payload, probe 64M int64 keys (80% hit), VectorBatches of 65536; output probe key + build payload as " \
               "host batches, copies included"
        cfg_extra = {"matches": d["matches"], "matches_ok": d["matches_ok"],
                     "build_kernel_ms_per_step": d["build_kernel_ms_per_step"]}
    kms = d["kernel_ms_per_step"]
    achieved = alg / (kms / 1e3) / 1e9 if kms > 0 else 0.0
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": d["kernel"] + " (all launches of a step, HIP events)", "kernel_ms": round(kms, 4),
            "algorithmic_bytes": alg}
    cfg = {"workload": work, "rows": d["rows"], "end_to_end": True,
           "kernel_share": round(kms / ms, 4) if ms > 0 else None}
    cfg.update(cfg_extra)
    cpu = None
    if not args.no_cpu:  # the oracle on the same configuration (the Python generator's rows of it)
        from omnioperator_amd import workloads
        if mode == "agg":
            import oracle
            w = workloads.agg_plain()
            spec = oracle.make_agg_spec([oracle.LONG], [(oracle.SUM, oracle.LONG, oracle.LONG),
                                                        (oracle.COUNT_ALL, oracle.LONG, oracle.LONG),
                                                        (oracle.MIN, oracle.LONG, oracle.LONG),
                                                        (oracle.MAX, oracle.LONG, oracle.LONG)])
            vc = oracle.Col(oracle.LONG, w.values)
            t0 = time.perf_counter()
            oracle.hash_agg(spec, [oracle.Col(oracle.LONG, w.keys)], [vc, None, vc, vc], w.batch_rows)
            cpu = {"value": len(w.keys) / (time.perf_counter() - t0), "unit": "rows/s", "cores": 1, "kind": "port",
                   "sample": f"the full configuration ({len(w.keys)} rows); host CPU {cpu_model()}"}
        else:
            cpu = cpu_baseline_join(workloads.join_inner(), args)
    return {"value": d["rows"] / (ms / 1e3), "ms_per_step": ms, "config": cfg, "roofline": roof,
            "cpu_baseline": cpu, "dtype": "int64"}


def bench_learned(args):
    """The learned advisor on the C5 task set: 512 seeded group-by and join tasks (rows log-uniform in
    [1e5, 2e8]), each run with every fixed strategy, the library's automatic rule and the advisor."""
    from omnioperator_amd import capi, learned
    capi.init(0)
    tasks = learned.sample_tasks(args.learned_tasks, seed=0xC5, min_rows=1e5, max_rows=args.learned_max_rows,
                                 join_share=0.5)
    adv = learned.Advisor()
    t0 = time.perf_counter()
    secs, rows, recs = learned.evaluate(tasks, adv, log=lambda m: print(m, file=sys.stderr, flush=True))
    wall = time.perf_counter() - t0
    adv.close()
    if args.learned_records:
        with open(args.learned_records, "w") as f:
            json.dump(recs, f)
    by_op = {op: {k: round(sum(r[k] for r in recs if r["op"] == op), 4) for k in learned.CONFIGS}
             for op in ("agg", "join")}
    fixed = {k: secs[k] for k in ("lds", "global", "radix")}
    best_fixed = min(fixed, key=fixed.get)
    alg = 0
    for r in recs:
        if r["op"] == "join":
            alg += int(r["distinct"]) * 16 + r["rows"] * 8
        else:
            alg += r["rows"] * 16 + r["groups"] * 24
    achieved = alg / secs["advisor"] / 1e9
    n_join = sum(1 for r in recs if r["op"] == "join")
    cfg = {"workload": f"learned strategy/capacity: {len(tasks)} seeded tasks ({len(tasks) - n_join} group-by, {n_join} "
                       f"inner join), rows log-uniform [1e5, {args.learned_max_rows:.0e}], distinct log-uniform "
                       "[10, 1e8] capped at rows, Zipf s U[0, 1.5], keys int32 / int64 / (int32, int64)",
           "tasks": len(tasks), "rows": rows, "seconds": {k: round(v, 4) for k, v in secs.items()},
           "best_fixed_strategy": best_fixed,
           "speedup_vs_best_fixed": round(secs[best_fixed] / secs["advisor"], 3),
           "speedup_vs_each_fixed": {k: round(v / secs["advisor"], 3) for k, v in fixed.items()},
           "speedup_vs_auto_rule": round(secs["auto"] / secs["advisor"], 3),
           "seconds_by_operator": by_op,
           "advisor_plan_seconds": round(sum(r.get("plan", 0.0) for r in recs), 4),
           "wall_seconds_incl_generation_and_checks": round(wall, 1),
           "algorithmic_bytes": alg}
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "whole task set (advisor)",
            "kernel_ms": round(secs["advisor"] * 1e3, 3), "algorithmic_bytes": alg}
    return {"value": rows / secs["advisor"], "ms_per_step": secs["advisor"] * 1e3, "config": cfg, "roofline": roof,
            "cpu_baseline": None, "dtype": "int64"}


def profiled_traffic(row_name: str, profile: str = "join_probe_c2.md"):
    """HBM bytes per launch from the copy-calibrated rocprofv3 FETCH_SIZE / WRITE_SIZE table of a profile under
    profiles/ (the row named `row_name`); None when it was not profiled."""
    path = os.path.join(ROOT, "profiles", profile)
    try:
        with open(path) as f:
            for line in f:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cells) == 4 and cells[0] == row_name and cells[3].isdigit():
                    return int(cells[3])
    except OSError:
        pass
    return None


TYPE_NAMES = {1: "INT", 2: "LONG", 7: "DECIMAL128"}


def kbytes_words(types):
    return sum(2 if t == 7 else 1 for t in types)


def join_key_columns(capi, w, mode: str):
    """(key types, build key columns, probe key columns) of the join workload for --join-keys: the int64 key
    itself; (INT, INT) = its high and low halves, the low half masked to 31 bits so that the packed int64 key
    (high << 32 | low) is the key with bit 31 cleared; (LONG, LONG) = the key and splitmix64 of it (a 128-bit
    packed key, joined through the key dictionary); DECIMAL128 = the key sign-extended to 128 bits."""
    def split(k):
        return [(k >> 32).astype(np.int32), (k & 0x7FFFFFFF).astype(np.int32)]

    def mix(k):
        z = k.view(np.uint64) + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return (z ^ (z >> np.uint64(31))).view(np.int64)

    def dec(k):
        return np.stack([k.view(np.uint64), (k >> 63).view(np.uint64)], 1)

    with np.errstate(over="ignore"):
        if mode == "single":
            return [capi.LONG], [w.build_keys], [w.probe_keys]
        if mode == "int_int":
            return [capi.INT, capi.INT], split(w.build_keys), split(w.probe_keys)
        if mode == "long_long":
            return [capi.LONG, capi.LONG], [w.build_keys, mix(w.build_keys)], [w.probe_keys, mix(w.probe_keys)]
        return [capi.DECIMAL128], [dec(w.build_keys)], [dec(w.probe_keys)]


def cpu_baseline_join(w, args, jt=0, bs=0, ktypes=None, bkcols=None, pkcols=None):
    """The oracle (C++ restatement of HashBuilderOperator + LookupJoinOperator, one core) on the full
    build side and a bounded probe sample; rows/s = (build + sampled probe rows) / wall time."""
    import oracle
    sample = min(len(w.probe_keys), args.cpu_probe_rows)
    if ktypes is None or len(ktypes) == 1 and ktypes[0] == oracle.LONG:
        bk = oracle.Col(oracle.LONG, w.build_keys)
        pk = oracle.Col(oracle.LONG, w.probe_keys[:sample])
    else:
        bk = [oracle.Col(t, v) for t, v in zip(ktypes, bkcols)]
        pk = [oracle.Col(t, v[:sample]) for t, v in zip(ktypes, pkcols)]
        if len(bk) == 1:
            bk, pk = bk[0], pk[0]
    bp = oracle.Col(oracle.INT, w.build_payload)
    t0 = time.perf_counter()
    if jt == 0:
        r = oracle.hash_join(bk, [bp], pk, w.batch_rows, w.batch_rows)
    else:
        r = oracle.hash_join_typed(bk, [bp], pk, w.batch_rows, w.batch_rows, jt, bs)
    t1 = time.perf_counter()
    return {"value": (len(w.build_keys) + sample) / (t1 - t0), "unit": "rows/s", "cores": 1, "kind": "port",
            "sample": f"full build ({len(w.build_keys)} rows) + first {sample} probe rows, "
                      f"{len(r.probe_rows)} output rows; host CPU {cpu_model()}"}


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--workload", default="batched",
                   choices=["join", "agg", "skewed", "batched", "learned", "partition", "agg_operator",
                            "join_operator"],
                   help="batched (default): the 20,000-batch mixed-cardinality stream, learned + batched form; join: "
                        "2M x 64M inner join; agg: 1M-row plain group-by; skewed: 200M-row composite-key group-by; "
                        "learned: the 512-task advisor set; partition: shuffle partition ids")
    p.add_argument("--skewed-rows", type=int, default=200_000_000)
    p.add_argument("--cpu-skewed-rows", type=int, default=4_000_000)
    p.add_argument("--batched-form", default="learned", choices=["learned", "fused", "learned_phases", "graph", "two_launches", "per_batch"])
    p.add_argument("--no-compare", action="store_true",
                   help="batched: skip the plain and fixed-strategy runs beside the learned form")
    p.add_argument("--tiles", default=None, help="batched: fixed block,lds_bytes,max_probe,part_lds_bytes "
                                                 "instead of the autotuner")
    p.add_argument("--learned-tasks", type=int, default=512)
    p.add_argument("--learned-max-rows", type=float, default=2e8)
    p.add_argument("--learned-records", default=None, help="write the per-task records of --workload learned here")
    p.add_argument("--agg-form", default="batches", choices=["batches", "per_batch"],
                   help="plain group-by: all batches in one add_batches call, or one add_batch per batch")
    p.add_argument("--dense", action="store_true", help="dense build keys (array-table variant)")
    p.add_argument("--partitions-out", type=int, default=200, help="partition workload: shuffle partitions")
    p.add_argument("--agg-strategy", default="radix", choices=["radix", "auto", "global", "lds"],
                   help="skewed workload: the group-by strategy (radix = partitioned, LDS tables per piece)")
    p.add_argument("--agg-rule", default="normal", choices=["normal", "array"],
                   help="agg workload: AggHashTableRule (array = direct-index map for the single integer key)")
    p.add_argument("--join-type", default="inner", choices=sorted(JOIN_TYPES),
                   help="join workload: join type (right_build_right keeps the build side; left keeps the probe side)")
    p.add_argument("--build-rows", type=int, default=2_000_000)
    p.add_argument("--probe-rows", type=int, default=64_000_000)
    p.add_argument("--cpu-probe-rows", type=int, default=8_000_000)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--probe-group", type=int, default=1, help="probe width (lookups in flight per lane)")
    p.add_argument("--join-keys", default="single", choices=["single", "int_int", "long_long", "decimal128"],
                   help="join workload: the key as one int64 column or as several key columns / a DECIMAL128")
    p.add_argument("--partitions", type=int, default=-1, help="radix-partitioned probe: -1 auto, 0 off, or P")
    p.add_argument("--join-autotune", action="store_true",
                   help="join workload: probe width and partition count from the autotuner (cached sweep)")
    p.add_argument("--calibrate-copy-mib", type=int, default=0,
                   help="only run 3 device-to-device copies of this many MiB (HBM counter calibration)")
    args = p.parse_args()
    if args.calibrate_copy_mib:
        from omnioperator_amd import capi
        capi.init(0)
        n = args.calibrate_copy_mib << 20
        src, dst = capi.DeviceArray(n, np.uint8), capi.DeviceArray(n, np.uint8)
        for _ in range(3):
            capi.check(capi.lib().omni_gpu_memcpy_d2d(dst.ptr, src.ptr, n))
        capi.synchronize()
        print(json.dumps({"copy_bytes": n, "copies": 3}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the path is measured on one GPU's "
              "share; sharding across GPUs (exchange of partial groups, broadcast build side) is not run.",
              file=sys.stderr)
        sys.exit(2)
    res = {"join": bench_join, "agg": bench_agg_plain, "skewed": bench_agg_skewed,
           "batched": bench_batched, "learned": bench_learned, "partition": bench_partition,
           "agg_operator": lambda a: bench_operator(a, "agg"),
           "join_operator": lambda a: bench_operator(a, "join")}[args.workload](args)
    single = args.workload == "learned"  # one pass over the task set; each task warmed by an untimed run
    line = {"metric": METRIC, "value": res["value"], "unit": "rows/s", "n_gpus": 1,
            "steps": 1 if single else args.steps, "warmup": 0 if single else args.warmup, "ms_per_step": res["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": res["dtype"], "data": "synthetic (seeded)", "config": res["config"],
            "roofline": res["roofline"], "cpu_baseline": res["cpu_baseline"]}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
