#!/usr/bin/env python3
"""Benchmark of the MI355X hash-join / hash-aggregation hot path.

Default workload (the configuration the headline metric is quoted on): inner hash join, build side
2,000,000 unique int64 keys + int32 payload, probe side 64,000,000 int64 keys (80% hit build keys,
20% miss), VectorBatches of 65,536 rows.  One step = build (31 build batches appended + table build)
+ probe of all 977 probe batches in one launch, writing (probe row id int64, build payload int32)
per match into device-resident output buffers.  Inputs are resident in HBM before timing starts.

Prints one JSON line with rows/s (build + probe rows per second), the probe kernel's roofline figure
(algorithmic bytes / average kernel duration from HIP events on the kernel's stream) and a CPU
baseline: the oracle (the C++ restatement of the reference CPU operators, one core) on a bounded
sample of the same workload.
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


def bench_join(args):
    from omnioperator_amd import capi, workloads

    capi.init(0)
    w = workloads.join_inner(build_rows=args.build_rows, probe_rows=args.probe_rows, dense=args.dense)
    nb, npr = len(w.build_keys), len(w.probe_keys)
    # device-resident inputs, cut into VectorBatches
    bkey = capi.DeviceColumn(capi.LONG, w.build_keys)
    bpay = capi.DeviceColumn(capi.INT, w.build_payload)
    probe = capi.DeviceColumn(capi.LONG, w.probe_keys)
    bspans = workloads.batches(nb, w.batch_rows)
    pspans = workloads.batches(npr, w.batch_rows)
    pcols, prows = capi.HashJoin.probe_arrays([probe.c(a) for a, _ in pspans], [b - a for a, b in pspans])
    cap = npr
    out_rows = capi.DeviceArray(cap * 8, np.int64, (cap,))
    out_pay = capi.DeviceArray(cap * 4, np.int32, (cap,))
    oarr = (capi.OutColumn * 1)()
    oarr[0].type_id, oarr[0].values, oarr[0].nulls = capi.INT, out_pay.ptr, None

    # per-batch column descriptors prepared once (the C++ operator path has no such host cost)
    bcols = [(bkey.c(a), (capi.Column * 1)(bpay.c(a)), b - a) for a, b in bspans]

    def step():
        j = capi.HashJoin(capi.LONG, [capi.INT])
        j.tune(args.probe_group, args.partitions)
        for kc, pc, n in bcols:
            j.add_build(kc, pc, n)
        j.finish()
        m = j.probe_batches_raw(pcols, prows, out_rows, oarr, cap)
        parts[0] = j.last_partitions()
        is_array[0] = j.info()["is_array"]
        j.close()
        return m

    parts, is_array = [0], [False]

    for _ in range(args.warmup):
        m = step()
    capi.synchronize()
    capi.profile_reset()
    capi.profile_enable(True)
    capi.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        m = step()
    capi.synchronize()
    t1 = time.perf_counter()
    capi.profile_enable(False)
    launches, probe_ms = capi.profile_get("join_probe")
    _, build_ms = capi.profile_get("join_build")
    assert m == w.hits, f"match count {m} != expected hits {w.hits}"
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    rows = nb + npr
    value = rows / (ms_per_step / 1e3)
    # the probe stage of one step: the probe kernel, or the partition chain (histogram, scan, scatter,
    # persistent probe) when the table is radix-partitioned
    avg_probe_ms = probe_ms / args.steps
    probe_bytes = npr * 8 + m * (8 + 4)
    achieved = probe_bytes / (avg_probe_ms / 1e3) / 1e9
    variant = (f"join_probe<{'true' if is_array[0] else 'false'}, {'false' if is_array[0] else 'true'}, false, "
               f"{args.probe_group}>") if parts[0] <= 1 else None
    traffic = profiled_traffic(variant, "dense" if args.dense else "sparse") if variant else None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "join_probe" if parts[0] <= 1 else
                          f"join_probe_* (radix-partitioned, {parts[0]} partitions: hist+scan+scatter+probe)",
                "kernel_ms": round(avg_probe_ms, 4), "algorithmic_bytes": probe_bytes}
    extra = {"build_kernel_ms": round(build_ms / max(1, args.steps), 4), "matches": m, "probe_group": args.probe_group,
             "probe_partitions": parts[0]}
    cpu = cpu_baseline_join(w, args) if not args.no_cpu else None
    cfg = {"workload": "inner hash join: build 2M unique int64 keys + int32 payload, probe 64M int64 keys "
                       "(80% hit), VectorBatches of 65536" if (nb, npr) == (2_000_000, 64_000_000) else
                       f"inner hash join: build {nb} + probe {npr} rows",
           "build_rows": nb, "probe_rows": npr, "batch_rows": w.batch_rows,
           "keys": "dense permutation" if args.dense else "sparse 64-bit (splitmix64)",
           "output": "probe rowid int64 + build payload int32"}
    cfg.update(extra)
    return {"value": value, "ms_per_step": ms_per_step, "config": cfg, "roofline": roofline,
            "cpu_baseline": cpu, "dtype": "int64"}


def profiled_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` on `workload` from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary in
    profiles/join_probe_c2.md (copy-calibrated); None when that kernel variant was not profiled."""
    path = os.path.join(ROOT, "profiles", "join_probe_c2.md")
    try:
        with open(path) as f:
            for line in f:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cells) == 6 and cells[0] == kernel and cells[1] == workload:
                    return int(cells[4])
    except OSError:
        pass
    return None


def cpu_baseline_join(w, args):
    """The oracle (C++ restatement of HashBuilderOperator + LookupJoinOperator, one core) on the full
    build side and a bounded probe sample; rows/s = (build + sampled probe rows) / wall time."""
    import oracle
    sample = min(len(w.probe_keys), args.cpu_probe_rows)
    bk = oracle.Col(oracle.LONG, w.build_keys)
    bp = oracle.Col(oracle.INT, w.build_payload)
    pk = oracle.Col(oracle.LONG, w.probe_keys[:sample])
    t0 = time.perf_counter()
    r = oracle.hash_join(bk, [bp], pk, w.batch_rows, w.batch_rows)
    t1 = time.perf_counter()
    return {"value": (len(w.build_keys) + sample) / (t1 - t0), "unit": "rows/s", "cores": 1, "kind": "port",
            "sample": f"full build ({len(w.build_keys)} rows) + first {sample} probe rows, "
                      f"{len(r.probe_rows)} matches; host CPU {cpu_model()}"}


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--workload", default="join", choices=["join"])
    p.add_argument("--dense", action="store_true", help="dense build keys (array-table variant)")
    p.add_argument("--build-rows", type=int, default=2_000_000)
    p.add_argument("--probe-rows", type=int, default=64_000_000)
    p.add_argument("--cpu-probe-rows", type=int, default=8_000_000)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--probe-group", type=int, default=1, help="probe width (lookups in flight per lane)")
    p.add_argument("--partitions", type=int, default=-1, help="radix-partitioned probe: -1 auto, 0 off, or P")
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
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU design is "
              "described in DESIGN.md and is not run.", file=sys.stderr)
        sys.exit(2)
    res = bench_join(args)
    line = {"metric": METRIC, "value": res["value"], "unit": "rows/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": res["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": res["dtype"], "data": "synthetic (seeded)", "config": res["config"],
            "roofline": res["roofline"], "cpu_baseline": res["cpu_baseline"]}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
