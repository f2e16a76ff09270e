"""The autotuner's sweep logic on the CPU (a fake workload whose candidates take known times), its cache, and --
on the GPU -- that every join candidate gives the same matches and the tuner picks one that runs."""
import time

import numpy as np
import pytest

from omnioperator_amd import autotune


@pytest.fixture
def cache(tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, "CACHE", str(tmp_path / "autotune.json"))
    return tmp_path


def test_tune_join_picks_the_fastest_and_caches(cache):
    cost = {(1, 0): 4, (2, 0): 3, (4, 0): 1, (8, 0): 5, (16, 0): 6, (1, -1): 2}
    seen = []

    def run(cfg):
        seen.append(cfg)
        if (cfg.probe_group, cfg.partitions) not in cost:
            raise ValueError("partition count too large for this table")
        time.sleep(cost[(cfg.probe_group, cfg.partitions)] * 0.002)

    key = autotune.join_key_of([2], 1 << 20, 1 << 25)
    best = autotune.tune_join(run, key)
    assert (best.probe_group, best.partitions) == (4, 0)
    swept = autotune._load()[key]["swept"]
    assert len(swept) == len(autotune.JOIN_CANDIDATES)
    assert sum(isinstance(t, str) for _, t in swept) == len(autotune.JOIN_CANDIDATES) - len(cost)
    n = len(seen)
    assert autotune.tune_join(run, key) == best  # from the cache: no new run
    assert len(seen) == n


def test_join_key_buckets():
    assert autotune.join_key_of([2], 2_000_000, 64_000_000) == autotune.join_key_of([2], 1_500_000, 48_000_000)
    assert autotune.join_key_of([2], 2_000_000, 64_000_000) != autotune.join_key_of([1, 2], 2_000_000, 64_000_000)
    assert autotune.join_key_of([2], 2_000_000, 64_000_000) != autotune.join_key_of([2], 2_000_000, 2_000_000)


@pytest.mark.gpu
def test_join_candidates_agree_and_tuner_picks_one(cache):
    from omnioperator_amd import capi, workloads
    w = workloads.join_inner(build_rows=300_000, probe_rows=3_000_000)
    bk = capi.DeviceColumn(capi.LONG, w.build_keys)
    bp = capi.DeviceColumn(capi.INT, w.build_payload)
    pk = capi.DeviceColumn(capi.LONG, w.probe_keys)
    n = len(w.probe_keys)
    out_rows = capi.DeviceArray(8 * n, np.int64, (n,))
    pay = capi.DeviceArray(4 * n, np.int32, (n,))
    oarr = (capi.OutColumn * 1)()
    oarr[0].type_id, oarr[0].values, oarr[0].nulls = capi.INT, pay.ptr, None
    results = {}

    def run(cfg):
        j = capi.HashJoin(capi.LONG, [capi.INT])
        j.tune(cfg.probe_group, cfg.partitions)
        j.add_build((capi.Column * 1)(bk.c()), (capi.Column * 1)(bp.c()), len(w.build_keys))
        j.finish()
        m = j.probe_batches_raw([pk.c()], [n], out_rows, oarr, n)
        rows = out_rows.numpy(m)
        j.close()
        results[cfg] = (m, int(rows.sum()), int(pay.numpy(m).astype(np.int64).sum()))

    best = autotune.tune_join(run, autotune.join_key_of([capi.LONG], len(w.build_keys), n), repeats=1)
    assert best in results
    assert len(set(results.values())) == 1  # every candidate that ran found the same matches
    assert next(iter(results.values()))[0] == w.hits
