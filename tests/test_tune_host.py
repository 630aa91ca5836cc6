"""The autotuner's selection logic (mpsim_amd.tune) with a synthetic timer on
CPU tensors: which pivot size, inner sweeps, SVD method and concurrency it
picks, and that the stored result is what the engine reads back."""
import types

import json

import numpy as np
import torch


def _fake_engine():
    return types.SimpleNamespace(torch=torch, dev=torch.device("cpu"),
                                 gate=lambda m: torch.from_numpy(np.asarray(m, np.complex64)),
                                 split_ways=4, split_min=8, split_rows=128)


def test_tuner_selection_and_readback(tmp_path):
    from mpsim_amd import tune

    def timer(eng, items, cfg, reps):
        chi, n = items[0].a.shape[0], len(items)
        if cfg.block:  # stage 1: 16-row pivots win up to 32 rows; 2 inner sweeps only at chi 8
            t = 1.0 + (0.0 if (cfg.block == 8) == (2 * chi <= 32) else 0.5)
            return t + (0.0 if cfg.inner_sweeps == (2 if chi == 8 else 1) else 0.1)
        if eng.split_rows != 1:  # stage 2: one-sided wins at 128 rows, Gram at 256
            return 1.0 + (0.2 if (cfg.method == 1) == (chi >= 128) else 0.0)
        # stage 3: three groups win on a layer from 128 rows (chi 64) up, not
        # at 64 rows; splitting loses below 8 problems
        if n >= 16:
            if chi < 64:
                return {1: 1.0, 2: 1.1, 3: 1.2, 4: 1.3}[eng.split_ways]
            return {1: 1.0, 2: 0.6, 3: 0.45, 4: 0.5}[eng.split_ways]
        return 1.0 if eng.split_ways == 1 else (1.2 if n < 8 else 0.5)

    path = str(tmp_path / "tuned.json")
    eng = _fake_engine()
    out = tune.tune(eng, chis=(8, 16, 32, 64, 128), reps=1, path=path, conc_chis=(32, 64, 128),
                    layer=31, batch_sizes=(2, 4, 8, 16), timer=timer)
    assert out["block_switch"] == 32 and out["inner_sweeps"] == 1
    assert out["method"] == 0
    assert out["split_ways"] == 3 and out["split_min"] == 8 and out["split_rows"] == 128
    assert (eng.split_ways, eng.split_min, eng.split_rows) == (4, 8, 128)  # restored
    stages = {r["stage"] for r in out["measurements"]}
    assert stages == {"pivot", "method", "split", "split_min"}
    cfg = tune.tuned_config(path)
    assert (cfg.block_switch, cfg.inner_sweeps, cfg.method) == (32, 1, 0)
    assert cfg.precondition == tune.DEFAULT_PRECONDITION == out["precondition"]
    assert tune.tuned_engine(path) == (3, 8, 128)
    # the QR preconditioning is not tuned: a stored choice survives a re-tune
    with open(path) as f:
        stored = json.load(f)
    stored["precondition"] = 0
    with open(path, "w") as f:
        json.dump(stored, f)
    out = tune.tune(eng, chis=(8, 16), reps=1, path=path, conc_chis=(32,), layer=31,
                    batch_sizes=(2,), timer=timer)
    assert out["precondition"] == 0 and tune.tuned_config(path).precondition == 0


def test_tuned_defaults_without_a_file(tmp_path):
    from mpsim_amd import tune
    missing = str(tmp_path / "none.json")
    cfg = tune.tuned_config(missing)
    assert (cfg.block_switch, cfg.inner_sweeps, cfg.method) == (0, 0, 0)
    assert cfg.precondition == tune.DEFAULT_PRECONDITION
    assert tune.tuned_engine(missing) == (tune.DEFAULT_SPLIT_WAYS, tune.DEFAULT_SPLIT_MIN,
                                          tune.DEFAULT_SPLIT_ROWS)


def test_tuner_turns_splitting_off_when_it_never_pays(tmp_path):
    from mpsim_amd import tune

    def timer(eng, items, cfg, reps):
        return 1.0 + (0.1 * (eng.split_ways - 1) if eng.split_rows == 1 else 0.0)

    out = tune.tune(_fake_engine(), chis=(8, 64), reps=1, path=str(tmp_path / "t.json"),
                    conc_chis=(32, 64), layer=4, timer=timer)
    assert out["split_ways"] == 1 and out["split_rows"] == tune.NO_SPLIT_ROWS
    assert "split_min" not in {r["stage"] for r in out["measurements"]}


def test_serving_lane_ways_follow_busy_lanes():
    """The four hardware queues go to the lanes that still have work: one
    each, the rest to the largest-bond lane, re-spread as lanes finish."""
    from mpsim_amd.serve import _Shared
    s = _Shared(1, 1, busy=[0, 1, 2])
    assert [s.ways(i) for i in range(3)] == [1, 1, 2]
    s.lane_done(2)
    assert [s.ways(i) for i in range(2)] == [1, 3]
    s.lane_done(0)
    assert s.ways(1) == 4
    s.lane_done(1)
    assert s.ways(1) == 1


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus N > 1 exits with a message before
    touching the device."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    res = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 2
    assert "one GPU only" in res.stderr
