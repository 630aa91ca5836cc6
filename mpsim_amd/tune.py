"""Autotuner for the Jacobi SVD and the concurrency of two-site calls.

    python -m mpsim_amd.tune            # on a GPU; writes mpsim_amd/tuned.json

Three stages, each timed on seeded synthetic two-site problems:

1. Pivot size and inner sweeps. For each bond-size class, on a batch that
   fills the device, every pair (pivot rows 2b in {16, 32}, inner
   eigensolver sweeps in {1, 2}); the pivot-size crossover becomes
   `block_switch` (problems with more rows than this use 32-row pivots) and
   the most frequent best inner-sweep count becomes `inner_sweeps`.
2. SVD method. For the classes of 128 rows and more, the Gram method (the
   library default for FP64 problems with 32-row pivots) against the
   one-sided rows method, with stage 1's settings; `method` is the faster
   one on the largest class (where the time goes).
3. Concurrency. A large call runs as `split_ways` groups on as many streams
   (Engine.two_site). On layer-shaped batches (one problem per disjoint
   pair of a 64-site brickwork layer) of each size class every split_ways in
   1..4 is timed. `split_ways` is the fastest on the largest class,
   `split_rows` the smallest class from which splitting beats one group on
   every larger class too, and `split_min` the smallest batch of that class
   from which splitting beats one group.

Engine picks the stored values up as its defaults (tuned_config,
tuned_engine). The library's compile-time tile shapes (theta tiles, 16-row
Jacobi blocks, Gram tiles) are not tuned here, and neither is the QR
preconditioning of the Gram method (`precondition`, default 1: problems of
saturated bonds): whether it saves sweeps depends on the spectrum of theta,
which synthetic problems do not reproduce, so its default comes from whole
circuit workloads (measured on the 64-qubit brickwork, the uneven batch and
the served stream) and a stored value overrides it.
"""
from __future__ import annotations

import json
import os
import time
from typing import Callable, Dict, List, Optional, Sequence, Tuple

import numpy as np

from .engine import SvdConfig

HERE = os.path.dirname(os.path.abspath(__file__))
TUNED = os.environ.get("MPSX_TUNED", os.path.join(HERE, "tuned.json"))

DEFAULT_PRECONDITION = 1  # QR-precondition the Gram problems of saturated bonds
DEFAULT_SPLIT_WAYS = 4   # the hardware queues a process gets
DEFAULT_SPLIT_MIN = 8
DEFAULT_SPLIT_ROWS = 128
NO_SPLIT_ROWS = 1 << 30  # no size class gained from splitting


def _load(path: Optional[str]) -> Dict:
    try:
        with open(path or TUNED) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def tuned_config(path: Optional[str] = None) -> SvdConfig:
    """SvdConfig with the stored tuning applied (library defaults if none)."""
    t = _load(path)
    return SvdConfig(block_switch=int(t.get("block_switch", 0)),
                     inner_sweeps=int(t.get("inner_sweeps", 0)),
                     method=int(t.get("method", 0)),
                     precondition=int(t.get("precondition", DEFAULT_PRECONDITION)))


def tuned_engine(path: Optional[str] = None) -> Tuple[int, int, int]:
    """(split_ways, split_min, split_rows) from the stored tuning (defaults
    if none)."""
    t = _load(path)
    ways = int(t.get("split_ways", DEFAULT_SPLIT_WAYS))
    smin = int(t.get("split_min", DEFAULT_SPLIT_MIN))
    rows = int(t.get("split_rows", DEFAULT_SPLIT_ROWS))
    return max(1, min(4, ways)), max(1, smin), max(1, rows)


def _problems(engine, chi: int, count: int, seed: int):
    from .engine import TwoSiteItem
    from .gates import haar_unitary
    torch = engine.torch
    rng = np.random.RandomState(seed)
    items = []
    g = engine.gate(haar_unitary(4, rng))
    for _ in range(count):
        a = (rng.randn(chi, 2, chi) + 1j * rng.randn(chi, 2, chi)) / np.sqrt(4 * chi)
        b = (rng.randn(chi, 2, chi) + 1j * rng.randn(chi, 2, chi)) / np.sqrt(2)
        items.append(TwoSiteItem(torch.from_numpy(a.astype(np.complex64)).to(engine.dev),
                                 torch.from_numpy(b.astype(np.complex64)).to(engine.dev),
                                 g, 2 * chi))
    return items


def measure(engine, items, cfg: SvdConfig, reps: int = 2) -> float:
    """Seconds per update of one call over `items` (one untimed call first)."""
    torch = engine.torch
    engine.two_site(items, cfg)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        engine.two_site(items, cfg)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps / len(items)


def _measure_split(engine, items, cfg: SvdConfig, ways: int, reps: int, timer: Callable) -> float:
    """The timer with the engine splitting every call `ways` ways."""
    old = engine.split_ways, engine.split_min, engine.split_rows
    engine.split_ways, engine.split_min, engine.split_rows = ways, 1, 1
    try:
        return timer(engine, items, cfg, reps)
    finally:
        engine.split_ways, engine.split_min, engine.split_rows = old


def tune(engine=None, chis: Sequence[int] = (8, 16, 32, 64, 128, 256), reps: int = 2,
         path: Optional[str] = TUNED, conc_chis: Sequence[int] = (64, 128, 256),
         layer: int = 31, batch_sizes: Sequence[int] = (2, 4, 8, 16),
         timer: Callable = measure) -> Dict:
    """Runs the three stages and (with `path`) stores the result. `timer` is
    the measurement (seconds per update of a call), replaceable for tests."""
    from .engine import Engine
    eng = engine or Engine.get()
    rows: List[Dict] = []
    # ---- stage 1: pivot size and inner sweeps
    for chi in chis:
        # enough problems for ~1024 pivot workgroups per round (a full device)
        count = int(max(1, min(64, -(-1024 // max(1, (2 * chi) // 16)))))
        items = _problems(eng, chi, count, seed=chi)
        best = None
        for block in (8, 16):
            for inner in (1, 2):
                t = timer(eng, items, SvdConfig(block=block, inner_sweeps=inner), reps)
                rec = {"stage": "pivot", "chi": chi, "rows": 2 * chi, "count": count, "block": block,
                       "inner_sweeps": inner, "us_per_update": round(1e6 * t, 2)}
                rows.append(rec)
                if best is None or t < best[0]:
                    best = (t, rec)
        best[1]["best"] = True
    bests = [r for r in rows if r.get("best")]
    switch = 0
    for r in sorted(bests, key=lambda r: r["rows"]):
        if r["block"] == 8:
            switch = r["rows"]
    inner = int(np.bincount([r["inner_sweeps"] for r in bests]).argmax())
    out: Dict = {"block_switch": max(switch, 16), "inner_sweeps": inner}
    base = SvdConfig(block_switch=out["block_switch"], inner_sweeps=inner)

    # ---- stage 2: Gram method against the one-sided rows method
    method, largest = 0, -1
    for chi in chis:
        if 2 * chi < 128:
            continue
        count = int(max(1, min(64, -(-1024 // max(1, (2 * chi) // 16)))))
        items = _problems(eng, chi, count, seed=chi)
        times = {}
        for m in (0, 1):
            cfg = SvdConfig(block_switch=base.block_switch, inner_sweeps=inner, method=m)
            times[m] = timer(eng, items, cfg, reps)
            rows.append({"stage": "method", "chi": chi, "rows": 2 * chi, "count": count,
                         "method": m, "us_per_update": round(1e6 * times[m], 2)})
        if chi > largest:
            largest, method = chi, min(times, key=times.get)
    out["method"] = method
    base = SvdConfig(block_switch=base.block_switch, inner_sweeps=inner, method=method)

    # ---- stage 3: concurrent groups of a split call
    ways_best: Dict[int, int] = {}
    for chi in sorted(conc_chis):
        items = _problems(eng, chi, layer, seed=1000 + chi)
        best = None
        for ways in (1, 2, 3, 4):
            t = _measure_split(eng, items, base, ways, reps, timer)
            rows.append({"stage": "split", "chi": chi, "rows": 2 * chi, "count": layer,
                         "split_ways": ways, "us_per_update": round(1e6 * t, 2)})
            if best is None or t < best[0]:
                best = (t, ways)
        ways_best[chi] = best[1]
    split_ways, split_rows, split_min = DEFAULT_SPLIT_WAYS, DEFAULT_SPLIT_ROWS, DEFAULT_SPLIT_MIN
    if ways_best:
        split_ways = ways_best[max(ways_best)]
        split_rows = NO_SPLIT_ROWS
        for chi in sorted(ways_best, reverse=True):  # largest class down, while splitting wins
            if ways_best[chi] == 1:
                break
            split_rows = 2 * chi
        if split_rows == NO_SPLIT_ROWS:
            split_ways = 1
        else:
            chi = split_rows // 2
            split_min = max(batch_sizes) + 1  # only larger batches, unless a size below wins
            for n in sorted(batch_sizes):
                items = _problems(eng, chi, n, seed=2000 + n)
                t1 = _measure_split(eng, items, base, 1, reps, timer)
                tw = _measure_split(eng, items, base, split_ways, reps, timer)
                rows.append({"stage": "split_min", "chi": chi, "count": n, "split_ways": split_ways,
                             "us_per_update_one_group": round(1e6 * t1, 2),
                             "us_per_update_split": round(1e6 * tw, 2)})
                if tw < t1:
                    split_min = n
                    break
    out["split_ways"] = split_ways
    out["split_rows"] = split_rows
    out["split_min"] = split_min
    out["precondition"] = int(_load(path).get("precondition", DEFAULT_PRECONDITION))
    out["measurements"] = rows
    if path:
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
    return out


if __name__ == "__main__":
    res = tune()
    print(json.dumps({k: v for k, v in res.items() if k != "measurements"}))
    for r in res["measurements"]:
        print(r)
