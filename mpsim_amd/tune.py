"""Autotuner for the Jacobi SVD launch parameters.

    python -m mpsim_amd.tune            # on a GPU; writes mpsim_amd/tuned.json

For each bond-size class it times, on a batch of seeded synthetic two-site
problems that fills the device, every candidate pair (pivot rows 2b in
{16, 32}, inner eigensolver sweeps in {1, 2}) and records the fastest. The
pivot-size crossover becomes `block_switch` (problems with more rows than
this use 32-row pivots) and the most frequent best inner-sweep count becomes
`inner_sweeps`. Engine picks the stored values up as its defaults.
"""
from __future__ import annotations

import json
import os
import time
from typing import Dict, List, Optional

import numpy as np

from .engine import SvdConfig

HERE = os.path.dirname(os.path.abspath(__file__))
TUNED = os.environ.get("MPSX_TUNED", os.path.join(HERE, "tuned.json"))


def tuned_config() -> SvdConfig:
    """SvdConfig with the stored tuning applied (library defaults if none)."""
    try:
        with open(TUNED) as f:
            t = json.load(f)
        return SvdConfig(block_switch=int(t.get("block_switch", 0)),
                         inner_sweeps=int(t.get("inner_sweeps", 0)))
    except (OSError, ValueError):
        return SvdConfig()


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
    torch = engine.torch
    engine.two_site(items, cfg)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        engine.two_site(items, cfg)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps / len(items)


def tune(engine=None, chis=(8, 16, 32, 64, 128, 256), reps: int = 2,
         path: Optional[str] = TUNED) -> Dict:
    from .engine import Engine
    eng = engine or Engine.get()
    rows: List[Dict] = []
    for chi in chis:
        # enough problems for ~1024 pivot workgroups per round (a full device)
        count = int(max(1, min(64, -(-1024 // max(1, (2 * chi) // 16)))))
        items = _problems(eng, chi, count, seed=chi)
        best = None
        for block in (8, 16):
            for inner in (1, 2):
                t = measure(eng, items, SvdConfig(block=block, inner_sweeps=inner), reps)
                rec = {"chi": chi, "rows": 2 * chi, "count": count, "block": block,
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
    out = {"block_switch": max(switch, 16), "inner_sweeps": inner, "measurements": rows}
    if path:
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
    return out


if __name__ == "__main__":
    res = tune()
    print(json.dumps({k: v for k, v in res.items() if k != "measurements"}))
    for r in res["measurements"]:
        print(r)
