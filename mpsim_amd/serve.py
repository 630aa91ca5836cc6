"""Single-GPU serving loop with continuous, layer-level batching.

Requests (seeded circuits with an arrival time) are admitted while their
exact device footprint fits the HBM budget and a slot is free. The footprint
is known before the run because every bond dimension follows from the
circuit alone (zero singular values are kept): the peak over layers of
sum_l chi_l*d*chi_{l+1}*8 bytes for the site tensors, plus the Jacobi
workspace of its largest two-site layer. Admitted circuits keep their gates
and tensors resident; each tick advances every active circuit by one layer
in one grouped launch sequence (mpsim_amd.batch.tick), so requests that
arrive while others run join at the next tick and their work overlaps on the
device. Finished requests free their budget immediately.

Latency of a request = completion time - arrival time (queueing included).
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from .batch import DeviceCircuit, prepare, tick, ws_bound
from .circuits import Circuit, Request, bond_plan
from .engine import Engine, SvdConfig


def workspace_bytes(circuit: Circuit) -> int:
    """Upper bound of the Jacobi workspace of the circuit's largest two-site
    layer (batch.ws_bound per update)."""
    d = circuit.d
    best = 0
    prev = [1] * (circuit.n + 1)
    for layer, bonds in zip(circuit.layers, bond_plan(circuit)):
        tot = 0
        for q, _ in layer:
            if len(q) == 2:
                i = min(q)
                M, N = d * prev[i], d * prev[i + 2]
                tot += ws_bound(M, N)
        best = max(best, tot)
        prev = [1] + list(bonds) + [1]
    return best


def footprint_bytes(circuit: Circuit) -> int:
    d = circuit.d
    peak = circuit.n * d * 8
    for b in bond_plan(circuit):
        full = [1] + b + [1]
        peak = max(peak, 8 * sum(full[i] * d * full[i + 1] for i in range(circuit.n)))
    # old and new site tensors coexist during a layer
    return 2 * peak + workspace_bytes(circuit)


@dataclass
class Served:
    rid: int
    arrival: float
    admitted: float = 0.0
    finished: float = 0.0
    updates: int = 0
    tensors: Optional[list] = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


@dataclass
class ServeReport:
    completed: int
    wall_s: float
    circuits_per_s: float
    updates_per_s: float
    p50_s: float
    p99_s: float
    max_active: int
    ticks: int
    results: Dict[int, Served] = field(default_factory=dict)


class MPSServer:
    def __init__(self, engine: Optional[Engine] = None, hbm_budget: int = 64 << 30,
                 max_active: int = 512, svd: Optional[SvdConfig] = None,
                 keep_results: bool = False):
        self.engine = engine or Engine.get()
        self.budget = int(hbm_budget)
        self.max_active = int(max_active)
        self.svd = svd
        self.keep_results = keep_results

    def run(self, requests: Sequence[Request], realtime: bool = True) -> ServeReport:
        """Serves `requests` (sorted by arrival). With realtime=False every
        request is treated as already arrived at t=0 (closed-loop batch)."""
        torch = self.engine.torch
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        active: List[DeviceCircuit] = []
        meta: Dict[int, Served] = {}
        used = 0
        fp: Dict[int, int] = {}
        ticks = 0
        max_active = 0
        done: List[Served] = []
        t0 = time.perf_counter()

        def now() -> float:
            return time.perf_counter() - t0

        while pending or active:
            t = now()
            # admission: arrived requests in order, while budget and slots allow
            while pending and (not realtime or pending[0].arrival <= t) \
                    and len(active) < self.max_active:
                req = pending[0]
                need = footprint_bytes(req.circuit)
                if need > self.budget:
                    raise ValueError(f"request {req.rid} needs {need} B > budget {self.budget} B")
                if used + need > self.budget:
                    break
                pending.popleft()
                used += need
                fp[req.rid] = need
                dc = prepare(req.circuit, self.engine)
                dc.rid = req.rid
                active.append(dc)
                meta[req.rid] = Served(req.rid, req.arrival if realtime else 0.0, admitted=t)
            max_active = max(max_active, len(active))
            if not active:
                if pending:
                    time.sleep(max(0.0, min(pending[0].arrival - now(), 0.01)))
                continue
            n2 = tick(self.engine, active, self.svd)
            ticks += 1
            finished = [dc for dc in active if dc.done]
            if finished:
                torch.cuda.current_stream(self.engine.dev).synchronize()
                tf = now()
                for dc in finished:
                    s = meta[dc.rid]
                    s.finished = tf
                    s.updates = dc.circuit.two_site_count()
                    if self.keep_results:
                        s.tensors = dc.mps.tensors
                    used -= fp.pop(dc.rid)
                    done.append(s)
                active = [dc for dc in active if not dc.done]
        torch.cuda.current_stream(self.engine.dev).synchronize()
        wall = now()
        lat = np.array([s.latency for s in done]) if done else np.zeros(1)
        ups = sum(s.updates for s in done)
        return ServeReport(
            completed=len(done), wall_s=wall, circuits_per_s=len(done) / wall,
            updates_per_s=ups / wall, p50_s=float(np.percentile(lat, 50)),
            p99_s=float(np.percentile(lat, 99)), max_active=max_active, ticks=ticks,
            results={s.rid: s for s in done})
