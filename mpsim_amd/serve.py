"""Single-GPU serving loop with continuous, step-level batching.

Requests (seeded circuits with an arrival time) are admitted while their
exact device footprint fits the HBM budget and a slot is free. The footprint
is known before the run because every bond dimension follows from the
circuit alone (zero singular values are kept): the peak over the steps of
its schedule (circuits.schedule: disjoint nearest-neighbour pairs, SWAP
networks for non-adjacent ones) of sum_l chi_l*d*chi_{l+1}*8 bytes for the
site tensors, plus the Jacobi workspace of its largest step of two-site
updates. Admitted circuits keep their gates and tensors resident; a tick
advances every active circuit of a lane by one step (one layer of a
brickwork circuit) in one grouped launch sequence (mpsim_amd.batch.tick),
so requests that arrive while others run join at the lane's next tick and
their work overlaps on the device. Finished requests free their budget
immediately.

Lanes. A tick lasts as long as its slowest two-site update, so one lockstep
loop over every request makes a 64-bond circuit wait for each step of a
512-bond one. Requests are therefore split by bond cap into lanes (by
default caps <= 128, <= 256 and the rest); each lane has its own stream,
engine (workspace and helper streams) and host thread, admits from its own
arrival queue under the shared HBM budget, and ticks at its own pace. The
lanes' kernels overlap on the GPU. Every two-site problem is computed by
workgroups that read only its own data and stops on its own convergence
flag, so a result is bitwise the same whichever lane, tick or mix it ran in.

Latency of a request = completion time - arrival time (queueing included).
"""
from __future__ import annotations

import threading
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .batch import DeviceCircuit, prepare, tick, ws_bound
from .circuits import Circuit, Request, bond_plan_steps, schedule
from .engine import Engine, SvdConfig


def workspace_bytes(circuit: Circuit) -> int:
    """Upper bound of the Jacobi workspace of the circuit's largest step of
    two-site updates (batch.ws_bound per update)."""
    d = circuit.d
    best = 0
    prev = [1] * (circuit.n + 1)
    steps, _ = schedule(circuit)
    for st, bonds in zip(steps, bond_plan_steps(circuit)):
        tot = 0
        for i, _, _, _ in st.twos:
            M, N = d * prev[i], d * prev[i + 2]
            tot += ws_bound(M, N)
        best = max(best, tot)
        prev = [1] + list(bonds) + [1]
    return best


def footprint_bytes(circuit: Circuit) -> int:
    d = circuit.d
    peak = circuit.n * d * 8
    for b in bond_plan_steps(circuit):
        full = [1] + b + [1]
        peak = max(peak, 8 * sum(full[i] * d * full[i + 1] for i in range(circuit.n)))
    # old and new site tensors coexist during a step
    return 2 * peak + workspace_bytes(circuit)


@dataclass
class Served:
    rid: int
    arrival: float
    admitted: float = 0.0
    finished: float = 0.0
    updates: int = 0
    lane: int = 0
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
    lanes: List[dict] = field(default_factory=list)


class _Shared:
    """State the lanes share: the HBM budget, the active count, the lanes
    that still have work and the results (one lock)."""

    def __init__(self, budget: int, max_active: int, busy: Sequence[int] = ()) -> None:
        self.lock = threading.Lock()
        self.budget = budget
        self.used = 0
        self.max_active = max_active
        self.active = 0
        self.peak_active = 0
        self.busy = set(busy)
        self.done: List[Served] = []
        self.error: Optional[BaseException] = None

    def lane_done(self, i: int) -> None:
        with self.lock:
            self.busy.discard(i)

    def ways(self, i: int, queues: int = 4) -> int:
        """Stream groups for lane i's split calls: the hardware queues spread
        over the lanes that still have work, one each and the rest to the
        largest-bond one."""
        with self.lock:
            busy = sorted(self.busy)
        if not busy or i != busy[-1]:
            return 1
        return max(1, queues - (len(busy) - 1))

    def try_admit(self, need: int) -> bool:
        with self.lock:
            if self.used + need > self.budget or self.active >= self.max_active:
                return False
            self.used += need
            self.active += 1
            self.peak_active = max(self.peak_active, self.active)
            return True

    def release(self, need: int, served: Served) -> None:
        with self.lock:
            self.used -= need
            self.active -= 1
            self.done.append(served)


class MPSServer:
    """Continuous batching of circuit requests on one GPU.

    lanes: ascending bond caps that bound each lane but the last (a request
    goes to the first lane whose cap is >= its maxsvals; maxsvals None counts
    as unbounded); () runs one lockstep lane for every request.
    lane_split_ways: how many concurrent stream groups a lane's large two-site
    calls may use (Engine.split_ways), per lane; by default (None) the four
    streams (the hardware queues a process gets) are spread, at every tick,
    over the lanes that still have requests: one each, the rest to the
    largest-bond one (so a lane left alone gets all four).
    """

    def __init__(self, engine: Optional[Engine] = None, hbm_budget: int = 64 << 30,
                 max_active: int = 512, svd: Optional[SvdConfig] = None,
                 keep_results: bool = False, lanes: Sequence[int] = (128, 256),
                 lane_split_ways: Optional[Sequence[int]] = None):
        self.engine = engine or Engine.get()
        self.budget = int(hbm_budget)
        self.max_active = int(max_active)
        self.svd = svd
        self.keep_results = keep_results
        self.caps = tuple(int(c) for c in lanes)
        n = len(self.caps) + 1
        if lane_split_ways is not None and len(lane_split_ways) != n:
            raise ValueError(f"lane_split_ways needs {n} entries")
        self.fixed_ways = None if lane_split_ways is None else [int(w) for w in lane_split_ways]
        self.split_ways = self.fixed_ways or [1] * n
        self._lane_engines: List[Engine] = []

    def lane_of(self, circuit: Circuit) -> int:
        cap = circuit.maxsvals
        for i, c in enumerate(self.caps):
            if cap is not None and cap <= c:
                return i
        return len(self.caps)

    def _engines(self) -> List[Engine]:
        """One engine per lane (own workspace and helper streams), made once
        and reused across runs."""
        torch = self.engine.torch
        n = len(self.caps) + 1
        while len(self._lane_engines) < n:
            if n == 1:
                eng = self.engine
            else:
                eng = Engine(self.engine.dev)
                eng.svd = self.engine.svd
            self._lane_engines.append(eng)
        for eng, ways in zip(self._lane_engines, self.split_ways):
            eng.split_ways = ways
        if not hasattr(self, "_streams"):
            self._streams = [torch.cuda.current_stream(self.engine.dev) if n == 1
                             else torch.cuda.Stream(self.engine.dev) for _ in range(n)]
        return self._lane_engines

    def run(self, requests: Sequence[Request], realtime: bool = True) -> ServeReport:
        """Serves `requests` to completion. With realtime=False every request
        is treated as already arrived at t=0 (closed-loop batch)."""
        torch = self.engine.torch
        for r in requests:
            need = footprint_bytes(r.circuit)
            if need > self.budget:
                raise ValueError(f"request {r.rid} needs {need} B > budget {self.budget} B")
        n = len(self.caps) + 1
        queues: List[deque] = [deque() for _ in range(n)]
        for r in sorted(requests, key=lambda q: q.arrival):
            queues[self.lane_of(r.circuit)].append(r)
        busy = [i for i in range(n) if queues[i]]
        if self.fixed_ways is None:
            ways = [1] * n
            if n == 1:
                ways = [self.engine.split_ways]
            elif busy:
                ways[busy[-1]] = max(1, 4 - (len(busy) - 1))
            self.split_ways = ways
        engines = self._engines()
        shared = _Shared(self.budget, self.max_active, busy)
        stats = [dict(lane=i, cap=(self.caps[i] if i < len(self.caps) else None), requests=len(q),
                      ticks=0, max_active=0) for i, q in enumerate(queues)]
        # every earlier allocation is complete before the lane streams start
        torch.cuda.synchronize(self.engine.dev)
        t0 = time.perf_counter()

        def lane(i: int) -> None:
            try:
                with torch.cuda.stream(self._streams[i]):
                    self._lane_loop(i, engines[i], queues[i], shared, stats[i], t0, realtime)
            except BaseException as e:  # re-raised in the calling thread
                with shared.lock:
                    shared.error = shared.error or e
            finally:
                shared.lane_done(i)

        if n == 1:
            lane(0)
        else:
            threads = [threading.Thread(target=lane, args=(i,), name=f"mpsx-lane{i}")
                       for i in range(n) if queues[i]]
            for t in threads:
                t.start()
            for t in threads:
                t.join()
        if shared.error is not None:
            raise shared.error
        torch.cuda.synchronize(self.engine.dev)
        wall = time.perf_counter() - t0
        done = shared.done
        lat = np.array([s.latency for s in done]) if done else np.zeros(1)
        ups = sum(s.updates for s in done)
        for st in stats:
            ls = np.array([s.latency for s in done if s.lane == st["lane"]])
            if ls.size:
                st["p50_s"] = float(np.percentile(ls, 50))
                st["p99_s"] = float(np.percentile(ls, 99))
        return ServeReport(
            completed=len(done), wall_s=wall, circuits_per_s=len(done) / wall,
            updates_per_s=ups / wall, p50_s=float(np.percentile(lat, 50)),
            p99_s=float(np.percentile(lat, 99)), max_active=shared.peak_active,
            ticks=sum(st["ticks"] for st in stats), results={s.rid: s for s in done},
            lanes=stats)

    def _lane_loop(self, i: int, eng: Engine, pending: deque, shared: _Shared, stats: dict,
                   t0: float, realtime: bool) -> None:
        torch = eng.torch
        stream = torch.cuda.current_stream(eng.dev)
        active: List[Tuple[DeviceCircuit, Served, int]] = []

        def now() -> float:
            return time.perf_counter() - t0

        while pending or active:
            if shared.error is not None:  # another lane failed: stop
                return
            t = now()
            # admission: arrived requests of this lane in order, while the
            # shared budget and slots allow
            while pending and (not realtime or pending[0].arrival <= t):
                req = pending[0]
                need = footprint_bytes(req.circuit)
                if not shared.try_admit(need):
                    break
                pending.popleft()
                dc = prepare(req.circuit, eng)
                active.append((dc, Served(req.rid, req.arrival if realtime else 0.0, admitted=t,
                                          lane=i), need))
            stats["max_active"] = max(stats["max_active"], len(active))
            if not active:
                if pending:
                    wait = pending[0].arrival - now() if realtime else 0.0
                    time.sleep(max(0.0005, min(wait, 0.01)))
                continue
            if self.fixed_ways is None and len(self.caps) > 0:
                eng.split_ways = shared.ways(i)
            tick(eng, [a[0] for a in active], self.svd)
            stats["ticks"] += 1
            finished = [a for a in active if a[0].done]
            if finished:
                stream.synchronize()
                tf = now()
                for dc, s, need in finished:
                    s.finished = tf
                    s.updates = dc.circuit.two_site_count()
                    if self.keep_results:
                        s.tensors = dc.mps.tensors
                    shared.release(need, s)
                active = [a for a in active if not a[0].done]
