"""Layer-batched execution of many circuits on one GPU.

The reference applies one gate at a time (core.py:1245, "TODO:
Parallelize"). Here every circuit advances one layer per tick, and all the
one-site gates of a tick (over all circuits) go to one grouped launch and
all the two-site updates of a tick (disjoint pairs of every circuit's
current layer) go to one grouped theta + Jacobi SVD launch sequence. Gate
matrices are uploaded once per circuit and stay resident.

`run_plain` is the one-gate-at-a-time path through the public MPS API, the
baseline the batched form is compared with.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np

from .circuits import Circuit
from .core import MPS
from .engine import Engine, SvdConfig, TwoSiteItem

MAX_GROUP = 65535


@dataclass
class DeviceCircuit:
    """A circuit whose gates are resident on the device, with its MPS."""
    circuit: Circuit
    mps: MPS
    gates: List[List[object]]          # per layer, per op: device matrix
    t: int = 0                         # next layer
    discarded: List[object] = field(default_factory=list)
    # per layer: [(site, gate)] one-site ops and [(i, gate, maxsvals)] pairs (i, i+1)
    plan: List[tuple] = field(default_factory=list)

    @property
    def done(self) -> bool:
        return self.t >= len(self.circuit.layers)


def check_layer(layer, n: int, where: str = "batched execution") -> None:
    """A layer runs as one grouped launch of its one-site gates followed by
    one of its two-site updates, so its ops must act on disjoint qudits (in
    range) and every pair must be nearest-neighbour (i, i+1)."""
    seen = set()
    for q, _ in layer:
        for s in q:
            if not 0 <= s < n:
                raise ValueError(f"{where}: qudit {s} out of range for {n} qudits")
            if s in seen:
                raise ValueError(f"{where}: qudit {s} appears in more than one op of a layer; "
                                 "split the layer so each qudit is touched once")
            seen.add(s)
        if len(q) == 2 and q[1] != q[0] + 1:
            raise ValueError(f"{where} takes nearest-neighbour pairs (i, i+1), got {tuple(q)}")
        if len(q) not in (1, 2):
            raise ValueError(f"{where}: ops act on one or two qudits, got {tuple(q)}")


def _layer_plan(circuit: Circuit, gates: List[List[object]]) -> List[tuple]:
    """Per layer: the one-site ops as (site, gate address) and the two-site
    ops as (left site, gate); checks each layer with check_layer. The gate
    matrices stay alive in DeviceCircuit.gates."""
    plan = []
    for layer, glayer in zip(circuit.layers, gates):
        check_layer(layer, circuit.n)
        one, two = [], []
        for (q, _), g in zip(layer, glayer):
            if len(q) == 1:
                one.append((q[0], int(g.data_ptr())))
            else:
                two.append((q[0], g))
        plan.append((one, two))
    return plan


def upload_gates(engine: Engine, circuit: Circuit) -> List[List[object]]:
    """One host->device copy of every gate matrix of the circuit."""
    torch = engine.torch
    mats = [np.asarray(m, dtype=np.complex64).ravel() for layer in circuit.layers for _, m in layer]
    if not mats:
        return [[] for _ in circuit.layers]
    flat = torch.from_numpy(np.concatenate(mats)).to(engine.dev)
    out, pos = [], 0
    for layer in circuit.layers:
        row = []
        for q, m in layer:
            dim = circuit.d ** len(q)
            row.append(flat[pos:pos + dim * dim].view(dim, dim))
            pos += dim * dim
        out.append(row)
    return out


def prepare(circuit: Circuit, engine: Optional[Engine] = None) -> DeviceCircuit:
    eng = engine or Engine.get()
    mps = MPS(circuit.n, circuit.d, device=eng.dev)
    gates = upload_gates(eng, circuit)
    return DeviceCircuit(circuit, mps, gates, plan=_layer_plan(circuit, gates))


def reset(dc: DeviceCircuit) -> None:
    """Back to |0...0> (gates stay resident)."""
    dc.mps = MPS(dc.circuit.n, dc.circuit.d, device=dc.mps.device)
    dc.t = 0
    dc.discarded = []


def ws_bound(C: int, R: int) -> int:
    """Upper bound of the Jacobi workspace bytes of one C x R problem:
    complex FP64 working rows [X | Y] (padded rows Cp <= C + 31, row stride
    <= R + Cp + 3, norms and order 12 Cp), plus the Gram method's
    Cp x max(Cp, R) Gram, two sets of 32x32 pivot rotations (16 Cp x 64
    bytes) and flags, and the mixed Gram method's second Cp x max(Cp, R)
    region (the FP32 Gram, then X1 = Y X)."""
    cp = C + 31
    return 16 * cp * (R + cp + 4) + 32 * cp * max(cp, R) + 16 * cp * 64 + 2560


def _ws_estimate(it: TwoSiteItem) -> int:
    """Upper bound of one problem's Jacobi workspace bytes."""
    cl, d, _ = it.a.shape
    cr = it.b.shape[2]
    C, R = (d * cr, d * cl) if it.absorb_left else (d * cl, d * cr)
    return ws_bound(C, R)


def tick(engine: Engine, dcs: Sequence[DeviceCircuit], svd: Optional[SvdConfig] = None,
         keep_discarded: bool = False, ws_budget: int = 8 << 30, want_s: bool = False,
         observer: Optional[Callable] = None) -> int:
    """Advances every unfinished circuit by one layer; returns the number of
    two-site updates performed. The two-site updates of the tick go to the
    library in as few calls as the workspace budget allows.

    want_s: every update also returns its kept singular values.
    observer(dc, i, item, result): called for every two-site update on sites
    (i, i+1) of circuit dc after its library call, while the input tensors
    (item.a, item.b) are still alive and before the MPS takes the outputs."""
    one, two, owners = [], [], []
    for dc in dcs:
        if dc.t >= len(dc.plan):
            continue
        c = dc.circuit
        ones, twos = dc.plan[dc.t]
        tensors = dc.mps._tensors
        d = c.d
        cap = None if c.maxsvals is None else int(c.maxsvals)
        for q, gp in ones:
            t = tensors[q]
            cl, _, cr = t.shape
            if cl and cr:  # an empty bond (maxsvals=0) leaves nothing to update
                p = t.data_ptr()  # site tensors are contiguous by construction
                one.append((p, p, gp, cl, cr, d, 0))
        for i, g in twos:
            a, b = tensors[i], tensors[i + 1]
            M, N = d * a.shape[0], d * b.shape[2]
            k = min(M, N) if cap is None else min(M, N, cap)
            two.append(TwoSiteItem(a, b, g, k, want_s=want_s))
            owners.append((dc, i))
        dc.t += 1
    for s in range(0, len(one), MAX_GROUP):
        engine.one_site_rows(one[s:s + MAX_GROUP])
    start = 0
    while start < len(two):
        end, acc = start, 0
        while end < len(two) and (end == start or acc + _ws_estimate(two[end]) <= ws_budget):
            acc += _ws_estimate(two[end])
            end += 1
        res = engine.two_site(two[start:end], svd)
        for (dc, i), it, r in zip(owners[start:end], two[start:end], res):
            if observer is not None:
                observer(dc, i, it, r)
            dc.mps._tensors[i] = r.a_out
            dc.mps._tensors[i + 1] = r.b_out
            if keep_discarded:
                dc.discarded.append(r.discarded)
        start = end
    return len(two)


def run_batched(dcs: Sequence[DeviceCircuit], engine: Optional[Engine] = None,
                svd: Optional[SvdConfig] = None, **tick_kw) -> int:
    """Ticks until every circuit is done (tick_kw go to every tick)."""
    eng = engine or Engine.get()
    total = 0
    while any(not dc.done for dc in dcs):
        total += tick(eng, dcs, svd, **tick_kw)
    return total


def run_plain(circuit: Circuit, device=None) -> MPS:
    """One gate per call through the public API (the unbatched baseline)."""
    mps = MPS(circuit.n, circuit.d, device=device)
    d = circuit.d
    kw = {} if circuit.maxsvals is None else {"maxsvals": int(circuit.maxsvals)}
    for q, m in circuit.ops:
        if len(q) == 1:
            mps.apply_one_qudit_gate(m, q[0])
        else:
            mps.apply_two_qudit_gate(np.asarray(m).reshape([d] * 4), q[0], q[1], **kw)
    return mps
