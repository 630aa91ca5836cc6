"""Layer-batched execution of many circuits on one GPU.

The reference applies one gate at a time (core.py:1245, "TODO:
Parallelize"). Here every circuit advances one step of its schedule per
tick (circuits.schedule: a layer of nearest-neighbour ops on disjoint
qudits is one step; non-adjacent ops become the reference's SWAP networks,
overlapping ops keep their order), and all the one-site gates of a tick
(over all circuits) go to one grouped launch and all the two-site updates of
a tick (disjoint pairs of every circuit's current step) go to one grouped
theta + Jacobi SVD launch sequence. Gate matrices are uploaded once per
circuit and stay resident.

`run_plain` is the one-gate-at-a-time path through the public MPS API, the
baseline the batched form is compared with.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np

from .circuits import Circuit, flip_pair, schedule, swap_matrix
from .core import MPS
from .engine import Engine, SvdConfig, TwoSiteItem

MAX_GROUP = 65535


@dataclass
class DeviceCircuit:
    """A circuit whose gates are resident on the device, with its MPS."""
    circuit: Circuit
    mps: MPS
    gates: List[List[object]]          # per layer, per op: device matrix (pair order applied)
    t: int = 0                         # next step of the schedule
    discarded: List[object] = field(default_factory=list)
    # per step of circuits.schedule: [(site, gate address)] one-site ops and
    # [(i, gate)] two-site updates on (i, i+1)
    plan: List[tuple] = field(default_factory=list)

    @property
    def done(self) -> bool:
        return self.t >= len(self.plan)


def op_matrix(qudits, mat, d: int) -> np.ndarray:
    """The complex64 matrix an op's update uses: a pair given as (j, i) with
    j > i acts on the sites in the order (i, j) with its edges reordered
    (core.py:1031-1033)."""
    m = np.asarray(mat)
    if len(qudits) == 2 and qudits[0] > qudits[1]:
        m = flip_pair(m, d)
    return np.ascontiguousarray(m, dtype=np.complex64)


def step_plan(circuit: Circuit, gates: List[List[object]], swap_gate) -> List[tuple]:
    """Per step of circuits.schedule: the one-site ops as (site, gate
    address) and the two-site updates as (left site, gate); SWAP hops use
    `swap_gate`. The gate matrices stay alive in DeviceCircuit.gates."""
    steps, _ = schedule(circuit)
    plan = []
    for st in steps:
        one = [(site, int(gates[t][o].data_ptr())) for site, t, o in st.ones]
        two = [(i, gates[t][o] if o >= 0 else swap_gate) for i, t, o, _ in st.twos]
        plan.append((one, two))
    return plan


def upload_gates(engine: Engine, circuit: Circuit) -> List[List[object]]:
    """One host->device copy of every gate matrix of the circuit (pairs
    given high-to-low flipped, see op_matrix)."""
    torch = engine.torch
    d = circuit.d
    mats = [op_matrix(q, m, d).ravel() for layer in circuit.layers for q, m in layer]
    if not mats:
        return [[] for _ in circuit.layers]
    flat = torch.from_numpy(np.concatenate(mats)).to(engine.dev)
    out, pos = [], 0
    for layer in circuit.layers:
        row = []
        for q, m in layer:
            dim = d ** len(q)
            row.append(flat[pos:pos + dim * dim].view(dim, dim))
            pos += dim * dim
        out.append(row)
    return out


def prepare(circuit: Circuit, engine: Optional[Engine] = None) -> DeviceCircuit:
    """Uploads the circuit's gates and plans its steps (raises ValueError on
    an op with qudits out of range, or not one or two distinct qudits)."""
    eng = engine or Engine.get()
    schedule(circuit)  # validates before anything is allocated
    mps = MPS(circuit.n, circuit.d, device=eng.dev)
    gates = upload_gates(eng, circuit)
    plan = step_plan(circuit, gates, eng.gate(swap_matrix(circuit.d)))
    return DeviceCircuit(circuit, mps, gates, plan=plan)


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
    region (the FP32 Gram, then X1 = Y X) or the QR-preconditioned one's
    X0 (Cp x R), ring of eight Householder panels (2048 Cp bytes),
    triangular factors and column order."""
    cp = C + 31
    return (16 * cp * (R + cp + 4) + 32 * cp * (cp if cp > R else R) + 16 * cp * 64 + 2048 * cp
            + 4 * R + 35840)


def _ws_estimate(it: TwoSiteItem) -> int:
    """Upper bound of one problem's Jacobi workspace bytes."""
    cl, d, _ = it.a.shape
    cr = it.b.shape[2]
    C, R = (d * cr, d * cl) if it.absorb_left else (d * cl, d * cr)
    return ws_bound(C, R)


def tick(engine: Engine, dcs: Sequence[DeviceCircuit], svd: Optional[SvdConfig] = None,
         keep_discarded: bool = False, ws_budget: int = 8 << 30, want_s: bool = False,
         observer: Optional[Callable] = None) -> int:
    """Advances every unfinished circuit by one step; returns the number of
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
    est = [_ws_estimate(it) for it in two] if len(two) > 1 else [0] * len(two)
    start = 0
    while start < len(two):
        end, acc = start, 0
        while end < len(two) and (end == start or acc + est[end] <= ws_budget):
            acc += est[end]
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
