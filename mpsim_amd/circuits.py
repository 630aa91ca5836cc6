"""Seeded synthetic circuits (the benchmark and serving workloads).

A circuit is a list of layers; a layer is a list of ops (qudits, matrix),
matrix being a d^k x d^k complex128 unitary in the reference's edge
convention (row index (o_i, o_j), column index (p_i, p_j) for qudits (i, j)).
Haar gates use mpsim/gates.py:269-284's recipe on one RandomState stream per
circuit. The two-qudit ops of a layer act on disjoint nearest-neighbour
pairs, so a layer is one batched two-site launch.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .gates import haar_unitary

Op = Tuple[Tuple[int, ...], np.ndarray]

CNOT = np.array([[1.0, 0, 0, 0], [0, 1.0, 0, 0], [0, 0, 0, 1.0], [0, 0, 1.0, 0]],
                dtype=np.complex128)
SWAP = np.array([[1.0, 0, 0, 0], [0, 0, 1.0, 0], [0, 1.0, 0, 0], [0, 0, 0, 1.0]],
                dtype=np.complex128)


def swap_matrix(d: int) -> np.ndarray:
    """SWAP of two d-level qudits as a d^2 x d^2 matrix (gates.py:204-212 for d = 2)."""
    m = np.zeros((d * d, d * d), dtype=np.complex128)
    for a in range(d):
        for b in range(d):
            m[b * d + a, a * d + b] = 1.0
    return m


def flip_pair(mat: np.ndarray, d: int) -> np.ndarray:
    """The gate of a pair given as (j, i) with j > i, for the sites in the
    order (i, j): edges reordered [1, 0, 3, 2] as core.py:1031-1033 does,
    i.e. SWAP G SWAP."""
    return np.asarray(mat).reshape(d, d, d, d).transpose(1, 0, 3, 2).reshape(d * d, d * d)


@dataclass
class Step:
    """One grouped launch's worth of work: one-site gates and two-site
    updates on disjoint sites. ones: (site, layer, op); twos: (left site i,
    layer, op, flipped) for an update on (i, i+1) with that op's gate
    (flipped: the op was given as (i+1, i)), or (i, layer, -1, False) for a
    SWAP hop of a non-adjacent op's network."""
    ones: List[Tuple[int, int, int]] = field(default_factory=list)
    twos: List[Tuple[int, int, int, bool]] = field(default_factory=list)


@dataclass
class Circuit:
    n: int
    d: int
    layers: List[List[Op]]
    maxsvals: Optional[int] = None
    seed: int = 0
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def ops(self) -> List[Op]:
        return [op for layer in self.layers for op in layer]

    @property
    def depth(self) -> int:
        return len(self.layers)

    def two_site_count(self) -> int:
        return sum(1 for op in self.ops if len(op[0]) == 2)


def brickwork_cnot_u2(n: int = 16, depth: int = 8, seed: int = 1) -> Circuit:
    """Plain method, CPU-runnable: every layer is a Haar U(2) on each qubit,
    then CNOT(i, i+1) on pairs starting at offset (layer mod 2);
    maxsvals=None (no truncation)."""
    rng = np.random.RandomState(seed)
    layers = []
    for t in range(depth):
        one = [((q,), haar_unitary(2, rng)) for q in range(n)]
        two = [((i, i + 1), CNOT.copy()) for i in range(t % 2, n - 1, 2)]
        layers.append(one)
        layers.append(two)
    return Circuit(n, 2, layers, None, seed, f"brickwork_cnot_u2_n{n}_d{depth}")


def brickwork_haar(n: int = 64, depth: int = 20, d: int = 2, maxsvals: Optional[int] = 256,
                   seed: int = 2) -> Circuit:
    """Layers of Haar-random d^2 x d^2 unitaries on (i, i+1), offset
    alternating between 0 and 1 (64 qubits / depth 20 / maxsvals 256 is the
    single-large-circuit workload; 48 qutrits / depth 16 / 1024 the precision
    stress)."""
    rng = np.random.RandomState(seed)
    layers = []
    for t in range(depth):
        layers.append([((i, i + 1), haar_unitary(d * d, rng)) for i in range(t % 2, n - 1, 2)])
    return Circuit(n, d, layers, maxsvals, seed, f"brickwork_haar_n{n}_d{depth}_q{d}_chi{maxsvals}")


def random_pairs_circuit(seed: int, n: int, depth: int, maxsvals: int) -> Circuit:
    """One circuit of the uneven batch: each layer is a Haar U(2) on every
    qubit, then Haar 4x4 gates on a random set of disjoint nearest-neighbour
    pairs (bonds scanned left to right, each taken with probability 1/2 unless
    its left neighbour bond was taken)."""
    rng = np.random.RandomState(seed)
    layers = []
    for _ in range(depth):
        layers.append([((q,), haar_unitary(2, rng)) for q in range(n)])
        two = []
        i = 0
        while i < n - 1:
            if rng.rand() < 0.5:
                two.append(((i, i + 1), haar_unitary(4, rng)))
                i += 2
            else:
                i += 1
        layers.append(two)
    return Circuit(n, 2, layers, maxsvals, seed, f"rand_pairs_s{seed}")


def uneven_batch(count: int = 512, base_seed: int = 3_000_000) -> List[Circuit]:
    """Batched uneven circuits: n ~ U[12,64], depth ~ U[4,24], chi cap from
    {32,64,128,256}; circuit c is seeded with base_seed + c."""
    out = []
    for c in range(count):
        r = np.random.RandomState(base_seed + c)
        n = int(r.randint(12, 65))
        depth = int(r.randint(4, 25))
        cap = int(r.choice([32, 64, 128, 256]))
        out.append(random_pairs_circuit(base_seed + c + 10_000_000, n, depth, cap))
    return out


@dataclass
class Request:
    rid: int
    arrival: float
    circuit: Circuit


def poisson_requests(rate: float = 200.0, duration: float = 60.0, seed: int = 5,
                     base_seed: int = 5_000_000, limit: Optional[int] = None) -> List[Request]:
    """Serving stream: Poisson arrivals at `rate`/s for `duration` s of qubit
    brickwork circuits, n ~ U[16,80], depth ~ U[4,32], maxsvals from
    {64,128,256,512}; request c is seeded with base_seed + c."""
    arr = np.random.RandomState(seed)
    t = 0.0
    out = []
    c = 0
    while True:
        t += float(arr.exponential(1.0 / rate))
        if t >= duration or (limit is not None and c >= limit):
            break
        r = np.random.RandomState(base_seed + c)
        n = int(r.randint(16, 81))
        depth = int(r.randint(4, 33))
        cap = int(r.choice([64, 128, 256, 512]))
        circ = brickwork_haar(n, depth, 2, cap, seed=base_seed + c + 10_000_000)
        out.append(Request(c, t, circ))
        c += 1
    return out


# ---------------------------------------------------------------- planning
def schedule(circuit: Circuit) -> Tuple[List[Step], List[int]]:
    """Expands the layers into steps of grouped launches, with the same
    result as applying every op in order (core.py:1221-1247).

    A two-qudit op on non-adjacent qudits (i, j), i < j (an op given as
    (j, i) is flipped first), becomes the reference's SWAP network
    (core.py:1035-1043, 1154-1219): SWAP hops (i, i+1) .. (j-2, j-1) carry
    qudit i next to j, the gate acts on (j-1, j), and hops (j-2, j-1) ..
    (i, i+1) carry it back; every hop is a full two-site update with the
    circuit's truncation. Each op covers the sites from its lowest to its
    highest qudit; an op starts once every earlier op of the layer that
    covers one of its sites has finished, so ops on disjoint sites share
    steps and overlapping ones keep their order. A layer of disjoint
    nearest-neighbour ops is one step. Layers follow each other.

    Returns (steps, layer_end) with layer_end[t] = number of steps after
    layer t."""
    steps: List[Step] = []
    ends: List[int] = []
    n = circuit.n
    for t, layer in enumerate(circuit.layers):
        base = len(steps)
        free: Dict[int, int] = {}  # site -> first step (relative to base) it is free
        for o, (q, _) in enumerate(layer):
            q = tuple(int(x) for x in q)
            if len(q) not in (1, 2) or any(not 0 <= x < n for x in q):
                raise ValueError(f"layer {t} op {o}: qudits {q} not one or two of 0..{n - 1}")
            if len(q) == 1:
                lo = hi = q[0]
                chain = [("one", q[0], o, False)]
            else:
                i, j = q
                if i == j:
                    raise ValueError(f"layer {t} op {o}: the two qudits are the same")
                flipped = j < i
                lo, hi = min(i, j), max(i, j)
                there = [("two", s_, -1, False) for s_ in range(lo, hi - 1)]
                chain = there + [("two", hi - 1, o, flipped)] + there[::-1]
            start = max(free.get(x, 0) for x in range(lo, hi + 1))
            while len(steps) < base + start + len(chain):
                steps.append(Step())
            for k, (kind, site, op, fl) in enumerate(chain):
                st = steps[base + start + k]
                if kind == "one":
                    st.ones.append((site, t, op))
                else:
                    st.twos.append((site, t, op, fl))
            for x in range(lo, hi + 1):
                free[x] = start + len(chain)
        ends.append(len(steps))
    return steps, ends


def bond_plan_steps(circuit: Circuit) -> List[List[int]]:
    """Bond dimensions after every step of schedule(circuit): an update on
    (i, i+1) sets bond i to k = min(d*chi_left, d*chi_right, maxsvals)
    because the reference keeps zero singular values (TensorNetwork split
    without an error cut)."""
    n, d = circuit.n, circuit.d
    bonds = [1] * (n + 1)  # bonds[0] and bonds[n] are the dummy end bonds
    out = []
    steps, _ = schedule(circuit)
    for st in steps:
        for i, _, _, _ in st.twos:
            k = min(d * bonds[i], d * bonds[i + 2])
            if circuit.maxsvals is not None:
                k = min(k, int(circuit.maxsvals))
            bonds[i + 1] = k
        out.append(bonds[1:n])
    return [list(b) for b in out]


def bond_plan(circuit: Circuit) -> List[List[int]]:
    """Bond dimensions after every layer (bond_plan_steps at each layer's
    last step), computed on the host from the circuit alone."""
    per_step = bond_plan_steps(circuit)
    _, ends = schedule(circuit)
    first = [1] * (circuit.n - 1)
    return [list(per_step[e - 1]) if e else list(first) for e in ends]


def peak_footprint_bytes(circuit: Circuit) -> int:
    """max over steps of sum_l chi_l * d * chi_{l+1} * 8 bytes (complex64)."""
    d = circuit.d
    best = circuit.n * d * 8
    for b in bond_plan_steps(circuit):
        full = [1] + b + [1]
        tot = sum(full[i] * d * full[i + 1] for i in range(circuit.n)) * 8
        best = max(best, tot)
    return best


def two_site_flops(chi_l: int, chi_m: int, chi_r: int, d: int, k: int) -> float:
    """Algorithmic real flops of one two-site update: theta contraction,
    gate, and the complex Golub-Reinsch SVD count with thin U and V
    (4 * (4 p^2 q + 8 p q^2 + 9 q^3), Golub & Van Loan 4th ed. Fig. 8.6.1)."""
    M, N = d * chi_l, d * chi_r
    p, q = max(M, N), min(M, N)
    return (8.0 * d * d * chi_l * chi_m * chi_r + 8.0 * d ** 4 * chi_l * chi_r
            + 4.0 * (4.0 * p * p * q + 8.0 * p * q * q + 9.0 * q ** 3))


def two_site_bytes(chi_l: int, chi_m: int, chi_r: int, d: int, k: int) -> float:
    """Algorithmic HBM bytes of one two-site update (inputs, gate, outputs,
    singular values)."""
    return 8.0 * (chi_l * d * chi_m + chi_m * d * chi_r + d ** 4 + chi_l * d * k + k * d * chi_r) + 4.0 * k


def circuit_work(circuit: Circuit) -> Tuple[float, float, int]:
    """(flops, bytes, updates) of every two-site update of a circuit, SWAP
    hops of non-adjacent ops included."""
    n, d = circuit.n, circuit.d
    bonds = [1] * (n + 1)
    fl = by = 0.0
    cnt = 0
    steps, _ = schedule(circuit)
    for st in steps:
        for i, _, _, _ in st.twos:
            cl, cm, cr = bonds[i], bonds[i + 1], bonds[i + 2]
            k = min(d * cl, d * cr)
            if circuit.maxsvals is not None:
                k = min(k, int(circuit.maxsvals))
            fl += two_site_flops(cl, cm, cr, d, k)
            by += two_site_bytes(cl, cm, cr, d, k)
            bonds[i + 1] = k
            cnt += 1
    return fl, by, cnt
