"""Timed CPU baseline: the reference-faithful restatement on the host cores.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): called by bench.py's
cpu_baseline leg. The reference itself cannot run here (TensorNetwork 0.2.1
is absent and core.py:561 uses np.complex, removed in numpy 1.24), so the
timed path is oracle.mps_oracle.OracleMPS, which follows
apply_two_qudit_gate (core.py:950-1161): theta contraction, gate,
np.linalg.svd (LAPACK gesdd), truncation, S absorbed right, the reference's
dtype promotion (Haar gates are complex128, so tensors become complex128)
and the per-gate full norm of core.py:1160-1161.

Sample: the interior of the 64-qubit brickwork at the bond-dimension cap
(every interior bond 256, the shape from layer 8 on); site tensors are
seeded Gaussian tensors of that shape (the cost of an update depends only on
the shapes), and `updates` Haar two-site gates are applied on disjoint
interior pairs with maxsvals=256.
"""
from __future__ import annotations

import os
import platform
import time
from typing import Dict

import numpy as np

from .mps_oracle import OracleMPS, haar_unitary_from_rng


def blas_threads() -> int:
    try:
        from threadpoolctl import threadpool_info
        n = [p.get("num_threads", 0) for p in threadpool_info() if p.get("user_api") == "blas"]
        if n:
            return int(max(n))
    except Exception:
        pass
    return int(os.environ.get("OPENBLAS_NUM_THREADS", os.cpu_count() or 1))


def _capped_mps(n: int, d: int, cap: int, mode: str, track_norms: bool,
                rng: np.random.RandomState) -> OracleMPS:
    mps = OracleMPS(n, d, dtype_mode=mode, track_norms=track_norms)
    bonds = [1] + [min(d ** (i + 1), d ** (n - 1 - i), cap) for i in range(n - 1)] + [1]
    dt = np.complex64 if mode == "c64" else np.complex128
    for i in range(n):
        shape = (bonds[i], d, bonds[i + 1])
        t = (rng.randn(*shape) + 1j * rng.randn(*shape)) / np.sqrt(2 * bonds[i] * d)
        mps.tensors[i] = t.astype(dt)
    return mps


def time_updates(updates: int = 8, n: int = 64, d: int = 2, cap: int = 256,
                 mode: str = "reference", track_norms: bool = True, seed: int = 2,
                 threads: int = 0) -> Dict:
    """`updates` two-site updates at the cap shape; threads > 0 limits the
    BLAS/LAPACK threads for the timed part (threadpoolctl)."""
    if threads:
        from threadpoolctl import threadpool_limits
        with threadpool_limits(limits=threads, user_api="blas"):
            return time_updates(updates, n, d, cap, mode, track_norms, seed)
    rng = np.random.RandomState(seed)
    mps = _capped_mps(n, d, cap, mode, track_norms, rng)
    mid = n // 2
    pairs = [(mid - 2 * (updates // 2) + 2 * j, mid - 2 * (updates // 2) + 2 * j + 1)
             for j in range(updates)]
    gates = [haar_unitary_from_rng(d * d, rng).reshape(d, d, d, d) for _ in pairs]
    t0 = time.perf_counter()
    for (i, j), g in zip(pairs, gates):
        mps.apply_two_qudit_gate(g, i, j, maxsvals=cap)
    dt = time.perf_counter() - t0
    return {"updates": updates, "seconds": dt, "updates_per_s": updates / dt}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def time_plain16(n: int = 16, depth: int = 8, seed: int = 1) -> Dict:
    """The 16-qubit plain workload in full, reference-faithful: every layer a
    Haar U(2) (gates.py:269-285 recipe) on each qubit, then CNOT on (i, i+1)
    at offset (layer mod 2), no truncation, tensors complex64 until the
    float64 CNOT promotes them, per-gate full norm (core.py:1160-1161)."""
    rng = np.random.RandomState(seed)
    layers = []
    cnot = np.array([[1.0, 0, 0, 0], [0, 1.0, 0, 0], [0, 0, 0, 1.0], [0, 0, 1.0, 0]])
    for t in range(depth):
        layers.append([((q,), haar_unitary_from_rng(2, rng)) for q in range(n)])
        layers.append([((i, i + 1), cnot) for i in range(t % 2, n - 1, 2)])
    mps = OracleMPS(n, 2, dtype_mode="reference", track_norms=True)
    t0 = time.perf_counter()
    for layer in layers:
        mps.apply_ops(layer)
    dt = time.perf_counter() - t0
    return {"seconds": dt, "updates": sum(len(l) for l in layers[1::2])}
