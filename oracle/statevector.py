"""FP64 dense state-vector simulator: the independent check for the oracle.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py). Textbook gate application
on a d^n complex128 vector, site 0 the most significant digit (the index order
the reference's wavefunction uses, core.py:483-500, pinned by
core_test.py:1189-1194). It shares no code with the MPS restatement.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np


def zero_state(n: int, d: int = 2) -> np.ndarray:
    psi = np.zeros(d ** n, dtype=np.complex128)
    psi[0] = 1.0
    return psi


def apply_gate(psi: np.ndarray, gate: np.ndarray, qudits: Sequence[int], n: int,
               d: int = 2) -> np.ndarray:
    """psi <- (G acting on `qudits`, in that order) psi.

    `gate` is a (d^k x d^k) matrix whose row/column index is the big-endian
    number formed by the k qudits in the order given (so for a two-qudit gate
    on (i, j), G[(o_i, o_j), (p_i, p_j)]), which is the reference's edge
    convention (core.py:986-992).
    """
    k = len(qudits)
    g = np.asarray(gate, dtype=np.complex128).reshape([d] * (2 * k))
    t = psi.reshape([d] * n)
    # contract gate input legs (k..2k-1) with the state's qudit axes
    t = np.tensordot(g, t, axes=(list(range(k, 2 * k)), list(qudits)))
    # result axes: gate outputs (k of them) then remaining state axes in order
    rest = [a for a in range(n) if a not in qudits]
    order = list(qudits) + rest
    inv = np.argsort(order)
    t = np.transpose(t, inv)
    return t.reshape(d ** n)


def run(ops, n: int, d: int = 2) -> np.ndarray:
    """ops: iterable of (qudits tuple, matrix)."""
    psi = zero_state(n, d)
    for qudits, mat in ops:
        psi = apply_gate(psi, mat, qudits, n, d)
    return psi


def density_matrix(psi: np.ndarray, indices: Sequence[int], n: int, d: int = 2) -> np.ndarray:
    """Partial trace of |psi><psi| keeping `indices`, in the order given:
    rho[(s_i1 .. s_im), (t_i1 .. t_im)] = sum over the other digits of
    psi[.. s ..] conj(psi[.. t ..]) (the convention of the density matrices
    the reference's tests compare against, core_test.py:1635-1711)."""
    idx = [int(i) for i in indices]
    rest = [a for a in range(n) if a not in idx]
    t = np.asarray(psi, dtype=np.complex128).reshape([d] * n)
    m = np.transpose(t, idx + rest).reshape(d ** len(idx), -1)
    return m @ m.conj().T


def marginals(psi: np.ndarray, n: int, d: int = 2) -> np.ndarray:
    """p[i, s]: probability of digit s on site i (squared amplitudes summed
    over the other digits)."""
    pr = (np.abs(np.asarray(psi)) ** 2).reshape([d] * n)
    return np.stack([pr.sum(axis=tuple(a for a in range(n) if a != i)) for i in range(n)])
