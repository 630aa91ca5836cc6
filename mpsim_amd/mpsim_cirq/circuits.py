"""Conversion of circuit operations to MPS operations.

Mirrors mpsim/mpsim_cirq/circuits.py:12-43: a k-qubit operation becomes an
MPSOperation whose tensor is the operation's unitary reshaped to [2]*2k.
Cirq is optional: any object with `qubits`, `_has_unitary_()` and
`_unitary_()` (a cirq.GateOperation included) is accepted.
"""
from __future__ import annotations

from typing import Dict

import numpy as np

from ..core import CannotConvertToMPSOperation, MPSOperation
from ..gates import Node


def mps_operation_from_gate_operation(gate_operation, qudit_to_index_map: Dict) -> MPSOperation:
    num_qudits = len(gate_operation.qubits)
    qudit_dimension = 2
    qudit_indices = tuple(qudit_to_index_map[q] for q in gate_operation.qubits)
    has = getattr(gate_operation, "_has_unitary_", None)
    if has is None or not has():
        raise CannotConvertToMPSOperation(
            f"Cannot convert operation {gate_operation} into an MPS Operation"
            " because the operation does not have a unitary.")
    tensor = np.reshape(gate_operation._unitary_(), [qudit_dimension] * 2 * num_qudits)
    return MPSOperation(Node(tensor), qudit_indices, qudit_dimension)
