"""MPSimulator: the simulate(options={"maxsvals": ...}) surface.

Mirrors mpsim/mpsim_cirq/simulator.py:14-88. The reference subclasses
cirq.SimulatesFinalState; here the same methods are provided directly so
that the simulator works with or without cirq installed. Accepted programs:
  - a cirq.Circuit (when cirq is importable): parameters are resolved,
    qubits ordered with QubitOrder.DEFAULT, ops converted by _unitary_();
  - any circuit object with all_qubits() and all_operations() (a cirq
    circuit without cirq installed, or a stand-in): qubits in sorted order
    (or qubit_order), ops converted as circuits.py:26-43 does;
  - a mpsim_amd.circuits.Circuit (seeded synthetic workloads), which may
    carry its own truncation cap (maxsvals) used when the options give none;
    with params, a parameter sweep: each entry is a Circuit of the same
    structure or the list of its op matrices, and all entries run as one
    batched CircuitProgram (kept and reused, via update_gates, for later
    sweeps of the same structure);
  - a sequence of MPSOperation, or of (qudit indices, matrix) pairs.
The options dict is forwarded to MPS.apply as keyword arguments
(simulator.py:86), e.g. {"maxsvals": 64} or {"fraction": 0.5}.
"""
from __future__ import annotations

from typing import Any, List, Optional, Sequence

import numpy as np

from ..circuits import Circuit
from ..core import MPS, MPSOperation
from ..gates import Node
from .circuits import mps_operation_from_gate_operation

try:  # optional dependency
    import cirq  # type: ignore
except Exception:  # pragma: no cover - cirq is not installed in this image
    cirq = None


def _ops_from_pairs(pairs, d: int) -> List[MPSOperation]:
    out = []
    for qudits, mat in pairs:
        k = len(qudits)
        out.append(MPSOperation(Node(np.asarray(mat).reshape([d] * 2 * k)), tuple(qudits), d))
    return out


class MPSimulator:
    """Final-state MPS simulator (simulator.py:14-30)."""

    def __init__(self, options: Optional[dict] = None, device=None):
        self._options = {} if options is None else options
        self._device = device
        self._programs: dict = {}  # structure -> CircuitProgram of parameter sweeps

    @property
    def options(self) -> dict:
        return self._options

    def simulate(self, program, param_resolver=None, qubit_order=None, initial_state=None) -> MPS:
        return self.simulate_sweep(program, param_resolver, qubit_order, initial_state)[0]

    def simulate_sweep(self, program, params=None, qubit_order=None,
                       initial_state: Any = None) -> List[MPS]:
        """simulator.py:32-88: one MPS per parameter resolver."""
        if cirq is not None and isinstance(program, cirq.Circuit):
            return self._simulate_cirq(program, params, qubit_order)
        if isinstance(program, Circuit) and params is not None:
            return self._sweep(program, params)
        if hasattr(program, "all_qubits") and hasattr(program, "all_operations"):
            qubits = list(program.all_qubits())
            ordered = list(qubit_order) if qubit_order is not None else sorted(qubits)
            index = {q: i for i, q in enumerate(ordered)}
            mps = MPS(nqudits=len(qubits), device=self._device)
            mps.apply([mps_operation_from_gate_operation(op, index)
                       for op in program.all_operations()], **self._options)
            return [mps]
        if isinstance(program, Circuit):
            mps = MPS(program.n, program.d, device=self._device)
            opts = dict(self._options)
            if program.maxsvals is not None and "maxsvals" not in opts and "fraction" not in opts:
                opts["maxsvals"] = program.maxsvals
            mps.apply(_ops_from_pairs(program.ops, program.d), **opts)
            return [mps]
        if isinstance(program, (list, tuple)):
            ops = list(program)
            if ops and not isinstance(ops[0], MPSOperation):
                ops = _ops_from_pairs(ops, 2)
            qudits = sorted({q for op in ops for q in op.qudit_indices})
            n = (max(qudits) + 1) if qudits else 0
            d = ops[0].qudit_dimension if ops else 2
            mps = MPS(nqudits=n, qudit_dimension=d, device=self._device)
            mps.apply(ops, **self._options)
            return [mps]
        raise ValueError(
            f"Program is of type {type(program)} but should be a cirq.Circuit, a "
            "mpsim_amd.circuits.Circuit or a sequence of operations.")

    def _sweep(self, program: Circuit, params) -> List[MPS]:
        """One final state per entry of params (same structure as program,
        new gate values), all run as one batched, captured schedule."""
        variants = []
        for p in params:
            if isinstance(p, Circuit):
                c = p
            else:
                mats = list(p)
                if len(mats) != len(program.ops):
                    raise ValueError(f"a parameter set has {len(mats)} matrices for "
                                     f"{len(program.ops)} ops")
                it = iter(mats)
                c = Circuit(program.n, program.d, [[(q, next(it)) for q, _ in layer]
                                                   for layer in program.layers],
                            program.maxsvals, program.seed, program.name)
            if [[q for q, _ in L] for L in c.layers] != [[q for q, _ in L] for L in program.layers] \
                    or (c.n, c.d) != (program.n, program.d):
                raise ValueError("every parameter set must have the program's structure")
            variants.append(c)
        if not variants:
            return []
        opts = dict(self._options)
        if set(opts) - {"maxsvals"}:  # options the captured schedule does not take
            return [self.simulate_sweep(c)[0] for c in variants]
        cap = opts.get("maxsvals", program.maxsvals)
        cap = None if cap is None else int(cap)
        variants = [Circuit(c.n, c.d, c.layers, cap, c.seed, c.name) for c in variants]
        from ..program import CircuitProgram
        key = (program.n, program.d, cap, len(variants),
               tuple(tuple(tuple(q) for q, _ in L) for L in program.layers))
        prog = self._programs.get(key)
        if prog is None:
            from ..engine import Engine
            prog = CircuitProgram(variants, Engine.get(self._device))
            self._programs[key] = prog
        else:
            prog.update_gates(variants)
        prog.run()
        return [mps.copy() for mps in prog.states()]

    def _simulate_cirq(self, program, params, qubit_order) -> List[MPS]:  # pragma: no cover
        resolvers = cirq.to_resolvers(params) if params is not None else [cirq.ParamResolver({})]
        order = cirq.QubitOrder.DEFAULT if qubit_order is None else qubit_order
        results = []
        for prs in resolvers:
            solved = cirq.resolve_parameters(program, prs)
            ordered = cirq.QubitOrder.as_qubit_order(order).order_for(solved.all_qubits())
            index = {q: i for i, q in enumerate(ordered)}
            mps = MPS(nqudits=len(solved.all_qubits()), device=self._device)
            ops = [mps_operation_from_gate_operation(op, index) for op in solved.all_operations()]
            mps.apply(ops, **self._options)
            results.append(mps)
        return results
