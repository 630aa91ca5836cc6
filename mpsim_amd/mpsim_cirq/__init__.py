"""Circuit front end (mpsim/mpsim_cirq counterpart)."""
from .circuits import mps_operation_from_gate_operation
from .simulator import MPSimulator

__all__ = ["MPSimulator", "mps_operation_from_gate_operation"]
