"""GPU tests of the gate-application error paths.

An MPS holds its site tensors on the device, so these need a GPU although
every check raises before a kernel runs. Restates the reference's error
tests (core_test.py:475-485 identical indices, 699-712 node moves against
their direction, 1028-1046 inner products of mismatched states) and the
checks of core.py:950-1030 (two-qudit gates), 753-790 (one-qudit gates) and
1249-1276 (operations on three or more qudits); a rejected call leaves the
state as it was.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_bad_gates_and_operations_raise_and_leave_the_state():
    import mpsim_amd as m
    mps = m.MPS(4)
    mps.h(0)
    mps.cnot(0, 1)
    before = mps.wavefunction()
    print("4 qubits: bad operations", flush=True)
    with pytest.raises(TypeError):
        mps.apply([m.hgate()])
    with pytest.raises(ValueError):
        mps.apply(m.MPSOperation(m.cnot(), 0))  # not valid: a 4-edge tensor on one qubit
    three = m.MPSOperation(np.eye(8, dtype=complex).reshape([2] * 6), (0, 1, 2))
    assert three.is_valid()
    with pytest.raises(ValueError, match="three or more"):
        mps.apply(three)
    print("4 qubits: bad two-qudit gates", flush=True)
    with pytest.raises(ValueError):
        mps.apply_two_qudit_gate(m.cnot(), 0, 0)
    with pytest.raises(ValueError):
        mps.cnot(1, 1)
    with pytest.raises(ValueError):
        mps.apply_two_qudit_gate(m.cnot(), 0, 4)
    with pytest.raises(ValueError):
        mps.apply_two_qudit_gate(m.cnot(), -1, 2)
    with pytest.raises(ValueError, match="four free edges"):
        mps.apply_two_qudit_gate(m.hgate(), 0, 1)
    with pytest.raises(ValueError, match="same dimension"):
        mps.apply_two_qudit_gate(np.zeros((2, 2, 3, 3), complex), 0, 1)
    with pytest.raises(ValueError, match="should have MPS qudit dimension"):
        mps.apply_two_qudit_gate(np.eye(9, dtype=complex).reshape(3, 3, 3, 3), 1, 2)
    print("4 qubits: bad one-qudit gates", flush=True)
    with pytest.raises(ValueError):
        mps.apply_one_qudit_gate(m.hgate(), 4)
    with pytest.raises(ValueError):
        mps.apply_one_qudit_gate(m.cnot(), 0)
    assert np.array_equal(mps.wavefunction(), before)


@pytest.mark.parametrize("n", [2, 3, 9])
def test_identical_indices_raise(n):
    """core_test.py:475-485."""
    import mpsim_amd as m
    mps = m.MPS(nqudits=n)
    with pytest.raises(ValueError):
        mps.apply_two_qudit_gate(m.cnot(), 0, 0)
    with pytest.raises(ValueError):
        mps.cnot(1, 1)


def test_node_moves_against_their_direction_raise():
    """core_test.py:699-712."""
    import mpsim_amd as m
    mps = m.MPS(nqudits=5)
    with pytest.raises(ValueError):
        mps.move_node_from_left_to_right(current_node_index=4, final_node_index=0)
    with pytest.raises(ValueError):
        mps.move_node_from_right_to_left(current_node_index=0, final_node_index=4)
    assert mps.bond_dimensions() == [1, 1, 1, 1]


def test_inner_product_of_mismatched_states_raises():
    """core_test.py:1028-1046."""
    import mpsim_amd as m
    with pytest.raises(ValueError):
        m.MPS(nqudits=5).inner_product(m.MPS(nqudits=6))
    with pytest.raises(ValueError):
        m.MPS(nqudits=5, qudit_dimension=2).inner_product(m.MPS(nqudits=5, qudit_dimension=3))
