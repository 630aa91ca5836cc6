"""Host tests of the gate-operation surface (no GPU).

Restates the reference's MPSOperation tests (core_test.py:26-67) and the
one-qudit MPS check (core_test.py:70-75, which raises before any device
memory is touched): properties, validity, node copies and tensors of
one- and two-qudit operations, local and non-local.
"""
import numpy as np
import pytest


def test_single_qubit_identity_operation():
    """core_test.py:26-35."""
    from mpsim_amd import MPSOperation, igate
    op = MPSOperation(igate(), qudit_indices=0, qudit_dimension=2)
    assert op.qudit_indices == (0,)
    assert op.qudit_dimension == 2
    assert op.num_qudits == 1
    assert op.is_valid()
    assert op.is_unitary()
    assert op.is_single_qudit_operation()
    assert not op.is_two_qudit_operation()


def test_node_copy_and_tensor_of_one_qubit_operation():
    """core_test.py:38-49: node(copy=True) is a copy with the same edges;
    tensor() returns the values; changing the copy leaves the operation."""
    from mpsim_amd import MPSOperation
    from mpsim_amd.gates import Node
    rng = np.random.RandomState(1)
    tensor = rng.randn(2, 2)
    node = Node(tensor)
    op = MPSOperation(node, qudit_indices=(0,), qudit_dimension=2)
    cp = op.node(copy=True)
    assert len(cp.edges) == len(node.edges)
    assert np.allclose(op.tensor(), tensor)
    cp.tensor[0, 0] += 1.0
    assert np.allclose(op.tensor(), tensor)
    assert op.node(copy=False) is node


def test_two_qubit_cnot_operation():
    """core_test.py:52-58."""
    from mpsim_amd import MPSOperation, cnot
    op = MPSOperation(cnot(), qudit_indices=(0, 1), qudit_dimension=2)
    assert op.qudit_indices == (0, 1)
    assert op.qudit_dimension == 2
    assert not op.is_single_qudit_operation()
    assert op.is_two_qudit_operation()
    assert op.is_valid() and op.is_unitary()
    assert op.tensor().shape == (4, 4)
    assert op.tensor(reshape_to_square_matrix=False).shape == (2, 2, 2, 2)


def test_two_qubit_nonlocal_cnot_operation():
    """core_test.py:61-67."""
    from mpsim_amd import MPSOperation, cnot
    op = MPSOperation(cnot(), qudit_indices=(0, 2), qudit_dimension=2)
    assert op.qudit_indices == (0, 2)
    assert op.is_valid()
    assert not op.is_single_qudit_operation()
    assert op.is_two_qudit_operation()


def test_invalid_and_non_unitary_operations():
    """is_valid (core.py:113-129): a tensor of shape (d,)*2k; a CNOT said to
    act on one qubit, or on qutrits, is not valid. is_unitary is False for a
    projector."""
    from mpsim_amd import MPSOperation, cnot, hgate
    assert not MPSOperation(cnot(), 0).is_valid()
    assert not MPSOperation(cnot(), (0, 1), qudit_dimension=3).is_valid()
    assert not MPSOperation(hgate(), (0, 1)).is_valid()
    proj = MPSOperation(np.array([[1, 0], [0, 0]], dtype=complex), 0)
    assert proj.is_valid() and not proj.is_unitary()
    assert proj.is_hermitian()
    assert "(0,)" in str(proj)


@pytest.mark.parametrize("d", [2, 3, 10, 20, 100])
def test_mps_of_one_qudit_raises(d):
    """core_test.py:70-75."""
    from mpsim_amd import MPS
    with pytest.raises(ValueError):
        MPS(nqudits=1, qudit_dimension=d)
