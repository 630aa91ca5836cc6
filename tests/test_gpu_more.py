"""GPU tests of the paths next to the gate path: non-unitary gates with
orthonormalisation (core.py:807-939), from_wavefunction (core.py:245-328),
dagger (502-505) and expectation (723-751). Known answers restated from the
reference's core_test.py; tolerances as in test_gpu_core.py (complex64).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def close(a, b):
    return np.allclose(a, b, rtol=1e-5, atol=1e-6)


def test_povm_product_state_without_ortho():
    """core_test.py:1288-1335."""
    import mpsim_amd as m
    pi0 = m.computational_basis_projector(0)
    n = 3
    mps = m.MPS(n)
    mps.apply([m.MPSOperation(m.hgate(), i) for i in range(n)])
    for i, nrm in enumerate((1 / np.sqrt(2), 0.5, 0.5 / np.sqrt(2))):
        mps.apply_one_qudit_gate(pi0, i, ortho_after_non_unitary=False,
                                 renormalize_after_non_unitary=False)
        assert np.isclose(mps.norm(), nrm)
        assert mps.bond_dimensions() == [1, 1]
    assert close(mps.wavefunction(), np.array([1] + [0] * 7) / np.sqrt(2) ** 3)


@pytest.mark.parametrize("right", [True, False])
def test_povm_bell_ortho_reduces_bond(right):
    """core_test.py:1338-1419."""
    import mpsim_amd as m
    pi0 = m.computational_basis_projector(0)
    mps = m.MPS(2)
    mps.apply([m.MPSOperation(m.hgate(), (0,)), m.MPSOperation(m.cnot(), (0, 1))])
    before = mps.wavefunction()
    if right:
        mps.orthonormalize_right_edge_of(0)
    else:
        mps.orthonormalize_left_edge_of(1)
    assert close(mps.wavefunction(), before) and mps.bond_dimensions() == [2]
    mps.apply_one_qudit_gate(pi0, 0 if right else 1, ortho_after_non_unitary=False,
                             renormalize_after_non_unitary=False)
    correct = np.array([1, 0, 0, 0]) / np.sqrt(2)
    assert close(mps.wavefunction(), correct) and mps.bond_dimensions() == [2]
    if right:
        mps.orthonormalize_right_edge_of(0)
    else:
        mps.orthonormalize_left_edge_of(1)
    assert np.isclose(mps.norm(), 1 / np.sqrt(2))
    assert close(mps.wavefunction(), correct)
    assert mps.bond_dimensions() == [1]


def test_non_unitary_gate_default_path_renormalizes():
    """core_test.py:1422-1469: with the defaults the state is orthonormalised
    and renormalised to its previous norm."""
    import mpsim_amd as m
    pi0 = m.computational_basis_projector(0)
    mps = m.MPS(3)
    mps.h(0)
    mps.cnot(0, 1)
    mps.cnot(1, 2)
    mps.apply_one_qudit_gate(pi0, 1)
    assert np.isclose(mps.norm(), 1.0)
    assert close(mps.wavefunction(), np.array([1] + [0] * 7))
    assert mps.bond_dimensions() == [1, 1]


def test_from_wavefunction_round_trips():
    """core_test.py:230-297: random qubit and qudit vectors."""
    import mpsim_amd as m
    rng = np.random.RandomState(1)
    for n in range(2, 8):
        wf = rng.rand(2 ** n)
        wf /= np.linalg.norm(wf)
        mps = m.MPS.from_wavefunction(wf, nqudits=n)
        assert close(mps.wavefunction(), wf)
        assert mps.bond_dimensions() == mps.max_bond_dimensions()
    for n in range(2, 5):
        for d in (3, 4):
            wf = rng.rand(d ** n) + 1j * rng.rand(d ** n)
            wf /= np.linalg.norm(wf)
            assert close(m.MPS.from_wavefunction(wf, n, d).wavefunction(), wf)
    with pytest.raises(TypeError):
        m.MPS.from_wavefunction({1, 2, 3, 4}, 2)
    with pytest.raises(ValueError):
        m.MPS.from_wavefunction([1.0, 0, 0, 0], 3)
    with pytest.raises(ValueError):
        m.MPS.from_wavefunction(np.eye(2), 2)


def test_dagger_and_inner_products_of_random_states():
    """core_test.py:1575-1585 and 995-1025."""
    import mpsim_amd as m
    rng = np.random.RandomState(10)
    for n in (2, 3, 5, 8):
        w1 = rng.randn(2 ** n) + 1j * rng.randn(2 ** n)
        w1 /= np.linalg.norm(w1)
        w2 = rng.randn(2 ** n) + 1j * rng.randn(2 ** n)
        w2 /= np.linalg.norm(w2)
        a, b = m.MPS.from_wavefunction(w1, n), m.MPS.from_wavefunction(w2, n)
        assert np.isclose(a.inner_product(b), np.inner(w1, w2.conj()), atol=1e-5)
        a.dagger()
        assert close(a.wavefunction(), w1.conj())


def test_expectation_two_qubits():
    """core_test.py:1545-1562."""
    import mpsim_amd as m
    mps = m.MPS(2)
    ref = mps.copy()
    h0 = m.MPSOperation(m.hgate(), 0)
    assert np.isclose(mps.expectation(h0), 1 / np.sqrt(2))
    assert mps == ref
    assert np.isclose(mps.expectation(m.MPSOperation(m.xgate(), 0)), 0.0, atol=1e-6)
    mps.apply(m.MPSOperation(m.xgate(), 0))
    assert np.isclose(mps.expectation(h0), -1 / np.sqrt(2))
    with pytest.raises(ValueError):
        mps.expectation(m.MPSOperation(np.array([[0, 1], [0, 0]], dtype=complex), 0))
