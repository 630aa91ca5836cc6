"""Host tests of the product's gate library, mpsim_amd.gates (no GPU).

Restates the reference's gates_test.py:21-83 against this package's gates
(the oracle's copy is pinned separately), plus the textbook matrices, the
reference's dtypes and node names, and rgate's documented quirk
(gates.py:141-164: the generator's Y and Z terms multiply, so the gate is
exp(-i theta m_x X)).
"""
import numpy as np
import pytest


def test_common_gates_are_unitary():
    """gates_test.py:21-27."""
    from mpsim_amd.gates import cnot, cphase, hgate, igate, is_unitary, xgate, ygate, zgate
    for gate in (igate(), hgate(), xgate(), ygate(), zgate(), cnot()):
        assert is_unitary(gate)
    for exp in np.linspace(start=0, stop=2 * np.pi, num=100):
        assert is_unitary(cphase(exp))


@pytest.mark.parametrize("state,dim", [(0, 2), (1, 2), (0, 3), (1, 3), (2, 3)])
def test_projectors(state, dim):
    """gates_test.py:30-47 and 65-75."""
    from mpsim_amd.gates import computational_basis_projector, is_projector, is_unitary
    p = computational_basis_projector(state, dim)
    want = np.zeros((dim, dim))
    want[state, state] = 1.0
    assert np.array_equal(p.tensor, want)
    assert is_projector(p)
    assert not is_unitary(p)
    assert str(p) == f"|{state}><{state}|"


def test_invalid_projectors_and_states():
    """gates_test.py:50-62 (and the same checks of computational_basis_state,
    gates.py:79-100)."""
    from mpsim_amd.gates import computational_basis_projector, computational_basis_state
    for fn in (computational_basis_projector, computational_basis_state):
        with pytest.raises(ValueError):
            fn(state=-1)
        with pytest.raises(ValueError):
            fn(state=2, dim=-1)
        with pytest.raises(ValueError):
            fn(state=10, dim=8)
    v = computational_basis_state(2, 4)
    assert np.array_equal(v.tensor, [0.0, 0.0, 1.0, 0.0]) and str(v) == "|2>"


def test_haar_random_unitary():
    """gates_test.py:78-83, and the tensor shape [d] * 2n; the same seed
    gives the same gate."""
    from mpsim_amd.gates import haar_random_unitary, is_unitary
    for n in (2, 3, 4):
        for d in (2, 3, 5):
            gate = haar_random_unitary(nqudits=n, qudit_dimension=d, seed=1)
            assert is_unitary(gate)
            assert gate.tensor.shape == tuple([d] * 2 * n)
    a = haar_random_unitary(2, 2, seed=7).tensor
    assert np.array_equal(a, haar_random_unitary(2, 2, seed=7).tensor)
    assert not np.array_equal(a, haar_random_unitary(2, 2, seed=8).tensor)


def test_textbook_matrices_dtypes_and_names():
    """The Pauli, Hadamard, CNOT, SWAP and controlled-phase matrices; one-qubit
    gates complex64, cnot and swap float64 (gates.py:104-232); the node names
    the reference gives them (zgate's node is "zmat")."""
    from mpsim_amd import gates as g
    s = 2 ** -0.5
    want = {"igate": [[1, 0], [0, 1]], "xgate": [[0, 1], [1, 0]], "ygate": [[0, -1j], [1j, 0]],
            "zgate": [[1, 0], [0, -1]], "hgate": [[s, s], [s, -s]]}
    for name, m in want.items():
        node = getattr(g, name)()
        assert node.tensor.dtype == np.complex64, name
        assert np.allclose(node.tensor, m, atol=1e-7), name
    assert g.zgate().name == "zmat" and g.hgate().name == "hgate"
    cx = g.cnot().tensor
    sw = g.swap().tensor
    assert cx.dtype == np.float64 and cx.shape == (2, 2, 2, 2)
    for a in range(2):
        for b in range(2):
            assert cx[a, a ^ b, a, b] == 1.0  # |a, b> -> |a, a xor b>
            assert sw[b, a, a, b] == 1.0      # |a, b> -> |b, a>
    assert cx.sum() == 4 and sw.sum() == 4
    cp = g.cphase(0.25).tensor.reshape(4, 4)
    assert np.allclose(cp, np.diag([1, 1, 1, 1j]), atol=1e-7)
    # gates are fresh copies: changing one leaves the next
    h = g.hgate()
    h.tensor[0, 0] = 7.0
    assert g.hgate().tensor[0, 0] != 7.0


def test_rgate_is_a_rotation_about_x():
    """rgate(seed) as the reference computes it (gates.py:141-164): angles
    theta, alpha, phi from np.random after seeding; the generator is
    m_x X + (m_y Y) * (m_z Z) elementwise, whose second term is zero, so the
    gate is exp(-i theta m_x X); seed 0 does not seed."""
    from scipy.linalg import expm
    from mpsim_amd.gates import is_unitary, rgate
    for seed in (1, 5, 123):
        for scale in (1.0, 0.25):
            np.random.seed(seed)
            theta, alpha, phi = np.random.rand(3) * 2 * np.pi
            mx = np.sin(alpha) * np.cos(phi)
            want = expm(-1j * theta * scale * mx * np.array([[0, 1], [1, 0]]))
            u = rgate(seed, angle_scale=scale).tensor
            assert is_unitary(u)
            assert np.allclose(u, want, atol=1e-6)
            # a rotation about x: equal diagonal, equal off-diagonal entries
            assert np.isclose(u[0, 0], u[1, 1]) and np.isclose(u[0, 1], u[1, 0])
