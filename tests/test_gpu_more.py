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


# ---------------------------------------------------------------- reduced states
def test_reduced_density_matrix_simple():
    """core_test.py:1588-1618 and 1621-1632."""
    import mpsim_amd as m
    zero, one = np.diag([1.0, 0.0]), np.diag([0.0, 1.0])
    mps = m.MPS(2)
    for i in (0, 1):
        assert close(mps.reduced_density_matrix(node_indices=i), zero)
        assert mps == m.MPS(2)
    mps.apply(m.MPSOperation(m.xgate(), 0))
    assert close(mps.reduced_density_matrix(0), one)
    assert close(mps.reduced_density_matrix(1), zero)
    mps.apply(m.MPSOperation(m.xgate(), 1))
    assert close(mps.reduced_density_matrix(0), one)
    assert close(mps.reduced_density_matrix(1), one)
    with pytest.raises(IndexError):
        mps.reduced_density_matrix(node_indices=-1)
    with pytest.raises(IndexError):
        mps.reduced_density_matrix(node_indices=22)
    with pytest.raises(ValueError):
        mps.reduced_density_matrix(node_indices=[0, 0])


def test_reduced_density_matrices_of_random_states():
    """core_test.py:1635-1711: random 2-, 3-, 5- and 8-qubit states, every
    kept set of the 3-qubit state in both orders, random kept sets of the
    larger ones; against the FP64 partial trace."""
    import itertools
    import mpsim_amd as m
    from oracle import statevector as sv
    rng = np.random.RandomState(3)
    for _ in range(5):
        w = rng.randn(4) + 1j * rng.randn(4)
        w /= np.linalg.norm(w)
        mps = m.MPS.from_wavefunction(w, nqudits=2)
        for i in (0, 1):
            assert close(mps.reduced_density_matrix(i), sv.density_matrix(w, [i], 2))
        assert close(mps.wavefunction(), w)
    w = rng.randn(8) + 1j * rng.randn(8)
    w /= np.linalg.norm(w)
    mps = m.MPS.from_wavefunction(w, nqudits=3)
    for size in (1, 2, 3):
        for sites in itertools.combinations(range(3), size):
            for order in (sites, tuple(reversed(sites))):
                assert close(mps.reduced_density_matrix(order), sv.density_matrix(w, order, 3))
    for n in (5, 8):
        w = rng.randn(2 ** n) + 1j * rng.randn(2 ** n)
        w /= np.linalg.norm(w)
        mps = m.MPS.from_wavefunction(w, nqudits=n)
        for _ in range(6):
            size = rng.randint(low=1, high=n)
            sites = [int(q) for q in rng.choice(range(n), size=size, replace=False)]
            assert close(mps.reduced_density_matrix(sites), sv.density_matrix(w, sites, n))
        assert close(mps.wavefunction(), w)


def test_reduced_density_matrix_qutrits_and_larger_bonds():
    """A 12-qubit brickwork state (bonds up to 64) and a 5-qutrit state:
    kept sets spread over the chain, against the FP64 partial trace."""
    import mpsim_amd as m
    from oracle import mps_oracle as mo
    from oracle import statevector as sv
    rng = np.random.RandomState(4)
    for n, d, sets in ((12, 2, ([3, 7, 10], [11, 0], [5, 6])), (5, 3, ([4, 1], [2]))):
        ops = [((q,), mo.haar_unitary_from_rng(d, rng)) for q in range(n)]
        for layer in range(6):
            ops += [((i, i + 1), mo.haar_unitary_from_rng(d * d, rng))
                    for i in range(layer % 2, n - 1, 2)]
        mps = m.MPS(n, d)
        for q, g in ops:
            if len(q) == 1:
                mps.apply_one_qudit_gate(m.Node(g), q[0])
            else:
                mps.apply_two_qudit_gate(m.Node(g.reshape([d] * 4)), *q)
        w = sv.run(ops, n, d)
        for sites in sets:
            assert close(mps.reduced_density_matrix(sites), sv.density_matrix(w, sites, n, d))


def test_sample_zero_and_uniform():
    """core_test.py:1714-1737."""
    import mpsim_amd as m
    for d in (2, 3, 5):
        samples = m.MPS(2, d).sample(nsamples=100)
        assert len(samples) == 100 and all(set(s) == {0} for s in samples)
    np.random.seed(1)
    n, nsamples = 3, 100
    mps = m.MPS(n)
    mps.apply([m.MPSOperation(m.hgate(), i) for i in range(n)])
    hist = mps.sample(nsamples=nsamples, as_hist=True, as_string=True)
    assert sum(hist.values()) == nsamples
    for freq in np.array(list(hist.values())) / nsamples:
        assert abs(freq - 1 / 2 ** n) < 1 / np.sqrt(nsamples)
    with pytest.raises(ValueError):
        mps.sample(0)
    with pytest.raises(ValueError):
        mps.sample(2.0)


def test_sample_marginals_and_draws():
    """The per-site distributions are the one-site marginals of the FP64
    state vector (the reference draws every site from its own marginal,
    core.py:667-669), and with the same numpy seed the draws equal those
    made from the FP64 marginals."""
    import mpsim_amd as m
    from oracle import mps_oracle as mo
    from oracle import statevector as sv
    rng = np.random.RandomState(6)
    n = 10
    ops = [((q,), mo.haar_unitary_from_rng(2, rng)) for q in range(n)]
    for layer in range(5):
        ops += [((i, i + 1), mo.haar_unitary_from_rng(4, rng)) for i in range(layer % 2, n - 1, 2)]
    mps = m.MPS(n)
    for q, g in ops:
        if len(q) == 1:
            mps.apply_one_qudit_gate(m.Node(g), q[0])
        else:
            mps.apply_two_qudit_gate(m.Node(g.reshape(2, 2, 2, 2)), *q)
    ref = sv.marginals(sv.run(ops, n), n)
    got = mps._marginal_probabilities()
    assert np.max(np.abs(got - ref)) < 1e-5
    np.random.seed(21)
    draws = mps.sample(50, as_string=True)
    np.random.seed(21)
    expect = ["".join(str(np.random.choice([0, 1], size=1, p=ref[i].astype(np.float32))[0])
                      for i in range(n)) for _ in range(50)]
    assert draws == expect


def test_edge_accessors():
    """core_test.py:165-227."""
    import mpsim_amd as m
    for n in (2, 5):
        for d in (2, 3):
            mps = m.MPS(n, d)
            for i in range(n):
                e = mps.get_free_edge_of(i, copy=False)
                assert e.is_dangling() and e.node1.name == f"q{i}" and e.dimension == d
    mps = m.MPS(3)
    assert mps.get_left_connected_edge_of(0) is None
    e = mps.get_left_connected_edge_of(1)
    assert not e.is_dangling() and (e.node1.name, e.node2.name) == ("q0", "q1")
    e = mps.get_left_connected_edge_of(2)
    assert (e.node1.name, e.node2.name) == ("q2", "q1")
    assert mps.get_right_connected_edge_of(2) is None
    mps = m.MPS(10)
    for i in range(1, 9):
        assert mps.get_right_connected_edge_of(i - 1) == mps.get_left_connected_edge_of(i)


def test_wavefunction_with_more_than_65535_row_tiles():
    """core.py:483-500 at 23 qubits: the last GEMM of the chain has 2^22 rows
    (131072 row tiles of 32); H on every site gives the uniform state."""
    import mpsim_amd as m
    n = 23
    mps = m.MPS(n)
    mps.h(-1)
    psi = mps.wavefunction()
    assert psi.shape == (2 ** n,)
    assert np.max(np.abs(psi - 2.0 ** (-n / 2))) < 1e-6


def test_tall_narrow_gemm_vs_torch():
    """The tall-narrow GEMM path (no conjugation, k <= 32, n <= 256: the late
    steps of the wavefunction chain) against a complex128 torch product,
    with alpha and accumulation, for ragged m, k and n; and the tile path
    (k = 33) on the same data."""
    import torch
    from mpsim_amd.engine import Engine
    eng = Engine.get()
    g = torch.Generator(device="cpu").manual_seed(5)
    for m, k, n in ((1, 1, 1), (1000, 2, 2), (70001, 7, 5), (3000, 32, 64), (257, 16, 256),
                    (513, 33, 3)):
        A = torch.randn(m, k, dtype=torch.complex64, generator=g)
        B = torch.randn(k, n, dtype=torch.complex64, generator=g)
        C0 = torch.randn(m, n, dtype=torch.complex64, generator=g)
        alpha = 0.75 - 0.5j
        for accum in (False, True):
            C = C0.clone().cuda()
            eng.gemm([dict(A=A.cuda(), B=B.cuda(), C=C, m=m, n=n, k=k, lda=k, ldb=n, ldc=n,
                           alpha=alpha, flags=4 if accum else 0)])
            ref = alpha * (A.to(torch.complex128) @ B.to(torch.complex128))
            if accum:
                ref = ref + C0.to(torch.complex128)
            err = (C.cpu().to(torch.complex128) - ref).abs().max().item()
            assert err <= 1e-5 * max(1.0, ref.abs().max().item()) * max(1, k) ** 0.5, (m, k, n, err)
