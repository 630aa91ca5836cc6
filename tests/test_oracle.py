"""CPU tests that pin the oracle (no GPU).

Each check compares the restatement with something other than itself: the
FP64 state vector, closed forms, numpy's LAPACK SVD through gauge-invariant
quantities, or the known answers of the reference's own tests (cited).
A planted-mistake test shows that these checks catch plausible errors.
"""
import numpy as np
import pytest

from oracle import mps_oracle as mo
from oracle import statevector as sv
from oracle.mps_oracle import OracleMPS, gauge_invariant_error, mps_overlap


def basis(n, bits, d=2):
    v = np.zeros(d ** n)
    v[int("".join(str(b) for b in bits), d)] = 1.0
    return v


# ---------------------------------------------------------------- init
def test_max_bond_dimensions_reference_values():
    """core_test.py:86-130."""
    assert mo.max_bond_dimensions(5, 2) == [2, 4, 4, 2]
    assert mo.max_bond_dimensions(7, 2) == [2, 4, 8, 8, 4, 2]
    assert mo.max_bond_dimensions(6, 2) == [2, 4, 8, 4, 2]
    assert mo.max_bond_dimensions(8, 2) == [2, 4, 8, 16, 8, 4, 2]
    assert mo.max_bond_dimensions(5, 4) == [4, 16, 16, 4]
    assert mo.max_bond_dimensions(7, 4) == [4, 16, 64, 64, 16, 4]
    assert mo.max_bond_dimensions(4, 10) == [10, 100, 10]
    assert mo.max_bond_dimensions(6, 10) == [10, 100, 1000, 100, 10]
    mps = OracleMPS(10)
    assert mps.max_bond_dimension_of(-1) == 2 and mps.max_bond_dimension_of(4) == 32


def test_init_errors_and_product_state():
    """core_test.py:70-76, 377-386."""
    for d in (2, 3, 10):
        with pytest.raises(ValueError):
            OracleMPS(1, d)
    for n in range(2, 7):
        for d in (2, 3, 4):
            wf = OracleMPS(n, d).wavefunction()
            assert wf.shape == (d ** n,)
            assert np.allclose(wf, basis(n, [0] * n, d))
    assert OracleMPS(5, 3).bond_dimensions() == [1, 1, 1, 1]


# ---------------------------------------------------------------- gates
def test_haar_recipe_phase_fix():
    """gates.py:269-284: Q * diag(R_jj/|R_jj|) is unitary and equals the
    unique QR factor whose R has a positive real diagonal. Re-factoring the
    same Gaussian draw checks that: Q^H G is upper triangular with a
    positive real diagonal (a missing or inverted phase fix fails)."""
    for nq, d, seed in ((1, 2, 1), (2, 2, 2), (2, 3, 3)):
        dim = d ** nq
        u = mo.haar_random_unitary(nq, d, seed=seed).reshape(dim, dim)
        assert np.allclose(u.conj().T @ u, np.eye(dim), atol=1e-12)
        rng = np.random.RandomState(seed)
        g = np.sum(rng.randn(dim, dim, 2) * np.array([1, 1j]), axis=-1) / np.sqrt(2)
        r = u.conj().T @ g
        assert np.allclose(np.tril(r, -1), 0, atol=1e-12)
        assert np.all(np.abs(np.imag(np.diag(r))) < 1e-12) and np.all(np.real(np.diag(r)) > 0)


def test_gate_dtypes_and_unitarity():
    """gates.py:104-110, 195-232 and gates_test.py:21-28."""
    assert mo.hgate().dtype == np.complex64 and mo.cnot().dtype == np.float64
    for g in (mo.igate(), mo.hgate(), mo.xgate(), mo.ygate(), mo.zgate(), mo.cnot()):
        assert mo.is_unitary(g)
    for e in np.linspace(0, 2 * np.pi, 20):
        assert mo.is_unitary(mo.cphase(e))
    for n in (2, 3):
        for d in (2, 3):
            assert mo.is_unitary(mo.haar_random_unitary(n, d, seed=1))


# ---------------------------------------------------------------- two-site
def test_cnot_truth_table_flip_and_swap():
    """core_test.py:409-472, 548-598."""
    for x0 in (0, 1):
        for x1 in (0, 1):
            m = OracleMPS(2)
            if x0:
                m.x(0)
            if x1:
                m.x(1)
            m.cnot(0, 1)
            assert np.allclose(m.wavefunction(), basis(2, (x0, x1 ^ x0)))
            m = OracleMPS(2)
            if x0:
                m.x(0)
            if x1:
                m.x(1)
            m.cnot(1, 0)
            assert np.allclose(m.wavefunction(), basis(2, (x0 ^ x1, x1)))
            m = OracleMPS(2)
            if x0:
                m.x(0)
            if x1:
                m.x(1)
            m.swap(0, 1)
            assert np.allclose(m.wavefunction(), basis(2, (x1, x0)))


@pytest.mark.parametrize("left", [True, False])
def test_bell_and_truncation_known_answers(left):
    """core_test.py:773-782, 905-929, 1084-1101, 1123-1174; README.md:45-57."""
    m = OracleMPS(2)
    m.h(0)
    m.cnot(0, 1, keep_left_canonical=left)
    assert np.allclose(m.wavefunction(), np.array([1, 0, 0, 1]) / np.sqrt(2))
    m = OracleMPS(2)
    m.h(0)
    m.cnot(0, 1, maxsvals=1, keep_left_canonical=left)
    assert np.allclose(m.wavefunction(), [1 / np.sqrt(2), 0, 0, 0])
    assert np.isclose(m.norm(), 1 / np.sqrt(2))
    m.renormalize()
    assert np.allclose(m.wavefunction(), [1, 0, 0, 0])
    m = OracleMPS(2)
    m.h(0)
    m.cnot(0, 1, fraction=0.5)
    assert np.allclose(m.wavefunction(), [1 / np.sqrt(2), 0, 0, 0])
    m = OracleMPS(2)
    m.h(0)
    m.cnot(0, 1, maxsvals=0)
    assert m.norm() == 0
    with pytest.raises(ValueError):
        m.renormalize()
    with pytest.raises(ValueError):
        OracleMPS(2).cnot(0, 1, maxsvals=1, fraction=0.5)
    with pytest.raises(ValueError):
        OracleMPS(2).cnot(0, 1, fraction=1.5)


def test_bond_growth_keeps_zero_singular_values():
    """core_test.py:932-973: a CNOT on a product state gives bond 2."""
    rng = np.random.RandomState(0)
    m = OracleMPS(4)
    for q in range(4):
        m.apply_one_qudit_gate(mo.haar_unitary_from_rng(2, rng), q)
    m.apply_two_qudit_gate(mo.cnot(), 0, 1, fraction=1)
    assert m.bond_dimensions() == [2, 1, 1]
    m2 = OracleMPS(4)
    m2.apply_two_qudit_gate(mo.cnot(), 0, 1, fraction=0.5)
    assert m2.bond_dimensions() == [1, 1, 1]


def test_nonlocal_ghz_and_qft():
    """core_test.py:1225-1258 (SWAP network)."""
    for n in range(3, 7):
        m = OracleMPS(n)
        m.h(0)
        for i in range(1, n):
            m.cnot(0, i)
        v = np.zeros(2 ** n)
        v[0] = v[-1] = 1 / np.sqrt(2)
        assert np.allclose(m.wavefunction(), v)
        m = OracleMPS(n)
        for i in range(n - 1, -1, -1):
            m.h(i)
            for j in range(i - 1, -1, -1):
                m.apply_two_qudit_gate(mo.cphase(2 ** (j - i)), j, i)
        assert np.allclose(m.wavefunction(), np.ones(2 ** n) / 2 ** (n / 2))


@pytest.mark.parametrize("mode", ["reference", "c128"])
def test_random_circuits_vs_statevector(mode):
    """simulator_test.py:274-305 restated: seeded random circuits with
    arbitrary-order pairs (j < i flips, non-adjacent SWAP networks)."""
    rng = np.random.RandomState(3)
    for n in (3, 5, 7):
        ops = []
        for _ in range(25):
            if rng.rand() < 0.5:
                ops.append(((int(rng.randint(n)),), mo.haar_unitary_from_rng(2, rng)))
            else:
                i, j = rng.choice(n, 2, replace=False)
                ops.append(((int(i), int(j)), mo.haar_unitary_from_rng(4, rng)))
        m = OracleMPS(n, dtype_mode=mode)
        m.apply_ops(ops)
        ref = sv.run(ops, n)
        assert np.max(np.abs(m.wavefunction() - ref)) < 1e-10


def test_truncation_is_eckart_young_per_update():
    """One truncated update keeps exactly the top-k singular triplets of
    theta: the discarded weight equals the sum of the dropped s^2 from
    numpy's SVD, and the kept state is the best rank-k approximation."""
    rng = np.random.RandomState(5)
    m = OracleMPS(6, dtype_mode="c128")
    for layer in range(4):
        for i in range(layer % 2, 5, 2):
            m.apply_two_qudit_gate(mo.haar_unitary_from_rng(4, rng).reshape(2, 2, 2, 2), i, i + 1)
    a, b = m.tensors[2], m.tensors[3]
    g = mo.haar_unitary_from_rng(4, rng).reshape(2, 2, 2, 2)
    theta = np.einsum("abpq,lpqr->labr", g, np.tensordot(a, b, axes=([2], [0])))
    M = theta.reshape(a.shape[0] * 2, 2 * b.shape[2])
    s = np.linalg.svd(M, compute_uv=False)
    k = 3
    m.apply_two_qudit_gate(g, 2, 3, maxsvals=k)
    assert np.isclose(m.discarded[-1], np.sum(s[k:] ** 2))
    new = np.tensordot(m.tensors[2], m.tensors[3], axes=([2], [0])).reshape(M.shape)
    assert np.isclose(np.linalg.norm(new - M), np.sqrt(np.sum(s[k:] ** 2)))


def test_inner_product_orientation():
    """core_test.py:995-1025: inner_product(other) = sum self * conj(other)."""
    rng = np.random.RandomState(9)
    n = 4
    ops_a = [((q,), mo.haar_unitary_from_rng(2, rng)) for q in range(n)] + \
        [((0, 1), mo.haar_unitary_from_rng(4, rng)), ((2, 3), mo.haar_unitary_from_rng(4, rng)),
         ((1, 2), mo.haar_unitary_from_rng(4, rng))]
    ops_b = [((q,), mo.haar_unitary_from_rng(2, rng)) for q in range(n)] + \
        [((1, 2), mo.haar_unitary_from_rng(4, rng))]
    a, b = OracleMPS(n), OracleMPS(n)
    a.apply_ops(ops_a)
    b.apply_ops(ops_b)
    wa, wb = sv.run(ops_a, n), sv.run(ops_b, n)
    assert np.isclose(a.inner_product(b), np.inner(wa, wb.conj()))
    assert np.isclose(mps_overlap(b.tensors, a.tensors), np.vdot(wb, wa))
    err, nr, nx = gauge_invariant_error(a.tensors, a.tensors)
    assert abs(err) < 1e-12 and np.isclose(nr, 1.0)


def test_reference_dtype_promotion():
    """Float64 CNOT and complex128 Haar gates promote the touched tensors
    (numpy rules), complex64 H keeps complex64."""
    m = OracleMPS(3)
    m.h(0)
    assert m.tensors[0].dtype == np.complex64
    m.cnot(0, 1)
    assert m.tensors[0].dtype == np.complex128 and m.tensors[2].dtype == np.complex64
    m = OracleMPS(3, dtype_mode="c64")
    m.cnot(0, 1)
    assert m.tensors[0].dtype == np.complex64


def test_planted_mistake_is_caught():
    """A plausible mistake planted in a copy of the oracle (transposed gate
    legs) must be caught by the state-vector comparison."""

    class Broken(OracleMPS):
        def _two_site(self, g, i, kwargs):
            super()._two_site(np.transpose(g, (1, 0, 2, 3)), i, kwargs)

    rng = np.random.RandomState(11)
    ops = [((0, 1), mo.haar_unitary_from_rng(4, rng)), ((1, 2), mo.haar_unitary_from_rng(4, rng))]
    good, bad = OracleMPS(3, dtype_mode="c128"), Broken(3, dtype_mode="c128")
    good.apply_ops(ops)
    bad.apply_ops(ops)
    ref = sv.run(ops, 3)
    assert np.max(np.abs(good.wavefunction() - ref)) < 1e-12
    assert np.max(np.abs(bad.wavefunction() - ref)) > 1e-3


def test_plain_config_oracle_vs_statevector():
    """The 16-qubit plain config in the oracle (reference dtype rules,
    maxsvals=None) equals the FP64 state vector to 1e-10."""
    from mpsim_amd.circuits import brickwork_cnot_u2
    c = brickwork_cnot_u2(16, 8, seed=1)
    m = OracleMPS(16, dtype_mode="reference")
    m.apply_ops(c.ops)
    assert max(m.bond_dimensions()) == 256
    assert np.max(np.abs(m.wavefunction() - sv.run(c.ops, 16))) < 1e-10


# ---------------------------------------------------------------- reduced states
def _random_state_ops(n, seed, depth=3):
    rng = np.random.RandomState(seed)
    ops = [((q,), mo.haar_unitary_from_rng(2, rng)) for q in range(n)]
    for layer in range(depth):
        ops += [((i, i + 1), mo.haar_unitary_from_rng(4, rng)) for i in range(layer % 2, n - 1, 2)]
    return ops


def test_density_matrix_closed_forms():
    """The state-vector partial trace: a product state |0>|1>|+> gives
    |01><01| on (0, 1) (index 1) and |10><10| on (1, 0) (index 2); a Bell
    pair gives I/2 on either qubit; keeping every site gives psi psi^H."""
    n = 3
    plus = np.array([1, 1]) / np.sqrt(2)
    psi = np.kron(np.kron([1, 0], [0, 1]), plus)
    r01 = sv.density_matrix(psi, (0, 1), n)
    r10 = sv.density_matrix(psi, (1, 0), n)
    assert np.allclose(r01, np.diag([0, 1, 0, 0])) and np.allclose(r10, np.diag([0, 0, 1, 0]))
    assert np.allclose(sv.density_matrix(psi, (2,), n), np.outer(plus, plus))
    bell = np.array([1, 0, 0, 1]) / np.sqrt(2)
    for q in (0, 1):
        assert np.allclose(sv.density_matrix(bell, (q,), 2), np.eye(2) / 2)
    w = sv.run(_random_state_ops(4, 1), 4)
    assert np.allclose(sv.density_matrix(w, range(4), 4), np.outer(w, w.conj()))
    assert np.allclose(sv.marginals(w, 4)[2], sv.density_matrix(w, (2,), 4).diagonal().real)


def test_rdm_known_answers():
    """core_test.py:1588-1618: |00>, |10>, |11> one-site matrices; and the
    index errors of core_test.py:1621-1632."""
    zero, one = np.diag([1.0, 0.0]), np.diag([0.0, 1.0])
    m = OracleMPS(2)
    for i in (0, 1):
        assert np.allclose(m.reduced_density_matrix(i), zero)
    m.x(0)
    assert np.allclose(m.reduced_density_matrix(0), one)
    assert np.allclose(m.reduced_density_matrix(1), zero)
    m.x(1)
    assert np.allclose(m.reduced_density_matrix(0), one)
    assert np.allclose(m.reduced_density_matrix(1), one)
    with pytest.raises(IndexError):
        m.reduced_density_matrix(-1)
    with pytest.raises(IndexError):
        m.reduced_density_matrix(22)
    with pytest.raises(ValueError):
        m.reduced_density_matrix([0, 0])


@pytest.mark.parametrize("d", [2, 3])
def test_rdm_restatement_vs_statevector(d):
    """core_test.py:1654-1711 restated: every kept set of size 1-3 of a
    random 5-site state, in sorted and reversed order, against the partial
    trace of the FP64 state vector; the one-site diagonals are the
    marginals."""
    import itertools
    n = 5
    rng = np.random.RandomState(7 + d)
    ops = [((q,), mo.haar_unitary_from_rng(d, rng)) for q in range(n)]
    for layer in range(3):
        ops += [((i, i + 1), mo.haar_unitary_from_rng(d * d, rng))
                for i in range(layer % 2, n - 1, 2)]
    m = OracleMPS(n, d, dtype_mode="c128")
    m.apply_ops(ops)
    w = sv.run(ops, n, d)
    for size in (1, 2, 3):
        for sites in itertools.combinations(range(n), size):
            for order in (sites, tuple(reversed(sites))):
                got = m.reduced_density_matrix(list(order))
                assert np.allclose(got, sv.density_matrix(w, order, n, d), atol=1e-12)
    assert np.allclose(m.marginals(), sv.marginals(w, n, d), atol=1e-12)


def test_rdm_planted_mistake_is_caught():
    """Swapping the ket and bra legs (returning rho^T) or ignoring the
    requested order must fail the state-vector comparison."""
    n = 4
    ops = _random_state_ops(n, 12)
    m = OracleMPS(n, dtype_mode="c128")
    m.apply_ops(ops)
    w = sv.run(ops, n)
    good = m.reduced_density_matrix([2, 0])
    ref = sv.density_matrix(w, (2, 0), n)
    assert np.allclose(good, ref, atol=1e-12)
    assert not np.allclose(good.T, ref, atol=1e-3)
    assert not np.allclose(m.reduced_density_matrix([0, 2]), ref, atol=1e-3)


def test_cpu_baseline_runs_on_a_small_sample():
    """The timed CPU baseline (the oracle at the cap shape; the 16-qubit
    plain workload) on small samples: the records carry their counts and
    positive rates (a 6-qubit depth-2 plain run has 3 + 2 CNOTs)."""
    from oracle.cpu_baseline import blas_threads, time_plain16, time_updates
    r = time_updates(updates=2, n=16, cap=32, threads=1)
    assert r["updates"] == 2 and r["seconds"] > 0 and r["updates_per_s"] > 0
    r = time_plain16(n=6, depth=2)
    assert r["updates"] == 5 and r["seconds"] > 0
    assert blas_threads() >= 1
