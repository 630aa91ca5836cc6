"""GPU parity tests of the MPS API (hot path through libmpsx).

Known answers restated from the reference's core_test.py (file:line cited per
test); random circuits checked against the FP64 state-vector oracle.
Tolerances: untruncated wavefunctions <= 1e-5 max-abs vs FP64 (complex64
storage). Closed forms use close(): np.allclose with atol=1e-6 instead of
numpy's default 1e-8, because the reference reaches its known answers in
complex128 (its float64 CNOT/SWAP promote every tensor) while this path
stores complex64, whose rounding is ~6e-8 relative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import statevector as sv
from oracle.mps_oracle import OracleMPS


def _mpsim():
    import mpsim_amd
    from mpsim_amd import _native
    _native.load()  # fail loudly if the library is missing
    return mpsim_amd


def close(a, b):
    return np.allclose(a, b, rtol=1e-5, atol=1e-6)


def _basis(n, bits):
    v = np.zeros(2 ** n)
    v[int("".join(str(b) for b in bits), 2)] = 1.0
    return v


def test_library_is_native():
    m = _mpsim()
    from mpsim_amd import _native
    lib = _native.load()
    import ctypes
    arch = ctypes.create_string_buffer(64)
    cus = ctypes.c_int(0)
    _native.check(lib.mpsx_device_info(0, arch, 64, ctypes.byref(cus)))
    assert arch.value.decode().startswith("gfx950"), arch.value
    assert cus.value > 0


def test_product_state_wavefunction():
    m = _mpsim()
    for n in range(2, 7):
        for d in (2, 3, 4):
            mps = m.MPS(n, d)
            wf = mps.wavefunction()
            ref = np.zeros(d ** n)
            ref[0] = 1
            assert close(wf, ref)  # core_test.py:377-386


def test_one_qubit_gates():
    """core_test.py:389-406."""
    m = _mpsim()
    from mpsim_amd.gates import one_state, plus_state, zero_state
    for gate, expected in ((m.xgate(), one_state), (m.hgate(), plus_state),
                           (m.zgate(), zero_state), (m.igate(), zero_state)):
        for n in range(2, 6):
            for j in range(n):
                mps = m.MPS(n)
                mps.apply_one_qudit_gate(gate, j)
                assert close(mps.get_node(j).tensor.reshape(2), expected)


def test_hadamard_all():
    """core_test.py:421-430."""
    m = _mpsim()
    n = 5
    mps = m.MPS(n)
    mps.apply_one_qudit_gate_to_all(m.hgate())
    assert close(mps.wavefunction(), np.ones(2 ** n) / 2 ** (n / 2))


def test_cnot_truth_table_and_flip():
    """core_test.py:409-472."""
    m = _mpsim()
    for x0, x1 in ((0, 0), (0, 1), (1, 0), (1, 1)):
        mps = m.MPS(2)
        if x0:
            mps.x(0)
        if x1:
            mps.x(1)
        mps.apply_two_qudit_gate(m.cnot(), 0, 1)
        assert close(mps.wavefunction(), _basis(2, (x0, x1 ^ x0)))
        mps = m.MPS(2)
        if x0:
            mps.x(0)
        if x1:
            mps.x(1)
        mps.cnot(1, 0)
        assert close(mps.wavefunction(), _basis(2, (x0 ^ x1, x1)))


@pytest.mark.parametrize("left", [True, False])
def test_cnot_five_qubits_all_pairs(left):
    """core_test.py:514-535."""
    m = _mpsim()
    n = 5
    for a in range(n - 1):
        mps = m.MPS(n)
        mps.x(a)
        mps.cnot(a, a + 1, keep_left_canonical=left)
        bits = [0] * n
        bits[a] = bits[a + 1] = 1
        assert close(mps.wavefunction(), _basis(n, bits))


@pytest.mark.parametrize("left", [True, False])
def test_bell_state(left):
    """core_test.py:773-782."""
    m = _mpsim()
    mps = m.MPS(2)
    mps.h(0)
    mps.cnot(0, 1, keep_left_canonical=left)
    assert close(mps.wavefunction(), np.array([1, 0, 0, 1]) / np.sqrt(2))
    assert mps.bond_dimensions() == [2]


def test_bell_truncation_known_answers():
    """core_test.py:905-929, 1084-1101, 1123-1174; README.md:45-57."""
    m = _mpsim()
    mps = m.MPS(2)
    mps.h(0)
    mps.cnot(0, 1, maxsvals=1)
    assert close(mps.wavefunction(), [1 / np.sqrt(2), 0, 0, 0])
    assert np.isclose(mps.norm(), 1 / np.sqrt(2))
    mps.renormalize()
    assert close(mps.wavefunction(), [1, 0, 0, 0])
    assert np.isclose(mps.norm(), 1.0)
    for to in np.linspace(0.1, 2.0, 7):
        mps.renormalize(to_norm=to)
        assert close(mps.wavefunction(), [to, 0, 0, 0])
    # fraction
    mps = m.MPS(2)
    mps.h(0)
    mps.cnot(0, 1, fraction=0.5)
    assert close(mps.wavefunction(), [1 / np.sqrt(2), 0, 0, 0])
    mps = m.MPS(2)
    mps.h(0)
    mps.cnot(0, 1, fraction=1)
    assert close(mps.wavefunction(), [1 / np.sqrt(2), 0, 0, 1 / np.sqrt(2)])
    # maxsvals = 0 -> zero norm, renormalize raises
    mps = m.MPS(2)
    mps.h(0)
    mps.cnot(0, 1, maxsvals=0)
    assert mps.norm() == 0
    with pytest.raises(ValueError):
        mps.renormalize()


def test_bond_dims_keep_zero_singular_values():
    """core_test.py:932-973: zero singular values are kept."""
    m = _mpsim()
    mps = m.MPS(2)
    mps.h(0)
    mps.cnot(0, 1)
    assert mps.bond_dimension_of(0) == 2
    mps.cnot(0, 1)
    assert mps.bond_dimension_of(0) == 2
    mps = m.MPS(4)
    mps.r(-1)
    mps.apply_two_qudit_gate(m.cnot(), 0, 1, fraction=1)
    assert mps.bond_dimensions() == [2, 1, 1]
    mps = m.MPS(4)
    mps.r(-1)
    mps.apply_two_qudit_gate(m.cnot(), 0, 1, fraction=0.5)
    assert mps.bond_dimensions() == [1, 1, 1]


@pytest.mark.parametrize("left", [True, False])
def test_three_cnots_is_swap(left):
    """core_test.py:854-873."""
    m = _mpsim()
    for n in range(2, 8):
        mps = m.MPS(n)
        mps.x(0)
        mps.cnot(0, 1, keep_left_canonical=left)
        mps.h(-1)
        mps.cnot(0, 1, keep_left_canonical=left)
        mps.h(-1)
        mps.cnot(0, 1)
        correct = np.zeros(2 ** n)
        correct[2 ** (n - 2)] = 1
        assert close(mps.wavefunction(), correct)


def test_swap_network_ghz_and_qft():
    """core_test.py:1225-1258: non-local gates through the SWAP network."""
    m = _mpsim()
    for n in range(3, 8):
        mps = m.MPS(n)
        mps.h(0)
        for i in range(1, n):
            mps.cnot(0, i)
        correct = np.zeros(2 ** n)
        correct[0] = correct[-1] = 1 / np.sqrt(2)
        assert close(mps.wavefunction(), correct)
    for n in range(3, 7):
        mps = m.MPS(n)
        for i in range(n - 1, -1, -1):
            mps.h(i)
            for j in range(i - 1, -1, -1):
                mps.apply_two_qudit_gate(m.cphase(2 ** (j - i)), j, i)
        assert close(mps.wavefunction(), np.ones(2 ** n) / 2 ** (n / 2))


def test_move_node_right_apply_move_back():
    """core_test.py:750-770."""
    m = _mpsim()
    for n in range(3, 8):
        mps = m.MPS(n)
        mps.x(0)
        mps.move_node_from_left_to_right(0, n - 2)
        mps.cnot(n - 2, n - 1)
        mps.move_node_from_right_to_left(n - 2, 0)
        correct = np.zeros(2 ** n)
        correct[2 ** (n - 1) + 1] = 1
        assert close(mps.wavefunction(), correct)


def test_inner_product_and_norm():
    """core_test.py:976-1045 (inner products against dense vectors)."""
    m = _mpsim()
    rng = np.random.RandomState(7)
    n = 5
    a, b = m.MPS(n), m.MPS(n)
    ops_a, ops_b = [], []
    for mps, ops in ((a, ops_a), (b, ops_b)):
        for layer in range(3):
            for q in range(n):
                u = m.gates.haar_unitary(2, rng)
                ops.append(((q,), u))
                mps.apply_one_qudit_gate(u, q)
            for i in range(layer % 2, n - 1, 2):
                u = m.gates.haar_unitary(4, rng)
                ops.append(((i, i + 1), u))
                mps.apply_two_qudit_gate(u.reshape(2, 2, 2, 2), i, i + 1)
    wa, wb = sv.run(ops_a, n), sv.run(ops_b, n)
    assert np.isclose(a.inner_product(b), np.inner(wa, wb.conj()), atol=1e-5)
    assert np.isclose(b.inner_product(a), np.inner(wb, wa.conj()), atol=1e-5)
    assert np.isclose(a.norm(), 1.0, atol=1e-5)


@pytest.mark.parametrize("n,depth,seed", [(6, 6, 11), (8, 8, 12), (10, 6, 13)])
def test_random_circuits_vs_statevector(n, depth, seed):
    """simulator_test.py:274-305 restated without cirq: seeded random
    brickwork circuits, untruncated, against the FP64 state vector."""
    m = _mpsim()
    from mpsim_amd.circuits import random_pairs_circuit
    circ = random_pairs_circuit(seed, n, depth, maxsvals=10 ** 9)
    mps = m.MPS(n)
    for qudits, mat in circ.ops:
        if len(qudits) == 1:
            mps.apply_one_qudit_gate(mat, qudits[0])
        else:
            mps.apply_two_qudit_gate(mat.reshape(2, 2, 2, 2), *qudits)
    ref = sv.run(circ.ops, n)
    err = np.max(np.abs(mps.wavefunction() - ref))
    assert err <= 1e-5, err


def test_plain_config_16_qubits_vs_statevector():
    """The CPU-runnable plain config: 16 qubits, brickwork depth 8 of U(2)
    layers and CNOTs, no truncation; <= 1e-5 max-abs vs FP64."""
    m = _mpsim()
    from mpsim_amd.circuits import brickwork_cnot_u2
    circ = brickwork_cnot_u2(16, 8, seed=1)
    mps = m.MPS(16)
    for qudits, mat in circ.ops:
        if len(qudits) == 1:
            mps.apply_one_qudit_gate(mat, qudits[0])
        else:
            mps.apply_two_qudit_gate(mat.reshape(2, 2, 2, 2), *qudits)
    ref = sv.run(circ.ops, 16)
    err = np.max(np.abs(mps.wavefunction() - ref))
    assert err <= 1e-5, err
    assert max(mps.bond_dimensions()) == 256


def _recorded_updates(m, circ, cap, left):
    """Runs circ's nearest-neighbour two-site ops through the public API
    (MPS.apply_two_qudit_gate) and, on a second MPS, through the same engine
    call with want_s, recording every update's kept singular values and
    discarded weight. Both runs must end bitwise equal."""
    import torch
    from mpsim_amd.engine import TwoSiteItem
    d = circ.d
    api = m.MPS(circ.n, d)
    rec_mps = m.MPS(circ.n, d)
    eng = rec_mps._engine
    rec = []
    for qudits, mat in circ.ops:
        i, j = qudits
        assert j == i + 1
        g = np.asarray(mat).reshape([d] * 4)
        api.apply_two_qudit_gate(g, i, j, maxsvals=cap, keep_left_canonical=left)
        a, b = rec_mps._tensors[i], rec_mps._tensors[j]
        k = min(d * a.shape[0], d * b.shape[2], cap)
        r = eng.two_site([TwoSiteItem(a, b, eng.gate(g.reshape(d * d, d * d)), k,
                                      absorb_left=not left, want_s=True)])[0]
        rec_mps._tensors[i], rec_mps._tensors[j] = r.a_out, r.b_out
        rec.append((r.s.cpu().numpy().astype(np.float64), float(r.discarded.cpu())))
    torch.cuda.synchronize()
    for x, y in zip(api.tensors, rec_mps.tensors):
        assert np.array_equal(x.cpu().numpy(), y.cpu().numpy())
    return api, rec


def _check_spectra(rec, ora):
    """Every update's kept singular values within 1e-4 relative plus 1e-6 of
    ||theta||_F (theta is formed from complex64 tensors in FP32; the FP64
    oracle forms it exactly), and its discarded weight within 1e-6
    ||theta||_F^2, against the FP64 oracle's split of the same update."""
    assert len(rec) == len(ora.spectra)
    worst = 0.0
    for (s, disc), s_ref, d_ref in zip(rec, ora.spectra, ora.discarded):
        assert s.shape == s_ref.shape
        fro = np.sqrt(np.sum(s_ref ** 2) + d_ref)
        err = np.abs(s - s_ref)
        assert np.all(err <= 1e-4 * s_ref + 1e-6 * fro)
        assert abs(disc - d_ref) <= 1e-6 * fro ** 2
        if s.size:
            worst = max(worst, float(np.max(err / np.maximum(s_ref, 1e-30 + 1e-3 * s_ref[0]))))
    return worst


@pytest.mark.parametrize("left", [True, False])
def test_truncated_spectra_vs_fp64_oracle(left):
    """Truncated brickwork (scaled-down single-large-circuit workload, 12
    qubits, maxsvals 8): the state matches the FP64 restatement
    gauge-invariantly (overlap error <= 1e-4, norms), and every update's kept
    singular values and discarded weight agree with the oracle's split
    (_check_spectra)."""
    m = _mpsim()
    from mpsim_amd.circuits import brickwork_haar
    from oracle.mps_oracle import gauge_invariant_error
    circ = brickwork_haar(12, 8, 2, maxsvals=8, seed=21)
    ora = OracleMPS(12, 2, dtype_mode="c128")
    for qudits, mat in circ.ops:
        ora.apply_two_qudit_gate(mat.reshape(2, 2, 2, 2), *qudits, maxsvals=8,
                                 keep_left_canonical=left)
    mps, rec = _recorded_updates(m, circ, 8, left)
    assert mps.bond_dimensions() == ora.bond_dimensions()
    err, nr, nx = gauge_invariant_error(ora.tensors, [t.cpu().numpy() for t in mps.tensors])
    assert err <= 1e-4, err
    assert abs(nr - nx) <= 1e-4 * nr
    print(f"left={left}: overlap error {err:.2e}, worst kept-value error {_check_spectra(rec, ora):.2e}")


def test_qutrit_untruncated_vs_statevector():
    """d = 3 (the precision-stress config's qudits), untruncated, 6 qutrits
    of 9x9 Haar brickwork, against the FP64 state vector (<= 1e-5)."""
    m = _mpsim()
    from mpsim_amd.circuits import brickwork_haar
    circ = brickwork_haar(6, 5, 3, maxsvals=None, seed=41)
    mps = m.MPS(6, 3)
    for qudits, mat in circ.ops:
        mps.apply_two_qudit_gate(mat.reshape(3, 3, 3, 3), *qudits)
    ref = sv.run(circ.ops, 6, 3)
    err = np.max(np.abs(mps.wavefunction() - ref))
    assert err <= 1e-5, err


@pytest.mark.parametrize("left", [True, False])
def test_qutrit_truncated_vs_fp64_oracle(left):
    """Scaled-down precision stress: 10 qutrits, depth 8, maxsvals 20;
    gauge-invariant overlap error <= 1e-4, and every update's kept singular
    values and discarded weight against the FP64 oracle's split
    (_check_spectra)."""
    m = _mpsim()
    from mpsim_amd.circuits import brickwork_haar
    from oracle.mps_oracle import gauge_invariant_error
    circ = brickwork_haar(10, 8, 3, maxsvals=20, seed=42)
    ora = OracleMPS(10, 3, dtype_mode="c128")
    for qudits, mat in circ.ops:
        ora.apply_two_qudit_gate(mat.reshape(3, 3, 3, 3), *qudits, maxsvals=20,
                                 keep_left_canonical=left)
    mps, rec = _recorded_updates(m, circ, 20, left)
    assert mps.bond_dimensions() == ora.bond_dimensions()
    err, nr, nx = gauge_invariant_error(ora.tensors, [t.cpu().numpy() for t in mps.tensors])
    assert err <= 1e-4, err
    assert abs(nr - nx) <= 1e-4 * nr
    print(f"qutrits left={left}: overlap error {err:.2e}, "
          f"worst kept-value error {_check_spectra(rec, ora):.2e}")


@pytest.mark.parametrize("d,n,depth", [(5, 4, 4), (6, 3, 3), (8, 3, 3), (16, 3, 2)])
def test_large_qudit_untruncated_vs_statevector(d, n, depth):
    """Qudit dimensions above 4 (core.py:1013-1028 accepts any d whose gate
    edges match the MPS; here the generic theta kernel): untruncated d^2 x d^2
    Haar brickwork after a d x d Haar gate on every qudit (the generic
    one-site kernel) against the FP64 state vector (<= 1e-5). A gate on a
    non-adjacent pair raises as in the reference, whose SWAP network applies
    the qubit SWAP (core.py:1188-1189, "SWAP is only for qubits")."""
    m = _mpsim()
    from mpsim_amd.circuits import brickwork_haar
    from mpsim_amd.gates import haar_unitary
    circ = brickwork_haar(n, depth, d, maxsvals=None, seed=50 + d)
    rng = np.random.RandomState(60 + d)
    ops = [((q,), haar_unitary(d, rng)) for q in range(n)] + list(circ.ops)
    mps = m.MPS(n, d)
    for qudits, mat in ops:
        if len(qudits) == 1:
            mps.apply_one_qudit_gate(mat, qudits[0])
        else:
            mps.apply_two_qudit_gate(mat.reshape(d, d, d, d), *qudits)
    ref = sv.run(ops, n, d)
    err = np.max(np.abs(mps.wavefunction() - ref))
    print(f"d={d} n={n}: max error vs FP64 state vector {err:.2e}")
    assert err <= 1e-5, err
    far = haar_unitary(d * d, np.random.RandomState(d)).reshape(d, d, d, d)
    with pytest.raises(ValueError, match="should have MPS qudit dimension"):
        mps.apply_two_qudit_gate(far, 0, n - 1)


@pytest.mark.parametrize("left", [True, False])
def test_large_qudit_truncated_vs_fp64_oracle(left):
    """d = 5 (generic theta kernel), 6 qudits, depth 6, maxsvals 12: bond
    dimensions equal to the FP64 oracle's, gauge-invariant overlap error and
    norm within 1e-4, and every update's kept singular values and discarded
    weight against the oracle's split (_check_spectra)."""
    m = _mpsim()
    from mpsim_amd.circuits import brickwork_haar
    from oracle.mps_oracle import gauge_invariant_error
    d, cap = 5, 12
    circ = brickwork_haar(6, 6, d, maxsvals=cap, seed=55)
    ora = OracleMPS(6, d, dtype_mode="c128")
    for qudits, mat in circ.ops:
        ora.apply_two_qudit_gate(mat.reshape(d, d, d, d), *qudits, maxsvals=cap,
                                 keep_left_canonical=left)
    mps, rec = _recorded_updates(m, circ, cap, left)
    assert mps.bond_dimensions() == ora.bond_dimensions()
    err, nr, nx = gauge_invariant_error(ora.tensors, [t.cpu().numpy() for t in mps.tensors])
    assert err <= 1e-4, err
    assert abs(nr - nx) <= 1e-4 * nr
    print(f"d=5 left={left}: overlap error {err:.2e}, "
          f"worst kept-value error {_check_spectra(rec, ora):.2e}")


def test_singular_values_of_one_update_vs_lapack():
    """Kept singular values of a single two-site update at a cap-size shape
    (theta 512 x 512 -> k = 256) against numpy's FP64 SVD: <= 1e-4 relative
    (the spec's bound) on every kept value."""
    m = _mpsim()
    import torch
    from mpsim_amd.engine import Engine, TwoSiteItem
    rng = np.random.RandomState(17)
    cl = cm = cr = 256
    A = ((rng.randn(cl, 2, cm) + 1j * rng.randn(cl, 2, cm)) / np.sqrt(4 * cm)).astype(np.complex64)
    B = ((rng.randn(cm, 2, cr) + 1j * rng.randn(cm, 2, cr)) / np.sqrt(2)).astype(np.complex64)
    G = m.gates.haar_unitary(4, rng)
    eng = Engine.get()
    res = eng.two_site([TwoSiteItem(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(),
                                    eng.gate(G), 256, want_s=True)])[0]
    theta = np.einsum("abpq,lpqr->labr", G.reshape(2, 2, 2, 2).astype(np.complex128),
                      np.tensordot(A.astype(np.complex128), B.astype(np.complex128), axes=([2], [0])))
    s_ref = np.linalg.svd(theta.reshape(2 * cl, 2 * cr), compute_uv=False)
    s = res.s.cpu().numpy().astype(np.float64)
    rel = np.max(np.abs(s - s_ref[:256]) / s_ref[:256])
    assert rel <= 1e-4, rel
    disc = float(res.discarded.cpu())
    assert abs(disc - np.sum(s_ref[256:] ** 2)) <= 1e-4 * np.sum(s_ref ** 2)


def test_two_site_shape_sweep_vs_lapack():
    """One grouped library call over 40 seeded two-site problems of mixed
    qudit dimension (d in {2,3,4,5,7}: specialised and generic theta kernels
    in the same call), bond dimensions 1..64 (including chain-end 1s), every
    kind of k (0, 1, a cut, all values) and both absorb directions, each
    against numpy's FP64 SVD of the same theta (formed from the complex64
    inputs): kept singular values within 1e-4 relative + 1e-6 ||theta||_F,
    discarded weight, the truncated product A'B' equal to the best rank-k
    approximation within 1e-4 ||theta||_F, and the unscaled factor
    orthonormal (core.py:1132-1145)."""
    m = _mpsim()
    import torch
    from mpsim_amd.engine import Engine, TwoSiteItem
    rng = np.random.RandomState(2024)
    eng = Engine.get()
    bonds = [1, 2, 3, 7, 16, 33, 64]
    cases, items = [], []
    for _ in range(40):
        d = int(rng.choice([2, 3, 4, 5, 7]))
        cl, cm, cr = (int(rng.choice(bonds)) for _ in range(3))
        M, N = d * cl, d * cr
        q = min(M, N)
        k = int(rng.choice([0, 1, q, int(rng.randint(1, q + 1))]))
        left = bool(rng.randint(2))
        A = ((rng.randn(cl, d, cm) + 1j * rng.randn(cl, d, cm)) / np.sqrt(2 * d * cm)).astype(np.complex64)
        B = ((rng.randn(cm, d, cr) + 1j * rng.randn(cm, d, cr)) / np.sqrt(2)).astype(np.complex64)
        G = m.gates.haar_unitary(d * d, rng)
        cases.append((d, cl, cm, cr, k, left, A, B, G))
        items.append(TwoSiteItem(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(),
                                 eng.gate(G), k, absorb_left=left, want_s=True))
    res = eng.two_site(items)
    torch.cuda.synchronize()
    for (d, cl, cm, cr, k, left, A, B, G), r in zip(cases, res):
        tag = (d, cl, cm, cr, k, left)
        theta = np.einsum("abpq,lpqr->labr", G.reshape(d, d, d, d).astype(np.complex128),
                          np.tensordot(A.astype(np.complex128), B.astype(np.complex128),
                                       axes=([2], [0]))).reshape(d * cl, d * cr)
        u, s_ref, vh = np.linalg.svd(theta, full_matrices=False)
        fro = np.linalg.norm(theta)
        a = r.a_out.cpu().numpy().reshape(d * cl, k)
        b = r.b_out.cpu().numpy().reshape(k, d * cr)
        s = r.s.cpu().numpy().astype(np.float64) if k else np.zeros(0)
        assert np.all(np.abs(s - s_ref[:k]) <= 1e-4 * s_ref[:k] + 1e-6 * fro), tag
        assert abs(float(r.discarded.cpu()) - np.sum(s_ref[k:] ** 2)) <= 1e-5 * fro ** 2, tag
        best = (u[:, :k] * s_ref[:k]) @ vh[:k]
        assert np.linalg.norm(a.astype(np.complex128) @ b - best) <= 1e-4 * max(fro, 1e-30), tag
        unscaled = b @ b.conj().T if left else a.conj().T @ a
        assert np.allclose(unscaled, np.eye(k), atol=1e-5), tag
