"""GPU tests of the MPSimulator surface (mpsim/mpsim_cirq/simulator.py),
parameter sweeps and the per-gate norm log.

cirq is not installed, so circuits are stand-in objects with the interface
the simulator uses (all_qubits(), all_operations(); operations with qubits,
_has_unitary_(), _unitary_()), built from the standard gate matrices; the
reference's own tests (simulator_test.py:77-94, 274-305) are restated on
them and checked against the FP64 state vector and the FP64 oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]])
Z = np.diag([1.0, -1.0]).astype(complex)
H = np.array([[1, 1], [1, -1]], dtype=complex) / np.sqrt(2)
S = np.diag([1, 1j])
T = np.diag([1, np.exp(1j * np.pi / 4)])
CNOT = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=complex)
CZ = np.diag([1, 1, 1, -1]).astype(complex)
SWAP = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=complex)
ISWAP = np.array([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]])


def fsim(theta, phi):
    c, s = np.cos(theta), np.sin(theta)
    return np.array([[1, 0, 0, 0], [0, c, -1j * s, 0], [0, -1j * s, c, 0],
                     [0, 0, 0, np.exp(-1j * phi)]])


# the gate domain of simulator_test.py:279-292 (CZPowGate() is CZ)
ONE = [X, Y, Z, H, S, T]
TWO = [CNOT, CZ, SWAP, CZ, ISWAP, fsim(0.2, 0.3)]


class Op:
    """Stand-in for a cirq.GateOperation."""

    def __init__(self, qubits, mat):
        self.qubits = tuple(qubits)
        self._mat = mat

    def _has_unitary_(self):
        return self._mat is not None

    def _unitary_(self):
        return self._mat


class Circ:
    """Stand-in for a cirq.Circuit (qubits are ints: their sort order is
    QubitOrder.DEFAULT's)."""

    def __init__(self, ops):
        self.ops = list(ops)

    def all_qubits(self):
        return frozenset(q for op in self.ops for q in op.qubits)

    def all_operations(self):
        return iter(self.ops)


def random_circuit(rng, nqubits, moments=25, density=0.999):
    """Moments of gates from the domain on random disjoint qubit sets
    (cirq.testing.random_circuit's shape: any pairs, any distance)."""
    ops = []
    for _ in range(moments):
        free = list(rng.permutation(nqubits))
        while free:
            two = len(free) >= 2 and rng.rand() < 0.5
            qs = [int(free.pop()) for _ in range(2 if two else 1)]
            if rng.rand() < density:
                dom = TWO if two else ONE
                ops.append(Op(qs, dom[rng.randint(len(dom))]))
    return Circ(ops)


def _sv(circ, n):
    from oracle import statevector as sv
    return sv.run([(op.qubits, op._unitary_()) for op in circ.ops], n)


def test_simulate_bell_state_with_truncation():
    """simulator_test.py:77-94: H then CNOT with options {"maxsvals": 1}
    gives [1, 0, 0, 0]/sqrt(2); without options the Bell state."""
    from mpsim_amd.mpsim_cirq import MPSimulator
    circ = Circ([Op([0], H), Op([0, 1], CNOT)])
    res = MPSimulator(options={"maxsvals": 1}).simulate(circ)
    assert np.allclose(res.wavefunction(), np.array([1, 0, 0, 0]) / np.sqrt(2), atol=1e-6)
    res = MPSimulator().simulate(circ)
    assert np.allclose(res.wavefunction(), np.array([1, 0, 0, 1]) / np.sqrt(2), atol=1e-6)
    # the same through an operation list and a Circuit
    from mpsim_amd.circuits import Circuit
    pairs = [((0,), H), ((0, 1), CNOT)]
    res = MPSimulator(options={"maxsvals": 1}).simulate(pairs)
    assert np.allclose(res.wavefunction(), np.array([1, 0, 0, 0]) / np.sqrt(2), atol=1e-6)
    res = MPSimulator().simulate(Circuit(2, 2, [[((0,), H)], [((0, 1), CNOT)]], maxsvals=1))
    assert np.allclose(res.wavefunction(), np.array([1, 0, 0, 0]) / np.sqrt(2), atol=1e-6)
    with pytest.raises(ValueError):  # an empty circuit has no qubits (simulator_test.py:16-19)
        MPSimulator().simulate(Circ([]))


@pytest.mark.parametrize("nqubits", [2, 4, 8])
def test_random_circuits(nqubits):
    """simulator_test.py:274-305: random circuits of 25 moments over the gate
    domain, no truncation: the final wavefunction equals the FP64 state
    vector to 1e-5."""
    from mpsim_amd.mpsim_cirq import MPSimulator
    rng = np.random.RandomState(1)
    sim = MPSimulator()
    for _ in range(10):
        circ = random_circuit(rng, nqubits)
        mps = sim.simulate(circ)
        assert np.max(np.abs(mps.wavefunction() - _sv(circ, nqubits))) <= 1e-5


def haar_circuit(rng, nqubits, moments):
    """Moments of Haar-random one- and two-qubit gates (gates.py:269-285
    recipe) on random disjoint qubit sets: any distance, either order."""
    from mpsim_amd.gates import haar_unitary
    ops = []
    for _ in range(moments):
        free = list(rng.permutation(nqubits))
        while free:
            two = len(free) >= 2 and rng.rand() < 0.5
            qs = [int(free.pop()) for _ in range(2 if two else 1)]
            ops.append(Op(qs, haar_unitary(4 if two else 2, rng)))
    return Circ(ops)


def test_truncated_random_circuits_vs_oracle():
    """options {"maxsvals": k} forwarded to every gate, SWAP-network hops
    included, on 8 qubits: bond dimensions, overlap error <= 1e-6 and norms
    against the FP64 oracle running the same ops with the same kwargs.
    Haar-random gates keep the spectra non-degenerate: with the Clifford
    gate domain above, truncation cuts through exactly degenerate Schmidt
    values, where the kept subspace is arbitrary (complex64 LAPACK ends
    0.99 from complex128 on such a circuit at maxsvals=2)."""
    from mpsim_amd.mpsim_cirq import MPSimulator
    from oracle.mps_oracle import OracleMPS, gauge_invariant_error
    rng = np.random.RandomState(3)
    for k in (4, 8, 16):
        circ = haar_circuit(rng, 8, 10)
        mps = MPSimulator(options={"maxsvals": k}).simulate(circ)
        ora = OracleMPS(8, 2, dtype_mode="c128")
        ora.apply_ops([(op.qubits, op._unitary_()) for op in circ.ops], maxsvals=k)
        assert mps.bond_dimensions() == ora.bond_dimensions()
        err, nr, nx = gauge_invariant_error(ora.tensors, [t.cpu().numpy() for t in mps.tensors])
        print(f"maxsvals {k}: overlap error {err:.2e}")
        assert err <= 1e-6 and abs(nr - nx) <= 1e-6 * nr, (k, err)


def test_parameter_sweep_runs_batched_and_matches_eager():
    """simulate_sweep(circuit, params): every parameter set (new gate values
    of the same structure) runs in one captured batch; each final state is
    bitwise the eager simulation of that set, and a second sweep of the
    same structure reuses the program (update_gates)."""
    from mpsim_amd.circuits import brickwork_haar
    from mpsim_amd.mpsim_cirq import MPSimulator
    base = brickwork_haar(10, 6, 2, maxsvals=8, seed=60)
    sets = [brickwork_haar(10, 6, 2, maxsvals=8, seed=61 + i) for i in range(4)]
    sim = MPSimulator()
    circuits = sets[:2] + [base, sets[2]]
    matrices = [[m for _, m in c.ops] for c in sets]
    for params, same_as in ((circuits, circuits), (matrices, sets)):
        res = sim.simulate_sweep(base, params=params)
        assert len(res) == 4 and len(sim._programs) == 1
        for mps, c in zip(res, same_as):
            eager = sim.simulate(c)
            for a, b in zip(mps.tensors, eager.tensors):
                assert tuple(a.shape) == tuple(b.shape)
                assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())


def test_per_gate_norm_log():
    """MPS(track_norms=True) appends the full norm after every two-qudit gate
    (core.py:1160-1161); a non-adjacent gate's SWAP hops are two-qudit gates
    of their own and log too. The norms match the FP64 oracle's to 1e-5."""
    import mpsim_amd as m
    from mpsim_amd.circuits import brickwork_haar
    from oracle.mps_oracle import OracleMPS
    circ = brickwork_haar(8, 6, 2, maxsvals=4, seed=70)
    ops = circ.ops + [((0, 5), CNOT), ((6, 2), ISWAP)]
    mps = m.MPS(8, track_norms=True)
    ora = OracleMPS(8, 2, dtype_mode="c128", track_norms=True)
    for q, g in ops:
        mps.apply_two_qudit_gate(np.asarray(g).reshape(2, 2, 2, 2), *q, maxsvals=4)
        ora.apply_two_qudit_gate(np.asarray(g).reshape(2, 2, 2, 2), *q, maxsvals=4)
    # each SWAP hop is an apply_two_qudit_gate call of its own and logs too
    hops = 2 * (5 - 0 - 1) + 2 * (6 - 2 - 1)
    assert len(mps._norms) == len(ora.norms) == len(ops) + hops
    assert np.allclose(mps._norms, ora.norms, rtol=1e-5, atol=0)
