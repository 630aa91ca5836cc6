"""Host tests of the batched schedule (circuits.schedule) against the FP64
state vector (no GPU).

The schedule turns a circuit's layers into steps of grouped launches: SWAP
networks for non-adjacent pairs (core.py:1035-1043, 1154-1219), pairs given
as (j, i) flipped (core.py:1031-1033), ops on disjoint sites sharing steps.
Applying the steps as written (SWAP hops, flipped gates, one-qudit gates) to
a dense state must give what applying the circuit's ops in order gives
(core.py:1221-1247; the reference's own non-local tests are core_test.py:
1225-1258 and simulator_test.py:113-144).
"""
import numpy as np
import pytest

from oracle import statevector as sv


def _random_layers(rng, n, d, depth, width):
    from mpsim_amd.gates import haar_unitary
    layers = []
    for _ in range(depth):
        layer = []
        for _ in range(width):
            if rng.rand() < 0.3:
                layer.append(((int(rng.randint(n)),), haar_unitary(d, rng)))
            else:
                i, j = (int(x) for x in rng.choice(n, size=2, replace=False))
                layer.append(((i, j), haar_unitary(d * d, rng)))
        layers.append(layer)
    return layers


def _run_steps(circuit, ignore_flip=False):
    from mpsim_amd.circuits import flip_pair, schedule, swap_matrix
    steps, ends = schedule(circuit)
    n, d = circuit.n, circuit.d
    psi = sv.zero_state(n, d)
    for st in steps:
        sites = []
        for site, t, o in st.ones:
            sites.append(site)
            psi = sv.apply_gate(psi, circuit.layers[t][o][1], (site,), n, d)
        for site, t, o, fl in st.twos:
            sites += [site, site + 1]
            if o < 0:
                g = swap_matrix(d)
            else:
                g = circuit.layers[t][o][1]
                g = flip_pair(g, d) if fl and not ignore_flip else g
            psi = sv.apply_gate(psi, g, (site, site + 1), n, d)
        assert len(sites) == len(set(sites)), "ops of one step share a site"
    assert ends[-1] == len(steps)
    return psi


@pytest.mark.parametrize("n,d,depth,width,seed", [(6, 2, 5, 4, 1), (7, 2, 4, 6, 2), (4, 3, 4, 3, 3),
                                                  (5, 2, 3, 8, 4), (2, 2, 3, 2, 5)])
def test_steps_apply_the_circuit(n, d, depth, width, seed):
    from mpsim_amd.circuits import Circuit
    rng = np.random.RandomState(seed)
    c = Circuit(n, d, _random_layers(rng, n, d, depth, width))
    want = sv.run(c.ops, n, d)
    got = _run_steps(c)
    assert np.abs(got - want).max() <= 1e-12


def test_nonlocal_cnot_and_ghz():
    """core_test.py:1225-1258: CNOT on (0, n-1) after H on 0 gives
    (|0..0> + |1 0..0 1>)/sqrt 2; CNOT(n-1, 0) flips qubit 0 (the gate
    library's matrices are complex64, as the reference's are)."""
    from mpsim_amd.circuits import Circuit
    from mpsim_amd.gates import cnot, hgate, xgate
    for n in (3, 5, 8):
        h = np.asarray(hgate().tensor)
        cx = np.asarray(cnot().tensor).reshape(4, 4)
        x = np.asarray(xgate().tensor)
        psi = _run_steps(Circuit(n, 2, [[((0,), h)], [((0, n - 1), cx)]]))
        want = np.zeros(2 ** n, complex)
        want[0] = want[2 ** (n - 1) + 1] = 2 ** -0.5
        assert np.abs(psi - want).max() <= 1e-7
        psi = _run_steps(Circuit(n, 2, [[((n - 1,), x)], [((n - 1, 0), cx)]]))
        want = np.zeros(2 ** n, complex)
        want[2 ** (n - 1) + 1] = 1.0
        assert np.abs(psi - want).max() <= 1e-7


def test_disjoint_ops_share_steps_and_swaps_return():
    """A layer of disjoint nearest-neighbour ops is one step; a pair at
    distance k costs 2 (k - 1) SWAP hops around its gate."""
    from mpsim_amd.circuits import Circuit, schedule
    u = np.eye(4, dtype=complex)
    steps, _ = schedule(Circuit(8, 2, [[((0, 1), u), ((2, 3), u), ((5, 4), u), ((7,), np.eye(2))]]))
    assert len(steps) == 1 and len(steps[0].twos) == 3 and len(steps[0].ones) == 1
    for k in (2, 3, 5):
        steps, _ = schedule(Circuit(8, 2, [[((1, 1 + k), u)]]))
        twos = [tw for st in steps for tw in st.twos]
        assert len(twos) == 2 * (k - 1) + 1
        assert sum(1 for tw in twos if tw[2] < 0) == 2 * (k - 1)


def test_check_catches_a_dropped_flip():
    """The comparison above fails when the steps ignore the flip of pairs
    given as (j, i)."""
    from mpsim_amd.circuits import Circuit
    from mpsim_amd.gates import haar_unitary
    rng = np.random.RandomState(8)
    c = Circuit(5, 2, [[((3, 1), haar_unitary(4, rng))], [((4, 3), haar_unitary(4, rng))]])
    want = sv.run(c.ops, 5, 2)
    assert np.abs(_run_steps(c) - want).max() <= 1e-12
    assert np.abs(_run_steps(c, ignore_flip=True) - want).max() > 1e-3
