"""GPU tests of the captured schedule (mpsim_amd/program.py): a batch of
uneven circuits run eagerly once, then replayed as one hipGraph, must give
bitwise the factors of the layer-batched eager executor (which is itself
checked against the oracle in test_gpu_batch.py). Replays are interleaved
with unrelated eager work and synchronisations, and each replay is checked.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _circuits(seed0, count=8):
    from mpsim_amd.circuits import random_pairs_circuit
    out = []
    for c in range(count):
        r = np.random.RandomState(seed0 + c)
        out.append(random_pairs_circuit(seed0 + 100 + c, int(r.randint(6, 24)),
                                        int(r.randint(3, 10)), int(r.choice([4, 8, 16, 32]))))
    return out


def _eager(circuits):
    from mpsim_amd import batch
    dcs = [batch.prepare(c) for c in circuits]
    batch.run_batched(dcs)
    return [[t.cpu().numpy() for t in dc.mps.tensors] for dc in dcs]


def _same(program, ref):
    for mps, tensors in zip(program.states(), ref):
        assert [tuple(t.shape) for t in mps.tensors] == [t.shape for t in tensors]
        for a, b in zip(mps.tensors, tensors):
            assert np.array_equal(a.cpu().numpy(), b)


@pytest.mark.parametrize("graph", [True, False])
def test_program_replays_match_eager_batches(graph):
    import torch
    from mpsim_amd.program import CircuitProgram
    circuits = _circuits(70_000)
    ref = _eager(circuits)
    prog = CircuitProgram(circuits, graph=graph)
    try:
        for rep in range(3):
            prog.run()
            assert prog.last_unconverged == 0
            _same(prog, ref)
            # unrelated eager work between replays
            x = torch.randn(4096, device="cuda")
            assert torch.isfinite(x.sum()).item()
            _eager(_circuits(80_000, 2))
            torch.cuda.synchronize()
        assert (prog.eager_runs, prog.replays) == ((1, 2) if graph else (3, 0))
    finally:
        prog.close()


def test_program_new_gate_values_replay():
    """update_gates writes new gate values of the same structure into the
    resident buffers; the replay equals an eager run of the new circuits."""
    from mpsim_amd.circuits import Circuit
    from mpsim_amd.program import CircuitProgram
    circuits = _circuits(90_000, 5)
    prog = CircuitProgram(circuits)
    try:
        prog.run()
        rng = np.random.RandomState(5)
        from mpsim_amd.gates import haar_unitary
        fresh = [Circuit(c.n, c.d, [[(q, haar_unitary(c.d ** len(q), rng)) for q, _ in layer]
                                    for layer in c.layers], c.maxsvals) for c in circuits]
        prog.update_gates(fresh)
        prog.run()
        assert prog.replays == 1 or prog.eager_runs == 2
        _same(prog, _eager(fresh))
    finally:
        prog.close()


def test_program_states_match_oracle():
    """The final states against the complex128 oracle (untruncated caps)."""
    from mpsim_amd.circuits import random_pairs_circuit
    from mpsim_amd.program import CircuitProgram
    from oracle.mps_oracle import OracleMPS, gauge_invariant_error
    circuits = [random_pairs_circuit(95_000 + c, 10 + 2 * c, 6, 64) for c in range(3)]
    prog = CircuitProgram(circuits)
    try:
        prog.run()
        prog.run()
        for c, mps in zip(circuits, prog.states()):
            o = OracleMPS(c.n, c.d, dtype_mode="c128")
            o.apply_ops(c.ops, maxsvals=c.maxsvals)
            err, nr, nx = gauge_invariant_error(o.tensors, [t.cpu().numpy() for t in mps.tensors])
            assert err < 1e-6 and abs(nx - nr) < 1e-5
    finally:
        prog.close()
