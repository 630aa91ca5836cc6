"""GPU tests of the layer-batched executor and the serving loop.

Every two-site problem in a grouped launch is computed by workgroups that
only read that problem's data, and its Jacobi iteration stops on its own
convergence flag, so a circuit's result does not depend on what else is in
the batch: these tests require bit-identical tensors alone and in a mix,
and agreement with the FP64 oracle within the truncated-state tolerance
(1 - |<ref|psi>|/(|ref||psi|) <= 1e-4).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same(ta, tb):
    return len(ta) == len(tb) and all(
        tuple(a.shape) == tuple(b.shape) and np.array_equal(a.cpu().numpy(), b.cpu().numpy())
        for a, b in zip(ta, tb))


def test_batched_equals_plain_path_bitwise():
    from mpsim_amd.batch import prepare, run_batched, run_plain
    from mpsim_amd.circuits import brickwork_haar, random_pairs_circuit
    for circ in (brickwork_haar(10, 6, 2, maxsvals=8, seed=31),
                 random_pairs_circuit(32, 9, 6, maxsvals=16)):
        dc = prepare(circ)
        run_batched([dc])
        plain = run_plain(circ)
        assert _same(dc.mps.tensors, plain.tensors)


def test_uneven_batch_vs_oracle():
    from oracle.mps_oracle import OracleMPS, gauge_invariant_error
    from mpsim_amd.batch import prepare, run_batched
    from mpsim_amd.circuits import random_pairs_circuit
    circs = [random_pairs_circuit(100 + i, n, depth, cap)
             for i, (n, depth, cap) in enumerate([(8, 5, 4), (12, 4, 8), (10, 7, 16), (6, 6, 32),
                                                  (14, 3, 8)])]
    dcs = [prepare(c) for c in circs]
    run_batched(dcs)
    for c, dc in zip(circs, dcs):
        ora = OracleMPS(c.n, c.d, dtype_mode="c128")
        ora.apply_ops(c.ops, maxsvals=c.maxsvals)
        assert dc.mps.bond_dimensions() == ora.bond_dimensions()
        err, nr, nx = gauge_invariant_error(ora.tensors, [t.cpu().numpy() for t in dc.mps.tensors])
        assert err <= 1e-4, (c.name, err)
        assert abs(nr - nx) <= 1e-4 * nr


def test_serving_result_same_alone_and_in_mix():
    from mpsim_amd.circuits import Request, poisson_requests
    from mpsim_amd.serve import MPSServer
    reqs = poisson_requests(rate=400.0, duration=0.05, limit=8)
    # keep the test small: shorten every circuit to at most 6 layers
    for r in reqs:
        r.circuit.layers = r.circuit.layers[:6]
        r.circuit.n = min(r.circuit.n, 40)
        r.circuit.layers = [[op for op in layer if max(op[0]) < r.circuit.n]
                            for layer in r.circuit.layers]
    srv = MPSServer(keep_results=True, hbm_budget=8 << 30)
    mixed = srv.run(reqs, realtime=False).results
    for r in reqs[:3]:
        alone = srv.run([Request(r.rid, 0.0, r.circuit)], realtime=False).results[r.rid]
        assert _same(alone.tensors, mixed[r.rid].tensors), r.rid


def test_admission_respects_budget():
    from mpsim_amd.circuits import Request, brickwork_haar
    from mpsim_amd.serve import MPSServer, footprint_bytes
    circs = [brickwork_haar(12, 4, 2, maxsvals=8, seed=s) for s in range(4)]
    need = footprint_bytes(circs[0])
    srv = MPSServer(hbm_budget=int(need * 2.5))
    rep = srv.run([Request(i, 0.0, c) for i, c in enumerate(circs)], realtime=False)
    assert rep.completed == 4 and rep.max_active == 2
    with pytest.raises(ValueError):
        MPSServer(hbm_budget=need // 2).run([Request(0, 0.0, circs[0])], realtime=False)
