"""GPU parity at the full size of every benchmark workload.

The 16-qubit untruncated circuit is checked against the FP64 state vector in
test_gpu_core.py. Here:
  * 64 qubits, brickwork depth 20, maxsvals 256 (the headline workload),
    layer-batched: against the complex128 oracle run of the same circuit
    (overlap, bond dimensions, every update's kept spectrum along the run),
    and every update of two layers against LAPACK in FP64 on the GPU's own
    inputs (kept values to 1e-4 relative, discarded weight);
  * the 512 uneven circuits: batched eagerly and as one captured hipGraph
    (replays bitwise equal to the eager schedule), every circuit against the
    oracle;
  * 48 qutrits, depth 16, maxsvals 1024 (theta up to 3072 x 3072): the full
    circuit with per-update checks (||theta||^2 = kept + discarded, finite
    outputs, bond dimensions) and the first 3072 x 3072 update against LAPACK
    in FP64 (kept values, discarded weight, reconstruction);
  * the first requests of the seeded 200 req/s serving stream at full size:
    each result bitwise the same alone and in the mix, the small-cap ones
    against the oracle.
The oracle is the FP64 restatement of the reference's algorithm
(oracle/mps_oracle.py); LAPACK is numpy's np.linalg.svd in complex128.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _theta64(a, b, g, d):
    """FP64 theta = G (A B) from the complex64 inputs of one update, as the
    M x N matrix with rows (l, o1) and columns (o2, r)."""
    A = np.asarray(a, dtype=np.complex128)
    B = np.asarray(b, dtype=np.complex128)
    G = np.asarray(g, dtype=np.complex128).reshape(d, d, d, d)
    t = np.einsum("abpq,lpqr->labr", G, np.tensordot(A, B, axes=([2], [0])))
    return t.reshape(A.shape[0] * d, d * B.shape[2])


def _host(x):
    return x.cpu().numpy()


@pytest.mark.timeout(600)
def test_large_64_qubits_chi256_vs_fp64():
    """The headline workload (64 qubits, depth 20 of Haar 4x4 brickwork,
    maxsvals 256, 630 updates), layer-batched as the benchmark runs it.

    1. Against the complex128 oracle run of the same circuit: bond
       dimensions equal; overlap error 1 - |<ref|psi>|/(|ref||psi|) <= 1e-4
       (the spec) and <= 1e-5 (complex64 LAPACK on this circuit: 1.4e-6);
       relative norm error <= 1e-5.
    2. Every update of layers 10 and 19 (63 updates at the cap, theta
       512 x 512 -> 256) against LAPACK in FP64 on the GPU's own inputs:
       every kept singular value within 1e-4 relative, the discarded weight
       within 1e-6 of ||theta||^2.
    3. Every update's kept spectrum along the run against the oracle's: the
       two runs drift apart as truncation errors propagate (complex64 LAPACK
       against the complex128 run reaches 2e-4 relative on values above
       1e-2 sigma_1 by the last layer), so the bound here is 2e-3 on those.
    """
    import torch
    from mpsim_amd.batch import prepare, tick
    from mpsim_amd.circuits import bond_plan, brickwork_haar
    from mpsim_amd.engine import Engine
    from oracle.mps_oracle import OracleMPS, gauge_invariant_error
    circ = brickwork_haar(64, 20, 2, maxsvals=256, seed=2)
    eng = Engine.get()
    dc = prepare(circ, eng)
    spectra, sampled = [], []

    def observe(dc_, i, it, r):
        spectra.append(_host(r.s).astype(np.float64))
        if dc_.t - 1 in (10, 19):
            sampled.append((_host(it.a), _host(it.b), _host(it.gate), spectra[-1],
                            float(r.discarded.cpu())))

    print("large: 64 qubits chi 256 on the GPU", flush=True)
    while not dc.done:
        tick(eng, [dc], want_s=True, observer=observe)
    torch.cuda.synchronize()
    mine = [_host(t) for t in dc.mps.tensors]
    assert all(np.all(np.isfinite(t)) for t in mine)
    assert dc.mps.bond_dimensions() == bond_plan(circ)[-1]
    assert len(sampled) == 63

    # 2. per update on the GPU's own inputs
    worst = 0.0
    for a, b, g, s, disc in sampled:
        th = _theta64(a, b, g, 2)
        s_ref = np.linalg.svd(th, compute_uv=False)
        k = s.size
        rel = np.abs(s - s_ref[:k]) / s_ref[:k]
        worst = max(worst, float(rel.max()))
        assert np.all(rel <= 1e-4), float(rel.max())
        fro2 = float(np.sum(s_ref ** 2))
        assert abs(disc - float(np.sum(s_ref[k:] ** 2))) <= 1e-6 * fro2
    print(f"large: worst kept-value error vs LAPACK on the same input: {worst:.2e}", flush=True)

    # 1. and 3. against the oracle run
    print("large: complex128 oracle run", flush=True)
    ora = OracleMPS(64, 2, dtype_mode="c128")
    ora.apply_ops(circ.ops, maxsvals=256)
    assert dc.mps.bond_dimensions() == ora.bond_dimensions()
    err, nr, nx = gauge_invariant_error(ora.tensors, mine)
    print(f"large: overlap error vs FP64 oracle {err:.3e}, norms {nr:.9f} {nx:.9f}", flush=True)
    assert err <= 1e-4 and err <= 1e-5
    assert abs(nr - nx) <= 1e-5 * nr
    drift = 0.0
    assert len(spectra) == len(ora.spectra) == 630
    for s, s_ref in zip(spectra, ora.spectra):
        big = s_ref >= 1e-2 * s_ref[0]
        drift = max(drift, float(np.max(np.abs(s[big] - s_ref[big]) / s_ref[big])))
    print(f"large: kept-spectrum drift along the run (sigma >= 1e-2 sigma_1): {drift:.2e}")
    assert drift <= 2e-3


@pytest.mark.timeout(600)
def test_uneven_512_circuits_batched_and_captured():
    """The 512 seeded uneven circuits (n in [12,64], depth in [4,24], chi cap
    in {32,64,128,256}): layer-batched eagerly, then as one CircuitProgram
    (first run eager, then two hipGraph replays with unrelated GPU work
    between them). Every replay is bitwise equal to the program's eager run
    and to the layer-batched run; every circuit's bond dimensions follow the
    host plan, and every circuit matches the complex128 oracle (overlap error
    <= 1e-4, relative norm error <= 1e-4)."""
    import torch
    from mpsim_amd.batch import prepare, run_batched
    from mpsim_amd.circuits import bond_plan, uneven_batch
    from mpsim_amd.engine import Engine
    from mpsim_amd.program import CircuitProgram
    from oracle.mps_oracle import OracleMPS, gauge_invariant_error
    circs = uneven_batch(512)
    eng = Engine.get()
    print("uneven: layer-batched run", flush=True)
    dcs = [prepare(c, eng) for c in circs]
    run_batched(dcs, eng)
    torch.cuda.synchronize()
    eager = [[_host(t) for t in dc.mps.tensors] for dc in dcs]
    del dcs
    for c, ts in zip(circs, eager):
        plan = bond_plan(c)
        want = plan[-1] if plan else [1] * (c.n - 1)
        assert [t.shape[2] for t in ts[:-1]] == want, c.name
        assert all(np.all(np.isfinite(t)) for t in ts)
    print("uneven: captured program", flush=True)
    prog = CircuitProgram(circs, eng)
    try:
        for rep in range(3):
            prog.run()
            assert prog.last_unconverged == 0
            for mps, ts in zip(prog.states(), eager):
                for a, b in zip(mps.tensors, ts):
                    assert tuple(a.shape) == b.shape and np.array_equal(_host(a), b)
            x = torch.randn(1 << 20, device="cuda")
            assert torch.isfinite(x.sum()).item()
            torch.cuda.synchronize()
        assert (prog.eager_runs, prog.replays) == (1, 2)
    finally:
        prog.close()
    print("uneven: oracle", flush=True)
    worst = 0.0
    for c, ts in zip(circs, eager):
        ora = OracleMPS(c.n, 2, dtype_mode="c128")
        ora.apply_ops(c.ops, maxsvals=c.maxsvals)
        err, nr, nx = gauge_invariant_error(ora.tensors, ts)
        assert err <= 1e-4 and abs(nr - nx) <= 1e-4 * nr, (c.name, err)
        worst = max(worst, err)
    print(f"uneven: worst overlap error vs FP64 oracle over 512 circuits {worst:.2e}")


@pytest.mark.timeout(900)
def test_qutrits_48_chi1024():
    """The precision stress: 48 qutrits, depth 16 of Haar 9x9 brickwork,
    maxsvals 1024 (theta up to 3072 x 3072, 376 updates). The FP64 oracle of
    the whole circuit takes hours on the host, so:
      * every update: outputs finite, and ||theta||_F^2 = sum of the kept
        s^2 + the discarded weight to 1e-5 relative (||theta||^2 from the
        inputs in complex128 on the GPU);
      * bond dimensions after the circuit equal the host plan;
      * the first 3072 x 3072 update against LAPACK in FP64 on the GPU's
        input: every kept singular value (k = 1024) within 1e-4 relative,
        the discarded weight within 1e-6 ||theta||^2, and
        ||theta - a_out b_out||^2 equal to the discarded weight to
        1e-5 ||theta||^2 (complex64 outputs).
    """
    import torch
    from mpsim_amd.batch import prepare, tick
    from mpsim_amd.circuits import bond_plan, brickwork_haar
    from mpsim_amd.engine import Engine
    d = 3
    circ = brickwork_haar(48, 16, d, maxsvals=1024, seed=4)
    eng = Engine.get()
    dc = prepare(circ, eng)
    first = []
    count = [0]
    worst = [0.0]

    def observe(dc_, i, it, r):
        a, b = it.a, it.b
        cl, _, cm = a.shape
        cr = b.shape[2]
        A = a.reshape(cl * d, cm).to(torch.complex128)
        B = b.reshape(cm, d * cr).to(torch.complex128).contiguous()
        G = it.gate.to(torch.complex128).reshape(d * d, d * d)
        # theta[l, (o1 o2), r] = sum_(p1 p2) G[(o1 o2), (p1 p2)] (A B)[l, (p1 p2), r]
        th = torch.matmul(G, (A @ B).reshape(cl, d * d, cr).contiguous())
        fro2 = float(th.abs().square().sum())
        s = r.s.to(torch.float64)
        disc = float(r.discarded)
        kept = float(s.square().sum())
        worst[0] = max(worst[0], abs(fro2 - kept - disc) / fro2)
        assert abs(fro2 - kept - disc) <= 1e-5 * fro2, (i, fro2, kept, disc)
        assert bool(torch.isfinite(r.a_out).all()) and bool(torch.isfinite(r.b_out).all())
        if not first and cl * d == 3072 and cr * d == 3072:
            th2 = th.reshape(cl * d, d * cr)
            k = r.a_out.shape[2]
            U = r.a_out.reshape(cl * d, k).to(torch.complex128)
            V = r.b_out.reshape(k, d * cr).to(torch.complex128)
            rec2 = float((th2 - U @ V).abs().square().sum())
            first.append((_host(it.a), _host(it.b), _host(it.gate), _host(s), disc, rec2, fro2))
        count[0] += 1

    print("qutrits: 48 qutrits chi 1024 on the GPU", flush=True)
    while not dc.done:
        tick(eng, [dc], want_s=True, observer=observe)
        print(f"qutrits: layer {dc.t} done", flush=True)
    torch.cuda.synchronize()
    assert count[0] == circ.two_site_count() == 376
    assert dc.mps.bond_dimensions() == bond_plan(circ)[-1]
    print(f"qutrits: worst |theta|^2 - kept - discarded: {worst[0]:.2e} relative", flush=True)
    assert first, "no 3072 x 3072 update"
    a, b, g, s, disc, rec2, fro2 = first[0]
    print("qutrits: LAPACK on the first 3072 x 3072 update", flush=True)
    th = _theta64(a, b, g, d)
    s_ref = np.linalg.svd(th, compute_uv=False)
    k = s.size
    assert k == 1024
    rel = np.abs(s - s_ref[:k]) / s_ref[:k]
    print(f"qutrits: kept-value error vs LAPACK {rel.max():.2e} (sigma_k/sigma_1 "
          f"{s_ref[k - 1] / s_ref[0]:.2e}), discarded {disc:.6e} vs {np.sum(s_ref[k:] ** 2):.6e}")
    assert np.all(rel <= 1e-4)
    assert abs(disc - float(np.sum(s_ref[k:] ** 2))) <= 1e-6 * fro2
    assert abs(rec2 - disc) <= 1e-5 * fro2


@pytest.mark.timeout(600)
def test_served_requests_full_size_alone_equals_mixed():
    """The first 8 requests of the seeded serving stream (Poisson 200 req/s;
    n in [16,80], depth in [4,32], maxsvals in {64,128,256,512}), served
    together with their real arrival times: the cap-512, cap-64 and two
    cap-128 requests give bitwise the same tensors when served alone; every
    result's bond dimensions follow the host plan; the cap-64 and cap-128
    ones match the complex128 oracle (overlap error <= 1e-4, or 3x the
    complex64 LAPACK error where the circuit itself amplifies rounding
    beyond that)."""
    from mpsim_amd.circuits import Request, bond_plan, poisson_requests
    from mpsim_amd.serve import MPSServer
    from oracle.mps_oracle import OracleMPS, gauge_invariant_error
    reqs = poisson_requests(200.0, 60.0, limit=8)
    assert [r.circuit.maxsvals for r in reqs] == [512, 64, 128, 256, 512, 128, 128, 64]
    srv = MPSServer(keep_results=True, hbm_budget=64 << 30)
    print("served: 8 requests in the mix", flush=True)
    mixed = srv.run(reqs, realtime=True).results
    assert sorted(mixed) == list(range(8))
    for r in reqs:
        ts = mixed[r.rid].tensors
        assert [t.shape[2] for t in ts[:-1]] == bond_plan(r.circuit)[-1]
    for rid in (0, 1, 2, 5):
        r = reqs[rid]
        print(f"served: request {rid} alone", flush=True)
        alone = srv.run([Request(rid, 0.0, r.circuit)], realtime=False).results[rid]
        for x, y in zip(alone.tensors, mixed[rid].tensors):
            assert tuple(x.shape) == tuple(y.shape) and np.array_equal(_host(x), _host(y)), rid
    for rid in (1, 2, 5, 7):
        c = reqs[rid].circuit
        ora = OracleMPS(c.n, 2, dtype_mode="c128")
        ora.apply_ops(c.ops, maxsvals=c.maxsvals)
        c64 = OracleMPS(c.n, 2, dtype_mode="c64")
        c64.apply_ops(c.ops, maxsvals=c.maxsvals)
        err, nr, nx = gauge_invariant_error(ora.tensors, [_host(t) for t in mixed[rid].tensors])
        err64, _, n64 = gauge_invariant_error(ora.tensors, c64.tensors)
        print(f"served: request {rid} (cap {c.maxsvals}, norm {nr:.3e}) overlap error vs FP64 "
              f"oracle {err:.2e} (complex64 LAPACK: {err64:.2e})")
        # Request 7 (51 qubits, depth 30, cap 64) truncates its state to a norm
        # of 2.7e-6 in the reference's non-canonical gauge, which amplifies
        # rounding: complex64 LAPACK itself ends 1.1e-2 from complex128 there.
        # The bound is the 1e-4 target, or 3x what complex64 LAPACK reaches on
        # the same circuit where that is worse.
        assert err <= max(1e-4, 3 * err64)
        assert abs(nr - nx) <= max(1e-4, 3 * abs(nr - n64) / nr) * nr
