"""Error against FP64 for every workload, as one JSON line (a report, not a
pytest module: run it on a GPU as `python tests/accuracy_report.py`).

The oracle is the checker: the 16-qubit untruncated circuit against the FP64
state vector (max-abs); the 64-qubit chi=256 circuit, a sample of the uneven
circuits, sampled served requests and a scaled qutrit circuit against the
FP64 restatement (gauge-invariant overlap error 1 - |<ref|psi>|/(|ref||psi|)
and relative norm error). The GPU path is the batched one (layer/step ticks
through the HIP library). The per-config bounds are asserted by the -m gpu
tests (test_gpu_configs.py, test_gpu_accuracy.py); this prints the numbers.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--depth", type=int, default=20, help="depth of the 64-qubit circuit")
    ap.add_argument("--sample", type=int, default=16, help="uneven circuits checked")
    args = ap.parse_args()
    import numpy as np
    import torch

    from mpsim_amd import _native
    from mpsim_amd.batch import prepare, run_batched
    from mpsim_amd.circuits import (brickwork_cnot_u2, brickwork_haar, poisson_requests,
                                    uneven_batch)
    from mpsim_amd.engine import Engine
    from oracle import statevector as sv
    from oracle.mps_oracle import OracleMPS, gauge_invariant_error

    _native.load()  # fail loudly without the HIP library
    eng = Engine.get()

    def gpu(circ):
        dc = prepare(circ, eng)
        run_batched([dc], eng)
        torch.cuda.synchronize()
        return [t.cpu().numpy() for t in dc.mps.tensors], dc

    def vs_oracle(circ):
        t0 = time.perf_counter()
        mine, _ = gpu(circ)
        ora = OracleMPS(circ.n, circ.d, dtype_mode="c128")
        ora.apply_ops(circ.ops, maxsvals=circ.maxsvals)
        err, nr, nx = gauge_invariant_error(ora.tensors, mine)
        return {"overlap_error": float(err), "norm_rel_error": abs(nr - nx) / nr,
                "seconds": round(time.perf_counter() - t0, 2)}

    out = {}
    c1 = brickwork_cnot_u2(16, 8, seed=1)
    _, dc1 = gpu(c1)
    psi = dc1.mps.wavefunction()
    out["plain16_maxabs_vs_fp64_statevector"] = float(np.max(np.abs(psi - sv.run(c1.ops, 16))))
    out["large64_chi256"] = vs_oracle(brickwork_haar(64, args.depth, 2, maxsvals=256, seed=2))
    out["uneven_sample"] = [dict(vs_oracle(c), name=c.name, n=c.n, cap=c.maxsvals)
                            for c in uneven_batch(args.sample)[:args.sample]]
    out["served_sample"] = [dict(vs_oracle(r.circuit), rid=r.rid, n=r.circuit.n,
                                 cap=r.circuit.maxsvals)
                            for r in [q for q in poisson_requests(limit=40)
                                      if q.circuit.maxsvals <= 128][:3]]
    out["qutrit_scaled_n12_d10_chi81"] = vs_oracle(brickwork_haar(12, 10, 3, maxsvals=81, seed=4))
    print(json.dumps({"report": "accuracy", "results": out}), flush=True)


if __name__ == "__main__":
    main()
