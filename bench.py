"""Benchmark: two-site gate updates/s at chi=256 on one MI355X.

Workload (the single-large-circuit config): 64 qubits, brickwork depth 20 of
Haar-random 4x4 unitaries on nearest neighbours (offset alternating), seed 2,
maxsvals=256, complex64. A step is one full circuit from |0...0> (630
two-site updates); all disjoint pairs of a layer go to one grouped theta +
Jacobi SVD launch sequence. Gate matrices are resident in HBM before the
timed region. value = updates / second over the K timed steps.

Also reported: the plain one-gate-at-a-time GPU path on the same circuit,
the dominant kernel's roofline (HIP events on its launch stream), the
reference-faithful CPU restatement timed on this host on a bounded sample,
the served stream (`served`: continuous batching, p50/p99, vs the plain
path) and the uneven batch of 512 circuits (`uneven`: eager step ticks, one
captured hipGraph, vs the plain path).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-plain]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "2-site gate updates/s @χ=256; served circuits/s + p99 latency vs plain GPU"
PEAK_FP32_MFMA = 157.3  # TFLOP/s, MI355X dense FP32 matrix
PEAK_FP64_MFMA = 78.6   # TFLOP/s, MI355X dense FP64 matrix
PEAK_HBM = 8000.0       # GB/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--cpu-updates", type=int, default=6)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--inner", type=int, default=0)
    ap.add_argument("--tol", type=float, default=0.0,
                    help="Jacobi orthogonality tolerance (0 = library default)")
    ap.add_argument("--svd-method", type=int, default=0,
                    help="0: Gram method for FP64 32-row pivots (default); 1: one-sided everywhere")
    ap.add_argument("--split-ways", type=int, default=0,
                    help="concurrent stream groups of a large two-site call (0 = engine default)")
    ap.add_argument("--fp32", action="store_true",
                    help="FP32 Jacobi working matrix (faster, ~1e-6 relative error per update)")
    ap.add_argument("--fp32-sweeps", type=int, default=0,
                    help="Gram method: > 0 adds an FP32 phase of at most this many sweeps")
    ap.add_argument("--precondition", type=int, default=-1,
                    help="Householder QR before the Gram method: 0 none, 1 saturated bonds, 2 every "
                         "Gram problem (-1 = engine default)")
    ap.add_argument("--workload", default="large",
                    choices=["large", "uneven", "served", "qutrit", "copy", "hbm"],
                    help="large (default, the metric line) | uneven | served | qutrit | copy (counter "
                         "calibration) | hbm (memory-bound kernels); errors against FP64: "
                         "tests/accuracy_report.py")
    ap.add_argument("--circuits", type=int, default=512, help="uneven: circuits in the batch")
    ap.add_argument("--plain-sample", type=int, default=16,
                    help="uneven/served: circuits timed one at a time on the plain path")
    ap.add_argument("--requests", type=int, default=40,
                    help="served leg: first N requests of the 200 req/s stream")
    ap.add_argument("--no-served", action="store_true", help="skip the served leg of the metric line")
    ap.add_argument("--no-uneven", action="store_true",
                    help="skip the uneven-batch object of the metric line")
    ap.add_argument("--lanes", default="128,256",
                    help="served: bond caps bounding the serving lanes ('' = one lockstep lane)")
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--duration", type=float, default=60.0)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--depth", type=int, default=0, help="qutrit: brickwork depth (0 = 16)")
    return ap.parse_args()


def emit(line):
    print(json.dumps(line), flush=True)


def base_line(metric, value, unit, args, ms_per_step, config, dtype=None):
    return {"metric": metric, "value": round(value, 4), "unit": unit, "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype or ("c64+f32" if getattr(args, "fp32", False) else "c64+f64"),
            "data": "synthetic (seeded Haar gates)", "config": config}


def run_copy(args, eng):
    """Calibration for the HBM counters: k_scale with factor 1 reads and
    writes exactly 8 bytes per complex element (512 MiB each way)."""
    import torch
    x = torch.zeros(64 << 20, dtype=torch.complex64, device=eng.dev)
    for _ in range(3):
        eng.scale([x], 1.0)
    torch.cuda.synchronize()
    emit({"workload": "copy", "bytes_read_per_launch": x.numel() * 8,
          "bytes_written_per_launch": x.numel() * 8})


def run_hbm(args, eng, svd):
    """Achieved bandwidth (and theta flop rate) of the kernels around the SVD,
    timed with HIP events on their launch stream (mpsx_profile_*):
      k_one_site  grouped one-site gate over 64 tensors [1024, 2, 1024]
                  (1 GiB, every element read and written once per launch);
      k_gemm      the wavefunction chain of a random 26-qubit MPS (bond cap 64,
                  2^26 complex64 output): bytes of every GEMM's operands and result;
      k_theta     theta = G (A B) of 31 updates at chi = 256 (FP32 FMA): flops
                  8 d^2 chi^3 + 8 d^4 chi^2 per update, bytes A, B read + theta written;
      k_gather    S-absorb gather of the same updates: k rows of Y and of X read
                  (FP64) and the complex64 factors written."""
    import numpy as np
    import torch
    from mpsim_amd import _native as nat
    from mpsim_amd.core import MPS
    from mpsim_amd.engine import TwoSiteItem
    from mpsim_amd.gates import haar_unitary
    rng = np.random.RandomState(7)
    out = {}

    def timed(fn, name, reps=5):
        fn()
        torch.cuda.synchronize()
        nat.profile_read(reset=True)
        nat.profile_enable(True)
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        nat.profile_enable(False)
        k = nat.profile_read(reset=True)
        ms, nl = k[name]
        return ms, nl

    # one-site gate
    n, chi, d = 64, 1024, 2
    ts = [torch.randn(chi, d, chi, dtype=torch.complex64, device=eng.dev) for _ in range(n)]
    g = eng.gate(haar_unitary(2, rng))
    items = [(t, t, g) for t in ts]
    ms, nl = timed(lambda: eng.one_site(items), "k_one_site")
    by = sum(16 * t.numel() for t in ts) + n * d * d * 8
    out["k_one_site"] = {"bytes_per_launch": by, "avg_us": round(1e3 * ms / nl, 2),
                         "gb_per_s": round(by / (ms / nl * 1e-3) / 1e9, 1)}
    del ts, items
    torch.cuda.empty_cache()

    # wavefunction GEMM chain
    nq, cap = 26, 64
    mps = MPS(nq, 2, device=eng.dev, _empty=True)
    bonds = [1] + [min(2 ** (i + 1), 2 ** (nq - i - 1), cap) for i in range(nq - 1)] + [1]
    mps._tensors = [(torch.randn(bonds[i], 2, bonds[i + 1], dtype=torch.complex64, device=eng.dev)
                     / np.sqrt(2 * bonds[i])) for i in range(nq)]
    reps = 3
    fn = lambda: mps._wavefunction_device()  # noqa: E731
    fn()
    torch.cuda.synchronize()
    nat.profile_read(reset=True)
    nat.profile_enable(True)
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    nat.profile_enable(False)
    kk = nat.profile_read(reset=True)
    ms = sum(kk.get(k, (0.0, 0))[0] for k in ("k_gemm", "k_gemm_rows"))
    nl = sum(kk.get(k, (0.0, 0))[1] for k in ("k_gemm", "k_gemm_rows"))
    # two-sided roofline per product: max(bytes / HBM peak, flops / FP32 vector
    # peak without packed FP32, 78.6 TFLOP/s)
    by, fl, bound, rows = 0, 0, 0.0, 2
    for i in range(1, nq):
        cl, cr = bonds[i], bonds[i + 1]
        b_i = 8 * (rows * cl + cl * 2 * cr + rows * 2 * cr)
        f_i = 8 * rows * cl * 2 * cr
        by += b_i
        fl += f_i
        bound += max(b_i / (PEAK_HBM * 1e9), f_i / 78.6e12)
        rows *= 2
    t = ms / reps * 1e-3
    out["k_gemm_wavefunction"] = {
        "launches_per_call": nl // reps, "bytes_per_call": by, "flops_per_call": fl,
        "ms_per_call": round(ms / reps, 3), "gb_per_s": round(by / t / 1e9, 1),
        "tflops": round(fl / t / 1e12, 2),
        "roofline_ms": round(bound * 1e3, 3), "frac_of_roofline": round(bound / t, 3),
        "by_kernel_ms_per_call": {k: round(v[0] / reps, 3) for k, v in kk.items()},
        "roofline_basis": "sum over the chain's products of max(bytes / 8 TB/s, flops / 78.6 "
                          "TFLOP/s FP32 vector)"}
    del mps

    # theta and gather of 31 chi=256 updates, as one group (a split call would
    # run four launches of each kernel concurrently, a quarter of the problems each)
    chi, cnt = 256, 31
    gate = eng.gate(haar_unitary(4, rng))
    its = [TwoSiteItem(torch.randn(chi, 2, chi, dtype=torch.complex64, device=eng.dev) / 32.0,
                       torch.randn(chi, 2, chi, dtype=torch.complex64, device=eng.dev) / 32.0,
                       gate, 2 * chi) for _ in range(cnt)]
    ways, eng.split_ways = eng.split_ways, 1
    eng.two_site(its, svd)
    torch.cuda.synchronize()
    nat.profile_read(reset=True)
    nat.profile_enable(True)
    for _ in range(2):
        eng.two_site(its, svd)
    torch.cuda.synchronize()
    nat.profile_enable(False)
    eng.split_ways = ways
    k = nat.profile_read(reset=True)
    assert k["k_theta"][1] == 2 and k["k_gather"][1] == 2, k
    zb = 8 if svd.fp32 else 16
    M = N = 2 * chi
    kk = 2 * chi  # the items keep every singular value
    th_ms, th_n = k["k_theta"]
    th_f = cnt * (8.0 * 4 * chi ** 3 + 8.0 * 16 * chi ** 2)
    th_b = cnt * (2 * 8 * chi * 2 * chi + zb * M * N)
    out["k_theta"] = {"problems": cnt, "avg_us": round(1e3 * th_ms / th_n, 2),
                      "tflops": round(th_f / (th_ms / th_n * 1e-3) / 1e12, 2),
                      "gb_per_s": round(th_b / (th_ms / th_n * 1e-3) / 1e9, 1)}
    ga_ms, ga_n = k["k_gather"]
    ga_b = cnt * ((zb + 8) * kk * (M + N))
    out["k_gather"] = {"problems": cnt, "k": kk, "avg_us": round(1e3 * ga_ms / ga_n, 2),
                       "bytes_per_launch": ga_b,
                       "gb_per_s": round(ga_b / (ga_ms / ga_n * 1e-3) / 1e9, 1)}
    emit({"workload": "hbm", "peak_hbm_gb_per_s": 8000, "peak_fp32_tflops": PEAK_FP32_MFMA,
          "kernels": out})


def run_qutrit(args, eng, svd):
    """Precision stress: 48 qutrits, brickwork depth 16 of 9x9 Haar gates,
    maxsvals=1024 (theta up to 3072 x 3072); one pass, layer-batched."""
    import torch
    from mpsim_amd import _native as nat
    from mpsim_amd.batch import prepare, run_batched
    from mpsim_amd.circuits import brickwork_haar, circuit_work
    depth = args.depth or 16
    circ = brickwork_haar(48, depth, 3, maxsvals=1024, seed=4)
    flops, _, updates = circuit_work(circ)
    svd_f, _ = svd_flops(circ, svd.block_switch or 64, gram_min_rows(svd))
    kset = round_kernels(svd)
    peak = PEAK_FP32_MFMA if svd.fp32 else PEAK_FP64_MFMA
    dc = prepare(circ, eng)
    torch.cuda.synchronize()
    nat.profile_read(reset=True)
    nat.profile_enable(True)
    t0 = time.perf_counter()
    run_batched([dc], eng, svd)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    nat.profile_enable(False)
    busy = nat.profile_busy(busy_kernels(svd))
    kern = nat.profile_read(reset=True)
    jr_n = round_launches(kern, kset)
    ach = svd_f["big"] / (busy * 1e-3) / 1e12 if busy else 0.0
    line = base_line("2-site gate updates/s, 48 qutrits chi<=1024", updates / dt, "2-site updates/s",
                     args, 1e3 * dt, {"workload": f"48 qutrits, brickwork depth {depth}, 9x9 Haar, "
                                      "maxsvals=1024", "updates": updates})
    line["roofline"] = {"kernel": "+".join(busy_kernels(svd)), "bound": "mfma", "achieved": round(ach, 3),
                        "peak": peak, "unit": "TFLOP/s",
                        "frac": round(ach / peak, 5), "rounds": jr_n,
                        "frac_of_fp32_mfma_peak": round(ach / PEAK_FP32_MFMA, 5),
                        "achieved_basis": "SVD flops of the updates with 32-row pivots / busy time "
                                          "of their round kernels"}
    line["extra"] = {"kernel_ms": {k: round(v[0], 2) for k, v in sorted(kern.items())},
                     "bond_dimensions_max": max(dc.mps.bond_dimensions()),
                     "algorithmic_tflops_per_s": round(flops / dt / 1e12, 3)}
    if not args.no_cpu:
        # the reference-faithful restatement on the host: one update at the cap
        # shape (theta 3072 x 3072 -> 1024, complex128, LAPACK SVD, the per-gate
        # full norm of the 48-site MPS)
        from oracle.cpu_baseline import blas_threads, cpu_model, time_updates
        r = time_updates(updates=1, n=48, d=3, cap=1024, mode="reference", track_norms=True)
        line["cpu_baseline"] = {
            "value": round(r["updates_per_s"], 5), "unit": "2-site updates/s", "cores": blas_threads(),
            "kind": "port", "cpu_model": cpu_model(),
            "sample": "one two-site update at the cap shape (site tensors [1024, 3, 1024], theta "
                      "3072 x 3072 -> 1024) of the 48-qutrit chain, reference-faithful restatement "
                      "(complex128, np.linalg.svd, S absorbed right, per-gate full norm)",
            "seconds": round(r["seconds"], 2)}
    emit(line)


def uneven_leg(args, eng, svd, steps: int, warmup: int):
    """The uneven-batch configuration (512 seeded circuits of n in [12,64],
    depth in [4,24], bond caps in {32,64,128,256}): grouped step ticks over the
    whole batch (eager), the same batch as one captured CircuitProgram (first
    run eager, then hipGraph replays), and the plain path (one circuit, one
    gate per call) on a sample. Returns (eager seconds per batch, extra)."""
    import torch
    from mpsim_amd.batch import prepare, reset, run_batched, run_plain
    from mpsim_amd.circuits import uneven_batch
    circs = uneven_batch(args.circuits)
    dcs = [prepare(c, eng) for c in circs]
    updates = sum(c.two_site_count() for c in circs)
    for _ in range(warmup):
        for dc in dcs:
            reset(dc)
        run_batched(dcs, eng, svd)
    torch.cuda.synchronize()
    times = []
    for _ in range(steps):
        for dc in dcs:
            reset(dc)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run_batched(dcs, eng, svd)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    tb = sum(times) / len(times)
    del dcs
    from mpsim_amd.program import CircuitProgram
    prog = CircuitProgram(circs, eng, svd)
    prog.run()
    prog.run()
    torch.cuda.synchronize()
    tg_list = []
    for _ in range(max(steps, 3)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        prog.run()
        torch.cuda.synchronize()
        tg_list.append(time.perf_counter() - t0)
    tg = sum(tg_list) / len(tg_list)
    graph_info = {"graph_circuits_per_s": round(len(circs) / tg, 4), "graph_replays": prog.replays,
                  "graph_eager_runs": prog.eager_runs, "graph_unconverged": prog.last_unconverged,
                  "graph_arena_gb": round(prog.arena_bytes / 2 ** 30, 3),
                  "graph_workspace_gb": round(prog.workspace.numel() / 2 ** 30, 3)}
    prog.close()
    del prog
    sample = circs[:args.plain_sample]
    run_plain(sample[0], device=eng.dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for c in sample:
        run_plain(c, device=eng.dev)
    torch.cuda.synchronize()
    tp = (time.perf_counter() - t0) / len(sample)
    value = len(circs) / tb
    extra = {"workload": f"{len(circs)} seeded circuits, n in [12,64], depth in [4,24], "
                         "chi cap in {32,64,128,256}",
             "circuits": len(circs), "two_site_updates": updates,
             "eager_circuits_per_s": round(value, 4),
             "plain_gpu_circuits_per_s": round(1.0 / tp, 4),
             "plain_sample": f"first {len(sample)} circuits, one at a time",
             "batched_vs_plain": round(value * tp, 3),
             "updates_per_s": round(updates / tb, 2), **graph_info,
             "graph_vs_plain": round(len(circs) / tg * tp, 3)}
    return tb, extra


def run_uneven(args, eng, svd):
    """Batched uneven circuits on their own line: value = eager circuits/s."""
    tb, extra = uneven_leg(args, eng, svd, args.steps, args.warmup)
    line = base_line("batched uneven circuits/s vs one-at-a-time GPU path", extra["eager_circuits_per_s"],
                     "circuits/s", args, 1e3 * tb,
                     {"workload": extra["workload"], "circuits": extra["circuits"],
                      "two_site_updates": extra["two_site_updates"]})
    line["extra"] = {k: v for k, v in extra.items()
                     if k not in ("workload", "circuits", "two_site_updates", "eager_circuits_per_s")}
    emit(line)


def _served_run(srv, reqs, eng):
    """One served run of `reqs` at their arrival times (continuous batching),
    then the plain GPU path (one circuit at a time, one gate per call through
    the public MPS API) timed on one request per bond-cap class, the one of
    median work among `reqs` of that class; each request is priced at its
    class's plain flop rate, which gives the plain path's time for the same
    requests. The sampled requests' served results are compared bitwise
    with their plain results."""
    import torch
    from mpsim_amd.batch import run_plain
    from mpsim_amd.circuits import circuit_work
    rep = srv.run(reqs, realtime=True)
    work = {r.rid: circuit_work(r.circuit)[0] for r in reqs}
    by_rid = {r.rid: r for r in reqs}
    plain, same = {}, True
    for cap in sorted({r.circuit.maxsvals for r in reqs}):
        mine = sorted((work[r.rid], r.rid) for r in reqs if r.circuit.maxsvals == cap)
        w, rid = mine[len(mine) // 2]
        circ = by_rid[rid].circuit
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        mps = run_plain(circ, device=eng.dev)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        plain[cap] = {"rid": rid, "n": circ.n, "depth": circ.depth, "gflop": round(w / 1e9, 2),
                      "seconds": round(dt, 3), "tflops": round(w / dt / 1e12, 4)}
        same = same and all(tuple(a.shape) == tuple(b.shape) and bool(torch.equal(a, b))
                            for a, b in zip(mps.tensors, rep.results[rid].tensors))
    t_plain = sum(work[r.rid] / (plain[r.circuit.maxsvals]["tflops"] * 1e12) for r in reqs)
    plain_cps = len(reqs) / t_plain
    return {
        "value": round(rep.circuits_per_s, 4), "unit": "circuits/s", "requests": len(reqs),
        "arrivals_within_s": round(max(r.arrival for r in reqs), 4),
        "wall_s": round(rep.wall_s, 3),
        "p50_latency_s": round(rep.p50_s, 4), "p99_latency_s": round(rep.p99_s, 4),
        "served_updates_per_s": round(rep.updates_per_s, 2),
        "served_tflops": round(sum(work.values()) / rep.wall_s / 1e12, 3),
        "plain_gpu_circuits_per_s": round(plain_cps, 5),
        "served_vs_plain": round(rep.circuits_per_s / plain_cps, 3),
        "plain_samples": plain,
        "served_equals_plain_bitwise": bool(same),
        "lanes": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in ln.items()}
                  for ln in rep.lanes],
        "max_active": rep.max_active, "ticks": rep.ticks,
    }


def serve_leg(args, eng, svd):
    """The served half of the metric, on the seeded serving stream (Poisson
    200 req/s over 60 s; n in [16,80], depth in [4,32], maxsvals in
    {64,128,256,512}), with requests offered at their real arrival times,
    i.e. far above capacity, and run to completion with continuous batching
    (MPSServer: size-class lanes, footprint admission; capacity = completed
    / wall time):
      * the first `--requests` requests of the stream (every bond cap);
      * the first `--requests` requests with maxsvals <= 256 (the chi <= 256
        served form the speed-up target is stated for).
    Each run's plain-path comparison is described in _served_run."""
    import torch
    from mpsim_amd.circuits import poisson_requests
    from mpsim_amd.serve import MPSServer
    lanes = tuple(int(x) for x in args.lanes.split(",") if x.strip())
    srv = MPSServer(eng, hbm_budget=int(args.budget_gb * (1 << 30)), svd=svd, keep_results=True,
                    lanes=lanes)
    stream = poisson_requests(args.rate, args.duration, limit=8 * args.requests)
    warm = [r for r in stream if r.circuit.depth <= 6][:6]
    srv.run(warm, realtime=False)
    torch.cuda.synchronize()
    out = {"metric": "served circuits/s (continuous batching, offered above capacity) with p50/p99 "
                     "latency",
           "workload": (f"seeded Poisson {args.rate:g} req/s stream, n in [16,80], depth in [4,32], "
                        "maxsvals in {64,128,256,512}, complex64; requests offered at their arrival "
                        "times and run to completion"),
           "offered_req_per_s": args.rate, "hbm_budget_gb": args.budget_gb,
           "plain_basis": ("plain path timed on one request per bond-cap class (median work); each "
                           "served request priced at its class's plain flop rate")}
    out["all_caps"] = _served_run(srv, stream[:args.requests], eng)
    small = [r for r in stream if r.circuit.maxsvals <= 256][:args.requests]
    out["chi_le_256"] = _served_run(srv, small, eng)
    out["value"] = out["all_caps"]["value"]
    out["unit"] = "circuits/s"
    return out


def run_served(args, eng, svd):
    """The served leg on its own line."""
    leg = serve_leg(args, eng, svd)
    a = leg["all_caps"]
    line = base_line(leg["metric"], leg["value"], "circuits/s", args,
                     1e3 * a["wall_s"] / max(a["requests"], 1),
                     {"workload": leg["workload"], "requests": a["requests"],
                      "hbm_budget_gb": args.budget_gb})
    line["extra"] = {k: v for k, v in leg.items() if k not in ("metric", "value", "unit", "workload")}
    emit(line)


def gram_min_rows(svd):
    """Rows above which the updates with 32-row pivots run the Gram method."""
    return 64 if not svd.fp32 and svd.method == 0 else 0


def round_kernels(svd):
    """Kernels of one Jacobi round of the updates with 32-row pivots: the
    eigensolve and update launches of the Gram method (the default; with
    fp32_sweeps > 0 also those of its FP32 phase), else the one-sided round
    kernel."""
    if not svd.fp32 and svd.method == 0:
        ks = ["k_hsolve", "k_hround", "k_hupdate"]
        return ks + [k + "32" for k in ks] if svd.fp32_sweeps > 0 else ks
    return [("k_jacobi_round32" if svd.fp32 else "k_jacobi_round64") + "<32>"]


def busy_kernels(svd):
    """Kernels whose launch intervals make the SVD's busy time: the round
    kernels, and with preconditioning the Householder QR that precedes the
    Gram rounds (it does part of the work that the SVD flops count)."""
    ks = round_kernels(svd)
    if svd.precondition and not svd.fp32 and svd.method == 0 and svd.fp32_sweeps <= 0:
        ks = ks + ["k_qr_order", "k_qr_permute", "k_qr_step"]
    return ks


def round_launches(kern, kset):
    """Rounds run: one eigensolve launch (k_hsolve* or k_hround*) per round
    of the Gram method, one launch per round of the one-sided method."""
    eig = [k for k in kset if k.startswith("k_hsolve") or k.startswith("k_hround")] or kset
    return sum(kern.get(k, (0.0, 0))[1] for k in eig)


def svd_flops(circuit, switch=64, min_rows=0):
    """Golub-Reinsch SVD flops (the SVD term of the per-update work) of the
    updates with 32-row pivots (the rows to orthogonalise, d*chi_l, exceed
    `switch`, and `min_rows`: 64 for the Gram method, which serves six or
    more 16-row blocks) and of the others, summed over the circuit, and the
    full per-update flops. Returns ({"big": flops, "small": flops}, full)."""
    from mpsim_amd.circuits import two_site_flops
    n, d = circuit.n, circuit.d
    bonds = [1] * (n + 1)
    svd = {"big": 0.0, "small": 0.0}
    full = 0.0
    for layer in circuit.layers:
        for q, _ in layer:
            if len(q) != 2:
                continue
            i = min(q)
            cl, cm, cr = bonds[i], bonds[i + 1], bonds[i + 2]
            k = min(d * cl, d * cr, circuit.maxsvals or 10 ** 9)
            f = two_site_flops(cl, cm, cr, d, k)
            M, N = d * cl, d * cr
            p, qq = max(M, N), min(M, N)
            svd["big" if M > max(switch, min_rows) else "small"] += 4.0 * (4.0 * p * p * qq + 8.0 * p * qq * qq + 9.0 * qq ** 3)
            full += f
            bonds[i + 1] = k
    return svd, full


def load_traffic():
    path = os.path.join(ROOT, "profiles", "traffic_jacobi_round.json")
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); this build has no "
              "multi-GPU path.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from mpsim_amd import _native as nat
    from mpsim_amd.batch import prepare, reset, run_batched, run_plain
    from mpsim_amd.circuits import brickwork_haar, circuit_work
    from mpsim_amd.engine import Engine, SvdConfig

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    eng = Engine.get("cuda:0")
    if args.split_ways:
        eng.split_ways = args.split_ways
    import dataclasses
    svd = dataclasses.replace(eng.svd, block=args.block or eng.svd.block,
                              inner_sweeps=args.inner or eng.svd.inner_sweeps, fp32=args.fp32,
                              method=args.svd_method, tol=args.tol or eng.svd.tol,
                              fp32_sweeps=args.fp32_sweeps,
                              precondition=(args.precondition if args.precondition >= 0
                                            else eng.svd.precondition))
    peak = PEAK_FP32_MFMA if args.fp32 else PEAK_FP64_MFMA
    if args.workload == "copy":
        return run_copy(args, eng)
    if args.workload == "hbm":
        return run_hbm(args, eng, svd)
    if args.workload == "uneven":
        return run_uneven(args, eng, svd)
    if args.workload == "served":
        return run_served(args, eng, svd)
    if args.workload == "qutrit":
        return run_qutrit(args, eng, svd)
    circ = brickwork_haar(64, 20, 2, maxsvals=256, seed=2)
    flops, bytes_, updates = circuit_work(circ)
    svd_f, _ = svd_flops(circ, svd.block_switch or 64, gram_min_rows(svd))
    kset = round_kernels(svd)

    dc = prepare(circ, eng)
    for _ in range(args.warmup):
        reset(dc)
        run_batched([dc], eng, svd)
    torch.cuda.synchronize()

    def timed_steps():
        ts = []
        for _ in range(args.steps):
            reset(dc)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            run_batched([dc], eng, svd)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return ts

    # value: the timed steps without instrumentation. Kernel timing brackets
    # every launch with two HIP events, which costs throughput (the stream
    # stalls on each event), so the roofline comes from a second pass of the
    # same steps with the library's event profiler on; both rates are
    # reported.
    times = timed_steps()
    nat.profile_read(reset=True)
    nat.profile_enable(True)
    ptimes = timed_steps()
    nat.profile_enable(False)
    jr_busy = nat.profile_busy(busy_kernels(svd))
    full = nat.profile_read_full(reset=True)
    kern = {k: (v["total_ms"], v["launches"]) for k, v in full.items()}
    total = sum(times)
    value = updates * args.steps / total
    ms_per_step = 1e3 * total / args.steps

    # The roofline unit is one Jacobi round of the updates with 32-row pivots
    # (k_hsolve / k_hround + k_hupdate for the Gram method, one
    # k_jacobi_round launch otherwise). Large two-site calls run as four concurrent groups on four
    # streams, so rounds overlap: throughput is the SVD flops over the busy
    # time of the round kernels (the union of all their launch intervals);
    # the per-round figure (flops per round / average round duration) is
    # reported beside it.
    dom = "+".join(kset)
    jr_ms = sum(kern.get(k, (0.0, 0))[0] for k in kset)
    # rounds: one eigensolve launch each (k_hsolve opens a sweep, k_hround runs
    # the next round's eigensolves beside the current round's update)
    jr_n = round_launches(kern, kset)
    dom_f = svd_f["big"] * args.steps
    achieved = dom_f / (jr_busy * 1e-3) / 1e12 if jr_busy > 0 else 0.0
    per_round = dom_f / (jr_ms * 1e-3) / 1e12 if jr_ms > 0 else 0.0
    traffic = load_traffic()
    tbytes = (traffic.get("bytes_per_round")
              if traffic and traffic.get("kernel") == "+".join(kset) else None)
    # executed FP64 MFMA flops per round (SQ_INSTS_VALU_MFMA_MOPS_F64 x 512 from a
    # counter pass) over the live busy time
    mc = (traffic or {}).get("mfma_counters")
    roofline = {
        "kernel": "+".join(busy_kernels(svd)),
        "bound": "mfma",
        "achieved": round(achieved, 3),
        "peak": peak,
        "unit": "TFLOP/s",
        "frac": round(achieved / peak, 5),
        # the same rate against the dense FP32 matrix peak (157.3 TFLOP/s), the
        # denominator of the FP32-MFMA formulation of this SVD
        "frac_of_fp32_mfma_peak": round(achieved / PEAK_FP32_MFMA, 5),
        "traffic": tbytes,
        "traffic_unit": "HBM bytes per round (calibrated FETCH_SIZE + WRITE_SIZE)",
        "achieved_basis": "algorithmic SVD flops of the updates with 32-row pivots / busy time of "
                          "their round kernels (union of launch intervals, HIP events on the launch "
                          "streams), from a profiled repeat of the timed steps",
        "profiled_pass_updates_per_s": round(updates * args.steps / sum(ptimes), 3),
        "busy_ms": round(jr_busy, 3),
        "concurrency": round(jr_ms / jr_busy, 3) if jr_busy > 0 else None,
        "achieved_per_round": round(per_round, 3),
        "rounds": jr_n,
        "avg_round_us": round(1e3 * jr_ms / max(jr_n, 1), 3),
        "avg_launch_us": {k: round(1e3 * kern[k][0] / max(kern[k][1], 1), 3)
                          for k in busy_kernels(svd) if k in kern},
        "executed_mfma_frac_counters": (
            round(mc["executed_flop_per_round"] * jr_n / (jr_busy * 1e-3) / 1e12 / peak, 4)
            if (mc and not svd.fp32 and jr_busy > 0 and mc.get("kernel") == dom) else None),
        "algorithmic_flops_per_round": dom_f / max(jr_n, 1),
        "algorithmic_unit": "SVD term 4(4p^2q+8pq^2+9q^3) of the updates with 32-row pivots",
    }

    extra = {
        "kernel_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in sorted(kern.items())},
        "round_kernels": dom,
        "sweeps_last_layer": eng.last_sweeps,
        "algorithmic_tflops_per_s": round(flops * args.steps / total / 1e12, 3),
        "algorithmic_hbm_gb_per_s": round(bytes_ * args.steps / total / 1e9, 3),
    }

    if not args.no_plain:
        run_plain(circ, device=eng.dev)  # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run_plain(circ, device=eng.dev)
        torch.cuda.synchronize()
        tp = time.perf_counter() - t0
        extra["plain_gpu_updates_per_s"] = round(updates / tp, 2)
        extra["batched_vs_plain"] = round((updates / tp) and value / (updates / tp), 3)

    served = None
    if not args.no_served:
        served = serve_leg(args, eng, svd)

    uneven = None
    if not args.no_uneven:
        # the uneven-batch configuration beside the metric: eager ticks, one
        # captured program, the plain path on a sample (2 eager passes, 3 replays)
        _, uneven = uneven_leg(args, eng, svd, steps=2, warmup=1)

    cpu = None
    if not args.no_cpu:
        from oracle.cpu_baseline import blas_threads, cpu_model, time_plain16, time_updates
        faithful = time_updates(updates=args.cpu_updates, mode="reference", track_norms=True)
        lean = time_updates(updates=args.cpu_updates, mode="c64", track_norms=False)
        one = time_updates(updates=2, mode="reference", track_norms=True, threads=1)
        c1 = time_plain16()
        cpu = {
            "value": round(faithful["updates_per_s"], 4),
            "unit": "2-site updates/s",
            "cores": blas_threads(),
            "kind": "port",
            "sample": (f"{args.cpu_updates} two-site updates on disjoint interior pairs of the "
                       "64-qubit chain at the chi=256 cap (theta 512x512, maxsvals=256), "
                       "reference-faithful restatement: complex128 after Haar promotion, "
                       "np.linalg.svd, S absorbed right, per-gate full norm"),
            "lean_c64_no_norm_updates_per_s": round(lean["updates_per_s"], 4),
            "one_thread_updates_per_s": round(one["updates_per_s"], 4),
            "plain16_full_circuit_s": round(c1["seconds"], 3),
            "plain16_full_circuit": ("the 16-qubit depth-8 untruncated circuit in full, reference-"
                                     "faithful (complex64 tensors promoted by the float64 CNOT, "
                                     "per-gate norm)"),
            "cpu_model": cpu_model(),
        }

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "2-site updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64+f32" if args.fp32 else "c64+f64",
        "data": "synthetic (seeded Haar gates)",
        "config": {"workload": "64-qubit brickwork depth 20, Haar 4x4 gates, maxsvals=256, "
                               "layer-batched",
                   "arithmetic": ("complex64 site tensors and theta; Jacobi working matrix, "
                                  "Gram and rotations in " + ("FP32" if args.fp32 else "FP64")), "qubits": 64, "depth": 20, "maxsvals": 256,
                   "updates_per_step": updates, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "served": served,
        "uneven": uneven,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
