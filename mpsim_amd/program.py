"""Static, replayable schedules of many circuits, captured as one hipGraph.

Because the reference keeps zero singular values, every bond dimension of a
unitary circuit follows from the circuit alone (k = min(d*chi_l, d*chi_r,
maxsvals) per update, core.py:1132-1137), so the whole batched schedule of
a set of circuits -- every launch, every tensor shape, every pointer -- is
known before anything runs. A CircuitProgram fixes it once:

  * every site of every circuit gets one buffer sized for the largest shape
    it will ever hold, all in one arena; two-site updates write their
    factors back into the buffers of the two sites they read (the library
    reads the inputs into its workspace before it writes any output);
  * the descriptor tables of every tick (one grouped one-site call and one
    or more grouped two-site calls per step) are built on the host and
    copied to the device once;
  * the first run is eager with the usual per-sweep convergence poll, and
    records how many Jacobi sweeps each two-site call needed;
  * the schedule is then captured into one hipGraph (mpsx_graph_*), each
    two-site call with check_every = 0 and its recorded sweep count plus a
    margin. Later runs reset the sites to |0...0> and replay the graph: no
    host work per launch and no host synchronisation until the end.

A replay is exact when every problem reached its convergence sweep within
the budget: a converged problem is skipped by the remaining sweeps, so its
factors are bitwise those of the eager run. The library counts problems
that did not converge on the device; a replay with a non-zero count is run
again eagerly and the graph is recaptured with the larger sweep counts.
The gate matrices stay resident; new gate values of the same circuit
structure can be written into them between runs (`update_gates`).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _native as nat
from .batch import _ws_estimate, op_matrix, upload_gates
from .circuits import Circuit, bond_plan_steps, schedule, swap_matrix
from .core import MPS
from .engine import MAX_OPS, Engine, SvdConfig, TwoSiteItem

_ALIGN = 32  # complex64 elements (256 bytes) between site buffers


@dataclass
class _Call:
    kind: str            # "one" or "two"
    host: np.ndarray     # descriptor table (kept alive: read at launch/capture time)
    dev_off: int         # byte offset of its device copy in the table buffer
    sweeps: int = 0      # two-site: sweeps the eager run needed


class CircuitProgram:
    """A batch of circuits compiled to one static schedule (one grouped call
    set per step of circuits.schedule, SWAP networks included).

    program = CircuitProgram(circuits)
    program.run()            # first run eager (records sweeps), then graph replays
    states = program.states()
    """

    def __init__(self, circuits: Sequence[Circuit], engine: Optional[Engine] = None,
                 svd: Optional[SvdConfig] = None, graph: bool = True, sweep_margin: int = 2,
                 ws_budget: int = 8 << 30) -> None:
        self.engine = eng = engine or Engine.get()
        torch = eng.torch
        self.circuits = list(circuits)
        self.svd = svd or eng.svd
        self.use_graph = graph
        self.margin = int(sweep_margin)
        self._graph = None
        self.replays = 0
        self.eager_runs = 0
        self.last_unconverged = 0

        # -- site buffers (sized for the largest shape over every step) --------
        self._steps = [schedule(c)[0] for c in self.circuits]
        self._plans = [bond_plan_steps(c) for c in self.circuits]
        self._site_off: List[List[int]] = []
        total = 0
        for c, plan in zip(self.circuits, self._plans):
            n, d = c.n, c.d
            biggest = [d] * n
            for bonds in plan:
                full = [1] + bonds + [1]
                for s in range(n):
                    biggest[s] = max(biggest[s], full[s] * d * full[s + 1])
            offs = []
            for s in range(n):
                offs.append(total)
                total += -(-biggest[s] // _ALIGN) * _ALIGN
            self._site_off.append(offs)
        self.arena = torch.empty(max(total, 1), dtype=torch.complex64, device=eng.dev)
        idx, val = [], []
        for c, offs in zip(self.circuits, self._site_off):
            for o in offs:
                idx += range(o, o + c.d)
                val += [1.0] + [0.0] * (c.d - 1)
        self._reset_idx = torch.tensor(idx, dtype=torch.int64, device=eng.dev)
        self._reset_val = torch.tensor(val, dtype=torch.complex64, device=eng.dev)
        self.arena_bytes = 8 * total

        # -- gates -------------------------------------------------------------
        self.gates = [upload_gates(eng, c) for c in self.circuits]
        self._swaps = {c.d: eng.gate(swap_matrix(c.d)) for c in self.circuits}

        # -- schedule (one grouped call set per step, circuits.schedule) ---------
        depth = max((len(st) for st in self._steps), default=0)
        ntwo = sum(sum(len(x.twos) for x in st) for st in self._steps)
        self.discarded = torch.zeros(max(ntwo, 1), dtype=torch.float64, device=eng.dev)
        self.unconverged = torch.zeros(1, dtype=torch.int32, device=eng.dev)
        self.two_site_updates = ntwo
        calls: List[_Call] = []
        ws_need, slot = 0, 0
        disc_ptr = int(self.discarded.data_ptr())
        for t in range(depth):
            one, two = [], []
            for ci, c in enumerate(self.circuits):
                if t >= len(self._steps[ci]):
                    continue
                d = c.d
                before = [1] + (self._plans[ci][t - 1] if t else [1] * (c.n - 1)) + [1]
                after = [1] + self._plans[ci][t] + [1]
                offs = self._site_off[ci]
                step = self._steps[ci][t]
                for site, lt, o in step.ones:
                    p = self._addr(offs[site])
                    one.append((p, p, nat.ptr(self.gates[ci][lt][o]), before[site], before[site + 1],
                                d, 0))
                for i, lt, o, _ in step.twos:
                    g = self.gates[ci][lt][o] if o >= 0 else self._swaps[d]
                    cl, cm, cr, k = before[i], before[i + 1], before[i + 2], after[i + 1]
                    two.append(dict(a=self._addr(offs[i]), b=self._addr(offs[i + 1]),
                                    gate=nat.ptr(g), cl=cl, cm=cm, cr=cr, d=d, k=k,
                                    disc=disc_ptr + 8 * slot))
                    slot += 1
            for s in range(0, len(one), MAX_OPS):
                calls.append(_Call("one", np.array(one[s:s + MAX_OPS], dtype=nat.SITE_OP), 0))
            start = 0
            while start < len(two):
                end, acc = start, 0
                while end < len(two) and (end == start or acc + self._ws_bytes(two[end]) <= ws_budget):
                    acc += self._ws_bytes(two[end])
                    end += 1
                tab = self._two_table(two[start:end])
                ws_need = max(ws_need, self._plan(tab))
                calls.append(_Call("two", tab, 0))
                start = end
        self.calls = calls
        self.workspace = torch.empty(max(ws_need, 1024), dtype=torch.uint8, device=eng.dev)
        # one device buffer holding every table (256-byte aligned)
        off = 0
        for cl in calls:
            cl.dev_off = off
            off += -(-cl.host.nbytes // 256) * 256
        blob = np.zeros(max(off, 256), dtype=np.uint8)
        for cl in calls:
            blob[cl.dev_off:cl.dev_off + cl.host.nbytes] = cl.host.view(np.uint8)
        self._tables = nat.to_device_table(blob, eng.dev)

    # ------------------------------------------------------------ building
    def _addr(self, off: int) -> int:
        return int(self.arena.data_ptr()) + 8 * off

    @staticmethod
    def _ws_bytes(p: dict) -> int:
        class _S:  # shape stand-in for the estimate
            def __init__(self, shape):
                self.shape = shape
        return _ws_estimate(TwoSiteItem(_S((p["cl"], p["d"], p["cm"])),
                                        _S((p["cm"], p["d"], p["cr"])), None, p["k"]))

    @staticmethod
    def _two_table(items: Sequence[dict]) -> np.ndarray:
        tab = np.zeros(len(items), dtype=nat.TWO_SITE_OP)
        for i, p in enumerate(items):
            tab[i]["a"], tab[i]["b"], tab[i]["gate"] = p["a"], p["b"], p["gate"]
            tab[i]["a_out"], tab[i]["b_out"] = p["a"], p["b"]  # in place
            tab[i]["discarded"] = p["disc"]
            tab[i]["chi_l"], tab[i]["chi_m"], tab[i]["chi_r"] = p["cl"], p["cm"], p["cr"]
            tab[i]["d"], tab[i]["k"] = p["d"], p["k"]
        return tab

    def _opts(self, sweeps: int) -> nat.SvdOpts:
        c = self.svd
        # graph form (sweeps > 0): fixed sweep count, device-side status
        return SvdConfig(block=c.block, max_sweeps=sweeps if sweeps else c.max_sweeps,
                         check_every=0 if sweeps else 1, inner_sweeps=c.inner_sweeps, tol=c.tol,
                         block_switch=c.block_switch, unconverged=int(self.unconverged.data_ptr()),
                         fp32=c.fp32, method=c.method, fp32_sweeps=c.fp32_sweeps,
                         precondition=c.precondition).to_c()

    def _plan(self, tab: np.ndarray) -> int:
        nbytes = ctypes.c_size_t(0)
        opts = self._opts(0)
        nat.check(self.engine.lib.mpsx_two_site_plan(tab.ctypes.data, len(tab), ctypes.byref(opts),
                                                     ctypes.byref(nbytes)))
        return int(nbytes.value)

    # ------------------------------------------------------------ running
    def _reset(self) -> None:
        self.arena.index_copy_(0, self._reset_idx, self._reset_val)
        self.unconverged.zero_()

    def _launch_all(self, stream: int, record: bool) -> None:
        lib = self.engine.lib
        base = int(self._tables.data_ptr())
        ws, wsn = int(self.workspace.data_ptr()), int(self.workspace.numel())
        for cl in self.calls:
            dptr = base + cl.dev_off
            if cl.kind == "one":
                nat.check(lib.mpsx_apply_one_site(stream, cl.host.ctypes.data, dptr, len(cl.host)))
                continue
            opts = self._opts(0 if record else cl.sweeps + self.margin)
            sweeps = ctypes.c_int(0)
            code = lib.mpsx_two_site(stream, cl.host.ctypes.data, dptr, len(cl.host),
                                     ctypes.byref(opts), ws, wsn, ctypes.byref(sweeps))
            if code == nat.MPSX_ENOTCONV:
                import warnings
                warnings.warn(lib.mpsx_last_error().decode(), nat.NotConvergedWarning)
            else:
                nat.check(code)
            if record:
                cl.sweeps = max(cl.sweeps, int(sweeps.value))

    def _capture(self) -> None:
        torch = self.engine.torch
        self._release_graph()
        torch.cuda.synchronize(self.engine.dev)
        side = torch.cuda.Stream(self.engine.dev)
        h = int(side.cuda_stream)
        lib = self.engine.lib
        nat.check(lib.mpsx_graph_begin(h))
        g = ctypes.c_void_p(0)
        try:
            self._launch_all(h, record=False)
        finally:
            code = lib.mpsx_graph_end(h, ctypes.byref(g))
        nat.check(code)
        self._graph = g
        self._side = side

    def _release_graph(self) -> None:
        if self._graph is not None and self._graph.value:
            nat.check(self.engine.lib.mpsx_graph_destroy(self._graph))
        self._graph = None

    def run(self) -> None:
        """Runs every circuit from |0...0> to its end and waits for it."""
        torch = self.engine.torch
        lib = self.engine.lib
        stream = self.engine.stream
        self._reset()
        if self._graph is None:
            self._launch_all(stream, record=True)
            self.eager_runs += 1
            torch.cuda.synchronize(self.engine.dev)
            self.last_unconverged = int(self.unconverged.item())
            if self.use_graph:
                self._capture()
            return
        nat.check(lib.mpsx_graph_launch(self._graph, stream))
        self.replays += 1
        self.last_unconverged = int(self.unconverged.item())
        if self.last_unconverged:
            # some problem needed more sweeps than recorded: redo eagerly
            self._release_graph()
            self.run()

    def update_gates(self, circuits: Sequence[Circuit]) -> None:
        """New gate values for circuits of the same structure (layers, qudit
        pairs, maxsvals): written into the resident gate buffers, which the
        captured graph reads at replay time."""
        if len(circuits) != len(self.circuits):
            raise ValueError("circuit count differs")
        for ci, c in enumerate(circuits):
            old = self.circuits[ci]
            if (c.n, c.d, c.maxsvals, [[q for q, _ in L] for L in c.layers]) != \
                    (old.n, old.d, old.maxsvals, [[q for q, _ in L] for L in old.layers]):
                raise ValueError(f"circuit {ci} has a different structure")
            for t, layer in enumerate(c.layers):
                for (q, m), g in zip(layer, self.gates[ci][t]):
                    g.copy_(self.engine.torch.from_numpy(op_matrix(q, m, c.d)).view(g.shape))
        self.circuits = list(circuits)

    # ------------------------------------------------------------ results
    def states(self) -> List[MPS]:
        """MPS views of the final states (they share the arena: the next
        run overwrites them; copy() to keep one)."""
        out = []
        for c, plan, offs in zip(self.circuits, self._plans, self._site_off):
            full = [1] + (plan[-1] if plan else [1] * (c.n - 1)) + [1]
            mps = MPS(c.n, c.d, device=self.engine.dev, _empty=True)
            mps._tensors = [self.arena[o:o + full[s] * c.d * full[s + 1]].view(full[s], c.d, full[s + 1])
                            for s, o in enumerate(offs)]
            out.append(mps)
        return out

    def close(self) -> None:
        self._release_graph()

    def __del__(self):
        try:
            self._release_graph()
        except Exception:
            pass
