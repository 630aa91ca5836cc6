"""Per-device executor: builds descriptor tables and calls libmpsx.

Every arithmetic step of the MPS path goes through here to the HIP library;
there is no other implementation to fall back on.
"""
from __future__ import annotations

import threading
import warnings
from collections import OrderedDict
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _native as nat

MAX_OPS = 65535  # descriptors per grouped launch (grid y / x limits of the kernels)


def _matrix_addr(x, rows: int, cols: int, ld: int, what: str) -> int:
    """Device address of a rows x cols matrix with row stride `ld` stored in
    a contiguous complex64 tensor at an element offset (x = tensor or
    (tensor, offset)); raises if any element would fall outside it."""
    t, off = x if isinstance(x, tuple) else (x, 0)
    if rows == 0 or cols == 0:
        return nat.ptr(t) + 8 * off if t.numel() else 0
    if cols > ld or off < 0 or off + (rows - 1) * ld + cols > t.numel():
        raise ValueError(f"gemm operand {what}: {rows}x{cols} (ld {ld}) at offset {off} "
                         f"exceeds a tensor of {t.numel()} elements")
    if t.element_size() != 8:
        raise ValueError(f"gemm operand {what} must be complex64")
    return nat.ptr(t) + 8 * off


SPLIT_ROWS = 128  # default split_rows: rows from which a problem counts toward splitting a call


def _rows(it) -> int:
    """Rows of the working matrix: the side orthogonalised (d*chi_l, or
    d*chi_r when S is absorbed left)."""
    if isinstance(it, SplitItem):
        m, n = it.mat.shape
        return n if it.absorb_left else m
    return it.a.shape[1] * (it.b.shape[2] if it.absorb_left else it.a.shape[0])


def _cost(it) -> float:
    """Relative Jacobi work of one problem: rows^2 * row length."""
    if isinstance(it, SplitItem):
        m, n = it.mat.shape
    else:
        m, n = it.a.shape[0] * it.a.shape[1], it.b.shape[1] * it.b.shape[2]
    c, r = (n, m) if it.absorb_left else (m, n)
    return float(c) * c * r


@dataclass
class SvdConfig:
    """Jacobi SVD knobs (see mpsx_svd_opts in include/mpsx.h)."""
    block: int = 0          # 0 = library default
    max_sweeps: int = 0
    check_every: int = 1
    inner_sweeps: int = 0
    tol: float = 0.0
    block_switch: int = 0   # 0 = library default (64 rows)
    unconverged: int = 0    # device address of an int32 non-convergence counter, or 0
    fp32: bool = False      # FP32 working matrix (faster, ~1e-6 relative error per update)
    method: int = 0         # 0: Gram method for FP64 32-row pivots; 1: one-sided everywhere
    fp32_sweeps: int = 0    # Gram method: > 0 adds an FP32 phase of at most this many sweeps
    precondition: int = 0   # Householder QR before the Gram method: 0 none, 1 saturated bonds, 2 all

    def to_c(self) -> nat.SvdOpts:
        return nat.SvdOpts(self.block, self.max_sweeps, self.check_every, self.inner_sweeps,
                           self.tol, self.block_switch, 1 if self.fp32 else 0,
                           self.unconverged or None, self.method, self.fp32_sweeps,
                           self.precondition)


@dataclass(slots=True)
class TwoSiteItem:
    a: "object"            # torch [chi_l, d, chi_m]
    b: "object"            # torch [chi_m, d, chi_r]
    gate: "object"         # torch complex64 [d^2, d^2] on device
    k: int
    absorb_left: bool = False
    want_s: bool = False


@dataclass
class SplitItem:
    """A plain M x N matrix to factorise like a two-site theta."""
    mat: "object"          # torch complex64 [M, N], contiguous
    k: int
    absorb_left: bool = False
    want_s: bool = False


class TwoSiteResult:
    """Factors of one update; `discarded` is a device float64 scalar view."""
    __slots__ = ("a_out", "b_out", "s", "_disc", "_i")

    def __init__(self, a_out, b_out, s, disc, i):
        self.a_out, self.b_out, self.s, self._disc, self._i = a_out, b_out, s, disc, i

    @property
    def discarded(self):
        return self._disc[self._i]


class Engine:
    """Owns the device, a gate cache and a grow-only workspace."""

    _engines: Dict[int, "Engine"] = {}
    _lock = threading.Lock()

    def __init__(self, dev):
        import torch
        self.torch = torch
        self.lib = nat.load()
        self.dev = nat.device(dev)
        self._ws = None
        self._gates: "OrderedDict[bytes, object]" = OrderedDict()
        from .tune import tuned_config, tuned_engine
        self.svd = tuned_config()
        self.last_sweeps = 0
        # concurrent groups of a split call (at most the 4 hardware queues a
        # process gets), the problems of >= split_rows rows a call needs to
        # be split, and that row count; from the stored tuning when there is
        # one (SPLIT_ROWS and 4 ways otherwise)
        self.split_ways, self.split_min, self.split_rows = tuned_engine()
        self._wss: list = []        # workspaces of the helper streams' groups
        self._helpers: list = []    # (stream, single-thread executor) per helper group
        self._keep: list = []       # device tables of the last call, alive until the next

    @classmethod
    def get(cls, dev=None) -> "Engine":
        d = nat.device(dev)
        idx = d.index if d.index is not None else 0
        with cls._lock:
            eng = cls._engines.get(idx)
            if eng is None:
                eng = cls(d)
                cls._engines[idx] = eng
            return eng

    # ------------------------------------------------------------ helpers
    @property
    def stream(self) -> int:
        return nat.stream_handle(self.dev)

    def workspace(self, nbytes: int, slot: int = 0):
        """Grow-only device workspace; slot g > 0 serves the g-th helper
        stream of split calls."""
        while len(self._wss) < slot:
            self._wss.append(None)
        cur = self._wss[slot - 1] if slot else self._ws
        if cur is None or cur.numel() < nbytes:
            grow = max(nbytes, 0 if cur is None else int(cur.numel() * 1.25))
            cur = self.torch.empty(max(grow, 1024), dtype=self.torch.uint8, device=self.dev)
            if slot:
                self._wss[slot - 1] = cur
            else:
                self._ws = cur
        return cur

    def gate(self, mat: np.ndarray):
        """Device complex64 copy of a gate matrix (cached by content)."""
        m = np.ascontiguousarray(np.asarray(mat, dtype=np.complex64))
        key = m.tobytes() + bytes(str(m.shape), "ascii")
        g = self._gates.get(key)
        if g is None:
            g = self.torch.from_numpy(m.copy()).to(self.dev)
            self._gates[key] = g
            if len(self._gates) > 8192:
                self._gates.popitem(last=False)
        else:
            self._gates.move_to_end(key)
        return g

    # ------------------------------------------------------------ one-site
    def one_site(self, items: Sequence[Tuple[object, object, object]]) -> None:
        """items: (tensor_in, tensor_out, gate_dev [d, d]); out may be in."""
        if not items:
            return
        rows = []
        for a, out, g in items:
            cl, d, cr = a.shape
            pa = nat.ptr(a)
            rows.append((pa, pa if out is a else nat.ptr(out), nat.ptr(g), cl, cr, d, 0))
        self.one_site_rows(rows)

    def one_site_rows(self, rows: Sequence[tuple]) -> None:
        """Grouped one-site launch from ready descriptor rows
        (a, out, gate, chi_l, chi_r, d, 0) with device addresses."""
        if not rows:
            return
        tab = np.array(rows, dtype=nat.SITE_OP)
        dtab = nat.to_device_table(tab, self.dev)
        nat.check(self.lib.mpsx_apply_one_site(self.stream, tab.ctypes.data, dtab.data_ptr(),
                                               len(rows)))
        del dtab

    # ------------------------------------------------------------ two-site
    def _prepare_two_site(self, items: Sequence[TwoSiteItem], cfg, slot: int = 0,
                          workspace: bool = True):
        """Outputs, planned host table, device table and workspace (None with
        workspace=False) of one library call (allocations on the calling
        thread's current stream)."""
        torch = self.torch
        n = len(items)
        # shapes and input pointers (one pass), then every output carved out of
        # one allocation and the descriptor table filled column by column
        cls_, ds, cms, crs, ks, pa, pb, pg, fl = ([] for _ in range(9))
        gate_ptr: Dict[int, int] = {}  # a batch shares few gate tensors
        for it in items:
            if isinstance(it, SplitItem):
                cl, cr = it.mat.shape
                d, cm = 1, 0
                pa.append(nat.ptr(it.mat)); pb.append(0); pg.append(0)
            else:
                cl, d, cm = it.a.shape
                cm2, d2, cr = it.b.shape
                if cm != cm2 or d != d2:
                    raise ValueError(f"site tensors do not share a bond: {tuple(it.a.shape)} {tuple(it.b.shape)}")
                g = gate_ptr.get(id(it.gate))
                if g is None:
                    g = gate_ptr[id(it.gate)] = nat.ptr(it.gate)
                pa.append(nat.ptr(it.a)); pb.append(nat.ptr(it.b)); pg.append(g)
            cls_.append(cl); ds.append(d); cms.append(cm); crs.append(cr); ks.append(it.k)
            fl.append(1 if it.absorb_left else 0)
        cl_a, d_a, cr_a, k_a = (np.asarray(x, dtype=np.int64) for x in (cls_, ds, crs, ks))
        na = cl_a * d_a * k_a
        nb = k_a * d_a * cr_a
        ns = np.array([it.k if it.want_s else 0 for it in items], dtype=np.int64)
        # one allocation per update holds a_out, b_out and S (complex64 slots at
        # 32-element boundaries; S as float32 pairs), so a surviving site tensor
        # pins only its own update's memory
        la = (na + 31) // 32 * 32
        lb = (nb + 31) // 32 * 32
        length = la + lb + (ns + 1) // 2
        disc = torch.empty(n, dtype=torch.float64, device=self.dev)
        bufs = [torch.empty(max(int(m), 1), dtype=torch.complex64, device=self.dev)
                for m in length.tolist()]
        ptr0 = np.array([b.data_ptr() for b in bufs], dtype=np.int64)  # fresh, contiguous
        tab = np.zeros(n, dtype=nat.TWO_SITE_OP)
        tab["a"], tab["b"], tab["gate"] = pa, pb, pg
        tab["a_out"] = np.where(na > 0, ptr0, 0)
        tab["b_out"] = np.where(nb > 0, ptr0 + 8 * la, 0)
        tab["s_out"] = np.where(ns > 0, ptr0 + 8 * (la + lb), 0)
        tab["discarded"] = int(disc.data_ptr()) + 8 * np.arange(n, dtype=np.int64)
        tab["chi_l"], tab["chi_m"], tab["chi_r"] = cls_, cms, crs
        tab["d"], tab["k"], tab["flags"] = ds, ks, fl
        outs: List[TwoSiteResult] = []
        for i, buf in enumerate(bufs):
            cl, d, cr, k = cls_[i], ds[i], crs[i], ks[i]
            a_out = buf.as_strided((cl, d, k), (d * k, k, 1), 0)
            b_out = buf.as_strided((k, d, cr), (d * cr, cr, 1), int(la[i]))
            s = (buf.view(torch.float32).as_strided((k,), (1,), 2 * int(la[i] + lb[i]))
                 if ns[i] else None)
            outs.append(TwoSiteResult(a_out, b_out, s, disc, i))
        import ctypes
        nbytes = ctypes.c_size_t(0)
        nat.check(self.lib.mpsx_two_site_plan(tab.ctypes.data, n, ctypes.byref(cfg),
                                              ctypes.byref(nbytes)))
        ws = self.workspace(nbytes.value, slot) if workspace else None
        dtab = nat.to_device_table(tab, self.dev)
        return outs, tab, dtab, ws

    def _run_two_site(self, tab, dtab, ws, cfg, stream: int) -> int:
        """The library call itself (blocking while it polls convergence);
        returns the sweeps it ran."""
        import ctypes
        sweeps = ctypes.c_int(0)
        code = self.lib.mpsx_two_site(stream, tab.ctypes.data, dtab.data_ptr(), len(tab),
                                      ctypes.byref(cfg), ws.data_ptr(), ws.numel(),
                                      ctypes.byref(sweeps))
        if code == nat.MPSX_ENOTCONV:
            warnings.warn(nat.load().mpsx_last_error().decode(), nat.NotConvergedWarning)
        else:
            nat.check(code)
        return sweeps.value

    def two_site(self, items: Sequence[TwoSiteItem], svd: Optional[SvdConfig] = None
                 ) -> List[TwoSiteResult]:
        """Batched two-site updates on the current stream. A call with at
        least `split_min` problems of `split_rows` or more rows runs as
        `split_ways` groups of equal work concurrently, group 0 on the current
        stream and each other group from its own worker thread on its own
        helper stream: the latency-bound pivot eigensolves of one group then
        overlap the matrix-core phases of the others. Each problem's
        arithmetic is the same either way."""
        n = len(items)
        if n == 0:
            return []
        cfg = (svd or self.svd).to_c()
        self._keep = []
        big = [i for i, it in enumerate(items) if _rows(it) >= self.split_rows]
        ways = max(1, min(int(self.split_ways), n))
        if ways < 2 or len(big) < self.split_min or cfg.check_every == 0:
            outs, tab, dtab, ws = self._prepare_two_site(items, cfg)
            self._keep = [dtab]
            self.last_sweeps = self._run_two_site(tab, dtab, ws, cfg, self.stream)
            return outs
        # groups of equal estimated work (rows^2 * row length per problem),
        # largest problems first, each to the least loaded group
        order = sorted(range(n), key=lambda i: -_cost(items[i]))
        grp = [[] for _ in range(ways)]
        load = [0.0] * ways
        for i in order:
            g = min(range(ways), key=lambda x: load[x])
            grp[g].append(i)
            load[g] += _cost(items[i])
        grp = [sorted(g) for g in grp if g]
        preps = []
        for slot, g in enumerate(grp):
            c = cfg if slot == 0 else (svd or self.svd).to_c()
            preps.append((c,) + self._prepare_two_site([items[i] for i in g], c, slot))
        self._keep = [p[3] for p in preps]
        torch = self.torch
        from concurrent.futures import ThreadPoolExecutor
        while len(self._helpers) < len(grp) - 1:
            self._helpers.append((torch.cuda.Stream(self.dev), ThreadPoolExecutor(max_workers=1)))
        cur = torch.cuda.current_stream(self.dev)
        fork = torch.cuda.Event()
        fork.record(cur)
        dev = self.dev
        futs = []
        for slot in range(1, len(grp)):
            hs, pool = self._helpers[slot - 1]
            hs.wait_event(fork)  # inputs and outputs are ready on the helper stream
            c, outs_g, tab, dtab, ws = preps[slot]

            def run(c=c, tab=tab, dtab=dtab, ws=ws, hs=hs):
                torch.cuda.set_device(dev)
                return self._run_two_site(tab, dtab, ws, c, int(hs.cuda_stream))

            futs.append(pool.submit(run))
        c, _, tab, dtab, ws = preps[0]
        sweeps = [self._run_two_site(tab, dtab, ws, c, int(cur.cuda_stream))]
        sweeps += [f.result() for f in futs]
        for slot in range(1, len(grp)):
            join = torch.cuda.Event()
            join.record(self._helpers[slot - 1][0])
            cur.wait_event(join)
        self.last_sweeps = max(sweeps)
        outs: List[TwoSiteResult] = [None] * n
        for g, p in zip(grp, preps):
            for j, i in enumerate(g):
                outs[i] = p[1][j]
        return outs

    # ------------------------------------------------------------ gemm
    def gemm(self, ops: Sequence[dict]) -> None:
        """ops: dicts with A, B, C, m, n, k, lda, ldb, ldc, flags, alpha
        (complex). A, B and C are complex64 tensors, or (tensor, element
        offset) pairs for a matrix that starts inside a tensor. Every matrix
        is checked to lie inside its tensor before anything is launched."""
        if not ops:
            return
        if len(ops) > MAX_OPS:  # one launch takes at most MAX_OPS products
            for s in range(0, len(ops), MAX_OPS):
                self.gemm(ops[s:s + MAX_OPS])
            return
        tab = np.zeros(len(ops), dtype=nat.GEMM_OP)
        for i, o in enumerate(ops):
            m, n, k, fl = int(o["m"]), int(o["n"]), int(o["k"]), int(o.get("flags", 0))
            # stored shapes: op(A) = A^H means A is stored k x m, else m x k
            ra, ca = (k, m) if fl & nat.GEMM_CONJ_A else (m, k)
            rb, cb = (n, k) if fl & nat.GEMM_CONJ_B else (k, n)
            pa = _matrix_addr(o["A"], ra, ca, o["lda"], "A")
            pb = _matrix_addr(o["B"], rb, cb, o["ldb"], "B")
            pc = _matrix_addr(o["C"], m, n, o["ldc"], "C")
            al = complex(o.get("alpha", 1.0))
            tab[i] = (pa, pb, pc, m, n, k, o["lda"], o["ldb"], o["ldc"], al.real, al.imag, fl, 0)
        dtab = nat.to_device_table(tab, self.dev)
        nat.check(self.lib.mpsx_gemm(self.stream, tab.ctypes.data, dtab.data_ptr(), len(ops)))
        self._last_gemm_table = dtab

    # ------------------------------------------------------------ scale
    def scale(self, tensors: Sequence[object], factor: complex, conj: bool = False) -> None:
        if not tensors:
            return
        tab = np.zeros(len(tensors), dtype=nat.SCALE_OP)
        f = complex(factor)
        for i, t in enumerate(tensors):
            tab[i] = (nat.ptr(t), t.numel(), f.real, f.imag, 1 if conj else 0, 0)
        dtab = nat.to_device_table(tab, self.dev)
        nat.check(self.lib.mpsx_scale(self.stream, tab.ctypes.data, dtab.data_ptr(), len(tensors)))
        self._last_scale_table = dtab
