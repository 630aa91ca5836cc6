"""MPS state and operations on device-resident complex64 tensors.

Drop-in counterpart of mpsim.MPS / mpsim.MPSOperation (mpsim/core.py) for
the gate-application path: same names, keyword arguments and exceptions.
Site i is a torch complex64 tensor [chi_l, d, chi_r] on the GPU (ends carry
a dummy bond of size 1); all arithmetic runs in libmpsx's HIP kernels.

Known, documented differences from the reference:
  - Site tensors are [chi_l, d, chi_r] (the reference's TensorNetwork axis
    order is internal and unobservable through the gate path).
  - Storage stays complex64; the reference promotes tensors to complex128
    after float64/complex128 gates (gates.py:195-212, 272-285).
  - The per-gate norm the reference appends to `_norms` after every
    two-qudit gate (core.py:1160-1161, "only for convenience in
    benchmarking") is off by default; MPS(track_norms=True) turns it on.
  - MPS.from_wavefunction splits with S absorbed right (U, S V^H); the
    reference's tn.split_node absorbs sqrt(S) into both sides. The state,
    norms and bond dimensions are the same; only the gauge of the site
    tensors differs.
  - `maxsvals` < 0 raises ValueError (the reference slices with it).
"""
from __future__ import annotations

from copy import deepcopy
from typing import List, Optional, Sequence, Tuple, Union

import numpy as np

from . import _native as nat
from .engine import Engine, SplitItem, TwoSiteItem
from .gates import (Edge, Node, as_tensor, cnot, haar_random_unitary, hgate, is_unitary, rgate,
                    swap, xgate)


class CannotConvertToMPSOperation(Exception):
    pass


class BondEdge:
    """The bond between sites j and j+1 of an MPS. Its two end nodes are
    named as the reference wires the bonds of a new MPS (core.py:221-229):
    (q_j, q_{j+1}), except the last bond of a chain of three or more sites,
    which is wired from the last site, (q_{n-1}, q_{n-2}). Two BondEdges are
    equal when they are the same bond of the same MPS."""

    def __init__(self, mps: "MPS", j: int) -> None:
        self._mps = mps
        self.index = j
        n = mps.nqudits
        self._ends = (j + 1, j) if (n >= 3 and j == n - 2) else (j, j + 1)

    @property
    def node1(self) -> Node:
        return self._mps.get_node(self._ends[0])

    @property
    def node2(self) -> Node:
        return self._mps.get_node(self._ends[1])

    @property
    def dimension(self) -> int:
        return int(self._mps.tensors[self.index].shape[2])

    def is_dangling(self) -> bool:
        return False

    def __eq__(self, other) -> bool:
        return (isinstance(other, BondEdge) and other._mps is self._mps
                and other.index == self.index)

    __hash__ = None

    def __repr__(self) -> str:
        return f"BondEdge({self.index}, dim={self.dimension})"


class MPSOperation:
    """A gate and the qudit indices it acts on (core.py:21-156)."""

    def __init__(self, node, qudit_indices: Union[int, Tuple[int, ...]], qudit_dimension: int = 2):
        self._node = node if isinstance(node, Node) or hasattr(node, "tensor") else Node(node)
        if isinstance(qudit_indices, (int, np.integer)):
            qudit_indices = (int(qudit_indices),)
        self._qudit_indices = tuple(int(q) for q in qudit_indices)
        self._qudit_dimension = int(qudit_dimension)

    @property
    def qudit_indices(self) -> Tuple[int, ...]:
        return self._qudit_indices

    @property
    def qudit_dimension(self) -> int:
        return self._qudit_dimension

    @property
    def num_qudits(self) -> int:
        return len(self._qudit_indices)

    def node(self, copy: bool = True):
        """core.py:86-95: a copy by default."""
        if not copy:
            return self._node
        return Node(np.array(as_tensor(self._node), copy=True), getattr(self._node, "name", None))

    def tensor(self, reshape_to_square_matrix: bool = True) -> np.ndarray:
        t = deepcopy(as_tensor(self._node))
        if reshape_to_square_matrix:
            dim = self._qudit_dimension ** self.num_qudits
            t = np.reshape(t, (dim, dim))
        return t

    def is_valid(self) -> bool:
        """core.py:113-129: shape (d,)*2k with all edges dangling."""
        d = self._qudit_dimension
        t = as_tensor(self._node)
        if t.shape != tuple([d] * 2 * self.num_qudits):
            return False
        nondangling = getattr(self._node, "has_nondangling_edge", None)
        if callable(nondangling) and nondangling():
            return False
        return True

    def is_unitary(self) -> bool:
        return is_unitary(self.tensor(reshape_to_square_matrix=True))

    def is_hermitian(self) -> bool:
        from .gates import is_hermitian
        return is_hermitian(self.tensor(reshape_to_square_matrix=True))

    def is_single_qudit_operation(self) -> bool:
        return self.num_qudits == 1

    def is_two_qudit_operation(self) -> bool:
        return self.num_qudits == 2

    def __str__(self) -> str:
        return f"Tensor {getattr(self._node, 'name', '')} on qudit(s) {self._qudit_indices}."


def max_bond_dimensions_for(n: int, d: int) -> List[int]:
    """core.py:235-242."""
    half = [d ** (i + 1) for i in range(n // 2)]
    out = half + list(reversed(half))
    if n % 2 == 0:
        out.remove(d ** (n // 2))
    return out


def parse_split_options(kwargs, max_bond_of_left: int) -> Tuple[Optional[int], bool]:
    """core.py:1104-1130: (maxsvals or None, keep_left_canonical)."""
    keep_left = kwargs["keep_left_canonical"] if "keep_left_canonical" in kwargs else True
    if "fraction" in kwargs and "maxsvals" in kwargs:
        raise ValueError("Only one of (fraction, maxsvals) can be provided as kwargs.")
    maxsvals = None
    if "fraction" in kwargs:
        fraction = kwargs.get("fraction")
        if not (0 <= fraction <= 1):
            raise ValueError("Keyword fraction must be between 0 and 1 but is", fraction)
        maxsvals = int(round(fraction * max_bond_of_left))
    if "maxsvals" in kwargs:
        maxsvals = int(kwargs.get("maxsvals"))
    if maxsvals is not None and maxsvals < 0:
        raise ValueError(f"maxsvals must be non-negative but is {maxsvals}")
    return maxsvals, bool(keep_left)


class MPS:
    """Matrix product state on one GPU (core.py:159-1423, gate path)."""

    def __init__(self, nqudits: int, qudit_dimension: int = 2, tensor_prefix: str = "q", *,
                 device=None, track_norms: bool = False, _empty: bool = False) -> None:
        if nqudits < 2:
            raise ValueError(f"Number of qudits must be greater than 2 but is {nqudits}.")
        self._nqudits = int(nqudits)
        self._qudit_dimension = int(qudit_dimension)
        self._prefix = tensor_prefix
        self._engine = Engine.get(device)
        self._track_norms = track_norms
        self._norms: List[float] = []
        self._max_bond_dimensions = max_bond_dimensions_for(self._nqudits, self._qudit_dimension)
        if _empty:
            self._tensors = []
            return
        torch = self._engine.torch
        # core.py:190-218: every site |0>, complex64.
        block = torch.zeros((self._nqudits, 1, self._qudit_dimension, 1), dtype=torch.complex64,
                            device=self._engine.dev)
        block[:, 0, 0, 0] = 1.0
        self._tensors = [block[i] for i in range(self._nqudits)]

    # ------------------------------------------------------------ properties
    @property
    def nqudits(self) -> int:
        return self._nqudits

    @property
    def qudit_dimension(self) -> int:
        return self._qudit_dimension

    @property
    def device(self):
        return self._engine.dev

    @property
    def tensors(self):
        """The device site tensors [chi_l, d, chi_r] (not copies)."""
        return self._tensors

    def bond_dimension_of(self, node_index: int) -> int:
        """core.py:339-361: dimension of the bond right of node_index."""
        if not self.is_valid():
            raise ValueError("MPS is invalid.")
        if node_index >= self._nqudits:
            raise ValueError(f"Index should be less than {self._nqudits} but is {node_index}.")
        if node_index == self._nqudits - 1 or node_index < -(self._nqudits - 1):
            raise ValueError(f"Node {node_index} has no right bond.")
        return int(self._tensors[node_index].shape[2])

    def bond_dimensions(self) -> List[int]:
        return [int(t.shape[2]) for t in self._tensors[:-1]]

    def max_bond_dimension_of(self, edge_index: int) -> int:
        if edge_index >= self._nqudits:
            raise ValueError(f"Edge index should be less than {self._nqudits} but is {edge_index}.")
        return self._max_bond_dimensions[edge_index]

    def max_bond_dimensions(self) -> List[int]:
        return self._max_bond_dimensions

    def is_valid(self) -> bool:
        """Structural check (core.py:388-419): n >= 2 tensors, physical legs
        of size d, neighbouring bonds matching, unit end bonds."""
        t = self._tensors
        if len(t) < 2:
            return False
        if t[0].shape[0] != 1 or t[-1].shape[2] != 1:
            return False
        for i, x in enumerate(t):
            if x.dim() != 3 or x.shape[1] != self._qudit_dimension:
                return False
            if i + 1 < len(t) and x.shape[2] != t[i + 1].shape[0]:
                return False
        return True

    def get_nodes(self, copy: bool = True) -> List[Node]:
        """Host copies of the site tensors as Nodes ([chi_l, d, chi_r])."""
        return [Node(t.detach().cpu().numpy(), name=self._prefix + str(i))
                for i, t in enumerate(self._tensors)]

    def get_node(self, node_index: int, copy: bool = True) -> Node:
        t = self._tensors[node_index]
        i = node_index % self._nqudits
        return Node(t.detach().cpu().numpy(), name=self._prefix + str(i))

    def get_free_edge_of(self, node_index: int, copy: bool = True) -> Edge:
        """core.py:443-451: the physical leg of a site (axis 1 of the
        [chi_l, d, chi_r] host copy)."""
        return self.get_node(node_index, copy).get_edge(1)

    def get_left_connected_edge_of(self, node_index: int) -> Optional["BondEdge"]:
        """core.py:453-466: the bond left of a site, None for site 0."""
        i = node_index % self._nqudits
        return None if i == 0 else BondEdge(self, i - 1)

    def get_right_connected_edge_of(self, node_index: int) -> Optional["BondEdge"]:
        """core.py:468-481: the bond right of a site, None for the last site."""
        i = node_index % self._nqudits
        return None if i == self._nqudits - 1 else BondEdge(self, i)

    # ------------------------------------------------------------ read-out
    def wavefunction(self) -> np.ndarray:
        """core.py:483-500: psi with site 0 the most significant digit."""
        if not self.is_valid():
            raise ValueError("MPS is not valid.")
        return self._wavefunction_device().cpu().numpy()

    def _wavefunction_device(self):
        torch = self._engine.torch
        d = self._qudit_dimension
        t0 = self._tensors[0]
        psi = t0.reshape(d, t0.shape[2])
        for t in self._tensors[1:]:
            cl, _, cr = t.shape
            rows = psi.shape[0]
            out = torch.empty((rows, d * cr), dtype=torch.complex64, device=self._engine.dev)
            self._engine.gemm([dict(A=psi, B=t, C=out, m=rows, n=d * cr, k=cl, lda=psi.shape[1],
                                    ldb=d * cr, ldc=d * cr)])
            psi = out.reshape(rows * d, cr)
        return psi.reshape(d ** self._nqudits)

    def dagger(self) -> None:
        """core.py:502-505: conjugate every tensor."""
        self._engine.scale([t for t in self._tensors if t.numel()], 1.0, conj=True)

    def inner_product(self, other: "MPS") -> complex:
        """core.py:507-561. Like the reference (which conjugates `other`,
        core.py:545-546) this returns sum_psi self(psi) * conj(other(psi))."""
        if other._nqudits != self._nqudits:
            raise ValueError("Cannot compute inner product: number of qudits must be equal.")
        if other._qudit_dimension != self._qudit_dimension:
            raise ValueError("Cannot compute inner product: qudit dimensions must be equal.")
        if not self.is_valid():
            raise ValueError("MPS is invalid.")
        if not other.is_valid():
            raise ValueError("Other MPS is invalid.")
        return complex(self._inner_device(other).item())

    def _inner_device(self, other: "MPS"):
        """E[y, x] = sum conj(other) self over the sites left of the cut."""
        torch = self._engine.torch
        dev = self._engine.dev
        d = self._qudit_dimension
        env = torch.ones((1, 1), dtype=torch.complex64, device=dev)
        for a, b in zip(self._tensors, other._tensors):
            xl, _, xr = a.shape
            yl, _, yr = b.shape
            f = torch.empty((yl, d * xr), dtype=torch.complex64, device=dev)
            e2 = torch.empty((yr, xr), dtype=torch.complex64, device=dev)
            # F[y, (p, x')] = sum_x E[y, x] a[x, p, x']
            self._engine.gemm([dict(A=env, B=a, C=f, m=yl, n=d * xr, k=xl, lda=xl, ldb=d * xr,
                                    ldc=d * xr)])
            # E'[y', x'] = sum_{y,p} conj(b[y, p, y']) F[y, p, x']
            self._engine.gemm([dict(A=b, B=f, C=e2, m=yr, n=xr, k=yl * d, lda=yr, ldb=xr,
                                    ldc=xr, flags=nat.GEMM_CONJ_A)])
            env = e2
        return env[0, 0]

    def norm(self) -> float:
        """core.py:563-565."""
        return float(np.sqrt(self.inner_product(self).real))

    def renormalize(self, to_norm: float = 1.0) -> None:
        """core.py:567-594."""
        if to_norm < 0.0:
            raise ValueError(f"Arg to_norm must be positive but is {to_norm}")
        if np.isclose(to_norm, 0.0, atol=1e-15):
            raise ValueError(f"Arg to_norm = {to_norm} is too close to numerical zero.")
        norm = self.norm()
        if np.isclose(norm, 0.0, atol=1e-15):
            raise ValueError("Norm of MPS is numerically zero, cannot renormalize.")
        f = (to_norm / norm) ** (1.0 / self._nqudits)
        self._engine.scale([t for t in self._tensors if t.numel()], f)

    # ------------------------------------------------------------ reduced states
    def reduced_density_matrix(self, node_indices: Union[int, Sequence[int]]) -> np.ndarray:
        """core.py:596-652: the density matrix of the kept sites with every
        other site traced out, shape (d^m, d^m). Rows are the kept ket legs
        and columns the kept bra legs, each a big-endian number over the
        sites in the order `node_indices` lists them (core.py:648-651)."""
        try:
            idx = tuple(iter(node_indices))
        except TypeError:
            idx = (node_indices,)
        idx = tuple(int(i) for i in idx)
        if len(set(idx)) < len(idx):
            raise ValueError("Node indices contains duplicates.")
        if min(idx) < 0 or max(idx) > self._nqudits - 1:
            raise IndexError("One or more invalid node indices.")
        kept = sorted(idx)
        rho = self._rdm_device(kept).cpu().numpy()
        m, d = len(kept), self._qudit_dimension
        pos = [kept.index(i) for i in idx]
        rho = rho.reshape([d] * (2 * m)).transpose(pos + [m + p for p in pos])
        return np.ascontiguousarray(rho).reshape(d ** m, d ** m)

    def _rdm_device(self, kept: Sequence[int]):
        """rho[a, b] over the kept sites (sorted), a the ket and b the bra
        multi-index, by one left-to-right sweep of grouped GEMMs.

        The environment after a prefix of sites is E[b, y, a, x]: b / a the
        bra / ket values of the kept sites so far, y / x the open bra / ket
        bond. Per site A[x, s, x']:
          F[b, y, a, (s, x')] = sum_x E[b, y, a, x] A[x, s, x']       (one GEMM)
          kept:   E'[(b, t), y', (a, s), x'] = sum_y conj(A[y, t, y']) F[b, y, a, s, x']
                  (one GEMM per b)
          traced: E'[b, y', a, x'] = sum_{y, s} conj(A[y, s, y']) F[b, y, a, s, x']
                  (one GEMM per (b, a), accumulated over s)
        """
        torch = self._engine.torch
        dev, eng, d = self._engine.dev, self._engine, self._qudit_dimension
        keep = set(kept)
        env = torch.ones((1, 1, 1, 1), dtype=torch.complex64, device=dev)
        kb = ka = 1
        for i, a in enumerate(self._tensors):
            xl, _, xr = a.shape
            f = torch.empty((kb, xl, ka, d * xr), dtype=torch.complex64, device=dev)
            eng.gemm([dict(A=env, B=a, C=f, m=kb * xl * ka, n=d * xr, k=xl, lda=xl,
                           ldb=d * xr, ldc=d * xr)])
            row_f = ka * d * xr
            if i in keep:
                nxt = torch.empty((kb, d, xr, ka, d, xr), dtype=torch.complex64, device=dev)
                eng.gemm([dict(A=a, B=(f, b * xl * row_f), C=(nxt, b * d * xr * ka * d * xr),
                               m=d * xr, n=row_f, k=xl, lda=d * xr, ldb=row_f, ldc=row_f,
                               flags=nat.GEMM_CONJ_A) for b in range(kb)])
                kb, ka = kb * d, ka * d
            else:
                nxt = torch.empty((kb, xr, ka, xr), dtype=torch.complex64, device=dev)
                for s in range(d):
                    eng.gemm([dict(A=(a, s * xr), B=(f, b * xl * row_f + c * d * xr + s * xr),
                                   C=(nxt, b * xr * ka * xr + c * xr), m=xr, n=xr, k=xl,
                                   lda=d * xr, ldb=row_f, ldc=ka * xr,
                                   flags=nat.GEMM_CONJ_A | (nat.GEMM_ACCUM if s else 0))
                              for b in range(kb) for c in range(ka)])
            env = nxt
        return env.reshape(kb, ka).transpose(0, 1)

    def _marginals_device(self):
        """p[i, s] = <psi| (|s><s| on site i) |psi> for every site at once:
        right environments by a right-to-left sweep, then one left-to-right
        sweep that splits each site's transfer step by physical value.

        With R~_i[y, x] the environment of the bonds right of site i-1 (bra
        y, ket x), Q_i = conj(R~_i) is what is stored:
          G[(x, s), y'] = sum_x' A[(x, s), x'] R~_{i+1}[y', x']  = A Q_{i+1}^H
          Q_i = A_(y, (s y')) G_(x, (s y'))^H
        Left sweep with L[y, x]:
          F = L A;  D_s[y', x'] = sum_y conj(A[y, s, y']) F[y, s, x'];
          p[i, s] = sum D_s[y', x'] R~_{i+1}[y', x'];  L' = sum_s D_s."""
        torch = self._engine.torch
        dev, eng, d, n = self._engine.dev, self._engine, self._qudit_dimension, self._nqudits
        c64 = torch.complex64
        q = [None] * (n + 1)
        q[n] = torch.ones((1, 1), dtype=c64, device=dev)
        for i in range(n - 1, 0, -1):
            a = self._tensors[i]
            xl, _, xr = a.shape
            g = torch.empty((xl * d, xr), dtype=c64, device=dev)
            eng.gemm([dict(A=a, B=q[i + 1], C=g, m=xl * d, n=xr, k=xr, lda=xr, ldb=xr, ldc=xr,
                           flags=nat.GEMM_CONJ_B)])
            q[i] = torch.empty((xl, xl), dtype=c64, device=dev)
            eng.gemm([dict(A=a, B=g, C=q[i], m=xl, n=xl, k=d * xr, lda=d * xr, ldb=d * xr,
                           ldc=xl, flags=nat.GEMM_CONJ_B)])
        env = torch.ones((1, 1), dtype=c64, device=dev)
        dsv = []
        for i, a in enumerate(self._tensors):
            xl, _, xr = a.shape
            f = torch.empty((xl, d * xr), dtype=c64, device=dev)
            eng.gemm([dict(A=env, B=a, C=f, m=xl, n=d * xr, k=xl, lda=xl, ldb=d * xr,
                           ldc=d * xr)])
            ds = torch.empty((d, xr, xr), dtype=c64, device=dev)
            eng.gemm([dict(A=(a, s * xr), B=(f, s * xr), C=(ds, s * xr * xr), m=xr, n=xr,
                           k=xl, lda=d * xr, ldb=d * xr, ldc=xr, flags=nat.GEMM_CONJ_A)
                      for s in range(d)])
            dsv.append(ds)
            if i + 1 < n:
                env = torch.empty((xr, xr), dtype=c64, device=dev)
                eng.gemm([dict(A=a, B=f, C=env, m=xr, n=xr, k=xl * d, lda=xr, ldb=xr, ldc=xr,
                               flags=nat.GEMM_CONJ_A)])
        p = torch.empty((n, d), dtype=c64, device=dev)
        # p[i, s] = conj(Q_{i+1})^T . vec(D_s): a 1 x 1 product over xr^2 terms
        eng.gemm([dict(A=q[i + 1], B=(dsv[i], s * q[i + 1].numel()), C=(p, i * d + s), m=1, n=1,
                       k=q[i + 1].numel(), lda=1, ldb=1, ldc=1, flags=nat.GEMM_CONJ_A)
                  for i in range(n) for s in range(d)])
        return p.real

    def _sample(self, as_string: bool = False, marginals: Optional[np.ndarray] = None):
        """core.py:654-682. The reference draws each site from the diagonal
        of its own one-site reduced density matrix of the unmeasured state
        (core.py:668 reads `self`, not the projected copy), so the sites of
        one sample are drawn independently from their marginals; this keeps
        that behaviour and the same sequence of np.random.choice calls."""
        if marginals is None:
            marginals = self._marginal_probabilities()
        states = list(range(self._qudit_dimension))
        string = [np.random.choice(states, size=1, p=marginals[i])[0]
                  for i in range(self._nqudits)]
        if as_string:
            return "".join(str(bit) for bit in string)
        return string

    def _marginal_probabilities(self) -> np.ndarray:
        """float32 [n, d] marginals with rounding-level negatives set to zero
        (a zero-probability value can come out as -1e-9 in FP32)."""
        p = self._marginals_device().cpu().numpy().astype(np.float32)
        p[(p < 0) & (p > -1e-6)] = 0.0
        return p

    def sample(self, nsamples: int, as_hist: bool = False, as_string: bool = False):
        """core.py:684-721: nsamples draws; as_hist gives {bitstring: count}
        (and implies as_string). The marginals are computed once on the
        device, since the state does not change between draws."""
        if not isinstance(nsamples, int):
            raise ValueError(f"Arg nsamples should be an int but is a {type(nsamples)}.")
        if nsamples <= 0:
            raise ValueError(f"Arg nsamples should be positive but is {nsamples}.")
        if as_hist:
            as_string = True
        marg = self._marginal_probabilities()
        raw = [self._sample(as_string, marg) for _ in range(nsamples)]
        if as_hist:
            hist = {}
            for bits in raw:
                hist[bits] = hist.get(bits, 0) + 1
            return hist
        return raw

    # ------------------------------------------------------------ one-site
    def apply_one_qudit_gate(self, gate, node_index: int, **kwargs) -> None:
        """core.py:753-845."""
        if not self.is_valid():
            raise ValueError("MPS is invalid.")
        if node_index not in range(self._nqudits):
            raise ValueError(f"Input tensor index={node_index} is out of bounds for"
                             f" an MPS on {self._nqudits} qudits.")
        nondangling = getattr(gate, "get_all_nondangling", None)
        g = as_tensor(gate)
        if g.ndim != 2 or (callable(nondangling) and len(nondangling()) != 0):
            raise ValueError("Single qudit gate must have two free edges and zero connected edges.")
        if g.shape[0] != g.shape[1]:
            raise ValueError("Gate edge dimensions must be equal.")
        if g.shape[0] != self._qudit_dimension:
            raise ValueError(f"Gate edges have dimension {g.shape[0]} "
                             f"but should have MPS qudit dimension = {self._qudit_dimension}")
        unitary = is_unitary(g)
        renorm = kwargs.get("renormalize_after_non_unitary") is not False
        ortho = kwargs.get("ortho_after_non_unitary") is not False
        norm = self.norm() if (not unitary and renorm) else None
        t = self._tensors[node_index]
        if t.numel():
            self._engine.one_site([(t, t, self._engine.gate(g))])
        if not unitary and ortho:  # core.py:828-841
            if node_index == 0:
                self.orthonormalize_right_edge_of(node_index)
            elif node_index == self._nqudits - 1:
                self.orthonormalize_left_edge_of(node_index)
            else:
                self.orthonormalize_right_edge_of(node_index)
                self.orthonormalize_left_edge_of(node_index)
        if norm is not None:
            self.renormalize(norm)

    # ------------------------------------------------------------ splits
    def _split_keep(self, mat, absorb_left: bool, threshold: float):
        """SVD of a device matrix keeping the values TensorNetwork keeps for
        max_truncation_err = threshold * norm(): the number of tails
        sqrt(sum of the j smallest s^2) that exceed the error (its numpy
        backend's rule for split_node_full_svd)."""
        M, N = mat.shape
        full = min(M, N)
        res = self._engine.two_site([SplitItem(mat, full, absorb_left, want_s=True)])[0]
        err = threshold * self.norm()
        s = res.s.double().cpu().numpy() if full else np.zeros(0)
        tails = np.sqrt(np.cumsum(np.square(s[::-1])))
        keep = min(full, int(np.count_nonzero(tails > err)))
        a = res.a_out.reshape(M, full)[:, :keep].contiguous()
        b = res.b_out.reshape(full, N)[:keep].contiguous()
        return a, b, keep

    def orthonormalize_right_edge_of(self, node_index: int, threshold: float = 1e-8) -> None:
        """core.py:847-892: node = U, right neighbour = S V^H . neighbour."""
        if not 0 <= node_index < self._nqudits - 1:
            raise ValueError("Invalid edge index.")
        torch = self._engine.torch
        t, nxt = self._tensors[node_index], self._tensors[node_index + 1]
        cl, d, cr = t.shape
        u, sv, k = self._split_keep(t.reshape(cl * d, cr), False, threshold)
        _, d2, cr2 = nxt.shape
        new = torch.empty((k, d2, cr2), dtype=torch.complex64, device=self._engine.dev)
        self._engine.gemm([dict(A=sv, B=nxt, C=new, m=k, n=d2 * cr2, k=cr, lda=cr,
                                ldb=d2 * cr2, ldc=d2 * cr2)])
        self._tensors[node_index] = u.reshape(cl, d, k)
        self._tensors[node_index + 1] = new

    def orthonormalize_left_edge_of(self, node_index: int, threshold: float = 1e-8) -> None:
        """core.py:894-939: node = V^H, left neighbour = neighbour . U S."""
        if not 0 < node_index <= self._nqudits - 1:
            raise ValueError("Invalid edge index.")
        torch = self._engine.torch
        t, prv = self._tensors[node_index], self._tensors[node_index - 1]
        cl, d, cr = t.shape
        us, v, k = self._split_keep(t.reshape(cl, d * cr), True, threshold)
        cl2, d2, _ = prv.shape
        new = torch.empty((cl2, d2, k), dtype=torch.complex64, device=self._engine.dev)
        self._engine.gemm([dict(A=prv, B=us, C=new, m=cl2 * d2, n=k, k=cl, lda=cl, ldb=k, ldc=k)])
        self._tensors[node_index] = v.reshape(k, d, cr)
        self._tensors[node_index - 1] = new

    @staticmethod
    def from_wavefunction(wavefunction, nqudits: int, qudit_dimension: int = 2,
                          tensor_prefix: str = "q", *, device=None) -> "MPS":
        """core.py:245-328: successive SVD splits of a state vector (site 0
        most significant), keeping every singular value."""
        if not isinstance(wavefunction, (list, tuple, np.ndarray)):
            raise TypeError("Invalid type for wavefunction.")
        wf = np.array(wavefunction)
        if wf.ndim != 1:
            raise ValueError("Invalid shape for wavefunction. Should be a vector.")
        if nqudits < 2:
            raise ValueError("At least two qudits are required.")
        d = qudit_dimension
        if wf.size != d ** nqudits:
            raise ValueError(
                "Mismatch between wavefunction, qudit_dimension, and nqudits. "
                f"Expected {d ** nqudits} elements in the wavefunction, but wavefunction has "
                f"{wf.size} elements.")
        mps = MPS(nqudits, d, tensor_prefix, device=device, _empty=True)
        eng = mps._engine
        rest = eng.torch.from_numpy(wf.astype(np.complex64)).to(eng.dev)
        chi = 1
        tensors = []
        for i in range(nqudits - 1):
            M = chi * d
            N = rest.numel() // M
            k = min(M, N)
            res = eng.two_site([SplitItem(rest.reshape(M, N), k)])[0]
            tensors.append(res.a_out.reshape(chi, d, k))
            rest = res.b_out.reshape(k * N)
            chi = k
        tensors.append(rest.reshape(chi, d, 1))
        mps._tensors = tensors
        return mps

    def expectation(self, observable: MPSOperation) -> float:
        """core.py:723-751: <psi|O|psi> for a Hermitian MPSOperation."""
        if not observable.is_hermitian():
            raise ValueError("Observable is not Hermitian.")
        if observable.qudit_dimension != self._qudit_dimension:
            raise ValueError("Dimension mismatch between observable and MPS.")
        other = self.copy()
        other.apply(observable)
        return float(self.inner_product(other).real)

    def apply_one_qudit_gate_to_all(self, gate) -> None:
        for i in range(self._nqudits):
            self.apply_one_qudit_gate(gate, i)

    # ------------------------------------------------------------ two-site
    def apply_two_qudit_gate(self, gate, node_index1: int, node_index2: int, **kwargs) -> None:
        """core.py:950-1161."""
        if not self.is_valid():
            raise ValueError("MPS is not valid.")
        if node_index1 not in range(self._nqudits) or node_index2 not in range(self._nqudits):
            raise ValueError(f"Input tensor indices={(node_index1, node_index2)} are out of "
                             f"bounds for an MPS on {self._nqudits} qudits.")
        if node_index1 == node_index2:
            raise ValueError("Node indices cannot be identical.")
        nondangling = getattr(gate, "get_all_nondangling", None)
        g = as_tensor(gate)
        d = self._qudit_dimension
        if g.ndim != 4 or (callable(nondangling) and len(nondangling()) != 0):
            raise ValueError("Two qubit gate must have four free edges and zero connected edges.")
        if len(set(g.shape)) != 1:
            raise ValueError("All gate edges must have the same dimension.")
        if g.shape[0] != d:
            raise ValueError(f"Gate edges have dimension {g.shape[0]} "
                             f"but should have MPS qudit dimension = {d}")
        # core.py:1031-1033: swap the roles of the two sites; like the
        # reference, a Node passed in is reordered in place.
        if node_index2 < node_index1:
            if isinstance(gate, Node):
                gate.reorder_edges([gate[1], gate[0], gate[3], gate[2]])
                g = gate.tensor
            else:
                g = np.transpose(g, (1, 0, 3, 2))
            node_index1, node_index2 = node_index2, node_index1
        invert = False
        if node_index1 < node_index2 - 1:
            invert = True
            original = node_index1
            self.move_node_from_left_to_right(node_index1, node_index2 - 1, **kwargs)
            node_index1 = node_index2 - 1
        self._two_site(g.reshape(d * d, d * d), node_index1, kwargs)
        if invert:
            self.move_node_from_right_to_left(node_index1, original, **kwargs)
        if self._track_norms:
            self._norms.append(self.norm())

    def _two_site(self, gmat: np.ndarray, i: int, kwargs) -> None:
        maxsvals, keep_left = parse_split_options(kwargs, self.max_bond_dimension_of(i))
        a, b = self._tensors[i], self._tensors[i + 1]
        d = self._qudit_dimension
        M, N = d * a.shape[0], d * b.shape[2]
        k = min(M, N) if maxsvals is None else min(M, N, maxsvals)
        item = TwoSiteItem(a, b, self._engine.gate(gmat), k, absorb_left=not keep_left)
        res = self._engine.two_site([item])[0]
        self._tensors[i] = res.a_out
        self._tensors[i + 1] = res.b_out

    def move_node_from_left_to_right(self, current_node_index: int, final_node_index: int,
                                     **kwargs) -> None:
        """core.py:1163-1190."""
        if current_node_index > final_node_index:
            raise ValueError("current_node_index should be smaller than final_node_index.")
        if current_node_index < 0:
            raise ValueError("current_node_index out of range.")
        if final_node_index >= self._nqudits:
            raise ValueError("final_node_index out of range.")
        while current_node_index < final_node_index:
            self.swap(current_node_index, current_node_index + 1, **kwargs)
            current_node_index += 1

    def move_node_from_right_to_left(self, current_node_index: int, final_node_index: int,
                                     **kwargs) -> None:
        """core.py:1192-1219."""
        if current_node_index < final_node_index:
            raise ValueError("current_node_index should be larger than final_node_index.")
        if current_node_index > self._nqudits:
            raise ValueError("current_node_index out of range.")
        if final_node_index < 0:
            raise ValueError("final_node_index out of range.")
        while current_node_index > final_node_index:
            self.swap(current_node_index - 1, current_node_index, **kwargs)
            current_node_index -= 1

    # ------------------------------------------------------------ dispatch
    def apply(self, operations, **kwargs) -> None:
        """core.py:1221-1247."""
        try:
            operations = iter(operations)
        except TypeError:
            operations = (operations,)
        for op in operations:
            self._apply_mps_operation(op, **kwargs)

    def _apply_mps_operation(self, operation: MPSOperation, **kwargs) -> None:
        """core.py:1249-1276."""
        if not isinstance(operation, MPSOperation):
            raise TypeError("Argument operation should be of type MPSOperation but is "
                            f"of type {type(operation)}.")
        if not operation.is_valid():
            raise ValueError("Input MPS Operation is not valid.")
        if operation.is_single_qudit_operation():
            self.apply_one_qudit_gate(operation.node(), *operation.qudit_indices, **kwargs)
        elif operation.is_two_qudit_operation():
            self.apply_two_qudit_gate(operation.node(), *operation.qudit_indices, **kwargs)
        else:
            raise ValueError("Only one-qudit and two-qudit gates are supported. "
                             "To apply a gate on three or more qudits, the gate must be "
                             "compiled into a sequence of one- and two-qudit gates.")

    # ------------------------------------------------------------ wrappers
    def x(self, index: int) -> None:
        if index == -1:
            self.apply_one_qudit_gate_to_all(xgate())
        else:
            self.apply_one_qudit_gate(xgate(), index)

    def h(self, index: int) -> None:
        if index == -1:
            self.apply_one_qudit_gate_to_all(hgate())
        else:
            self.apply_one_qudit_gate(hgate(), index)

    def r(self, index, seed: Optional[int] = None, angle_scale: float = 1.0) -> None:
        if index == -1:
            for i in range(self._nqudits):
                self.apply_one_qudit_gate(rgate(seed, angle_scale), i)
        else:
            self.apply_one_qudit_gate(rgate(seed, angle_scale), index)

    def cnot(self, a: int, b: int, **kwargs) -> None:
        self.apply_two_qudit_gate(cnot(), a, b, **kwargs)

    def haar_random(self, qudit1_index: int, qudit2_index: int, **kwargs) -> None:
        gate = haar_random_unitary(nqudits=2, qudit_dimension=self._qudit_dimension)
        self.apply_two_qudit_gate(gate, qudit1_index, qudit2_index, **kwargs)

    def sweep_haar_random_left_to_right(self, **kwargs) -> None:
        for i in range(0, self._nqudits - 1, 2):
            self.haar_random(i, i + 1, keep_left_canonical=True, **kwargs)

    def sweep_haar_random_right_to_left(self, **kwargs) -> None:
        for i in range(self._nqudits - 2, 0, -2):
            self.haar_random(i - 1, i, keep_left_canonical=False, **kwargs)

    def sweep_cnots_left_to_right(self, **kwargs) -> None:
        for i in range(0, self._nqudits - 1, 2):
            self.cnot(i, i + 1, keep_left_canonical=True, **kwargs)

    def sweep_cnots_right_to_left(self, **kwargs) -> None:
        for i in range(self._nqudits - 2, 0, -2):
            self.cnot(i - 1, i, keep_left_canonical=False, **kwargs)

    def swap(self, a: int, b: int, **kwargs) -> None:
        if b < a:
            a, b = b, a
        self.apply_two_qudit_gate(swap(), a, b, **kwargs)

    # ------------------------------------------------------------ misc
    def copy(self) -> "MPS":
        return self.__copy__()

    def __copy__(self) -> "MPS":
        new = MPS(self._nqudits, self._qudit_dimension, self._prefix, device=self._engine.dev,
                  track_norms=self._track_norms, _empty=True)
        new._tensors = [t.clone() for t in self._tensors]
        new._norms = list(self._norms)
        return new

    def __eq__(self, other) -> bool:
        """core.py:1389-1418."""
        if not isinstance(other, MPS):
            return False
        if self is other:
            return True
        if not self.is_valid():
            raise ValueError("MPS is invalid and cannot be compared to another MPS.")
        if not other.is_valid():
            raise ValueError("Other MPS is invalid.")
        if other._qudit_dimension != self._qudit_dimension or other._nqudits != self._nqudits:
            return False
        for a, b in zip(self._tensors, other._tensors):
            if tuple(a.shape) != tuple(b.shape):
                return False
            if not np.allclose(a.cpu().numpy(), b.cpu().numpy()):
                return False
        return True

    __hash__ = None

    def __str__(self) -> str:
        return "----".join(f"{self._prefix}{i}{tuple(t.shape)}" for i, t in enumerate(self._tensors))
