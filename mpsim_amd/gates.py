"""Gate library with the reference's names, matrices and dtypes.

Mirrors mpsim/gates.py. Gates are returned as `Node` objects: a small
stand-in for the tensornetwork.Node the reference uses, exposing what the
gate-application API reads (`tensor`, `edges`, `get_edge(i).dimension`,
`reorder_edges`, dangling-edge queries, `name`). Anything with a `.tensor`
attribute (a real tensornetwork.Node included) or a plain array is accepted
wherever a gate is expected.
"""
from __future__ import annotations

from copy import deepcopy
from typing import List, Optional, Sequence, Union

import numpy as np


class Edge:
    """A dangling edge of a gate node."""

    __slots__ = ("node1", "axis", "name")

    def __init__(self, node: "Node", axis: int):
        self.node1 = node
        self.axis = axis
        self.name = str(axis)

    @property
    def dimension(self) -> int:
        return int(self.node1.tensor.shape[self.axis])

    def is_dangling(self) -> bool:
        return True

    def __repr__(self) -> str:
        return f"Edge({self.node1.name!r}[{self.axis}], dim={self.dimension})"


class Node:
    """Free tensor node (all edges dangling), as gates are in the reference."""

    def __init__(self, tensor, name: Optional[str] = None):
        self.tensor = np.asarray(tensor)
        self.name = name if name is not None else "__unnamed_node__"
        self.edges = [Edge(self, i) for i in range(self.tensor.ndim)]

    def get_edge(self, i: int) -> Edge:
        return self.edges[i]

    def __getitem__(self, i: int) -> Edge:
        return self.edges[i]

    def get_all_edges(self) -> List[Edge]:
        return list(self.edges)

    def get_all_dangling(self) -> List[Edge]:
        return list(self.edges)

    def get_all_nondangling(self) -> List[Edge]:
        return []

    def has_nondangling_edge(self) -> bool:
        return False

    @property
    def shape(self):
        return self.tensor.shape

    def reorder_edges(self, edge_order: Sequence[Edge]) -> "Node":
        """Permutes the tensor axes to the given edge order, in place (the
        reference mutates gate nodes this way, core.py:1031-1033)."""
        perm = [e.axis for e in edge_order]
        if sorted(perm) != list(range(self.tensor.ndim)):
            raise ValueError("edge_order must be a permutation of the node's edges")
        self.tensor = np.transpose(self.tensor, perm)
        self.edges = [Edge(self, i) for i in range(self.tensor.ndim)]
        return self

    def copy(self) -> "Node":
        return Node(self.tensor.copy(), self.name)

    def __str__(self) -> str:
        return self.name

    def __repr__(self) -> str:
        return f"Node({self.name!r}, shape={self.tensor.shape})"


def as_tensor(gate) -> np.ndarray:
    """The tensor of a Node / tensornetwork.Node / array-like."""
    t = getattr(gate, "tensor", gate)
    return np.asarray(t)


# ------------------------------------------------------------------ checks
def _square(gate) -> np.ndarray:
    if not isinstance(gate, (np.ndarray, Node)) and not hasattr(gate, "tensor"):
        raise TypeError("Invalid type for gate.")
    g = as_tensor(gate)
    if g.ndim > 2:
        if len(set(g.shape)) != 1:
            raise ValueError("Gate shape should be of the form (d, d, ..., d).")
        dim = int(np.sqrt(g.size))
        g = g.reshape(dim, dim)
    return g


def is_unitary(gate) -> bool:
    """gates.py:15-33: allclose(U^H U, I, atol=1e-5)."""
    g = _square(gate)
    return bool(np.allclose(g.conj().T @ g, np.identity(g.shape[0]), atol=1e-5))


def is_hermitian(gate) -> bool:
    """gates.py:36-52."""
    g = _square(gate)
    return bool(np.allclose(g.conj().T, g, atol=1e-5))


def is_projector(gate) -> bool:
    """gates.py:55-69: rank one."""
    g = _square(gate)
    return bool(np.linalg.matrix_rank(g) == 1)


# ------------------------------------------------------------------ states
zero_state = np.array([1.0, 0.0], dtype=np.complex64)
one_state = np.array([0.0, 1.0], dtype=np.complex64)
# gates.py:74 under numpy 1.x casting rules (complex64); numpy>=2 would
# promote the float64 scalar product to complex128.
plus_state = ((zero_state + one_state) / np.sqrt(2)).astype(np.complex64)


def computational_basis_state(state: int, dim: int = 2) -> Node:
    """gates.py:77-100."""
    if state < 0:
        raise ValueError(f"Argument state should be positive but is {state}.")
    if dim < 0:
        raise ValueError(f"Argument dim should be positive but is {dim}.")
    if state >= dim:
        raise ValueError(f"Requires state < dim but state = {state} and dim = {dim}.")
    v = np.zeros((dim,))
    v[state] = 1.0
    return Node(v, name=f"|{state}>")


# ------------------------------------------------------------------ gates
_hmatrix = (np.array([[1.0, 1.0], [1.0, -1.0]]) / np.sqrt(2)).astype(np.complex64)
_imatrix = np.array([[1.0, 0.0], [0.0, 1.0]], dtype=np.complex64)
_xmatrix = np.array([[0.0, 1.0], [1.0, 0.0]], dtype=np.complex64)
_ymatrix = np.array([[0.0, -1j], [1j, 0.0]], dtype=np.complex64)
_zmatrix = np.array([[1.0, 0.0], [0.0, -1.0]], dtype=np.complex64)
_cnot_matrix = np.array([[1.0, 0, 0, 0], [0, 1.0, 0, 0], [0, 0, 0, 1.0], [0, 0, 1.0, 0]]).reshape(2, 2, 2, 2)
_swap_matrix = np.array([[1.0, 0, 0, 0], [0, 0, 1.0, 0], [0, 1.0, 0, 0], [0, 0, 0, 1.0]]).reshape(2, 2, 2, 2)


def igate() -> Node:
    return Node(deepcopy(_imatrix), name="igate")


def xgate() -> Node:
    return Node(deepcopy(_xmatrix), name="xgate")


def ygate() -> Node:
    return Node(deepcopy(_ymatrix), name="ygate")


def zgate() -> Node:
    return Node(deepcopy(_zmatrix), name="zmat")


def hgate() -> Node:
    return Node(deepcopy(_hmatrix), name="hgate")


def rgate(seed: Optional[int] = None, angle_scale: float = 1.0) -> Node:
    """gates.py:141-164, including its quirks: `if seed:` ignores seed=0 and
    the generator multiplies m_y*Y*m_z*Z elementwise (which is zero) instead
    of adding, so the gate is expm(-i theta m_x X)."""
    from scipy.linalg import expm
    if seed:
        np.random.seed(seed)
    theta, alpha, phi = np.random.rand(3) * 2 * np.pi
    mx = np.sin(alpha) * np.cos(phi)
    my = np.sin(alpha) * np.sin(phi)
    mz = np.cos(alpha)
    theta *= angle_scale
    unitary = expm(-1j * theta * (mx * _xmatrix + my * _ymatrix * mz * _zmatrix))
    return Node(unitary)


def computational_basis_projector(state: int, dim: int = 2) -> Node:
    """gates.py:168-191."""
    if state < 0:
        raise ValueError(f"Argument state should be positive but is {state}.")
    if dim < 0:
        raise ValueError(f"Argument dim should be positive but is {dim}.")
    if state >= dim:
        raise ValueError(f"Requires state < dim but state = {state} and dim = {dim}.")
    p = np.zeros((dim, dim))
    p[state, state] = 1.0
    return Node(p, name=f"|{state}><{state}|")


def cnot() -> Node:
    """float64 (2,2,2,2), gates.py:195-203,216-217."""
    return Node(deepcopy(_cnot_matrix), name="cnot")


def swap() -> Node:
    """float64 (2,2,2,2), gates.py:204-212,220-221."""
    return Node(deepcopy(_swap_matrix), name="swap")


def cphase(exp: float) -> Node:
    """gates.py:224-232."""
    m = np.array([[1.0, 0, 0, 0], [0, 1.0, 0, 0], [0, 0, 1.0, 0],
                  [0, 0, 0, np.exp(1j * 2 * np.pi * exp)]], dtype=np.complex64)
    return Node(m.reshape(2, 2, 2, 2), name="cphase")


def random_two_qubit_gate(seed: Optional[int] = None) -> Node:
    """gates.py:235-245 (scipy's unitary_group)."""
    from scipy.stats import unitary_group
    if seed:
        np.random.seed(seed)
    u = unitary_group.rvs(dim=4)
    return Node(deepcopy(u.reshape(2, 2, 2, 2)), name="R2Q")


def haar_unitary(dim: int, rng: np.random.RandomState) -> np.ndarray:
    """Haar-random dim x dim unitary by Mezzadri's recipe (math-ph/0609050),
    drawing from `rng` in the order mpsim/gates.py:269-284 does: randn(dim,
    dim, 2) -> (re + i im)/sqrt(2) -> QR -> scale column j by R_jj/|R_jj|."""
    g = rng.randn(dim, dim, 2)
    z = (g[..., 0] + 1j * g[..., 1]) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    ph = np.diagonal(r) / np.abs(np.diagonal(r))
    return q * ph[None, :]


def haar_random_unitary(nqudits: int = 2, qudit_dimension: int = 2, name: str = "Haar",
                        seed: Optional[int] = None) -> Node:
    """gates.py:248-286: fresh RandomState(seed), tensor of shape [d]*2n."""
    rng = np.random.RandomState(seed)
    u = haar_unitary(qudit_dimension ** nqudits, rng)
    return Node(u.reshape([qudit_dimension] * 2 * nqudits), name=name)
