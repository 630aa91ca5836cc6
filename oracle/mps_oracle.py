"""Plain-numpy restatement of mpsim's MPS gate-application path.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): used by tests/, by
__graft_entry__.smoke() as the checker and by bench.py's cpu_baseline leg.

Tensor layout differs from the reference on purpose: every site tensor is
[chi_left, d, chi_right] with dummy bonds of size 1 at the two chain ends
(the reference keeps TensorNetwork nodes with axes [phys, left, right] in the
interior and [phys, bond] at the ends, core.py:190-218). The axis order of a
site tensor is not observable through any gauge-invariant quantity; the
wavefunction, norms, bond dimensions and singular values are identical.

dtype modes
  "reference": numpy promotion exactly as the reference gets it. Tensors start
               complex64 (core.py:194,206,214); a float64 CNOT/SWAP
               (gates.py:195-212) or a complex128 Haar gate (gates.py:272-285)
               promotes every tensor it touches to complex128.
  "c64":       everything forced to complex64 (the lean CPU baseline).
  "c128":      everything forced to complex128 (the FP64 oracle).
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

# ---------------------------------------------------------------------------
# Gates (mpsim/gates.py)
# ---------------------------------------------------------------------------

# gates.py:104-110: one-qubit gate matrices are complex64.
# numpy>=2 would promote float64-scalar * complex64 to complex128; the
# reference (numpy 1.x value-based casting) gets complex64.
_H = (np.array([[1.0, 1.0], [1.0, -1.0]]) / np.sqrt(2)).astype(np.complex64)
_I = np.array([[1.0, 0.0], [0.0, 1.0]], dtype=np.complex64)
_X = np.array([[0.0, 1.0], [1.0, 0.0]], dtype=np.complex64)
_Y = np.array([[0.0, -1j], [1j, 0.0]], dtype=np.complex64)
_Z = np.array([[1.0, 0.0], [0.0, -1.0]], dtype=np.complex64)

# gates.py:195-212: CNOT and SWAP are float64 (no dtype given), reshaped (2,2,2,2).
_CNOT = np.array([[1.0, 0, 0, 0], [0, 1.0, 0, 0], [0, 0, 0, 1.0], [0, 0, 1.0, 0]])
_SWAP = np.array([[1.0, 0, 0, 0], [0, 0, 1.0, 0], [0, 1.0, 0, 0], [0, 0, 0, 1.0]])


def hgate():
    return _H.copy()


def xgate():
    return _X.copy()


def ygate():
    return _Y.copy()


def zgate():
    return _Z.copy()


def igate():
    return _I.copy()


def cnot():
    """CNOT as a (2,2,2,2) float64 tensor, gates.py:195-203,216-217."""
    return _CNOT.reshape(2, 2, 2, 2).copy()


def swap():
    """SWAP as a (2,2,2,2) float64 tensor, gates.py:204-212,220-221."""
    return _SWAP.reshape(2, 2, 2, 2).copy()


def cphase(exp: float):
    """diag(1,1,1,exp(2*pi*i*exp)) complex64, gates.py:224-232."""
    m = np.diag([1.0, 1.0, 1.0, np.exp(1j * 2 * np.pi * exp)]).astype(np.complex64)
    return m.reshape(2, 2, 2, 2)


def haar_unitary_from_rng(dim: int, rng: np.random.RandomState) -> np.ndarray:
    """Mezzadri's recipe as gates.py:269-284 performs it, on a caller's RNG.

    Draws randn(dim, dim, 2), forms (re + i*im)/sqrt(2), takes numpy's QR and
    multiplies column j of Q by R_jj/|R_jj|. Returns a complex128 matrix.
    """
    units = np.array([1, 1j])
    mat = np.sum(rng.randn(dim, dim, 2) * units, axis=-1) / np.sqrt(2)
    q, r = np.linalg.qr(mat)
    diag = np.diag(r).copy()
    diag /= np.abs(diag)
    return q * diag


def haar_random_unitary(nqudits: int = 2, qudit_dimension: int = 2,
                        seed: Optional[int] = None) -> np.ndarray:
    """gates.py:248-286: a fresh RandomState(seed) per call, tensor [d]*2n."""
    rng = np.random.RandomState(seed)
    dim = qudit_dimension ** nqudits
    return haar_unitary_from_rng(dim, rng).reshape([qudit_dimension] * 2 * nqudits)


def is_unitary(gate: np.ndarray) -> bool:
    """gates.py:15-33: reshape to square, allclose(U^H U, I, atol=1e-5)."""
    g = np.asarray(gate)
    if g.ndim > 2:
        if len(set(g.shape)) != 1:
            raise ValueError("Gate shape should be of the form (d, d, ..., d).")
        dim = int(np.sqrt(g.size))
        g = g.reshape(dim, dim)
    return bool(np.allclose(g.conj().T @ g, np.identity(g.shape[0]), atol=1e-5))


# ---------------------------------------------------------------------------
# TensorNetwork 0.2.1 SVD split, numpy backend (third-party; restated)
# ---------------------------------------------------------------------------

def tn_svd_split(theta: np.ndarray, max_singular_values: Optional[int]):
    """Semantics of tn.split_node_full_svd on the numpy backend as called at
    core.py:1132-1137 (no max_truncation_err): np.linalg.svd with
    full_matrices=False, keep k = min(max_singular_values or len(s), len(s))
    values including exact zeros, cast s to theta's dtype.

    Returns (u[:, :k], s[:k], vh[:k, :], s[k:]).
    """
    u, s, vh = np.linalg.svd(theta, full_matrices=False)
    keep = s.size if max_singular_values is None else int(max_singular_values)
    keep = min(keep, s.size)
    keep = max(keep, 0)
    s = s.astype(theta.dtype)
    return u[:, :keep], s[:keep], vh[:keep, :], s[keep:]


# ---------------------------------------------------------------------------
# MPS (mpsim/core.py)
# ---------------------------------------------------------------------------

def max_bond_dimensions(n: int, d: int) -> List[int]:
    """core.py:235-242: [d^1 .. d^(n//2)] + reversed, with the first d^(n//2)
    removed when n is even."""
    half = [d ** (i + 1) for i in range(n // 2)]
    out = half + list(reversed(half))
    if n % 2 == 0:
        out.remove(d ** (n // 2))
    return out


class OracleMPS:
    """CPU restatement of mpsim.MPS (core.py:159-1423) for the gate path."""

    def __init__(self, nqudits: int, qudit_dimension: int = 2,
                 dtype_mode: str = "reference", track_norms: bool = False):
        # core.py:184-187
        if nqudits < 2:
            raise ValueError(f"Number of qudits must be greater than 2 but is {nqudits}.")
        if dtype_mode not in ("reference", "c64", "c128"):
            raise ValueError(dtype_mode)
        d = qudit_dimension
        self.n = nqudits
        self.d = d
        self.mode = dtype_mode
        self.track_norms = track_norms
        base = np.complex128 if dtype_mode == "c128" else np.complex64
        # core.py:190-218: |0> on every site, complex64.
        self.tensors = []
        for _ in range(nqudits):
            t = np.zeros((1, d, 1), dtype=base)
            t[0, 0, 0] = 1.0
            self.tensors.append(t)
        self._max_bond = max_bond_dimensions(nqudits, d)
        self.norms: List[float] = []  # core.py:243, appended at core.py:1161
        self.discarded: List[float] = []
        self.spectra: List[np.ndarray] = []  # kept singular values of every update (FP64)

    # -- dtype handling ---------------------------------------------------
    def _cast_gate(self, g: np.ndarray) -> np.ndarray:
        if self.mode == "c64":
            return g.astype(np.complex64)
        if self.mode == "c128":
            return g.astype(np.complex128)
        return g

    def _cast_tensor(self, t: np.ndarray) -> np.ndarray:
        if self.mode == "c64":
            return t.astype(np.complex64)
        if self.mode == "c128":
            return t.astype(np.complex128)
        return t

    # -- bookkeeping ------------------------------------------------------
    def bond_dimensions(self) -> List[int]:
        return [t.shape[2] for t in self.tensors[:-1]]

    def max_bond_dimension_of(self, edge_index: int) -> int:
        """core.py:371-382 (negative indices count from the right)."""
        if edge_index >= self.n:
            raise ValueError("edge index out of range")
        return self._max_bond[edge_index]

    def max_bond_dimensions(self) -> List[int]:
        return self._max_bond

    # -- one-site ---------------------------------------------------------
    def apply_one_qudit_gate(self, gate: np.ndarray, i: int) -> None:
        """core.py:753-826 for unitary gates: gate edge 1 contracts with the
        site's physical leg, gate edge 0 becomes the new physical leg.
        A'[l, o, r] = sum_p G[o, p] A[l, p, r]."""
        if i not in range(self.n):
            raise ValueError("index out of bounds")
        g = self._cast_gate(np.asarray(gate))
        if g.shape != (self.d, self.d):
            raise ValueError("gate shape")
        a = self.tensors[i]
        new = np.einsum("op,lpr->lor", g, a)
        self.tensors[i] = self._cast_tensor(new)

    # -- two-site ---------------------------------------------------------
    def _split_options(self, i: int, j: int, kwargs) -> Tuple[Optional[int], bool]:
        """core.py:1104-1130."""
        keep_left = kwargs.get("keep_left_canonical", True) \
            if "keep_left_canonical" in kwargs else True
        if "fraction" in kwargs and "maxsvals" in kwargs:
            raise ValueError("Only one of (fraction, maxsvals) can be provided as kwargs.")
        maxsvals = None
        if "fraction" in kwargs:
            fraction = kwargs["fraction"]
            if not (0 <= fraction <= 1):
                raise ValueError("Keyword fraction must be between 0 and 1")
            maxsvals = int(round(fraction * self.max_bond_dimension_of(min(i, j))))
        if "maxsvals" in kwargs:
            maxsvals = int(kwargs["maxsvals"])
        return maxsvals, bool(keep_left)

    def apply_two_qudit_gate(self, gate: np.ndarray, i: int, j: int, **kwargs) -> None:
        """core.py:950-1161."""
        d = self.d
        if i not in range(self.n) or j not in range(self.n):
            raise ValueError("indices out of bounds")
        if i == j:
            raise ValueError("Node indices cannot be identical.")
        g = np.asarray(gate)
        if g.shape == (d * d, d * d):
            g = g.reshape(d, d, d, d)
        if g.shape != (d, d, d, d):
            raise ValueError("gate shape")
        # core.py:1031-1033: reorder edges [1,0,3,2] when j < i
        if j < i:
            g = g.transpose(1, 0, 3, 2)
            i, j = j, i
        # core.py:1035-1043: SWAP network for non-adjacent sites
        invert = False
        if i < j - 1:
            invert = True
            original = i
            self.move_node_from_left_to_right(i, j - 1, **kwargs)
            i = j - 1
        self._two_site(g, i, kwargs)
        # core.py:1154-1158
        if invert:
            self.move_node_from_right_to_left(i, original, **kwargs)
        # core.py:1160-1161 (benchmarking convenience in the reference)
        if self.track_norms:
            self.norms.append(self.norm())

    def _two_site(self, g: np.ndarray, i: int, kwargs) -> None:
        """core.py:1045-1152 on adjacent sites (i, i+1)."""
        maxsvals, keep_left = self._split_options(i, i + 1, kwargs)
        g = self._cast_gate(g)
        a, b = self.tensors[i], self.tensors[i + 1]
        chi_l, d, _ = a.shape
        _, _, chi_r = b.shape
        # core.py:1060: theta = A . B over the shared bond
        t = np.tensordot(a, b, axes=([2], [0]))            # [l, p1, p2, r]
        # core.py:1064-1068: gate edges 2,3 on the sites, 0,1 become new legs
        theta = np.einsum("abpq,lpqr->labr", g, t)        # [l, o1, o2, r]
        theta = self._cast_tensor(theta)
        m = theta.reshape(chi_l * d, d * chi_r)             # rows (l,o1), cols (o2,r)
        u, s, vh, rest = tn_svd_split(m, maxsvals)
        k = s.size
        self.discarded.append(float(np.sum(np.abs(rest) ** 2)))
        self.spectra.append(np.abs(s).astype(np.float64))
        # core.py:1139-1145: absorb S right (default) or left
        if keep_left:
            left = u
            right = s[:, None] * vh
        else:
            left = u * s[None, :]
            right = vh
        self.tensors[i] = self._cast_tensor(left.reshape(chi_l, d, k))
        self.tensors[i + 1] = self._cast_tensor(right.reshape(k, d, chi_r))

    def cnot(self, a: int, b: int, **kwargs) -> None:
        """core.py:1324-1328."""
        self.apply_two_qudit_gate(cnot(), a, b, **kwargs)

    def swap(self, a: int, b: int, **kwargs) -> None:
        """core.py:1376-1380."""
        if b < a:
            a, b = b, a
        self.apply_two_qudit_gate(swap(), a, b, **kwargs)

    def h(self, i: int) -> None:
        """core.py:1292-1303 (-1 applies to all)."""
        if i == -1:
            for k in range(self.n):
                self.apply_one_qudit_gate(hgate(), k)
        else:
            self.apply_one_qudit_gate(hgate(), i)

    def x(self, i: int) -> None:
        """core.py:1279-1290."""
        if i == -1:
            for k in range(self.n):
                self.apply_one_qudit_gate(xgate(), k)
        else:
            self.apply_one_qudit_gate(xgate(), i)

    def move_node_from_left_to_right(self, cur: int, final: int, **kwargs) -> None:
        """core.py:1163-1190."""
        if cur > final:
            raise ValueError("current_node_index should be smaller than final_node_index.")
        if cur < 0:
            raise ValueError("current_node_index out of range.")
        if final >= self.n:
            raise ValueError("final_node_index out of range.")
        while cur < final:
            self.swap(cur, cur + 1, **kwargs)
            cur += 1

    def move_node_from_right_to_left(self, cur: int, final: int, **kwargs) -> None:
        """core.py:1192-1219."""
        if cur < final:
            raise ValueError("current_node_index should be larger than final_node_index.")
        if cur > self.n:
            raise ValueError("current_node_index out of range.")
        if final < 0:
            raise ValueError("final_node_index out of range.")
        while cur > final:
            self.swap(cur - 1, cur, **kwargs)
            cur -= 1

    # -- read-out ---------------------------------------------------------
    def wavefunction(self) -> np.ndarray:
        """core.py:483-500: left-to-right contraction, site 0 most significant."""
        psi = self.tensors[0].reshape(self.d, -1)         # [p0, r]  (chi_l = 1)
        for t in self.tensors[1:]:
            cl, d, cr = t.shape
            psi = psi @ t.reshape(cl, d * cr)
            psi = psi.reshape(-1, cr)
        return psi.reshape(self.d ** self.n)

    def inner_product(self, other: "OracleMPS") -> complex:
        """core.py:507-561: <self|other> by a left-to-right transfer sweep.
        The reference conjugates `other` (core.py:545-546) and contracts it
        against `self`, so the value is sum conj(other) * self over the
        physical legs. This restatement keeps that orientation."""
        if other.n != self.n or other.d != self.d:
            raise ValueError("shape mismatch")
        env = np.ones((1, 1), dtype=np.complex128)        # [self bond, other bond]
        for a, b in zip(self.tensors, other.tensors):
            # env[x, y] a[x, p, x'] conj(b[y, p, y'])
            t = np.tensordot(env, a, axes=([0], [0]))      # [y, p, x']
            env = np.tensordot(t, np.conj(b), axes=([0, 1], [0, 1]))  # [x', y']
        return complex(env[0, 0])

    def norm(self) -> float:
        """core.py:563-565."""
        return float(np.sqrt(self.inner_product(self).real))

    def renormalize(self, to_norm: float = 1.0) -> None:
        """core.py:567-594: scale every tensor by (to_norm/norm)^(1/n)."""
        if to_norm < 0.0:
            raise ValueError(f"Arg to_norm must be positive but is {to_norm}")
        if np.isclose(to_norm, 0.0, atol=1e-15):
            raise ValueError("Arg to_norm is too close to numerical zero.")
        nrm = self.norm()
        if np.isclose(nrm, 0.0, atol=1e-15):
            raise ValueError("Norm of MPS is numerically zero, cannot renormalize.")
        f = (to_norm / nrm) ** (1.0 / self.n)
        self.tensors = [f * t for t in self.tensors]

    def reduced_density_matrix(self, node_indices) -> np.ndarray:
        """core.py:596-652. Validation as core.py:610-620; then a ket/bra
        sweep from the left (core.py:622-646): at each site the bra copy's
        tensor is the conjugate, the physical legs of traced sites are joined
        and those of kept sites stay open; the open legs are finally ordered
        ket legs then bra legs, each in the order node_indices gives
        (core.py:628-632, 648), and reshaped to (d^m, d^m) (core.py:649-651)."""
        try:
            idx = tuple(iter(node_indices))
        except TypeError:
            idx = (node_indices,)
        if len(set(idx)) < len(idx):
            raise ValueError("Node indices contains duplicates.")
        if min(idx) < 0 or max(idx) > self.n - 1:
            raise IndexError("One or more invalid node indices.")
        d = self.d
        # env axes: [open ket legs..., open bra legs..., ket bond, bra bond]
        env = np.ones((1, 1), dtype=np.complex128)
        opened: List[int] = []
        for i, a in enumerate(self.tensors):
            a = a.astype(np.complex128)
            m = len(opened)
            t = np.tensordot(env, a, axes=([2 * m], [0]))        # [..., ybond, s, x']
            if i in idx:
                t = np.tensordot(t, np.conj(a), axes=([2 * m], [0]))  # [.., s, x', t, y']
                # move s to the ket block end and t to the bra block end
                order = (list(range(m)) + [2 * m] + list(range(m, 2 * m)) + [2 * m + 2]
                         + [2 * m + 1, 2 * m + 3])
                env = np.transpose(t, order)
                opened.append(i)
            else:
                env = np.tensordot(t, np.conj(a), axes=([2 * m, 2 * m + 1], [0, 1]))
        m = len(opened)
        env = env.reshape([d] * (2 * m))
        pos = [opened.index(i) for i in idx]
        rho = np.transpose(env, pos + [m + p for p in pos])
        return rho.reshape(d ** m, d ** m)

    def marginals(self) -> np.ndarray:
        """Diagonal of every one-site reduced density matrix, the
        distribution core.py:667-668 draws each site from."""
        return np.stack([self.reduced_density_matrix(i).diagonal().real
                         for i in range(self.n)])

    # -- ops ------------------------------------------------------------------
    def apply_ops(self, ops: Sequence, **kwargs) -> None:
        """core.py:1221-1276: sequential dispatch by arity. Each op is
        (qudits tuple, matrix (d^k x d^k))."""
        for qudits, mat in ops:
            if len(qudits) == 1:
                self.apply_one_qudit_gate(np.asarray(mat).reshape(self.d, self.d), qudits[0])
            elif len(qudits) == 2:
                self.apply_two_qudit_gate(np.asarray(mat).reshape([self.d] * 4),
                                          qudits[0], qudits[1], **kwargs)
            else:
                raise ValueError("Only one-qudit and two-qudit gates are supported.")


def mps_overlap(a_tensors: Sequence[np.ndarray], b_tensors: Sequence[np.ndarray]) -> complex:
    """<a|b> = sum conj(a) b for two MPS given as [chi_l, d, chi_r] lists (FP64)."""
    env = np.ones((1, 1), dtype=np.complex128)
    for a, b in zip(a_tensors, b_tensors):
        a = a.astype(np.complex128)
        b = b.astype(np.complex128)
        t = np.tensordot(env, b, axes=([1], [0]))          # [xa, p, yb']
        env = np.tensordot(np.conj(a), t, axes=([0, 1], [0, 1]))  # [xa', yb']
    return complex(env[0, 0])


def gauge_invariant_error(ref_tensors, tensors) -> Tuple[float, float, float]:
    """Returns (1 - |<ref|x>|/(|ref||x|), |ref|, |x|) computed in FP64."""
    rr = mps_overlap(ref_tensors, ref_tensors).real
    xx = mps_overlap(tensors, tensors).real
    rx = abs(mps_overlap(ref_tensors, tensors))
    nr, nx = math.sqrt(max(rr, 0.0)), math.sqrt(max(xx, 0.0))
    if nr == 0.0 or nx == 0.0:
        return (0.0 if nr == nx else 1.0), nr, nx
    return 1.0 - rx / (nr * nx), nr, nx
