"""Host restatement of the QR preconditioning scheme of csrc/qr.hip (no GPU).

The reference has no counterpart (its split calls LAPACK's gesdd through
TensorNetwork, core.py:1132-1137); the GPU tests check the kernels' factors
against LAPACK's SVD. Here the same steps run in numpy on small problems so
that the formulas themselves are pinned by what the mathematics fixes:

* column order: norms from 32-row parts, decreasing, ties by index, NaN last;
* reflectors in the kernels' scaled form: the unscaled column x, w = x below
  the pivot row q and 1/sc at q (v = sc w), g' = w^H a,
  a -= tau |sc|^2 g' w, v_m^H v_j = conj(sc_m) sc_j conj(g'_m);
* T by the recurrence T[k][j] = -tau_j sum_{m<j} T[k][m] v_m^H v_j, checked
  against the explicit product of the sixteen reflectors;
* the block update A -= V (T^H (V^H A)), with the strips of Y columns taking
  four panels at a time;
* results: Q^H unitary, Q^H X P = R, R exactly zero left of its diagonal (what
  the Gram product's block skip relies on), |diag R| equal to numpy's QR.
"""
import numpy as np
import pytest

DEFER = 4


def column_order(X):
    """k_qr_colnorm + k_qr_order: squared norms summed over 32-row parts,
    NaN as -1, rank by (norm descending, index ascending)."""
    C, R = X.shape
    parts = [np.sum(np.abs(X[r0:r0 + 32]) ** 2, axis=0) for r0 in range(0, C, 32)]
    nrm = np.sum(parts, axis=0) if parts else np.zeros(R)
    nrm = np.where(np.isnan(nrm), -1.0, nrm)
    order = np.empty(R, dtype=int)
    for c in range(R):
        rank = int(np.sum((nrm > nrm[c]) | ((nrm == nrm[c]) & (np.arange(R) < c))))
        order[rank] = c
    return order


def factor_panel(S, rf, ncols):
    """Factors the strip S (rows 0..C-1, its 16 columns) at rows rf.. in the
    kernels' scaled form. Returns (strip with beta on the diagonal and zeros
    below, V rows 0..C-1, taus, Ys with Ys[j][m] = v_m^H v_j)."""
    C = S.shape[0]
    a = S.copy()
    taus = np.zeros(16)
    scs = np.ones(16, complex)
    betas = np.zeros(16, complex)
    Ys = np.zeros((16, 16), complex)
    for j in range(ncols):
        q = rf + j
        x = a[:, j]
        s2 = float(np.sum(np.abs(x[q + 1:]) ** 2))
        al = x[q]
        tau, beta, sc, isc = 0.0, al, 1.0 + 0j, 1.0 + 0j
        if s2 > 0.0:
            aa = abs(al)
            xn = np.sqrt(aa * aa + s2)
            u = al / aa if aa > 0.0 else 1.0 + 0j
            beta = -u * xn
            tau = 1.0 + aa / xn
            sc = np.conj(u) / (aa + xn)
            isc = u * (aa + xn)
        taus[j], scs[j], betas[j] = tau, sc, beta
        w = np.zeros(C, complex)
        w[q + 1:] = x[q + 1:]
        w[q] = isc
        g = w.conj() @ a  # g'_k for every column of the strip
        for k in range(16):
            if k > j:
                a[:, k] -= tau * abs(sc) ** 2 * g[k] * w
            elif k < j:
                Ys[j, k] = np.conj(scs[k]) * sc * np.conj(g[k])
    V = np.zeros((C, 16), complex)
    out = a.copy()
    for j in range(ncols):
        q = rf + j
        V[q, j] = 1.0
        V[q + 1:, j] = scs[j] * a[q + 1:, j]
        out[q, j] = betas[j]
        out[q + 1:, j] = 0.0
    return out, V, taus, Ys


def build_t(taus, Ys, ncols):
    """build_t: T[k][k] = tau_k, T[k][j] = -tau_j sum_{m<j} T[k][m] Ys[j][m]."""
    T = np.zeros((16, 16), complex)
    for k in range(16):
        for j in range(16):
            if j == k:
                T[k, j] = taus[j] if j < ncols else 0.0
            elif k < j < ncols:
                T[k, j] = -taus[j] * sum(T[k, m] * Ys[j, m] for m in range(j))
    return T


def block_qr(X, defer=DEFER):
    """[X P | I] -> [R | Q^H] by 16-column panels as the step launches run
    them: launch `panel` applies panel `panel` to every strip right of it
    (strips of Y columns only: panels k0..panel at every `defer`-th launch
    and the last) and factors panel + 1 in strip 0."""
    C, R = X.shape
    order = column_order(X)
    ncol = R + C
    width = -(-ncol // 16) * 16
    Z = np.zeros((C, width), complex)
    Z[:, :R] = X[:, order]
    Z[:, R:R + C] = np.eye(C)
    nref = min(C - 1, R)
    npan = -(-nref // 16)
    V, T = {}, {}
    for panel in range(-1, npan):
        if panel >= 0:
            for c0 in range(16 * (panel + 1), ncol, 16):
                ystrip = c0 >= R
                if ystrip and panel % defer != defer - 1 and panel != npan - 1:
                    continue
                k0 = panel - panel % defer if ystrip else panel
                for k in range(k0, panel + 1):
                    A = Z[:, c0:c0 + 16]
                    W = V[k].conj().T @ A
                    Z[:, c0:c0 + 16] = A - V[k] @ (T[k].conj().T @ W)
        if panel + 1 < npan:
            rf = 16 * (panel + 1)
            ncols = min(16, nref - rf)
            strip, v, taus, Ys = factor_panel(Z[:, rf:rf + 16], rf, ncols)
            Z[:, rf:rf + 16] = strip
            V[panel + 1], T[panel + 1] = v, build_t(taus, Ys, ncols)
    return Z[:, :R], Z[:, R:R + C], order, V, T


def _matrix(rng, C, R):
    return rng.standard_normal((C, R)) + 1j * rng.standard_normal((C, R))


@pytest.mark.parametrize("C,R", [(64, 64), (100, 37), (37, 100), (96, 200), (200, 40), (81, 81),
                                 (150, 17), (17, 150), (33, 33), (2, 5), (1, 4)])
def test_block_qr_factors(C, R):
    rng = np.random.default_rng(C * 1000 + R)
    X = _matrix(rng, C, R)
    X[:, 3 % R] = 0.0  # a zero column: ties in the order, tau = 0 at its step
    Rm, Qh, order, _, _ = block_qr(X)
    P = X[:, order]
    assert np.abs(Qh @ Qh.conj().T - np.eye(C)).max() <= 1e-12
    assert np.abs(Qh @ P - Rm).max() <= 1e-12 * np.abs(X).max() * max(C, R)
    # exactly zero left of the diagonal (row i, columns < i)
    for i in range(C):
        assert np.all(Rm[i, :min(i, R)] == 0.0), i
    n = min(C, R)
    ref = np.linalg.qr(P)[1]
    assert np.allclose(np.abs(np.diag(Rm)[:n]), np.abs(np.diag(ref)[:n]), rtol=1e-10, atol=1e-12)
    # the preconditioned Gram matrix is the original one rotated by Q
    G = X @ X.conj().T
    assert np.abs(Rm @ Rm.conj().T - Qh @ G @ Qh.conj().T).max() <= 1e-12 * np.abs(G).max() * C


def test_panel_is_the_product_of_its_reflectors():
    """I - V T V^H equals H_0 H_1 ... H_15, H_j = I - tau_j v_j v_j^H, each
    H_j Hermitian and unitary, and its conjugate transpose maps the strip to
    the factored one (beta on the diagonal, zeros below)."""
    rng = np.random.default_rng(3)
    C = 60
    S = _matrix(rng, C, 16)
    strip, V, taus, Ys = factor_panel(S, 0, 16)
    T = build_t(taus, Ys, 16)
    prod = np.eye(C, dtype=complex)
    for j in range(16):
        H = np.eye(C) - taus[j] * np.outer(V[:, j], V[:, j].conj())
        assert np.abs(H @ H.conj().T - np.eye(C)).max() <= 1e-13
        assert np.abs(H - H.conj().T).max() <= 1e-15
        prod = prod @ H
    assert np.abs(prod - (np.eye(C) - V @ T @ V.conj().T)).max() <= 1e-13
    # Ys is the Gram matrix of the reflectors
    G = V.conj().T @ V
    for j in range(16):
        for m in range(j):
            assert abs(Ys[j, m] - G[m, j]) <= 1e-12 * abs(G[m, j]) + 1e-13
    # Q_p^H applied to the strip gives the factored strip
    assert np.abs((np.eye(C) - V @ T.conj().T @ V.conj().T) @ S - strip).max() <= 1e-12 * np.abs(S).max()


def test_deferred_y_strips_change_nothing():
    """Taking the Y columns' reflectors four panels at a time gives the
    factors of one panel at a time (the same updates in the same order)."""
    rng = np.random.default_rng(9)
    X = _matrix(rng, 150, 120)
    R1, Q1 = block_qr(X, defer=4)[:2]
    R2, Q2 = block_qr(X, defer=1)[:2]
    assert np.abs(R1 - R2).max() <= 1e-12 * np.abs(R2).max()
    assert np.abs(Q1 - Q2).max() <= 1e-13


def test_rank_deficient_and_zero_matrices():
    """A rank-one matrix (every reflector after the first has a zero column
    below its pivot, tau = 0) and an all-zero one: Q^H stays unitary, R
    upper trapezoidal, one nonzero row for the rank-one matrix."""
    rng = np.random.default_rng(4)
    u = rng.standard_normal((70, 1)) + 1j * rng.standard_normal((70, 1))
    v = rng.standard_normal((1, 50)) + 1j * rng.standard_normal((1, 50))
    for X in (u @ v, np.zeros((40, 30), complex)):
        Rm, Qh, order, _, _ = block_qr(X)
        assert np.abs(Qh @ Qh.conj().T - np.eye(X.shape[0])).max() <= 1e-12
        assert np.abs(Qh @ X[:, order] - Rm).max() <= 1e-12 * max(1.0, np.abs(X).max()) * 70
        assert np.abs(Rm[1:]).max() <= 1e-12 * max(1.0, np.abs(X).max()) * 70


def test_column_order_ties_and_nan():
    """Norm descending, ties by index ascending, a NaN norm last; the order
    is a permutation whatever the data."""
    X = np.zeros((70, 5), complex)
    X[:, 0] = 1.0
    X[0, 1] = np.nan
    X[:, 2] = 2.0
    X[:, 3] = 2.0
    order = column_order(X)
    assert list(order) == [2, 3, 0, 4, 1]
    assert sorted(order) == list(range(5))


def test_checks_catch_planted_mistakes():
    """The checks above fail for plausible slips in the scheme: T built from
    conjugated v_m^H v_j, the block update with T instead of T^H, and a
    reflector scale without its conjugate."""
    rng = np.random.default_rng(5)
    C = 48
    S = _matrix(rng, C, 16)
    strip, V, taus, Ys = factor_panel(S, 0, 16)
    prod = np.eye(C, dtype=complex)
    for j in range(16):
        prod = prod @ (np.eye(C) - taus[j] * np.outer(V[:, j], V[:, j].conj()))
    bad_t = build_t(taus, Ys.conj(), 16)
    assert np.abs(prod - (np.eye(C) - V @ bad_t @ V.conj().T)).max() > 1e-3
    T = build_t(taus, Ys, 16)
    assert np.abs((np.eye(C) - V @ T @ V.conj().T) @ S - strip).max() > 1e-3
    # sc = u / (|x_q| + ||x||) instead of conj(u) / (...): v is wrong
    x = S[:, 0]
    aa, xn = abs(x[0]), np.linalg.norm(x)
    u = x[0] / aa
    v = np.concatenate([[1.0], x[1:] * u / (aa + xn)])
    H = np.eye(C) - (1.0 + aa / xn) * np.outer(v, v.conj())
    assert np.abs((H @ x)[1:]).max() > 1e-3
    v_good = np.concatenate([[1.0], x[1:] * np.conj(u) / (aa + xn)])
    H_good = np.eye(C) - (1.0 + aa / xn) * np.outer(v_good, v_good.conj())
    assert np.abs((H_good @ x)[1:]).max() <= 1e-13 * xn
    assert abs((H_good @ x)[0] - (-u * xn)) <= 1e-13 * xn


def gram_first_block(wave, bj):
    """k_hgram's first 16-column block for wave `wave` (blocks wave*16,
    wave*16 + 64, ...) on tile (I, bj) of a QR-preconditioned problem."""
    cs = wave * 16
    if 16 * bj >= cs + 16:
        cs += 64 * ((16 * bj - 16 - cs) // 64 + 1)
    return cs


@pytest.mark.parametrize("R", [16, 37, 90, 256, 500])
def test_gram_block_skip(R):
    """The Gram product of a preconditioned problem starts each wave at its
    first block that is not left of row block bj: every block skipped is
    zero in R's rows 16 bj .., and none that is not is skipped, so the tile
    sums exactly the nonzero terms the full loop would."""
    rng = np.random.default_rng(R)
    C = 96
    Rm = block_qr(_matrix(rng, C, R))[0]
    for bj in range(C // 16):
        rows = Rm[16 * bj:16 * bj + 16]
        for wave in range(4):
            cs = gram_first_block(wave, bj)
            seq = list(range(wave * 16, R, 64))
            skipped = [c0 for c0 in seq if c0 < cs]
            assert all(c0 + 16 <= 16 * bj for c0 in skipped)
            assert all(np.all(rows[:, c0:c0 + 16] == 0.0) for c0 in skipped)
            assert cs in seq[len(skipped):len(skipped) + 1] or cs >= R
            nxt = [c0 for c0 in seq if c0 >= cs]
            assert not nxt or nxt[0] + 16 > 16 * bj
