"""GPU tests of the QR-preconditioned Gram method (plan.method 3).

The preconditioner changes where the Jacobi starts (H = R R^H, Y = Q^H from
a Householder QR of theta with its columns ordered by norm), not what it
converges to: every factor must meet the same bars as the plain Gram
method's, against numpy's FP64 SVD of the same theta (core.py:1132-1145:
kept singular values, discarded weight, the best rank-k product, an
orthonormal unscaled factor), and circuits must agree with the FP64 oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _theta(d, A, B, G):
    cl, cr = A.shape[0], B.shape[2]
    return np.einsum("abpq,lpqr->labr", G.reshape(d, d, d, d).astype(np.complex128),
                     np.tensordot(A.astype(np.complex128), B.astype(np.complex128),
                                  axes=([2], [0]))).reshape(d * cl, d * cr)


def _check(theta, k, left, r, tag):
    u, s_ref, vh = np.linalg.svd(theta, full_matrices=False)
    fro = np.linalg.norm(theta)
    M, N = theta.shape
    a = r.a_out.cpu().numpy().reshape(M, k)
    b = r.b_out.cpu().numpy().reshape(k, N)
    s = r.s.cpu().numpy().astype(np.float64)
    assert np.all(np.abs(s - s_ref[:k]) <= 1e-4 * s_ref[:k] + 1e-6 * fro), tag
    assert abs(float(r.discarded.cpu()) - np.sum(s_ref[k:] ** 2)) <= 1e-5 * fro ** 2, tag
    best = (u[:, :k] * s_ref[:k]) @ vh[:k]
    assert np.linalg.norm(a.astype(np.complex128) @ b - best) <= 1e-4 * max(fro, 1e-30), tag
    unscaled = b @ b.conj().T if left else a.conj().T @ a
    assert np.allclose(unscaled, np.eye(k), atol=1e-5), tag
    return s


# (chi_l, chi_m, chi_r, d, k, absorb_left): square 512 (4 row chunks a
# wave), 500 rows (a partial last chunk), 600 and 1024 rows (8 chunks a wave;
# 1024 is the largest preconditioned), 1040 rows (beyond the limit: the plain
# Gram method in the same call), wide, tall with a partial last panel,
# rank-deficient (chi_m = 2), qutrits, absorb-left orientations
SHAPES = [(256, 256, 256, 2, 256, False), (250, 250, 250, 2, 200, False), (300, 300, 300, 2, 300, False),
          (512, 512, 512, 2, 512, False), (520, 520, 520, 2, 600, True),
          (60, 40, 200, 2, 100, False), (200, 40, 45, 2, 90, False),
          (100, 2, 100, 2, 50, False), (48, 48, 48, 3, 100, True),
          (45, 64, 200, 2, 90, True), (130, 130, 130, 2, 260, False)]


def test_preconditioned_two_site_vs_lapack_and_plain_gram():
    """One grouped call of the shapes above with every Gram problem
    preconditioned (precondition 2) against numpy's FP64 SVD, and its
    singular values against the plain Gram method's on the same inputs
    (both converge to the FP64 tolerance: within 2e-6 relative + 1e-7
    ||theta||_F)."""
    import torch
    from mpsim_amd import gates
    from mpsim_amd.engine import Engine, SvdConfig, TwoSiteItem
    eng = Engine.get()
    rng = np.random.RandomState(77)
    cases, items = [], []
    for (cl, cm, cr, d, k, left) in SHAPES:
        A = ((rng.randn(cl, d, cm) + 1j * rng.randn(cl, d, cm)) / np.sqrt(2 * d * cm)).astype(np.complex64)
        B = ((rng.randn(cm, d, cr) + 1j * rng.randn(cm, d, cr)) / np.sqrt(2)).astype(np.complex64)
        G = gates.haar_unitary(d * d, rng)
        cases.append((cl, cm, cr, d, k, left, A, B, G))
        items.append(TwoSiteItem(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), eng.gate(G), k,
                                 absorb_left=left, want_s=True))
    base = eng.svd
    pre = SvdConfig(block_switch=base.block_switch, inner_sweeps=base.inner_sweeps, precondition=2)
    plain = SvdConfig(block_switch=base.block_switch, inner_sweeps=base.inner_sweeps)
    res = eng.two_site(items, pre)
    torch.cuda.synchronize()
    ref = eng.two_site(items, plain)
    torch.cuda.synchronize()
    for (cl, cm, cr, d, k, left, A, B, G), r, r0 in zip(cases, res, ref):
        tag = (cl, cm, cr, d, k, left)
        print("case", tag, flush=True)
        theta = _theta(d, A, B, G)
        s = _check(theta, k, left, r, tag)
        s0 = r0.s.cpu().numpy().astype(np.float64)
        assert np.all(np.abs(s - s0) <= 2e-6 * s0 + 1e-7 * np.linalg.norm(theta)), tag


def test_preconditioned_split_with_zero_rows_and_columns():
    """Plain splits through the preconditioned method: exact zero columns
    (norm order ties, zero reflectors), a zero row block, a rank-one
    matrix, and an all-zero matrix (the unscaled factor must stay
    orthonormal, the kept values zero)."""
    import torch
    from mpsim_amd.engine import Engine, SplitItem, SvdConfig
    eng = Engine.get()
    rng = np.random.RandomState(5)
    mats = []
    m = (rng.randn(300, 200) + 1j * rng.randn(300, 200)).astype(np.complex64)
    m[:, 17:60] = 0
    m[250:, :] = 0
    mats.append((m, 150))
    u = rng.randn(256, 1) + 1j * rng.randn(256, 1)
    v = rng.randn(1, 256) + 1j * rng.randn(1, 256)
    mats.append(((u @ v).astype(np.complex64), 40))
    mats.append((np.zeros((128, 96), np.complex64), 32))
    items = [SplitItem(torch.from_numpy(x).cuda(), k, want_s=True) for x, k in mats]
    base = eng.svd
    cfg = SvdConfig(block_switch=base.block_switch, inner_sweeps=base.inner_sweeps, precondition=2)
    res = eng.two_site(items, cfg)
    torch.cuda.synchronize()
    for (x, k), r in zip(mats, res):
        print("split", x.shape, k, flush=True)
        _check(x.astype(np.complex128), k, False, r, x.shape)


def test_preconditioned_circuits_vs_oracle():
    """Brickwork circuits whose bonds saturate at the cap (the layers
    precondition 1 selects) run batched with precondition 1 and agree with
    the FP64 oracle (1 - overlap <= 1e-4) and bond for bond."""
    from oracle.mps_oracle import OracleMPS, gauge_invariant_error
    from mpsim_amd.batch import prepare, run_batched
    from mpsim_amd.circuits import brickwork_haar
    from mpsim_amd.engine import Engine, SvdConfig
    eng = Engine.get()
    base = eng.svd
    cfg = SvdConfig(block_switch=base.block_switch, inner_sweeps=base.inner_sweeps, precondition=1)
    circs = [brickwork_haar(16, 12, 2, maxsvals=64, seed=3), brickwork_haar(12, 10, 2, maxsvals=48, seed=4)]
    dcs = [prepare(c) for c in circs]
    run_batched(dcs, svd=cfg)
    for c, dc in zip(circs, dcs):
        ora = OracleMPS(c.n, c.d, dtype_mode="c128")
        ora.apply_ops(c.ops, maxsvals=c.maxsvals)
        assert dc.mps.bond_dimensions() == ora.bond_dimensions()
        err, nr, nx = gauge_invariant_error(ora.tensors, [t.cpu().numpy() for t in dc.mps.tensors])
        print(c.name, "err", err, flush=True)
        assert err <= 1e-4, (c.name, err)
        assert abs(nr - nx) <= 1e-4 * nr


def test_preconditioned_result_independent_of_batch():
    """A problem's factors are bitwise the same alone and in a call with a
    larger preconditioned problem (each runs the QR variant of its own size:
    up to 512 rows, or above), so served and batched results do not depend
    on the mix."""
    import torch
    from mpsim_amd import gates
    from mpsim_amd.engine import Engine, SvdConfig, TwoSiteItem
    eng = Engine.get()
    rng = np.random.RandomState(91)

    def item(cl, cm, cr, k):
        A = ((rng.randn(cl, 2, cm) + 1j * rng.randn(cl, 2, cm)) / np.sqrt(4 * cm)).astype(np.complex64)
        B = ((rng.randn(cm, 2, cr) + 1j * rng.randn(cm, 2, cr)) / np.sqrt(2)).astype(np.complex64)
        return TwoSiteItem(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(),
                           eng.gate(gates.haar_unitary(4, rng)), k, want_s=True)

    small, large = item(200, 200, 200, 150), item(480, 480, 480, 300)
    cfg = SvdConfig(block_switch=eng.svd.block_switch, inner_sweeps=eng.svd.inner_sweeps, precondition=1)
    alone = eng.two_site([small], cfg)[0]
    torch.cuda.synchronize()
    mixed = eng.two_site([large, small], cfg)[1]
    torch.cuda.synchronize()
    for x, y in ((alone.a_out, mixed.a_out), (alone.b_out, mixed.b_out), (alone.s, mixed.s)):
        assert np.array_equal(x.cpu().numpy(), y.cpu().numpy())
