"""Per-value accuracy of the two Jacobi SVD methods against LAPACK in FP64.

theta = U diag(sigma) V^H with Haar-random U, V and a geometric spectrum
sigma_i = 10^(-6 i / (n - 1)) (1 down to 1e-6), stored complex64 and
factorised as a plain split (the same SVD every two-site update runs). The
reference values are np.linalg.svd of exactly that complex64 matrix in
FP64, so the comparison isolates the SVD.

Bounds (|s_gpu - s_ref| per kept value; ||theta||_F = sqrt(sum sigma^2)):
  one-sided rows (method 1): 1e-4 * sigma_i on every value. Rotating the rows
      in FP64 to tol 1e-7 gives relative accuracy per value.
  Gram method (method 0, the default above 64 rows): 1e-4 * sigma_i +
      2e-6 * ||theta||_F. It rotates H = X X^H, whose FP64 entries carry
      rounding of 1e-16..1e-14 ||theta||^2 and whose pairs below
      1e-12 ||theta||^2 are not rotated, so sigma^2 is known to about
      1e-12 ||theta||^2 absolute: sigma_i below ~1e-2 ||theta|| lose relative
      accuracy. The truncation is unaffected (discarded weight to 1e-10
      ||theta||^2), and so is the state: the factors are X~ = Y X with Y
      unitary to FP64.
  QR-preconditioned Gram method (precondition 2): the same bound. Its H =
      R R^H = Q^H (X X^H) Q carries the same rounding, and X~ = (W Q^H) X0
      with W Q^H unitary to FP64.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _haar(n, rng):
    z = (rng.randn(n, n) + 1j * rng.randn(n, n)) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


@pytest.mark.parametrize("n", [128, 512])
def test_graded_spectrum_per_value(n):
    import dataclasses
    import torch
    from mpsim_amd.engine import Engine, SplitItem
    rng = np.random.RandomState(1000 + n)
    sigma = 10.0 ** (-6.0 * np.arange(n) / (n - 1))
    theta = ((_haar(n, rng) * sigma) @ _haar(n, rng).conj().T).astype(np.complex64)
    t64 = theta.astype(np.complex128)
    s_ref = np.linalg.svd(t64, compute_uv=False)
    fro = float(np.linalg.norm(t64))
    eng = Engine.get()
    dev = torch.from_numpy(theta).cuda()
    for method, pre, floor in ((1, 0, 0.0), (0, 0, 2e-6), (0, 2, 2e-6)):
        for k in (n, n // 2):
            r = eng.two_site([SplitItem(dev, k, want_s=True)],
                             dataclasses.replace(eng.svd, method=method, precondition=pre))[0]
            s = r.s.cpu().numpy().astype(np.float64)
            err = np.abs(s - s_ref[:k])
            big = s_ref[:k] >= 1e-2 * fro
            print(f"n={n} method={method} precondition={pre} k={k}: max rel err {np.max(err / s_ref[:k]):.2e} "
                  f"(sigma >= 1e-2 |theta|: {np.max(err[big] / s_ref[:k][big]):.2e}), "
                  f"max abs err / |theta| {np.max(err) / fro:.2e}")
            assert np.all(err <= 1e-4 * s_ref[:k] + floor * fro), (method, pre, k)
            disc = float(r.discarded.cpu())
            assert abs(disc - np.sum(s_ref[k:] ** 2)) <= 1e-10 * fro ** 2
            a = r.a_out.cpu().numpy().astype(np.complex128).reshape(n, k)
            b = r.b_out.cpu().numpy().astype(np.complex128).reshape(k, n)
            rec = np.linalg.norm(t64 - a @ b)
            # complex64 outputs: reconstruction to their rounding (~1e-7 |theta|)
            assert abs(rec - np.sqrt(np.sum(s_ref[k:] ** 2))) <= 1e-6 * fro, (method, k, rec)
            # the unscaled factor is an isometry (absorb right: U = a)
            assert np.linalg.norm(a.conj().T @ a - np.eye(k)) <= 1e-5
