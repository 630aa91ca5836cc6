"""mpsx_two_site_static (the two-site update on the library's fixed device
workspace, for callers that cannot allocate) against mpsx_two_site (caller
workspace) on the same planned batch, through the C ABI: the factors,
singular values and discarded weights must be bitwise equal (same kernels;
converged problems exit every later sweep at once). Two calls in a row, the
second with a smaller batch, so it runs on what the first left in the fixed
workspace; a batch larger than the fixed workspace is rejected before
anything is enqueued.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _items(eng, rng, shapes):
    import torch
    from mpsim_amd import gates
    from mpsim_amd.engine import SplitItem, TwoSiteItem
    out = []
    for cl, cm, cr, k, left in shapes:
        if cm == 0:  # plain split of a cl x cr matrix
            M = ((rng.randn(cl, cr) + 1j * rng.randn(cl, cr)) / np.sqrt(cr)).astype(np.complex64)
            out.append(SplitItem(torch.from_numpy(M).cuda(), k, absorb_left=left, want_s=True))
            continue
        A = ((rng.randn(cl, 2, cm) + 1j * rng.randn(cl, 2, cm)) / np.sqrt(4 * cm)).astype(np.complex64)
        B = ((rng.randn(cm, 2, cr) + 1j * rng.randn(cm, 2, cr)) / np.sqrt(2)).astype(np.complex64)
        G = gates.haar_unitary(4, rng)
        out.append(TwoSiteItem(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), eng.gate(G),
                               k, absorb_left=left, want_s=True))
    return out


def _run(eng, items, static):
    import torch
    from mpsim_amd import _native as nat
    lib = nat.load()
    cfg = eng.svd.to_c()
    outs, tab, dtab, ws = eng._prepare_two_site(items, cfg, workspace=not static)
    sweeps = ctypes.c_int(0)
    torch.cuda.synchronize()
    if static:
        rc = lib.mpsx_two_site_static(eng.stream, tab.ctypes.data, dtab.data_ptr(), len(tab),
                                      ctypes.byref(cfg), ctypes.byref(sweeps))
    else:
        rc = lib.mpsx_two_site(eng.stream, tab.ctypes.data, dtab.data_ptr(), len(tab),
                               ctypes.byref(cfg), ws.data_ptr(), ws.numel(), ctypes.byref(sweeps))
    nat.check(rc)
    torch.cuda.synchronize()
    return [(o.a_out.cpu().numpy(), o.b_out.cpu().numpy(), o.s.cpu().numpy(),
             float(o.discarded.cpu())) for o in outs], sweeps.value


def test_static_workspace_matches_caller_workspace():
    from mpsim_amd.engine import Engine
    eng = Engine.get()
    rng = np.random.RandomState(71)
    # first batch: two theta 512 x 512 updates on the Gram path (12.5 MiB of
    # workspace each, 17.5 MiB for the QR-preconditioned one of saturated
    # bonds), an absorb-left one, 16-row pivots, a 3-pivot Gram problem and a
    # plain split; second batch: smaller
    first = [(256, 256, 256, 256, False), (256, 128, 256, 200, True), (40, 32, 40, 64, False),
             (8, 8, 8, 16, False), (48, 48, 48, 60, True), (100, 0, 60, 60, False)]
    second = [(64, 64, 64, 100, False), (4, 4, 4, 8, True)]
    for shapes in (first, second):
        print(f"batch {shapes}", flush=True)
        items = _items(eng, rng, shapes)
        ref, sw_ref = _run(eng, items, static=False)
        got, sw = _run(eng, items, static=True)
        assert sw >= sw_ref
        for (a0, b0, s0, d0), (a1, b1, s1, d1) in zip(ref, got):
            assert np.array_equal(a0, a1) and np.array_equal(b0, b1)
            assert np.array_equal(s0, s1) and d0 == d1


def test_static_workspace_rejects_a_batch_that_does_not_fit():
    from mpsim_amd import _native as nat
    from mpsim_amd.engine import Engine
    eng = Engine.get()
    rng = np.random.RandomState(72)
    items = _items(eng, rng, [(256, 256, 256, 256, False)] * 6)  # 75 MiB (105.2 preconditioned)
    lib = nat.load()
    cfg = eng.svd.to_c()
    outs, tab, dtab, _ = eng._prepare_two_site(items, cfg, workspace=False)
    sweeps = ctypes.c_int(0)
    rc = lib.mpsx_two_site_static(eng.stream, tab.ctypes.data, dtab.data_ptr(), len(tab),
                                  ctypes.byref(cfg), ctypes.byref(sweeps))
    assert rc == nat.MPSX_EWORKSPACE
    assert "fixed workspace" in lib.mpsx_last_error().decode()
    assert sweeps.value == 0
