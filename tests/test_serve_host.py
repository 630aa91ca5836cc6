"""Host tests of the serving loop's bookkeeping (no GPU): admission under the
HBM budget and the active cap, the split of the hardware queues over the
lanes that have work, lane choice by bond cap, and the device footprint a
request is admitted with."""
import numpy as np
import pytest


def test_admission_respects_budget_and_active_cap():
    from mpsim_amd.serve import Served, _Shared
    sh = _Shared(budget=100, max_active=2)
    assert sh.try_admit(60)
    assert not sh.try_admit(41)  # over the budget
    assert sh.try_admit(40)
    assert not sh.try_admit(0)  # two active: the cap
    assert (sh.used, sh.active, sh.peak_active) == (100, 2, 2)
    sh.release(60, Served(rid=0, arrival=0.0))
    assert (sh.used, sh.active) == (40, 1)
    assert sh.try_admit(60)
    assert sh.peak_active == 2 and [s.rid for s in sh.done] == [0]


def test_hardware_queues_spread_over_busy_lanes():
    """One queue per lane with work, the rest to the largest-bond one; a
    lane left alone takes all four."""
    from mpsim_amd.serve import _Shared
    sh = _Shared(budget=1, max_active=1, busy=(0, 1, 2))
    assert [sh.ways(i) for i in range(3)] == [1, 1, 2]
    sh.lane_done(2)
    assert [sh.ways(i) for i in range(3)] == [1, 3, 1]
    sh.lane_done(0)
    assert sh.ways(1) == 4
    sh.lane_done(1)
    assert [sh.ways(i) for i in range(3)] == [1, 1, 1]


def test_lane_of_by_bond_cap():
    from mpsim_amd.circuits import brickwork_haar
    from mpsim_amd.serve import MPSServer
    srv = MPSServer(engine=object(), lanes=(128, 256))
    caps = {64: 0, 128: 0, 129: 1, 256: 1, 512: 2, None: 2}
    for cap, lane in caps.items():
        assert srv.lane_of(brickwork_haar(8, 2, 2, maxsvals=cap, seed=1)) == lane, cap
    one = MPSServer(engine=object(), lanes=())
    assert one.lane_of(brickwork_haar(8, 2, 2, maxsvals=64, seed=1)) == 0
    with pytest.raises(ValueError):
        MPSServer(engine=object(), lanes=(128,), lane_split_ways=(1, 2, 3))


@pytest.mark.parametrize("n,depth,cap", [(12, 8, 16), (16, 12, 64), (20, 6, None)])
def test_footprint_covers_every_step(n, depth, cap):
    """The admitted footprint is twice the largest state over the schedule's
    steps (old and new site tensors coexist) plus the largest step's Jacobi
    workspace, and it grows with the bond cap."""
    from mpsim_amd.circuits import bond_plan_steps, brickwork_haar
    from mpsim_amd.serve import footprint_bytes, workspace_bytes
    c = brickwork_haar(n, depth, 2, maxsvals=cap, seed=3)
    states = [8 * sum(a * 2 * b for a, b in zip([1] + bonds, bonds + [1])) for bonds in bond_plan_steps(c)]
    peak = max([n * 2 * 8] + states)
    fp = footprint_bytes(c)
    assert fp == 2 * peak + workspace_bytes(c)
    assert workspace_bytes(c) > 0
    # the site tensors the FP64 oracle ends with (complex64 on the device)
    # fit in half of it
    from oracle.mps_oracle import OracleMPS
    ora = OracleMPS(n, 2, dtype_mode="c128")
    ora.apply_ops(c.ops, **({} if cap is None else {"maxsvals": cap}))
    assert 2 * 8 * sum(t.size for t in ora.tensors) <= fp - workspace_bytes(c)
    if cap is not None:
        smaller = brickwork_haar(n, depth, 2, maxsvals=max(1, cap // 2), seed=3)
        assert footprint_bytes(smaller) <= fp
