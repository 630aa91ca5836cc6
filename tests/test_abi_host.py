"""CPU tests of the C ABI surface and the host-side logic (no GPU needed).

The library is loaded and every symbol include/mpsx.h declares is resolved;
the planning call (pure host code) is exercised; nothing launches a kernel.
"""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "mpsx.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mpsx_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from mpsim_amd import _native as nat
    lib = nat.load()
    syms = header_symbols()
    assert len(syms) >= 10
    for s in syms:
        assert hasattr(lib, s), s
    assert set(nat.EXPORTS) <= set(syms)
    assert lib.mpsx_version() >= 100


def test_struct_layouts_match():
    from mpsim_amd import _native as nat
    lib = nat.load()
    assert lib.mpsx_sizeof(1) == nat.TWO_SITE_OP.itemsize == 152
    assert lib.mpsx_sizeof(0) == nat.SITE_OP.itemsize
    assert lib.mpsx_sizeof(2) == nat.GEMM_OP.itemsize
    assert lib.mpsx_sizeof(3) == nat.SCALE_OP.itemsize
    assert lib.mpsx_sizeof(4) == ctypes.sizeof(nat.SvdOpts)
    assert lib.mpsx_sizeof(9) == 0


def _table(rows):
    from mpsim_amd import _native as nat
    tab = np.zeros(len(rows), dtype=nat.TWO_SITE_OP)
    for i, (cl, cm, cr, d, k, fl) in enumerate(rows):
        for f in ("a", "b", "gate", "a_out", "b_out"):
            tab[i][f] = 0x1000
        tab[i]["chi_l"], tab[i]["chi_m"], tab[i]["chi_r"] = cl, cm, cr
        tab[i]["d"], tab[i]["k"], tab[i]["flags"] = d, k, fl
    return tab


def _plan(tab, opts=None):
    from mpsim_amd import _native as nat
    lib = nat.load()
    o = opts or nat.SvdOpts(0, 0, 1, 0, 0.0)
    nb = ctypes.c_size_t(0)
    rc = lib.mpsx_two_site_plan(tab.ctypes.data, len(tab), ctypes.byref(o), ctypes.byref(nb))
    return rc, nb.value, lib.mpsx_last_error().decode()


def test_plan_layout_and_prefixes():
    tab = _table([(1, 1, 1, 2, 2, 0), (256, 256, 256, 2, 256, 0), (4, 8, 2, 3, 6, 1)])
    rc, nbytes, _ = _plan(tab)
    assert rc == 0
    p = tab
    # orientation: absorb-left problems orthogonalise theta^T's rows
    assert (p[0]["C"], p[0]["R"]) == (2, 2)
    assert (p[1]["C"], p[1]["R"]) == (512, 512)
    assert (p[2]["C"], p[2]["R"]) == (6, 12)
    for r in p:
        assert r["nb"] % 2 == 0 and r["Cp"] == r["nb"] * r["bw"] and r["Cp"] >= r["C"]
        assert r["LD"] >= r["R"] + r["Cp"] and r["ws_off"] % 256 == 0
    assert list(p["tile_base"]) == [0, 1, 65]
    assert list(p["row_base"]) == [0, int(p[0]["Cp"]), int(p[0]["Cp"] + p[1]["Cp"])]
    # gather workgroups: 32x32 tiles of the M x k factor, then one per row of the k x N one
    units = [-(-M // 32) * -(-k // 32) + k for M, k in ((2, 2), (512, 256))]
    assert list(p["out_base"]) == [0, units[0], units[0] + units[1]]
    assert nbytes > p[2]["ws_off"]


@pytest.mark.parametrize("cl,cr,d", [(40, 40, 2), (256, 256, 2), (171, 171, 3), (171, 40, 3),
                                     (33, 300, 2), (500, 17, 3)])
def test_workspace_bound_covers_the_plan(cl, cr, d):
    """batch.ws_bound (the per-problem estimate batching and admission use)
    is at least the region the library plans for the problem (the distance
    between two copies' ws_off), for both orientations and every SVD method
    (mixed Gram, FP64 Gram, one-sided rows; C = 513 = 16 nb - 31 is the
    tightest padding)."""
    from mpsim_amd import _native as nat
    from mpsim_amd.batch import ws_bound
    for flags in (0, 1):
        for method, fp32_sweeps, pre in ((0, 0, 0), (0, 12, 0), (1, 0, 0), (0, 0, 2)):
            k = min(d * cl, d * cr)
            tab = _table([(cl, 8, cr, d, k, flags)] * 2)
            o = nat.SvdOpts(0, 0, 1, 0, 0.0)
            o.method = method
            o.fp32_sweeps = fp32_sweeps
            o.precondition = pre
            rc, _, err = _plan(tab, o)
            assert rc == 0, err
            size = int(tab[1]["ws_off"] - tab[0]["ws_off"])
            assert ws_bound(int(tab[0]["C"]), int(tab[0]["R"])) >= size, (flags, method, size)


def test_workspace_bound_covers_random_plans():
    """batch.ws_bound against the planned region on 300 random shapes (both
    orientations, qudits 2-4, up to 1200 rows), for each SVD method and
    preconditioning setting."""
    from mpsim_amd import _native as nat
    from mpsim_amd.batch import ws_bound
    rng = np.random.RandomState(21)
    shapes = []
    for _ in range(300):
        d = int(rng.choice([2, 3, 4]))
        cl, cr = (int(x) for x in rng.randint(1, 1200 // d, size=2))
        shapes.append((cl, int(rng.randint(1, 600)), cr, d, int(rng.randint(0, 2))))
    for method, fp32_sweeps, pre in ((0, 0, 0), (0, 12, 0), (1, 0, 0), (0, 0, 1), (0, 0, 2)):
        o = nat.SvdOpts(0, 0, 1, 0, 0.0)
        o.method, o.fp32_sweeps, o.precondition = method, fp32_sweeps, pre
        for cl, cm, cr, d, flags in shapes:
            tab = _table([(cl, cm, cr, d, min(d * cl, d * cr), flags)] * 2)
            rc, _, err = _plan(tab, o)
            assert rc == 0, err
            size = int(tab[1]["ws_off"] - tab[0]["ws_off"])
            assert ws_bound(int(tab[0]["C"]), int(tab[0]["R"])) >= size, (cl, cm, cr, d, flags, method, pre)


@pytest.mark.parametrize("row,msg", [
    ((4, 4, 4, 2, 9, 0), "k=9"),
    ((4, 4, 4, 17, 2, 0), "d=17"),
    ((4, 4, 4, 0, 2, 0), "d=0"),
    ((-1, 4, 4, 2, 2, 0), "negative"),
])
def test_plan_rejects_bad_descriptors(row, msg):
    rc, _, err = _plan(_table([row]))
    assert rc == 1 and msg in err


def test_plan_generic_qudit_theta_tiles():
    """Qudit dimensions 5..16 are planned (core.py:1013-1028 accepts any d
    whose gate edges match the MPS); their theta runs one workgroup per
    (l, tile of r values), 64 r values for d <= 8 and 16 above, so the
    theta workgroup prefix advances by chi_l * ceil(chi_r / tile)."""
    rows = [(7, 5, 70, 5, 35, 0), (3, 9, 20, 9, 27, 1), (2, 2, 2, 16, 32, 0), (33, 4, 40, 2, 66, 0)]
    tab = _table(rows)
    rc, _, err = _plan(tab)
    assert rc == 0, err
    tiles = [7 * 2, 3 * 2, 2 * 1, 2 * 2]
    assert list(tab["tile_base"]) == [0, tiles[0], tiles[0] + tiles[1], sum(tiles[:3])]
    assert (tab[0]["C"], tab[0]["R"]) == (35, 350) and (tab[1]["C"], tab[1]["R"]) == (180, 27)


def test_plan_qr_preconditioning_rule():
    """precondition 1 marks (plan.method 3) the FP64 Gram problems whose
    middle bond is at least the smaller outer bond; 2 every Gram problem up
    to 1024 padded rows and 4096 columns, plain splits included; the
    one-sided rows problems and the mixed method are left alone."""
    from mpsim_amd import _native as nat
    rows = [(256, 256, 256, 2, 256, 0), (256, 128, 256, 2, 256, 0), (64, 300, 128, 2, 128, 1),
            (520, 520, 520, 2, 1040, 0), (32, 32, 32, 2, 32, 0), (300, 0, 200, 1, 200, 0),
            (512, 512, 512, 2, 512, 0)]
    want = {0: [1, 1, 1, 1, 0, 1, 1], 1: [3, 1, 3, 1, 0, 1, 3], 2: [3, 3, 3, 1, 0, 3, 3]}
    for pre, methods in want.items():
        tab = _table(rows)
        tab[5]["gate"] = 0  # plain split
        o = nat.SvdOpts(0, 0, 1, 0, 0.0)
        o.precondition = pre
        rc, _, err = _plan(tab, o)
        assert rc == 0, err
        assert list(tab["method"]) == methods, pre
        if pre == 2:  # the preconditioned problem's region holds X0, panels and order
            plain = _table(rows[:1] * 2)
            _plan(plain)
            pre_tab = _table(rows[:1] * 2)
            _plan(pre_tab, o)
            extra = int(pre_tab[1]["ws_off"]) - int(plain[1]["ws_off"])
            assert extra >= 16 * 512 * 512 + 2 * 16 * 16 * 512
    # tall (C > R) and preconditioned: the Jacobi pairs R's blocks (at least
    # six), Z and Y keep C's padded rows
    for pre, want_nb in ((1, 6), (0, 26)):
        tab = _table([(200, 200, 45, 2, 90, 0), (200, 200, 20, 2, 40, 0), (45, 64, 200, 2, 90, 1)])
        o = nat.SvdOpts(0, 0, 1, 0, 0.0)
        o.precondition = pre
        assert _plan(tab, o)[0] == 0
        assert (tab[0]["C"], tab[0]["R"], tab[0]["Cp"], tab[0]["nb"]) == (400, 90, 416, want_nb)
        assert (tab[1]["Cp"], tab[1]["nb"]) == (416, 26)  # R's blocks would be too few
        assert (tab[2]["C"], tab[2]["R"], tab[2]["Cp"], tab[2]["nb"]) == (400, 90, 416, want_nb)
    o = nat.SvdOpts(0, 0, 1, 0, 0.0)
    o.precondition = 1
    o.fp32_sweeps = 8
    tab = _table(rows[:1])
    assert _plan(tab, o)[0] == 0 and tab[0]["method"] == 2
    o = nat.SvdOpts(0, 0, 1, 0, 0.0)
    o.precondition = 3
    rc, _, err = _plan(_table(rows[:1]), o)
    assert rc == 1 and "precondition" in err


def test_qr_plans_meet_the_kernels_assumptions():
    """Every QR-preconditioned plan (random shapes, both orientations, up to
    the size limits) has what the QR kernels take for granted: padded rows a
    multiple of 32 and at most 1024, at most 4096 columns, an even block
    count, and column-norm parts (one FP64 value per column and 32 rows)
    that fit the V ring of 8 x 16 x 16 bytes per padded row."""
    from mpsim_amd import _native as nat
    rng = np.random.RandomState(12)
    rows = []
    for _ in range(400):
        d = int(rng.choice([2, 3, 4]))
        cl, cr = (int(x) for x in rng.randint(1, 1200 // d, size=2))
        cm = int(rng.randint(1, 600))
        rows.append((cl, cm, cr, d, int(rng.randint(1, 1 + d * min(cl, cr))), int(rng.randint(0, 2))))
    rows += [(512, 512, 512, 2, 512, 0), (100, 100, 2048, 2, 200, 0), (2048, 100, 100, 2, 200, 1),
             (340, 340, 341, 3, 1020, 0)]
    o = nat.SvdOpts(0, 0, 1, 0, 0.0)
    o.precondition = 2
    seen = 0
    for lo in range(0, len(rows), 50):
        tab = _table(rows[lo:lo + 50])
        rc, _, err = _plan(tab, o)
        assert rc == 0, err
        for r in tab[tab["method"] == 3]:
            seen += 1
            C, R, Cp, nb = int(r["C"]), int(r["R"]), int(r["Cp"]), int(r["nb"])
            assert Cp % 32 == 0 and C <= Cp <= 1024 and R <= 4096, (C, R, Cp)
            assert nb % 2 == 0 and nb >= 6 and 16 * nb <= Cp, (C, R, Cp, nb)
            assert -(-C // 32) * R * 8 <= 8 * 16 * 16 * Cp, (C, R, Cp)
    assert seen > 100


def test_plan_rejects_bad_options():
    from mpsim_amd import _native as nat
    rc, _, err = _plan(_table([(4, 4, 4, 2, 8, 0)]), nat.SvdOpts(5, 0, 1, 0, 0.0))
    assert rc == 1 and "block" in err


def test_two_site_rejects_unplanned_table():
    from mpsim_amd import _native as nat
    lib = nat.load()
    tab = _table([(4, 4, 4, 2, 8, 0)])
    tab[0]["ws_off"] = 12345
    o = nat.SvdOpts(0, 0, 1, 0, 0.0)
    sweeps = ctypes.c_int(0)
    rc = lib.mpsx_two_site(None, tab.ctypes.data, tab.ctypes.data, 1, ctypes.byref(o),
                           ctypes.c_void_p(0x1000), 1 << 20, ctypes.byref(sweeps))
    assert rc == 1 and "plan" in lib.mpsx_last_error().decode()


def _run_unlaunchable(tab, opts=None):
    """mpsx_two_site with a 1 KiB workspace: every check runs, and a table
    that passes them stops at the workspace check (MPSX_EWORKSPACE) before
    any HIP call."""
    from mpsim_amd import _native as nat
    lib = nat.load()
    o = opts or nat.SvdOpts(0, 0, 1, 0, 0.0)
    sweeps = ctypes.c_int(0)
    rc = lib.mpsx_two_site(None, tab.ctypes.data, tab.ctypes.data, len(tab), ctypes.byref(o),
                           ctypes.c_void_p(0x1000), 1024, ctypes.byref(sweeps))
    return rc, lib.mpsx_last_error().decode()


@pytest.mark.parametrize("field,value", [("flags", 1), ("chi_l", 3), ("chi_r", 5), ("d", 3),
                                         ("k", 100), ("chi_m", 1)])
def test_two_site_rejects_table_edited_after_planning(field, value):
    """The library re-plans every op from (d, chi_l, chi_r, k, flags) and the
    options and rejects a plan that does not match, also for the last op of
    the table (whose edits change no later prefix). k is re-checked against
    min(d chi_l, d chi_r); chi_m is not part of the plan (only the theta
    kernel reads it) and is accepted."""
    from mpsim_amd import _native as nat
    tab = _table([(4, 4, 2, 2, 4, 0), (8, 4, 3, 2, 6, 0)])
    rc, _, err = _plan(tab)
    assert rc == 0, err
    rc, err = _run_unlaunchable(tab.copy())
    assert rc == nat.MPSX_EWORKSPACE, err
    for i in (0, 1):
        bad = tab.copy()
        bad[i][field] = value
        rc, err = _run_unlaunchable(bad)
        if field == "chi_m":
            assert rc == nat.MPSX_EWORKSPACE, err
        else:
            assert rc == nat.MPSX_EINVAL and ("stale" in err or "k=" in err), (i, field, err)


def test_two_site_rejects_plan_for_other_options():
    """A table planned with one SVD method or precision is rejected when the
    call asks for another."""
    from mpsim_amd import _native as nat
    tab = _table([(64, 64, 64, 2, 64, 0)])
    rc, _, err = _plan(tab)
    assert rc == 0 and int(tab[0]["method"]) == 1, err
    rows = nat.SvdOpts(0, 0, 1, 0, 0.0)
    rows.method = 1
    rc, err = _run_unlaunchable(tab, rows)
    assert rc == nat.MPSX_EINVAL and "stale" in err, err
    mixed = nat.SvdOpts(0, 0, 1, 0, 0.0)
    mixed.fp32_sweeps = 12
    rc, err = _run_unlaunchable(tab, mixed)
    assert rc == nat.MPSX_EINVAL and "stale" in err, err
    rc, _, err = _plan(tab, mixed)
    assert rc == 0 and int(tab[0]["method"]) == 2, err
    f32 = nat.SvdOpts(0, 0, 1, 0, 0.0, 0, 1)
    rc, err = _run_unlaunchable(tab, f32)
    assert rc == nat.MPSX_EINVAL and "stale" in err, err


def test_haar_generator_is_the_reference_recipe_bit_for_bit():
    from oracle import mps_oracle as mo
    from mpsim_amd import gates
    for nq, d, seed in ((1, 2, 0), (2, 2, 7), (2, 3, 11)):
        a = gates.haar_random_unitary(nq, d, seed=seed).tensor
        b = mo.haar_random_unitary(nq, d, seed=seed)
        assert np.array_equal(a, b)


def test_bond_plan_matches_oracle_growth():
    """The host-side shape plan (zero singular values kept) equals the bond
    dimensions the oracle reaches on the same circuit."""
    from oracle.mps_oracle import OracleMPS
    from mpsim_amd.circuits import bond_plan, brickwork_haar, random_pairs_circuit
    for circ in (brickwork_haar(10, 6, 2, maxsvals=8, seed=4),
                 random_pairs_circuit(8, 9, 5, maxsvals=4)):
        plan = bond_plan(circ)
        m = OracleMPS(circ.n, circ.d, dtype_mode="c128")
        for layer, bonds in zip(circ.layers, plan):
            m.apply_ops(layer, maxsvals=circ.maxsvals)
            assert m.bond_dimensions() == bonds


def test_split_option_parsing():
    """core.py:1104-1130."""
    from mpsim_amd.core import max_bond_dimensions_for, parse_split_options
    assert parse_split_options({}, 4) == (None, True)
    assert parse_split_options({"maxsvals": 3.7}, 4) == (3, True)
    assert parse_split_options({"fraction": 0.5}, 4) == (2, True)
    assert parse_split_options({"fraction": 0.5}, 2) == (1, True)
    assert parse_split_options({"fraction": 0.25}, 2) == (0, True)   # banker's rounding
    assert parse_split_options({"keep_left_canonical": False}, 4) == (None, False)
    with pytest.raises(ValueError):
        parse_split_options({"fraction": 0.5, "maxsvals": 1}, 4)
    with pytest.raises(ValueError):
        parse_split_options({"fraction": -0.1}, 4)
    assert max_bond_dimensions_for(6, 10) == [10, 100, 1000, 100, 10]


def test_workloads_are_seeded_and_shaped():
    from mpsim_amd.circuits import (brickwork_cnot_u2, brickwork_haar, circuit_work,
                                    poisson_requests, uneven_batch)
    c1 = brickwork_cnot_u2()
    assert c1.two_site_count() == 60 and c1.maxsvals is None
    c2 = brickwork_haar()
    assert c2.two_site_count() == 630
    fl, by, cnt = circuit_work(c2)
    assert cnt == 630 and fl > 630 * 1e6
    a, b = uneven_batch(4), uneven_batch(4)
    assert [x.n for x in a] == [x.n for x in b]
    assert all(12 <= x.n <= 64 and 4 <= x.depth // 2 <= 24 for x in a)
    assert np.array_equal(a[0].ops[0][1], b[0].ops[0][1])
    reqs = poisson_requests(duration=0.2)
    assert all(r.arrival < 0.2 for r in reqs) and len(reqs) > 10


def test_per_update_work_accounting():
    """The chi=256, d=2 figures the benchmark uses (11.82 GFLOP, 4.19 MB)."""
    from mpsim_amd.circuits import two_site_bytes, two_site_flops
    f = two_site_flops(256, 256, 256, 2, 256)
    assert abs(f / 1e9 - 11.82) < 0.01
    assert abs(two_site_bytes(256, 256, 256, 2, 256) / 1e6 - 4.19) < 0.01


def test_simulator_rejects_bad_programs():
    from mpsim_amd.mpsim_cirq import MPSimulator
    with pytest.raises(ValueError):
        MPSimulator().simulate(42)


def _rr_pair(p, r, n):
    """The round-robin pairing of common.h (rr_item / rr_pair)."""
    def item(pos):
        return 0 if pos == 0 else ((pos - 1 + r) % (n - 1)) + 1
    a, b = item(p), item(n - 1 - p)
    return (min(a, b), max(a, b))


@pytest.mark.parametrize("nb", [2, 4, 6, 10, 32, 34, 192])
def test_gram_tile_enumeration(nb):
    """The Gram method's work split (k_hgram, k_hsolve, k_hupdate in
    twosite.hip): k_hgram's (slot j, y) grid enumerates every upper 16x16 block
    tile of H once, and in every round the tiles k_hsolve writes (the three of
    each pivot) and those k_hupdate writes (the four of each unordered pair of
    pivots, pair (P, P + y + 1 mod n) for y < n/2, P < n/2 when 2(y + 1) = n)
    cover every stored tile exactly once."""
    n = nb // 2
    upper = {(i, j) for i in range(nb) for j in range(i, nb)}
    seen = []
    for j in range(n):
        for y in range(nb + 1):
            rem, bi = j * (nb + 1) + y, 0
            while rem >= nb - bi:
                rem -= nb - bi
                bi += 1
            seen.append((bi, bi + rem))
    assert sorted(seen) == sorted(upper)
    for r in range(nb - 1):
        piv = [_rr_pair(P, r, nb) for P in range(n)]
        assert sorted(x for pr in piv for x in pr) == list(range(nb))
        written = []
        for P in range(n):
            I, J = piv[P]
            written += [(I, I), (I, J), (J, J)]
            for y in range(n // 2):
                dd = y + 1
                if 2 * dd == n and P >= n // 2:
                    continue
                for bi in piv[P]:
                    for bj in piv[(P + dd) % n]:
                        written.append((min(bi, bj), max(bi, bj)))
        assert sorted(written) == sorted(upper)


def _rr_partner(item, r, n):
    """rr_partner of twosite.hip: the inverse of rr_item, then the opposite slot."""
    pos = 0 if item == 0 else ((item - 1 - r) % (n - 1)) + 1
    opp = n - 1 - pos
    return 0 if opp == 0 else ((opp - 1 + r) % (n - 1)) + 1


@pytest.mark.parametrize("nb", [6, 8, 10, 32, 34, 192])
def test_gram_pipelined_round_split(nb):
    """The pipelined Gram round: the update tiles marked critical (some block
    tile of the pivot-pair tile is a pivot pair of the next round, found with
    rr_partner; k_hround's eigensolve workgroups update them, its part 2
    skips them) together with the pivot tiles the eigensolves write hold
    every tile the next round's eigensolves read, and each critical tile is
    updated by exactly one eigensolve workgroup of the next round (hnext_body:
    the tile between the pivots that held its two blocks), so no two
    workgroups of one launch touch the same tile."""
    n = nb // 2
    for r in range(2 * (nb - 1)):  # rounds past nb - 2 wrap to the next sweep's pairing
        piv = [_rr_pair(P, r % (nb - 1), nb) for P in range(n)]
        nxt = [_rr_pair(P, (r + 1) % (nb - 1), nb) for P in range(n)]
        for I, J in nxt:
            assert _rr_partner(I, r + 1, nb) == J and _rr_partner(J, r + 1, nb) == I
        have = set()
        for I, J in piv:
            have |= {(I, I), (I, J), (J, J)}
        for P in range(n):
            I, J = piv[P]
            for y in range(n // 2):
                dd = y + 1
                if 2 * dd == n and P >= n // 2:
                    continue
                I2, J2 = piv[(P + dd) % n]
                a, b = _rr_partner(I, r + 1, nb), _rr_partner(J, r + 1, nb)
                if a in (I2, J2) or b in (I2, J2):
                    have |= {(min(x, z), max(x, z)) for x in (I, J) for z in (I2, J2)}
        for I, J in nxt:
            assert {(I, I), (I, J), (J, J)} <= have, (nb, r, I, J)
        # one eigensolve workgroup per pivot of round r + 1; each critical tile
        # (a pair of pivots of round r) must be claimed by exactly one of them
        def pivot(item, rr):
            pos = 0 if item == 0 else ((item - 1 - rr) % (nb - 1)) + 1
            return min(pos, nb - 1 - pos)
        critical = set()
        for P in range(n):
            for y in range(n // 2):
                dd = y + 1
                if 2 * dd == n and P >= n // 2:
                    continue
                I, J = piv[P]
                I2, J2 = piv[(P + dd) % n]
                if {_rr_partner(I, r + 1, nb), _rr_partner(J, r + 1, nb)} & {I2, J2}:
                    critical.add(frozenset((P, (P + dd) % n)))
        claimed = []
        for In, Jn in nxt:
            pa, pb = pivot(In, r % (nb - 1)), pivot(Jn, r % (nb - 1))
            assert pa != pb, (nb, r, In, Jn)  # no pair repeats in consecutive rounds
            claimed.append(frozenset((pa, pb)))
        assert len(claimed) == len(set(claimed)) and set(claimed) == critical, (nb, r)


def test_gram_next_round_tiles_unshared():
    """For every even block count from 6 up, the tile between the two pivots
    of round r that hold the blocks of a pivot of round r + 1 differs for all
    pivots of round r + 1 (so hnext_body's workgroups own their tiles); with
    four blocks one tile holds both, which is why plan_method gives such
    problems the one-sided method."""
    def piv_of(nb, r):
        out = {}
        for P in range(nb // 2):
            for x in _rr_pair(P, r, nb):
                out[x] = P
        return out

    for nb in range(4, 258, 2):
        shared = False
        for r in range(nb - 1):
            cur, nxt = piv_of(nb, r), (r + 1) % (nb - 1)
            tiles = [frozenset((cur[In], cur[Jn])) for In, Jn in
                     (_rr_pair(P, nxt, nb) for P in range(nb // 2))]
            assert all(len(t) == 2 for t in tiles), (nb, r)
            shared |= len(set(tiles)) != len(tiles)
        assert shared == (nb == 4), nb


def test_schedule_keeps_the_order_of_overlapping_ops():
    """circuits.schedule: a layer of nearest-neighbour ops on disjoint qudits
    is one step; a non-adjacent pair becomes the reference's SWAP network
    (core.py:1035-1043, 1154-1219: hops right, the gate, hops back); ops
    that share a site (two pairs on one qudit, a one-site gate inside a
    network) run in list order; invalid ops raise before anything runs."""
    from mpsim_amd.circuits import Circuit, bond_plan, schedule
    u2, u4 = np.eye(2), np.eye(4)
    steps, ends = schedule(Circuit(6, 2, [[((0, 1), u4), ((2, 3), u4), ((4,), u2)]]))
    assert len(steps) == 1 and ends == [1]
    assert steps[0].twos == [(0, 0, 0, False), (2, 0, 1, False)] and steps[0].ones == [(4, 0, 2)]
    lay = [((0, 3), u4), ((4, 5), u4), ((1,), u2), ((1, 2), u4), ((5, 1), u4)]
    steps, ends = schedule(Circuit(6, 2, [lay], 8))
    twos = [[(i, o, f) for i, _, o, f in st.twos] for st in steps]
    ones = [[(q, o) for q, _, o in st.ones] for st in steps]
    assert twos[:5] == [[(0, -1, False), (4, 1, False)], [(1, -1, False)], [(2, 0, False)],
                        [(1, -1, False)], [(0, -1, False)]]
    assert ones[5] == [(1, 2)] and twos[5] == []
    assert twos[6] == [(1, 3, False)]                       # (1, 2) after the one-site gate on 1
    assert twos[7:] == [[(1, -1, False)], [(2, -1, False)], [(3, -1, False)], [(4, 4, True)],
                        [(3, -1, False)], [(2, -1, False)], [(1, -1, False)]]
    assert ends == [len(steps)]
    # bond dimensions follow the sequential application (the oracle's rule)
    from oracle.mps_oracle import OracleMPS
    ora = OracleMPS(6, 2)
    ora.apply_ops([(q, np.eye(2 ** len(q))) for q, _ in lay], maxsvals=8)
    assert bond_plan(Circuit(6, 2, [lay], 8))[-1] == ora.bond_dimensions()
    for bad in ([((0, 6), u4)], [((2, 2), u4)], [((0, 1, 2), np.eye(8))], [((-1,), u2)]):
        with pytest.raises(ValueError):
            schedule(Circuit(6, 2, [bad]))


def test_program_passes_the_svd_method():
    """CircuitProgram plans and launches with the SvdConfig it was given,
    method included (a one-sided request must not silently run the Gram
    method)."""
    import types
    import torch
    from mpsim_amd.engine import SvdConfig
    from mpsim_amd.program import CircuitProgram
    stub = types.SimpleNamespace(svd=SvdConfig(method=1, fp32=False, block_switch=16),
                                 unconverged=torch.zeros(1, dtype=torch.int32))
    for sweeps in (0, 7):
        o = CircuitProgram._opts(stub, sweeps)
        assert o.method == 1 and o.block_switch == 16
    tab = _table([(64, 64, 64, 2, 64, 0)])
    rc, _, err = _plan(tab, CircuitProgram._opts(stub, 0))
    assert rc == 0 and int(tab[0]["method"]) == 0, err


def test_gate_operation_conversion():
    """circuits.py:26-43: a k-qubit operation with a unitary becomes an
    MPSOperation on the mapped qudit indices whose tensor is the unitary
    reshaped to [2]*2k; one without a unitary raises
    CannotConvertToMPSOperation."""
    from mpsim_amd.core import CannotConvertToMPSOperation
    from mpsim_amd.mpsim_cirq.circuits import mps_operation_from_gate_operation

    class Op:
        def __init__(self, qubits, mat):
            self.qubits, self._mat = tuple(qubits), mat

        def _has_unitary_(self):
            return self._mat is not None

        def _unitary_(self):
            return self._mat

    u = np.arange(16).reshape(4, 4).astype(complex)
    op = mps_operation_from_gate_operation(Op(["b", "a"], u), {"a": 0, "b": 3})
    assert op.qudit_indices == (3, 0) and op.num_qudits == 2
    assert op.tensor(reshape_to_square_matrix=False).shape == (2, 2, 2, 2)
    assert np.array_equal(op.tensor(), u)
    op1 = mps_operation_from_gate_operation(Op(["a"], np.eye(2)), {"a": 1})
    assert op1.qudit_indices == (1,) and op1.is_valid()
    with pytest.raises(CannotConvertToMPSOperation):
        mps_operation_from_gate_operation(Op(["a"], None), {"a": 0})
