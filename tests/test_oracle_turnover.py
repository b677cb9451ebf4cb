"""Pins of the oracle's actin turnover (oracle/softbox_oracle_turnover.c) against closed forms, invariants and
the reference harnesses' own scripted checks.

  - aging: the forward-Euler cascade against the analytic two-step ODE (AgingHarness gate A,
    AgingHarness.java:188-230) and an exactly conserved sum;
  - growthAtp's mixing rule, cofilin accumulation's geometric closed form, the dissolve threshold;
  - growth and depolymerisation hold end2 and move end1 by one monomer radius; a split keeps the outer ends
    of the old segment (parent end2, child end1) and leaves one monomer radius between parent and child;
  - recomputeDrag against the slender-body drag of the clamped length;
  - the allocator gives accepted requests the lowest free slots in rank order, clamped at the free count;
  - DepolyHarness.checkSlotRecycle and checkLinkBreak (DepolyHarness.java:323-409) as scripted there;
  - applyDeath's scatter equals every slot clearing its own links to dying slots (the GPU's form) on
    reciprocal chains, adjacent deaths included;
  - the cadence keeps the monomer ledger exact (ActinPool conservation, AgingHarness.conservationOk) and the
    chain topology reciprocal over many cadences, and grows / depolymerises at the stated probabilities.
"""
import math

import numpy as np
import pytest

from _common import small_network
from oracle import ffi
from softbox_amd import scenes
from softbox_amd.physics import rod_drag_si
from softbox_amd.turnover import TurnoverStore

H = 0.0027  # actinMonoRadius (Constants.java:37), um
f32 = np.float32


def test_aging_matches_the_two_step_ode():
    kH, kD, bc = 0.3, 1.0, 1.0e-3
    C = 4
    nuc = np.zeros(3 * C, np.float32)
    nuc[:C] = 1.0
    state = np.array([0, 0, -1, 0], np.int32)
    nuc[3] = 0.25
    nuc[C + 3] = 0.5
    nuc[2 * C + 3] = 0.25
    pp = np.array([kH * bc, kD * bc], np.float32)
    worst = 0.0
    t_done = 0
    for T in (1.0, 3.0, 5.0, 10.0):
        n = int(round(T / bc))
        for _ in range(n - t_done):
            ffi.call("sbo_aging_age", C, state, nuc, pp, 1)
        t_done = n
        a = math.exp(-kH * T)
        b = kH / (kD - kH) * (math.exp(-kH * T) - math.exp(-kD * T))
        worst = max(worst, abs(nuc[0] - a), abs(nuc[C] - b), abs(nuc[2 * C] - (1 - a - b)))
        assert abs(float(nuc[0]) + float(nuc[C]) + float(nuc[2 * C]) - 1.0) < 2e-6
    assert worst < 5e-4, worst
    assert nuc[2] == 1.0 and nuc[C + 2] == 0.0  # a FREE slot does not age


def test_growth_atp_mixes_one_fresh_monomer():
    C = 3
    nuc = np.array([0.2, 0.0, 0.5, 0.3, 0.0, 0.1, 0.5, 1.0, 0.4], np.float32)
    grew = np.array([1, 0, 1], np.int32)
    mc = np.array([10, 7, 4], np.int32)
    want = nuc.copy()
    for s in (0, 2):
        M = mc[s]
        w, inv = f32(M - 1) / f32(M), f32(1.0) / f32(M)
        want[s] = want[s] * w + inv
        want[C + s] *= w
        want[2 * C + s] *= w
    ffi.call("sbo_aging_growth_atp", C, grew, mc, nuc)
    assert np.array_equal(nuc, want)
    assert abs(nuc[0] + nuc[C] + nuc[2 * C] - 1.0) < 1e-6


def test_cofilin_accumulates_geometrically_and_dissolves_past_the_ratio():
    C = 2
    state = np.zeros(C, np.int32)
    nuc = np.zeros(3 * C, np.float32)
    nuc[2 * C:] = [0.8, 0.4]
    cof = np.zeros(C, np.float32)
    p = 0.01
    pp = np.array([p, 0.5], np.float32)
    for _ in range(200):
        ffi.call("sbo_cofilin_accumulate", C, state, nuc, cof, pp, 1)
    want = np.array([0.8, 0.4]) * (1 - (1 - p) ** 200)
    assert np.allclose(cof, want, atol=2e-5)
    flag = np.zeros(C, np.int32)
    ret = np.zeros(C, np.int32)
    mc = np.array([40, 50], np.int32)
    ffi.call("sbo_cofilin_dissolve", C, state, mc, cof, flag, ret, pp, 1)
    assert flag.tolist() == [int(cof[0] > 0.5), int(cof[1] > 0.5)] == [1, 0] and ret.tolist() == [40, 0]


def _ends(c, u, M):
    L = (M + 1) * H
    return c - 0.5 * L * u, c + 0.5 * L * u


def test_grow_and_depoly_hold_end2():
    C = 2
    g = np.random.default_rng(1)
    u = g.standard_normal((3, C))
    u /= np.linalg.norm(u, axis=0)
    coord = (g.random((3, C)) * 5).astype(np.float32)
    uv = u.astype(np.float32)
    mc = np.array([40, 40], np.int32)
    e1, e2 = _ends(coord.astype(np.float64), uv.astype(np.float64), mc)
    seed = np.array([0, -1], np.int32)
    grew = np.zeros(C, np.int32)
    cflat, uflat = coord.ravel().copy(), uv.ravel().copy()
    ffi.call("sbo_grow", C, seed, mc, cflat, uflat, grew, np.array([1.0, H, 64], np.float32), 5, 3, 0, 0, 1)
    assert grew.tolist() == [1, 0] and mc.tolist() == [41, 40]
    n1, n2 = _ends(cflat.reshape(3, C).astype(np.float64), uv.astype(np.float64), mc)
    assert np.abs(n2[:, 0] - e2[:, 0]).max() < 1e-6                       # end2 (barbed, held) fixed
    assert abs(np.linalg.norm(n1[:, 0] - e1[:, 0]) - H) < 1e-6            # end1 extends by one radius
    # depoly: the pointed tip (end1 free) loses one monomer; end2 held
    state = np.zeros(C, np.int32)
    e1s = np.array([-1, 5], np.int32)
    ret, death = np.zeros(C, np.int32), np.zeros(C, np.int32)
    before = cflat.copy()
    m_before = mc.copy()
    ffi.call("sbo_depoly", C, state, mc, cflat, uflat, e1s, None, ret, death,
             np.array([1.0, H, 3], np.float32), np.array([0.0, 0.0], np.float32), 7, 3, 0, 0, 1)
    assert ret.tolist() == [1, 0] and mc.tolist() == [m_before[0] - 1, m_before[1]]
    b1, b2 = _ends(before.reshape(3, C).astype(np.float64), uv.astype(np.float64), m_before)
    a1, a2 = _ends(cflat.reshape(3, C).astype(np.float64), uv.astype(np.float64), mc)
    assert np.abs(a2[:, 0] - b2[:, 0]).max() < 1e-6 and abs(np.linalg.norm(a1[:, 0] - b1[:, 0]) - H) < 1e-6
    # below the floor the segment dies en masse
    mc[0] = 2
    ffi.call("sbo_depoly", C, state, mc, cflat, uflat, e1s, None, ret, death,
             np.array([1.0, H, 3], np.float32), np.array([0.0, 0.0], np.float32), 8, 3, 0, 0, 1)
    assert death.tolist() == [1, 0] and ret.tolist() == [2, 0]


def _chain_store(C):
    z = lambda: np.full(C, -1, np.int32)  # noqa: E731
    return z(), z(), z(), z()


def test_split_keeps_the_outer_ends_and_rewires_three_slots():
    # slots: 0 tip G (M = 70, node-held), 1 its outward neighbour Mold (G.end1 -> Mold.end2), 2 free
    C, RC = 3, 3
    e1s, e1d, e2s, e2d = _chain_store(C)
    e1s[0], e1d[0] = 1, 1
    e2s[1], e2d[1] = 0, 0
    u = np.array([0.6, 0.0, 0.8])
    coord = np.zeros((3, C), np.float32)
    coord[:, 0] = [1.0, 2.0, 3.0]
    coord[:, 1] = coord[:, 0] - (70 + 1) * H * 0.5 * u - 37 * H * 0.5 * u
    coord[:, 2] = [9.0, 9.0, 9.0]
    uv = np.repeat(u[:, None], C, axis=1).astype(np.float32)
    yv = np.repeat(np.array([0.0, 1.0, 0.0])[:, None], C, axis=1).astype(np.float32)
    mc = np.array([70, 36, 0], np.int32)
    seedNode = np.array([0, -1, -1], np.int32)
    state = np.array([0, 0, -1], np.int32)
    acc = np.zeros(C, np.int32)
    req = [np.zeros(3 * RC, np.float32) for _ in range(3)]
    cflat, uflat, yflat = coord.ravel().copy(), uv.ravel().copy(), yv.ravel().copy()
    sp = np.array([H, 64, 32], np.float32)
    e1_old, e2_old = _ends(coord.astype(np.float64)[:, 0], u, 70)
    ffi.call("sbo_mark_splits", C, seedNode, mc, cflat, uflat, yflat, acc, req[0], req[1], req[2], RC, sp, 1)
    assert acc.tolist() == [1, 0, 0]
    fc, fo, fl, ro = (np.zeros(C + 1, np.int32) for _ in range(4))
    bts, brs = np.ones(C, np.float32), np.ones(C, np.float32)
    bts[2] = brs[2] = 0.0
    ffi.call("sbo_birth_alloc", C, C, state, fc, fo, fl, acc, ro, req[0], req[1], req[2], RC, cflat, uflat, yflat,
             bts, brs, np.array([1.0, 0.5], np.float32))
    assert state.tolist() == [0, 0, 0] and bts[2] == 1.0 and brs[2] == 0.5
    assert acc.tolist() == [0, 0, 0] and fc[:C].tolist() == [0, 0, 0]  # csrScan clears what it counted
    ffi.call("sbo_split_wire", C, C, ro, fl, fo, mc, cflat, uflat, e1s, e1d, e2s, e2d, seedNode, sp)
    assert mc.tolist() == [38, 36, 32] and seedNode.tolist() == [0, -1, -1]
    c2 = cflat.reshape(3, C).astype(np.float64)
    p1, p2 = _ends(c2[:, 0], u, 38)
    c1, c2e = _ends(c2[:, 2], u, 32)
    assert np.abs(p2 - e2_old).max() < 2e-6          # the parent keeps the barbed end (node side)
    assert np.abs(c1 - e1_old).max() < 2e-6          # the child keeps the old pointed end
    assert np.abs(c2e - (p1 + H * u)).max() < 2e-6   # one monomer radius between them
    # G.end1 <-> C.end2, C.end1 <-> Mold.end2, Mold.end2 -> C.end1
    assert (e1s[0], e1d[0], e2s[2], e2d[2]) == (2, 1, 0, 0)
    assert (e1s[2], e1d[2], e2s[1], e2d[1]) == (1, 1, 2, 0)


def test_recompute_drag_is_the_slender_body_drag_of_the_clamped_length():
    C = 4
    mc = np.array([10, 10, 50, 0], np.int32)
    e1s = np.array([-1, 0, 0, -1], np.int32)
    e2s = np.array([1, 1, 1, -1], np.int32)   # slot 0 is a chain end, 1 and 2 interior, 3 a lone slot
    seg = np.zeros(C, np.float32)
    gT, gR, dT, dR = (np.zeros(3 * C, np.float32) for _ in range(4))
    aeta, kT = 0.1, 4.11e-21
    ts = TurnoverStore(C)
    ts.set_drag(aeta, kT)
    ffi.call("sbo_recompute_drag", C, mc, seg, e1s, e2s, gT, gR, dT, dR, ts.dragParams, 1)
    h = float(f32(H))
    for s, clampM in ((0, 32), (1, 30), (2, 50), (3, 32)):
        L = (mc[s] + 1) * h
        assert seg[s] == f32(L)
        g = rod_drag_si(max(L, clampM * h), float(f32(0.0035)), float(f32(aeta)))
        assert abs(gT[s] / g[0] - 1) < 1e-6 and abs(gT[C + s] / g[1] - 1) < 1e-6
        assert abs(gR[s] / g[3] - 1) < 1e-6 and abs(gR[C + s] / g[4] - 1) < 1e-6
        assert abs(dT[s] * g[0] / float(f32(kT)) - 1) < 1e-6


def test_allocator_grants_the_lowest_free_slots_in_rank_order():
    C = 10
    state = np.array([0, -1, 0, -1, -1, 0, 0, -1, 0, 0], np.int32)
    acc = np.array([0, 1, 0, 0, 1, 1, 0, 1, 1, 1], np.int32)   # 6 requests, 4 free slots
    fc, fo, fl, ro = (np.zeros(C + 1, np.int32) for _ in range(4))
    req = [np.arange(3 * C, dtype=np.float32) + 100 * k for k in range(3)]
    coord, uv, yv = (np.zeros(3 * C, np.float32) for _ in range(3))
    bts, brs = np.zeros(C, np.float32), np.zeros(C, np.float32)
    ffi.call("sbo_birth_alloc", C, C, state, fc, fo, fl, acc, ro, req[0], req[1], req[2], C, coord, uv, yv, bts,
             brs, np.array([1.0, 1.0], np.float32))
    assert fl[:4].tolist() == [1, 3, 4, 7] and fo[C] == 4 and ro[C] == 6
    # requests 1, 4, 5, 7 (ranks 0..3) get slots 1, 3, 4, 7; requests 8, 9 find none
    for r, slot in ((1, 1), (4, 3), (5, 4), (7, 7)):
        assert coord[slot] == req[0][r] and coord[C + slot] == req[0][C + r]
    assert (state >= 0).all()


def _apply_death(C, state, mc, seedNode, e1s, e1d, e2s, e2d, flag):
    bts, brs = np.ones(C, np.float32), np.ones(C, np.float32)
    ffi.call("sbo_apply_death", C, state, mc, seedNode, e1s, e1d, e2s, e2d, bts, brs, flag)
    return bts, brs


def _chain3():
    e1s, e1d, e2s, e2d = _chain_store(3)
    # A(0).end1 <-> B(1).end2 ; B.end1 <-> C(2).end2 (C the pointed tip)  (DepolyHarness.chain3, :396-404)
    e1s[0], e1d[0], e2s[1], e2d[1] = 1, 1, 0, 0
    e1s[1], e1d[1], e2s[2], e2d[2] = 2, 1, 1, 0
    return e1s, e1d, e2s, e2d


def test_link_break_of_the_reference_harness():
    # DepolyHarness.checkLinkBreak (DepolyHarness.java:351-393): (a) pointed-tip death, (b) interior death
    for dying in (2, 1):
        e1s, e1d, e2s, e2d = _chain3()
        state, mc, seedNode = np.zeros(3, np.int32), np.full(3, 32, np.int32), np.full(3, -1, np.int32)
        flag = np.zeros(3, np.int32)
        flag[dying] = 1
        bts, brs = _apply_death(3, state, mc, seedNode, e1s, e1d, e2s, e2d, flag)
        assert state[dying] < 0 and bts[dying] == 0 and brs[dying] == 0 and mc[dying] == 0
        if dying == 2:
            assert e1s[1] == -1 and e1s[0] == 1 and e2s[1] == 0
        else:
            assert e1s[0] == -1 and e2s[2] == -1 and e2s[1] == -1 and e1s[1] == -1


def test_slot_recycle_of_the_reference_harness():
    """DepolyHarness.checkSlotRecycle (DepolyHarness.java:323-348): a slot freed by death this cadence is the
    one a split claims in the same cadence."""
    C = 4
    fs_n = scenes.free_rods(n=C, seed=2, grid=False, periodic=True)
    ts = TurnoverStore(C)
    fs = fs_n.fs
    fs.end1NbrSlot[:] = -1
    fs.end1NbrSide[:] = -1
    fs.end2NbrSlot[:] = -1
    fs.end2NbrSide[:] = -1
    ts.monomerCount[:] = [2, 64, 32, 32]
    ts.seedNode[:] = [-1, 0, -1, -1]
    fs.end1NbrSlot[1], fs.end1NbrSide[1] = 2, 1
    fs.end2NbrSlot[2], fs.end2NbrSide[2] = 1, 0
    ts.interval = 1
    ts.depolyParams[0] = 1.0          # every pointed tip shrinks (slot 0 is below the floor: it dies)
    ts.proxy = 0
    ts.growthOn = 0
    ts.agingParams[:] = 0.0
    fs_n.turn = ts
    o = ffi.OracleNet(fs_n)
    o.step(0, 1)                      # the cadence fires at step 0, before the step's mechanics
    st = o.turn.ts
    assert o.turn.last[3] == 1 and o.turn.last[5] == 1
    assert st.filState[0] >= 0 and st.monomerCount[0] == 32      # reclaimed by the split child
    assert o.fs.end1NbrSlot[1] == 0 and st.monomerCount[1] == 32


def _random_chains(g, C, n_chain):
    """Reciprocal chains over a random permutation of C slots (end2 -> next.end1)."""
    e1s, e1d, e2s, e2d = _chain_store(C)
    perm = g.permutation(C)
    cuts = np.sort(g.choice(np.arange(1, C), n_chain - 1, replace=False))
    for chain in np.split(perm, cuts):
        for a, b in zip(chain[:-1], chain[1:]):
            e2s[a], e2d[a] = b, 0
            e1s[b], e1d[b] = a, 1
    return e1s, e1d, e2s, e2d


def test_death_scatter_equals_each_slot_clearing_its_own_links():
    """The GPU applies a death's link break as every slot clearing its own links to dying slots; on reciprocal
    chains that is the reference's scatter exactly, adjacent deaths included."""
    g = np.random.default_rng(5)
    for trial in range(50):
        C = 40
        e1s, e1d, e2s, e2d = _random_chains(g, C, 6)
        flag = (g.random(C) < 0.3).astype(np.int32)
        want = [a.copy() for a in (e1s, e1d, e2s, e2d)]
        state, mc, seedNode = np.zeros(C, np.int32), np.full(C, 9, np.int32), np.zeros(C, np.int32)
        _apply_death(C, state, mc, seedNode, *want, flag)
        got = [a.copy() for a in (e1s, e1d, e2s, e2d)]
        for s in range(C):
            if flag[s]:
                if e2s[s] >= 0:
                    got[2][s] = got[3][s] = -1
                if e1s[s] >= 0:
                    got[0][s] = got[1][s] = -1
                continue
            if e2s[s] >= 0 and flag[e2s[s]]:
                got[2][s] = got[3][s] = -1
            if e1s[s] >= 0 and flag[e1s[s]]:
                got[0][s] = got[1][s] = -1
        for a, b in zip(got, want):
            assert np.array_equal(a, b), trial
        assert (state[flag == 1] < 0).all() and (seedNode[flag == 1] < 0).all()


def _reciprocal(fs, live):
    """Every link of a live slot leads to a live slot whose link at the named end leads back."""
    for s in np.flatnonzero(live):
        for slot, side in ((fs.end1NbrSlot[s], fs.end1NbrSide[s]), (fs.end2NbrSlot[s], fs.end2NbrSide[s])):
            if slot < 0:
                continue
            if not live[slot]:
                return False
            back = fs.end1NbrSlot[slot] if side == 0 else fs.end2NbrSlot[slot]
            if back != s:
                return False
    return True


def test_cadence_conserves_monomers_and_keeps_chains_reciprocal():
    sc = small_network(seed=3, n_fil=60, links=120, eta=0.1)
    t = scenes.with_turnover(sc, spare=200, growth_rate=10.0, speedup=200.0, cofilin_ratio=0.35, interval=5)
    t.turn.monomerCount[t.turn.seedNode >= 0] = 60     # tips near the split threshold
    first = np.flatnonzero(t.fs.end1NbrSlot[: sc.fs.n] < 0)
    t.turn.monomerCount[first[::2]] = 4                # some pointed tips near death
    o = ffi.OracleNet(t)
    m0 = t.turn.active_monomers()
    totals = np.zeros(6, np.int64)
    for k in range(60):
        o.step(5 * k, 5)
        ts = o.turn.ts
        live = ts.filState >= 0
        taken, returned = o.turn.ledger
        assert ts.active_monomers() == m0 + taken - returned, k
        assert _reciprocal(o.fs, live), k
        assert (ts.monomerCount[~live] == 0).all()
        totals += np.array(o.turn.last)
    grew, ret, sev, deaths, dissolves, splits = totals
    assert grew > 100 and deaths > 0 and dissolves > 0 and splits > 0, totals
    assert abs(o.turn.conc - (float(t.turn.pool[0]) + (returned - taken) * t.turn.uMPerMonomer)) < 1e-9


def test_growth_fires_at_its_probability():
    """GrowthSystem.grow: each node-held tip below maxMon grows with P = onRate conc biochemDeltaT per cadence."""
    sc = small_network(seed=4, n_fil=400, links=0, eta=0.1)
    t = scenes.with_turnover(sc, spare=0, growth_rate=10.0, speedup=20.0, interval=1)
    ts = t.turn
    ts.depolyParams[0] = 0.0
    ts.depolyRateParams[:] = 0.0
    ts.uMPerMonomer = 0.0                               # the pool stays put: P is constant
    P = float(f32(ts.onRate * ts.pool[0] * ts.biochemDeltaT))
    o = ffi.OracleNet(t)
    n_tip = int((ts.seedNode >= 0).sum())
    grew = 0
    for k in range(20):
        o.step(k, 1)
        grew += o.turn.last[0]
    n = 20 * n_tip
    assert abs(grew - n * P) < 5 * math.sqrt(n * P * (1 - P)), (grew, n * P)
