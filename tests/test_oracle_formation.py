"""Pins of the oracle's crosslinker formation (CPU only).

- FilIDSystem's pointer doubling against a plain walk of each chain to its end2 terminal.
- The grid candidate query (gridFormCount/Scan/Emit) against the reference's own O(N^2) reference
  filFilCandidates (CrosslinkerSystem.java:405-431, restated in the oracle) and against an independent
  numpy brute force: the same pairs in the same lexicographic (i, j) order, box and periodic.
- The FORM_MAXC cap keeping the smallest partners, and the request-capacity clamp.
- formGates on hand-built geometry: the closest approach of two skew segments in closed form, the
  alignment aperture and its modes, orientSame, the jitter bounds, and P_form's empirical rate.
- formAdmit's rules (one new link per segment, saturation, spacing) and the scan-rank allocator
  (lowest free slots in rank order, overflow clamp, fresh rings).
"""
import math

import numpy as np
import pytest

from oracle import ffi
from softbox_amd import scenes
from softbox_amd.scenes import ACTIN_MONO_DIAM


def _filid(fs, filState=None):
    n = fs.n
    out = np.zeros(n, np.int32)
    scratch = np.zeros(n, np.int32)
    ffi.call("sbo_filid", n, filState, fs.end2NbrSlot, out, scratch)
    return out


def _walk_filid(end2, filState=None):
    n = len(end2)
    out = np.zeros(n, np.int64)
    for s in range(n):
        if filState is not None and filState[s] < 0:
            out[s] = -s - 2
            continue
        cur, seen = s, 0
        while True:
            nx = end2[cur]
            if nx < 0 or nx == cur or (filState is not None and filState[nx] < 0):
                break
            cur = nx
            seen += 1
            assert seen <= n
        out[s] = cur
    return out


def test_filid_is_the_chain_terminal():
    sc = scenes.network(n_fil=37, segs=10, box=3.0, n_links=0, seed=2, grid=False)
    fs = sc.fs
    assert np.array_equal(_filid(fs), _walk_filid(fs.end2NbrSlot))
    # irregular topology: chains of uneven length, a lone rod, FREE slots that cut chains
    n = 200
    g = np.random.default_rng(0)
    end2 = np.full(n, -1, np.int32)
    perm = g.permutation(n)
    pos = 0
    while pos < n:
        L = int(g.integers(1, 40))
        chain = perm[pos:pos + L]
        end2[chain[:-1]] = chain[1:]
        pos += L
    state = np.zeros(n, np.int32)
    state[g.random(n) < 0.1] = -1
    fs2 = type(fs)(n)
    fs2.end2NbrSlot[:] = end2
    assert np.array_equal(_filid(fs2, state), _walk_filid(end2, state))
    assert np.array_equal(_filid(fs2), _walk_filid(end2))
    # rounds: ceil(log2 n) rounded up to even (FilIDSystem.java:31-34)
    for k, want in ((1, 2), (2, 2), (3, 2), (4, 2), (5, 4), (16, 4), (17, 6), (200000, 18)):
        assert ffi.load().sbo_filid_rounds(k) == want, k


def _formation_scene(periodic, n_fil=60, box=1.4, seed=1, grab=0.04):
    sc = scenes.network(n_fil=n_fil, segs=10, box=box, n_links=40, seed=seed, periodic=periodic, seed_frac=1.0)
    p = dict(sc.p, **scenes.formation_params(box, 0.1, sc.fs.n, periodic, grab=grab))
    return sc, p


def _grid_candidates(sc, p, req_cap=None):
    fs = sc.fs
    n = fs.n
    rc = req_cap or int(p["formReqCap"])
    filID = _filid(fs)
    gp = np.array(p["formGridParams"], np.float32)
    gd = np.array(p["formGridDims"], np.int32)
    cells = int(gd[3])
    cell = np.zeros(n, np.int32)
    per = int(p["periodic"])
    bl = np.array(p["boxLen"], np.float32)
    ffi.call("sbo_body_cell", n, n, fs.coord, gp, gd, per, cell)
    cnt, off, cont = np.zeros(cells, np.int32), np.zeros(cells + 1, np.int32), np.zeros(n, np.int32)
    ffi.call("sbo_grid_build_serial", n, n, cell, gd, cnt, off, cont)
    fp = np.array(p["formParams"], np.float32)
    cc, cb = np.zeros(n, np.int32), np.zeros(n + 1, np.int32)
    ra, rb = np.zeros(rc, np.int32), np.zeros(rc, np.int32)
    counts = np.zeros(4, np.int32)
    ffi.call("sbo_grid_form_count", n, fs.coord, fs.segLength, filID, gp, gd, off, cont, cc, fp, per, bl)
    ffi.call("sbo_grid_form_scan", n, cc, cb, rc, ra, rb, counts)
    ffi.call("sbo_grid_form_emit", n, fs.coord, fs.segLength, filID, gp, gd, off, cont, cb, cc, rc, ra, rb, fp,
             per, bl)
    return ra, rb, int(counts[0]), filID


def _numpy_pairs(fs, filID, grab, periodic, box):
    n = fs.n
    c = fs.plane("coord").astype(np.float64)
    L = fs.segLength.astype(np.float64)
    pairs = []
    for i in range(n):
        d = c[:, i:i + 1] - c[:, i + 1:]
        if periodic:
            d = d - box * np.rint(d / box)
        cd = np.sqrt((d * d).sum(0))
        ok = (cd <= 0.5 * L[i] + 0.5 * L[i + 1:] + grab) & (filID[i + 1:] != filID[i])
        pairs.extend((i, i + 1 + int(j)) for j in np.flatnonzero(ok))
    return pairs


@pytest.mark.parametrize("periodic", [False, True])
def test_grid_candidates_equal_the_reference_brute_force(periodic):
    sc, p = _formation_scene(periodic)
    fs = sc.fs
    ra, rb, total, filID = _grid_candidates(sc, p)
    assert total > 20
    rc = int(p["formReqCap"])
    ba, bb = np.zeros(rc, np.int32), np.zeros(rc, np.int32)
    counts = np.zeros(4, np.int32)
    ffi.call("sbo_fil_fil_candidates", fs.n, fs.coord, fs.segLength, filID, rc, ba, bb,
             np.array(p["formParams"], np.float32), counts, int(p["periodic"]), np.array(p["boxLen"], np.float32))
    assert int(counts[0]) == total
    assert np.array_equal(ra, ba) and np.array_equal(rb, bb)     # order and the -1 tail included
    want = _numpy_pairs(fs, filID, 0.04, periodic, 1.4)
    assert list(zip(ra[:total].tolist(), rb[:total].tolist())) == want
    assert np.all(ra[total:] == -1) and np.all(rb[total:] == -1)


def test_candidate_cap_and_request_clamp():
    """300 short segments of distinct filaments at one point: segment i has 299 - i partners; the region
    keeps the FORM_MAXC = 256 smallest (CrosslinkerSystem.java:583-593).  A request capacity below the
    total clamps the list (gridFormScan :524-527, gridFormEmit :540-542)."""
    n = 300
    fs = scenes.FilamentStore(n)
    fs.segLength[:] = 0.01
    fs.plane("coord")[:] = 0.5
    fs.end1NbrSlot[:] = fs.end2NbrSlot[:] = -1
    fs.end1NbrSide[:] = fs.end2NbrSide[:] = -1
    p = dict(scenes.formation_params(2.0, 0.01, n, False), periodic=0, boxLen=[0.0] * 3)
    sc = scenes.Scene("cluster", fs, None, p, 2.0, 0)
    ra, rb, total, _ = _grid_candidates(sc, p, req_cap=n * 256)
    want_counts = [min(299 - i, 256) for i in range(n)]
    assert total == sum(want_counts)
    for i in (0, 10, 43, 44, 298):
        b = sum(want_counts[:i])
        k = want_counts[i]
        assert np.all(ra[b:b + k] == i)
        assert np.array_equal(rb[b:b + k], np.arange(i + 1, i + 1 + k))
    ra2, rb2, total2, _ = _grid_candidates(sc, p, req_cap=1000)
    assert total2 == 1000
    assert np.array_equal(ra2[:1000], ra[:1000]) and np.array_equal(rb2[:1000], rb[:1000])


# ---------------------------------------------------------------- formGates
def _planar(rows):
    return np.ascontiguousarray(np.asarray(rows, np.float64).T.ravel(), np.float32)


def _gates(segs, pairs, max_angle=0.6, grab=0.04, p_form=1.0, jitter=0.0, mode=0, step=7, seed=3):
    """segs: [(centre, u, length)]; pairs: [(a, b)] -> (gatePass, loc1, loc2, orient)."""
    n = len(segs)
    c = _planar([s[0] for s in segs])
    u = _planar([s[1] for s in segs])
    L = np.array([s[2] for s in segs], np.float32)
    e1 = (c.reshape(3, n) - 0.5 * L * u.reshape(3, n)).astype(np.float32).ravel()
    e2 = (c.reshape(3, n) + 0.5 * L * u.reshape(3, n)).astype(np.float32).ravel()
    ra = np.array([a for a, _ in pairs], np.int32)
    rb = np.array([b for _, b in pairs], np.int32)
    k = len(pairs)
    l1, l2 = np.zeros(k, np.float32), np.zeros(k, np.float32)
    orient, gate = np.zeros(k, np.int32), np.zeros(k, np.int32)
    fp = np.array([max_angle, grab * grab, 0.027, 10, p_form, jitter, grab], np.float32)
    ffi.call("sbo_form_gates", n, k, c, u, e1, e2, L, ra, rb, l1, l2, orient, gate, fp, step, seed, mode, 0, 0, 0,
             np.zeros(3, np.float32))
    return gate, l1, l2, orient


def _skew(theta, d, sign=1.0):
    """Segment a along x through the origin; b through (0, 0, d) at angle theta in a plane parallel to xy
    (antiparallel when sign < 0).  Their common perpendicular is the z axis: closest approach d, at the
    middle of both."""
    a = ((0.0, 0.0, 0.0), (1.0, 0.0, 0.0), 0.1)
    b = ((0.0, 0.0, d), (sign * math.cos(theta), sign * math.sin(theta), 0.0), 0.1)
    return [a, b]


def test_gates_closest_approach_closed_form():
    for d, want in ((0.03, 1), (0.0399, 1), (0.0401, 0), (0.08, 0)):
        gate, l1, l2, orient = _gates(_skew(0.3, d), [(0, 1)])
        assert gate[0] == want, d
        if want:
            # the crossing sits at the middle of both: loc = half the length (jitter 0)
            assert abs(l1[0] - 0.05) < 2e-7 and abs(l2[0] - 0.05) < 2e-7
            assert orient[0] == 1


def test_gates_alignment_modes_and_orientation():
    par, anti = _skew(0.3, 0.02), _skew(0.3, 0.02, sign=-1.0)
    wide = _skew(0.9, 0.02)
    # b along (cos, sin, 0) with theta = 0.3: the cross product r1 x r2 points to +z (kept); the reference's
    # "parallel" test rejects a cross product with no positive component (:634), as the antiparallel
    # pair here has (r1 x r2).z < 0
    g, *_, o = _gates(par, [(0, 1)], mode=0)
    assert g[0] == 1 and o[0] == 1
    g, *_, o = _gates(par, [(0, 1)], mode=-1)
    assert g[0] == 0
    g, *_ = _gates(anti, [(0, 1)], mode=0)
    assert g[0] == 0                                     # the reference's sign quirk of the parallel test
    anti_up = _skew(-0.3, 0.02, sign=-1.0)               # antiparallel with (r1 x r2).z > 0
    g, *_, o = _gates(anti_up, [(0, 1)], mode=0)
    assert g[0] == 1 and o[0] == 0
    g, *_ = _gates(anti_up, [(0, 1)], mode=1)
    assert g[0] == 0
    g, *_ = _gates(anti_up, [(0, 1)], mode=-1)
    assert g[0] == 1
    g, *_ = _gates(wide, [(0, 1)], mode=0)               # 0.9 rad outside the 0.6 aperture
    assert g[0] == 0


def test_gates_jitter_bounds_and_pform_rate():
    k = 4000
    segs = _skew(0.3, 0.02) * 1
    pairs = [(0, 1)] * k
    jit = 2.0 * ACTIN_MONO_DIAM
    gate, l1, l2, _ = _gates(segs, pairs, jitter=jit, p_form=0.25, step=11)
    # every candidate index draws its own jitter: uniform in [loc - jit, loc + jit]
    assert np.all(np.abs(l1 - 0.05) <= jit * (1 + 1e-6)) and np.all(np.abs(l2 - 0.05) <= jit * (1 + 1e-6))
    assert abs(float(np.mean(l1)) - 0.05) < 4 * jit / math.sqrt(3 * k)
    assert np.std(l1) > 0.5 * jit / math.sqrt(3)
    rate = gate.mean()
    assert abs(rate - 0.25) < 5 * math.sqrt(0.25 * 0.75 / k)


# ---------------------------------------------------------------- admission + allocation
def test_admission_rules():
    """Candidates 0..4 all pass their gates.  Segment 2 is claimed by candidates 1 and 3: only the lower
    index (1) may form, and only if its other segment also names it; segment 5 is saturated; candidate 4
    sits within minSep of an existing link on segment 7."""
    nSeg, C = 10, 4
    ra = np.array([0, 2, 4, 2, 7], np.int32)
    rb = np.array([1, 3, 5, 6, 8], np.int32)
    l1 = np.array([0.05, 0.05, 0.05, 0.05, 0.050], np.float32)
    l2 = np.full(5, 0.05, np.float32)
    gate = np.ones(5, np.int32)
    minc = np.zeros(nSeg, np.int32)
    ffi.call("sbo_form_admit_reduce", 5, 99, nSeg, ra, rb, gate, minc)
    assert minc.tolist() == [0, 0, 1, 1, 2, 2, 3, 4, 4, 99]
    act = np.zeros(nSeg, np.int32)
    act[5] = 10                                          # saturated (maxLinksOnSeg 10)
    state = np.array([0, 0, -1, 0], np.int32)
    fa = np.array([7, 9, 7, 9], np.int32)
    fb = np.array([9, 9, 9, 9], np.int32)
    lc1 = np.array([0.06, 0.0, 0.05, 0.0], np.float32)   # an active link 0.01 from candidate 4's site on 7
    lc2 = np.zeros(C, np.float32)
    acc = np.zeros(5, np.int32)
    fp = np.array([0.6, 0.0016, 0.027, 10, 1.0, 0.0, 0.04], np.float32)
    ffi.call("sbo_form_admit", 5, C, ra, rb, l1, l2, gate, minc, act, state, fa, fb, lc1, lc2, acc, fp)
    assert acc.tolist() == [1, 1, 0, 0, 0]
    # the same link but dead: it no longer blocks candidate 4
    state[0] = -1
    ffi.call("sbo_form_admit", 5, C, ra, rb, l1, l2, gate, minc, act, state, fa, fb, lc1, lc2, acc, fp)
    assert acc.tolist() == [1, 1, 0, 0, 1]


def test_allocator_ranks_free_slots_and_clamps():
    C = 6
    state = np.array([0, -1, 0, -1, -1, 0], np.int32)    # free slots 1, 3, 4
    accept = np.array([0, 1, 1, 0, 1, 1, 1], np.int32)   # 5 accepted: only the first 3 ranks can form
    K = len(accept)
    ra = np.arange(K, dtype=np.int32) + 100
    rb = np.arange(K, dtype=np.int32) + 200
    l1 = np.arange(K, dtype=np.float32) * 0.01
    l2 = l1 + 0.5
    orient = np.array([1, 0, 1, 1, 0, 1, 0], np.int32)
    fa, fb = np.full(C, 9, np.int32), np.full(C, 9, np.int32)
    lc1, lc2 = np.zeros(C, np.float32), np.zeros(C, np.float32)
    sh = np.full(C * 10, 0.7, np.float32)
    sp = np.full(C, 4, np.int32)
    los = np.full(C, 5, np.int32)
    th = np.full(C * 5, 0.3, np.float32)
    tp = np.full(C, 2, np.int32)
    free_count, free_off, free_list = np.zeros(C, np.int32), np.zeros(C + 1, np.int32), np.zeros(C, np.int32)
    rank = np.zeros(K + 1, np.int32)
    formed = ffi.call("sbo_form_allocate", K, C, accept, ra, rb, l1, l2, orient, free_count, free_off, free_list,
                      rank, fa, fb, lc1, lc2, state, sh, sp, los, th, tp)
    assert formed == 3
    assert free_list[:3].tolist() == [1, 3, 4]
    # requests 1, 2, 4 (ranks 0, 1, 2) land in slots 1, 3, 4; requests 5, 6 overflow
    for req, slot in ((1, 1), (2, 3), (4, 4)):
        assert fa[slot] == 100 + req and fb[slot] == 200 + req and state[slot] == 0
        assert lc1[slot] == l1[req] and lc2[slot] == l2[req] and los[slot] == orient[req]
        assert np.all(sh[slot * 10:(slot + 1) * 10] == 0) and sp[slot] == 0
        assert np.all(th[slot * 5:(slot + 1) * 5] == 0) and tp[slot] == 0
    for slot in (0, 2, 5):                               # occupied slots untouched
        assert fa[slot] == 9 and np.all(sh[slot * 10:(slot + 1) * 10] == np.float32(0.7))


def test_network_with_formation_forms_links_and_stays_consistent():
    """The whole pass inside the oracle's network step: links form every 100 steps into freed slots, each
    formed link joins two segments of different filaments at most grab apart, no segment exceeds the
    saturation bound, and no segment gains two links in one pass."""
    sc = scenes.network(n_fil=120, segs=10, box=2.0, n_links=600, seed=5, periodic=True, eta=0.1,
                        bell=(5.0, 50.0, 4.26), seed_frac=1.0)
    sc.p.update(scenes.formation_params(2.0, 0.1, sc.fs.n, True, on_rate=4000.0))
    o = ffi.OracleNet(sc)
    state0 = o.xl.linkState.copy()
    formed_total = 0
    filID = None
    for t in range(0, 301):
        before = o.xl.linkState.copy()
        fa_before, fb_before = o.xl.linkFilA.copy(), o.xl.linkFilB.copy()
        o.step(t, 1)
        if t % 100 == 0:
            k = o.form.s.lastFormed
            formed_total += k
            slots = o.form.a["freeList"][:k]              # rank r formed into freeList[r]
            assert np.all(before[slots] < 0)
            # a link formed in this pass may also die in this step's unbind; none rose otherwise
            new = (before < 0) & (o.xl.linkState >= 0)
            assert set(np.flatnonzero(new).tolist()) <= set(slots.tolist())
            filID = o.form.a["filID"]
            segs = np.concatenate([o.xl.linkFilA[slots], o.xl.linkFilB[slots]])
            assert len(np.unique(segs)) == len(segs)       # one new link per segment per pass
            assert np.all(filID[o.xl.linkFilA[slots]] != filID[o.xl.linkFilB[slots]])
        else:
            assert not np.any((before < 0) & (o.xl.linkState >= 0))
    assert formed_total > 0
    act = np.bincount(np.concatenate([o.xl.linkFilA[o.xl.linkState >= 0], o.xl.linkFilB[o.xl.linkState >= 0]]),
                      minlength=o.fs.n)
    assert act.max() <= 10 + 1                            # saturation is checked against start-of-pass counts
    assert not np.array_equal(state0, o.xl.linkState)
