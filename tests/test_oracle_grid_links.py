"""Oracle pins for the grid/broad phase, the CSR inverse and the crosslink laws (CPU only).

Broad phase: the candidate set equals the O(N^2) brute force (SpatialGrid.bruteForce, the
reference's own gate, BroadPhaseHarness.java:419-433).  CSR: the textbook stable counting sort
invariants.  Crosslinks: the reference's analytic gates (CrosslinkerHarness.java:275-347, 445-527).
"""
import math

import numpy as np
import pytest

from _common import assert_same, small_network
from oracle import ffi
from softbox_amd import scenes
from softbox_amd.stores import CrosslinkerStore, FilamentStore


def _grid_run(sc):
    fs, p = sc.fs, sc.p
    n = fs.n
    view = np.zeros(3 * n, np.float32)
    rad = np.zeros(n, np.float32)
    ffi.call("sbo_publish_filaments", n, n, fs.coord, fs.segLength, np.float32(p["actinRadius"]), view, rad)
    gp = np.array(p["gridParams"], np.float32)
    gd = np.array(p["gridDims"], np.int32)
    bl = np.array(p["boxLen"], np.float32)
    per = int(p["periodic"])
    cell = np.zeros(n, np.int32)
    ffi.call("sbo_body_cell", n, n, view, gp, gd, per, cell)
    cells = int(gd[3])
    cnt, off, cont = np.zeros(cells, np.int32), np.zeros(cells + 1, np.int32), np.zeros(n, np.int32)
    ffi.call("sbo_grid_build_serial", n, n, cell, gd, cnt, off, cont)
    mc = int(p["maxCand"])
    cand, cc = np.full(n * mc, -1, np.int32), np.zeros(n, np.int32)
    ffi.call("sbo_broad_phase", n, n, view, rad, cell, off, cont, gd, gp, per, bl, mc, cand, cc)
    bcand, bcc = np.full(n * mc, -1, np.int32), np.zeros(n, np.int32)
    ffi.call("sbo_brute_force", n, n, view, rad, gp, per, bl, mc, bcand, bcc)
    return dict(view=view, rad=rad, cell=cell, off=off, cont=cont, cand=cand, cc=cc, bcand=bcand, bcc=bcc, mc=mc)


def _pairs(cand, cc, mc):
    out = set()
    for i in np.flatnonzero(cc):
        for j in cand[i * mc: i * mc + min(cc[i], mc)]:
            out.add((int(i), int(j)))
    return out


@pytest.mark.parametrize("periodic", [False, True])
def test_broad_phase_equals_brute_force(periodic):
    sc = small_network(seed=21, n_fil=250, links=0, periodic=periodic)
    # jitter so rods straddle cell faces and (periodic) the box faces
    g = np.random.default_rng(3)
    sc.fs.plane("coord")[:] += (g.standard_normal((3, sc.fs.n)) * 0.05).astype(np.float32)
    r = _grid_run(sc)
    assert_same(r["cc"], r["bcc"], "candCount vs brute")
    assert _pairs(r["cand"], r["cc"], r["mc"]) == _pairs(r["bcand"], r["bcc"], r["mc"])
    assert r["cc"].sum() > sc.fs.n
    if periodic:
        # some pairs are only neighbours through a periodic face
        v = r["view"].reshape(3, -1)
        box = sc.box
        cross = [(i, j) for (i, j) in _pairs(r["cand"], r["cc"], r["mc"])
                 if np.abs(v[:, i] - v[:, j]).max() > 0.5 * box]
        assert cross


def test_grid_csr_invariants():
    sc = small_network(seed=22, n_fil=300, links=0, periodic=False)
    r = _grid_run(sc)
    off, cont, cell = r["off"], r["cont"], r["cell"]
    n = sc.fs.n
    assert off[0] == 0 and off[-1] == n and np.all(np.diff(off) >= 0)
    assert np.array_equal(np.sort(cont), np.arange(n))                 # every body exactly once
    for c in np.flatnonzero(np.diff(off)):
        seg = cont[off[c]: off[c + 1]]
        assert np.all(cell[seg] == c) and np.all(np.diff(seg) > 0)     # right cell, ascending (stable)


def test_broad_phase_overflow_keeps_counting():
    """A dense cluster: candCount exceeds MAX_CAND, the slice holds the first MAX_CAND in stencil order
    (SpatialGrid.java:396-398) and the brute reference agrees on the count."""
    sc = scenes.free_rods(n=600, length=0.1, box=0.3, seed=5, periodic=False)
    sc.p.update(scenes.grid_for(0.3, 0.0535, 0.04, False, margin=0.2), gridOn=1)
    sc.p["maxCand"] = 64
    r = _grid_run(sc)
    assert r["cc"].max() > 64
    assert_same(r["cc"], r["bcc"], "overflowing counts")
    i = int(np.argmax(r["cc"]))
    got = r["cand"][i * 64:(i + 1) * 64]
    assert len(set(got.tolist())) == 64 and np.all(got > i)


def test_csr_inverse_stable_and_skips_negative_keys():
    g = np.random.default_rng(4)
    nk, ni = 300, 5000
    key = g.integers(-2, nk, ni).astype(np.int32)
    cnt, off, lst = np.zeros(nk, np.int32), np.zeros(nk + 1, np.int32), np.zeros(ni, np.int32)
    ffi.call("sbo_csr_build", nk, ni, key, cnt, off, lst)
    valid = np.flatnonzero(key >= 0)
    assert off[-1] == valid.size
    # stable counting sort == argsort by key with a stable algorithm
    want = valid[np.argsort(key[valid], kind="stable")]
    assert_same(lst[: valid.size], want, "CSR list")
    assert_same(np.diff(off), np.bincount(key[valid], minlength=nk).astype(np.int32), "CSR counts")


# ---------------------------------------------------------------- crosslink laws
def _two_rods(sep, dt=1e-4, eta=0.1):
    """Two parallel single-rod filaments along x, separated along z (CrosslinkerHarness.buildScene
    geometry): one link between their midpoints."""
    fs = FilamentStore(2)
    for k in range(2):
        fs.plane("uVec")[:, k] = (1, 0, 0)
        fs.plane("yVec")[:, k] = (0, 1, 0)
    fs.plane("coord")[:, 1] = (0, 0, np.float32(sep))
    from softbox_amd.physics import set_rod_drag

    set_rod_drag(fs, 0.0891, 0.0035, eta)
    scenes.derive_host(fs)
    xl = CrosslinkerStore(1, 2)
    xl.set_link(0, 0, 0.5 * float(fs.segLength[0]), 1, 0.5 * float(fs.segLength[1]))
    xl.linkOrientSame[0] = 1
    xl.compute_fil_link_ct()
    return fs, xl


def _link_step(fs, xl, rest, dt, frac=0.4):
    n, L = fs.n, xl.nLinks
    xlp = np.array([rest, frac, dt], np.float32)
    ffi.call("sbo_zero_accumulators", n, fs.forceSum, fs.torqueSum)
    ffi.call("sbo_link_forces", n, L, fs.coord, fs.uVec, fs.end1, fs.bTransGam, xl.linkFilA, xl.linkFilB, xl.loc1,
             xl.loc2, xl.filLinkCt, xl.xlinkData, xlp)
    cA, oA, iA = np.zeros(n, np.int32), np.zeros(n + 1, np.int32), np.zeros(L, np.int32)
    ffi.call("sbo_csr_build", n, L, xl.linkFilA, cA, oA, iA)
    ffi.call("sbo_seg_gather_a", n, oA, iA, xl.xlinkData, xl.linkState, fs.forceSum, fs.torqueSum)
    ffi.call("sbo_csr_build", n, L, xl.linkFilB, cA, oA, iA)
    ffi.call("sbo_seg_gather_b", n, oA, iA, xl.xlinkData, xl.linkState, fs.forceSum, fs.torqueSum)
    ffi.call("sbo_integrate", n, fs.coord, fs.uVec, fs.yVec, fs.forceSum, fs.torqueSum, fs.randForce, fs.randTorque,
             fs.bTransGam, fs.bRotGam, np.float32(dt))
    ffi.call("sbo_derive", n, fs.coord, fs.uVec, fs.yVec, fs.zVec, fs.end1, fs.end2, fs.segLength)


def test_link_exact_rest_hold():
    """CrosslinkerHarness.java:276-284: sep = restLength, Brownian off -> zero force, zero motion."""
    rest = 0.0125
    fs, xl = _two_rods(rest)
    c0 = fs.coord.copy()
    for _ in range(50):
        _link_step(fs, xl, rest, 1e-4)
    # |p2-p1| rounds to restLength within float; stretch = max(len-rest, 0) may be a few ulps
    assert np.abs(fs.coord - c0).max() < 1e-9


@pytest.mark.parametrize("dt", [1e-4, 1e-5])
def test_link_stretch_relaxation_decay(dt):
    """CrosslinkerHarness.java:300-347: per-step decay k = fracMove*(gamma_par/gamma_perp)*L with L the
    current separation (v1's non-normalised linkVec), independent of dt, within 0.1%."""
    rest = 0.0125
    eps0 = 0.05 * rest
    fs, xl = _two_rods(rest + eps0, dt)
    gpar, gperp = float(fs.plane("bTransGam")[0, 0]), float(fs.plane("bTransGam")[1, 0])
    eps = [float(fs.plane("coord")[2, 1] - fs.plane("coord")[2, 0]) - rest]
    for _ in range(30):
        _link_step(fs, xl, rest, dt)
        eps.append(float(fs.plane("coord")[2, 1] - fs.plane("coord")[2, 0]) - rest)
    ks, kp = [], []
    for e, e1 in zip(eps[:-1], eps[1:]):
        if e < 1e-5 or e1 <= 0:
            continue
        ks.append((e - e1) / e)
        kp.append(0.4 * (gpar / gperp) * (e + rest))
    assert len(ks) > 5
    assert abs(np.mean(ks) / np.mean(kp) - 1) < 1e-3


def test_link_payload_equal_and_opposite_and_momentum():
    sc = small_network(seed=23, n_fil=120, links=400)
    fs, xl = sc.fs, sc.xl
    g = np.random.default_rng(6)
    fs.plane("coord")[:] += (g.standard_normal((3, fs.n)) * 0.03).astype(np.float32)
    scenes.derive_host(fs)
    n, L = fs.n, xl.nLinks
    ffi.call("sbo_link_forces", n, L, fs.coord, fs.uVec, fs.end1, fs.bTransGam, xl.linkFilA, xl.linkFilB, xl.loc1,
             xl.loc2, xl.filLinkCt, xl.xlinkData, xl.xlParams)
    d = xl.xlinkData.reshape(-1, 12)
    assert np.abs(d[:, 0:3]).max() > 0
    assert_same(d[:, 6:9], -d[:, 0:3], "B force = -A force")
    F = np.zeros(3 * n, np.float32)
    T = np.zeros(3 * n, np.float32)
    ffi.call("sbo_link_brute_gather", n, L, xl.linkFilA, xl.linkFilB, xl.xlinkData, xl.linkState, F, T)
    F2, T2 = np.zeros_like(F), np.zeros_like(T)
    cA, oA, iA = np.zeros(n, np.int32), np.zeros(n + 1, np.int32), np.zeros(L, np.int32)
    cB, oB, iB = np.zeros(n, np.int32), np.zeros(n + 1, np.int32), np.zeros(L, np.int32)
    ffi.call("sbo_csr_build", n, L, xl.linkFilA, cA, oA, iA)
    ffi.call("sbo_seg_gather_a", n, oA, iA, xl.xlinkData, xl.linkState, F2, T2)
    ffi.call("sbo_csr_build", n, L, xl.linkFilB, cB, oB, iB)
    ffi.call("sbo_seg_gather_b", n, oB, iB, xl.xlinkData, xl.linkState, F2, T2)
    # two-pass gather: A-side sum then B-side sum (two float roundings) vs the brute one-pass sum
    assert np.allclose(F2, F, rtol=1e-6, atol=1e-30)
    tot = F.reshape(3, n).astype(np.float64).sum(1)
    assert np.abs(tot).max() < 1e-5 * np.abs(F).max()


@pytest.mark.parametrize("th", [0.5, 0.2])
def test_link_torsion_closed_form(th):
    """Constant geometry: after the 5-slot ring fills, |T| = spring * angle and the axis is
    unit(uA x uB) (parallel) / unit(uA x -uB) with angle |acos - pi| (antiparallel).  The angle is
    v1's fastAcos (CrosslinkerSystem.java:372-376): exact acos for |dot| <= 0.95, sqrt(2(1-dot))
    above (th = 0.2 lands there)."""
    fs = FilamentStore(2)
    fs.plane("uVec")[:, 0] = (1, 0, 0)
    fs.plane("uVec")[:, 1] = (math.cos(th), math.sin(th), 0)
    xl = CrosslinkerStore(2, 2)
    xl.set_link(0, 0, 0.0, 1, 0.0)
    xl.set_link(1, 0, 0.0, 1, 0.0)
    xl.linkOrientSame[:] = (1, 0)
    spring = 1e-19
    tp = np.array([spring, 1.0], np.float32)
    for _ in range(5):
        xl.xlinkData[:] = 0.0
        ffi.call("sbo_link_torsion", 2, 2, fs.uVec, xl.linkFilA, xl.linkFilB, xl.linkState, xl.linkOrientSame,
                 xl.torqueMagHist, xl.torqueMagPlace, xl.xlinkData, tp)
    d = xl.xlinkData.reshape(2, 12)
    dot = float(np.float32(math.cos(th)))
    ang = math.acos(dot) if dot <= 0.95 else math.sqrt(2 * (1 - dot))
    assert np.allclose(d[0, 3:6], [0, 0, spring * ang], rtol=1e-5, atol=1e-30)
    assert np.allclose(d[0, 9:12], [0, 0, -spring * ang], rtol=1e-5, atol=1e-30)
    # antiparallel: axis uA x (-uB) = -z, angle pi - fastAcos(dot)
    assert np.allclose(d[1, 3:6], [0, 0, -spring * (math.pi - ang)], rtol=1e-5, atol=1e-30)


@pytest.mark.parametrize("rng", [0, 1])
def test_link_empirical_off_rate(rng):
    """CrosslinkerHarness.java:496-527: frozen pose, a population of links at a fixed strain; after the
    boxcar fills, deaths/step = (c0 + c1 exp(c2 s)) dt within 5%."""
    rest = 0.0125
    N, dt, warm, M = 20000, 1e-4, 30, 1500
    for s in (0.0, 1.0):
        fs, _ = _two_rods(rest * (1 + s))
        xl = CrosslinkerStore(N, 2)
        for k in range(N):
            xl.set_link(k, 0, 0.5 * float(fs.segLength[0]), 1, 0.5 * float(fs.segLength[1]))
        offp = np.array([1.0, 1.0, 2.0, dt, rest], np.float32)
        deaths = active_steps = 0
        active = N
        for t in range(warm + M):
            ffi.call("sbo_link_unbind", 2, N, fs.uVec, fs.end1, xl.linkFilA, xl.linkFilB, xl.loc1, xl.loc2,
                     xl.linkState, xl.strainHist, xl.strainPlace, offp, t, 1234 + int(s * 1000), rng, 0)
            after = int((xl.linkState >= 0).sum())
            if t >= warm:
                deaths += active - after
                active_steps += active
            active = after
        p_emp = deaths / active_steps
        p_exp = (1.0 + math.exp(2.0 * s)) * dt
        assert abs(p_emp / p_exp - 1) < 0.05, (s, p_emp, p_exp, deaths)
