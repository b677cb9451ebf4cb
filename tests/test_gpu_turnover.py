"""Actin turnover on the GPU against the oracle, bit for bit.

  - every per-System entry point (aging, depoly / depolyProxy, applyDeath, grow, growthAtp, cofilin
    accumulate / dissolve / fresh, markSplits, the filament-slot allocator, splitWire, splitInheritNuc,
    recomputeDrag), chained as one cadence would chain them, both random streams;
  - the fused cadence (sbx_turnover_cadence) over 30 cadences: every state array, the pool and its ledger;
  - the network step with turnover at its cadence (and with formation, and the actomyosin step): pose,
    forces, links and turnover state every few cadences, eager and captured.
"""
import ctypes as C

import numpy as np
import pytest

from _common import assert_same, small_network
from oracle import ffi
from softbox_amd import _lib, scenes
from softbox_amd.systems import RNG_PHILOX, RNG_WANG

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from softbox_amd.context import Context

    c = Context(0)
    yield c
    c.close()


def _scene(seed=3, spare=150, speedup=50.0, periodic=True, rng=RNG_PHILOX, interval=5, links=120, n_fil=60,
           formation=False):
    sc = small_network(seed=seed, n_fil=n_fil, links=links, eta=0.1, periodic=periodic, rng=rng)
    if formation:
        sc.p.update(scenes.formation_params(sc.box, 0.1, sc.fs.n, periodic, on_rate=400.0, interval=10))
    t = scenes.with_turnover(sc, spare=spare, growth_rate=10.0, speedup=speedup, cofilin_ratio=0.35,
                             interval=interval)
    ts = t.turn
    g = np.random.default_rng(seed)
    tips = np.flatnonzero(ts.seedNode >= 0)
    ts.monomerCount[tips] = g.integers(55, 64, tips.size)        # tips near the split threshold
    pointed = np.flatnonzero((t.fs.end1NbrSlot < 0) & (ts.filState >= 0))
    ts.monomerCount[pointed[::3]] = g.integers(2, 6, pointed[::3].size)   # pointed tips near death
    n = t.fs.n
    nf = g.random((3, n))
    nf[2] *= 3.0                                                         # ADP-rich, aged segments
    nf /= nf.sum(0, keepdims=True)
    ts.nucFrac[:] = nf.astype(np.float32).ravel()
    ts.cofFrac[:] = (g.random(n) * 0.36).astype(np.float32)             # a few near the dissolve ratio
    return t


def _arrays(t):
    fs, ts = t.fs, t.turn
    n = fs.n
    a = dict(coord=fs.coord.copy(), uVec=fs.uVec.copy(), yVec=fs.yVec.copy(), segLength=fs.segLength.copy(),
             bTransGam=fs.bTransGam.copy(), bRotGam=fs.bRotGam.copy(), bTransDiff=fs.bTransDiff.copy(),
             bRotDiff=fs.bRotDiff.copy(), bts=fs.brownTransScale.copy(), brs=fs.brownRotScale.copy(),
             e1s=fs.end1NbrSlot.copy(), e1d=fs.end1NbrSide.copy(), e2s=fs.end2NbrSlot.copy(),
             e2d=fs.end2NbrSide.copy(), filState=ts.filState.copy(), mc=ts.monomerCount.copy(),
             seedNode=ts.seedNode.copy(), nucFrac=ts.nucFrac.copy(), cofFrac=ts.cofFrac.copy())
    g = np.random.default_rng(11)
    nf = g.random((3, n)).astype(np.float32)
    nf /= nf.sum(0, keepdims=True)
    a["nucFrac"] = nf.ravel().astype(np.float32)
    a["cofFrac"] = (g.random(n) * 0.5).astype(np.float32)
    for k in ("returnedMon", "deathFlag", "grewFlag", "sevFlag", "sevRet", "acceptFlag"):
        a[k] = np.zeros(n, np.int32)
    for k in ("freeCount", "freeOffsets", "freeList", "rankOffsets"):
        a[k] = np.zeros(n + 1, np.int32)
    for k in ("reqCoord", "reqUVec", "reqYVec"):
        a[k] = np.zeros(3 * n, np.float32)
    return a


def _fp(v):
    return np.ascontiguousarray(v, np.float32)


@pytest.mark.parametrize("rng", [RNG_PHILOX, RNG_WANG])
def test_turnover_systems_match_oracle(ctx, rng):
    from softbox_amd.context import ptr

    t = _scene(rng=rng)
    ts = t.turn
    n = t.fs.n
    H = _arrays(t)
    D = {k: ctx.to_device(v) for k, v in H.items()}
    P = lambda k: ptr(D[k])  # noqa: E731
    lib = ctx.lib
    step, seed = 40, 9

    def check(what):
        ctx.sync()
        for k, v in H.items():
            assert_same(ctx.to_host(D[k]), v, f"{what}: {k}")

    fp = {k: _fp(getattr(ts, k)) for k in ("agingParams", "depolyParams", "depolyRateParams", "growParams",
                                           "splitParams", "dragParams", "cofilinParams", "birthParams")}
    fp["growParams"][0] = 0.4
    fp["depolyParams"][0] = 0.3
    F = lambda k: fp[k].ctypes.data_as(_lib.f32p)  # noqa: E731

    ctx.check(lib.sbx_aging_age(ctx.h, P("filState"), P("nucFrac"), F("agingParams"), 1, n))
    ffi.call("sbo_aging_age", n, H["filState"], H["nucFrac"], fp["agingParams"], 1)
    check("age")
    for proxy in (False, True):
        ctx.check(lib.sbx_depoly(ctx.h, P("filState"), P("mc"), P("coord"), P("uVec"), P("e1s"),
                                 P("nucFrac") if proxy else None, P("returnedMon"), P("deathFlag"), F("depolyParams"),
                                 F("depolyRateParams"), step + proxy, seed, 1, n, rng, 0))
        ffi.call("sbo_depoly", n, H["filState"], H["mc"], H["coord"], H["uVec"], H["e1s"],
                 H["nucFrac"] if proxy else None, H["returnedMon"], H["deathFlag"], fp["depolyParams"],
                 fp["depolyRateParams"], step + proxy, seed, rng, 0, 1)
        check(f"depoly proxy={proxy}")
    assert H["deathFlag"].sum() > 0 and H["returnedMon"].sum() > H["deathFlag"].sum()
    ctx.check(lib.sbx_apply_death(ctx.h, P("filState"), P("mc"), P("seedNode"), P("e1s"), P("e1d"), P("e2s"),
                                  P("e2d"), P("bts"), P("brs"), P("deathFlag"), n))
    ffi.call("sbo_apply_death", n, H["filState"], H["mc"], H["seedNode"], H["e1s"], H["e1d"], H["e2s"], H["e2d"],
             H["bts"], H["brs"], H["deathFlag"])
    check("applyDeath")
    ctx.check(lib.sbx_growth_grow(ctx.h, P("seedNode"), P("mc"), P("coord"), P("uVec"), P("grewFlag"),
                                  F("growParams"), step, seed, 1, n, rng, 0))
    ffi.call("sbo_grow", n, H["seedNode"], H["mc"], H["coord"], H["uVec"], H["grewFlag"], fp["growParams"], step,
             seed, rng, 0, 1)
    check("grow")
    assert H["grewFlag"].sum() > 0
    ctx.check(lib.sbx_aging_growth_atp(ctx.h, P("grewFlag"), P("mc"), P("nucFrac"), n))
    ffi.call("sbo_aging_growth_atp", n, H["grewFlag"], H["mc"], H["nucFrac"])
    check("growthAtp")
    ctx.check(lib.sbx_cofilin_accumulate(ctx.h, P("filState"), P("nucFrac"), P("cofFrac"), F("cofilinParams"), 1, n))
    ffi.call("sbo_cofilin_accumulate", n, H["filState"], H["nucFrac"], H["cofFrac"], fp["cofilinParams"], 1)
    check("cofilinAccumulate")
    ctx.check(lib.sbx_cofilin_dissolve(ctx.h, P("filState"), P("mc"), P("cofFrac"), P("sevFlag"), P("sevRet"),
                                       F("cofilinParams"), 1, n))
    ffi.call("sbo_cofilin_dissolve", n, H["filState"], H["mc"], H["cofFrac"], H["sevFlag"], H["sevRet"],
             fp["cofilinParams"], 1)
    check("cofilinDissolve")
    assert H["sevFlag"].sum() > 0
    ctx.check(lib.sbx_apply_death(ctx.h, P("filState"), P("mc"), P("seedNode"), P("e1s"), P("e1d"), P("e2s"),
                                  P("e2d"), P("bts"), P("brs"), P("sevFlag"), n))
    ffi.call("sbo_apply_death", n, H["filState"], H["mc"], H["seedNode"], H["e1s"], H["e1d"], H["e2s"], H["e2d"],
             H["bts"], H["brs"], H["sevFlag"])
    check("applyDeath (dissolve)")
    # force some tips to the threshold so splits happen
    tips = np.flatnonzero(H["seedNode"] >= 0)[:40]
    H["mc"][tips] = 64
    D["mc"].copy_(ctx.to_device(H["mc"]))
    ctx.check(lib.sbx_growth_mark_splits(ctx.h, P("seedNode"), P("mc"), P("coord"), P("uVec"), P("yVec"),
                                         P("acceptFlag"), P("reqCoord"), P("reqUVec"), P("reqYVec"), F("splitParams"),
                                         1, n))
    ffi.call("sbo_mark_splits", n, H["seedNode"], H["mc"], H["coord"], H["uVec"], H["yVec"], H["acceptFlag"],
             H["reqCoord"], H["reqUVec"], H["reqYVec"], n, fp["splitParams"], 1)
    check("markSplits")
    ctx.check(lib.sbx_birth_alloc(ctx.h, P("filState"), P("freeCount"), P("freeOffsets"), P("freeList"),
                                  P("acceptFlag"), P("rankOffsets"), P("reqCoord"), P("reqUVec"), P("reqYVec"),
                                  P("coord"), P("uVec"), P("yVec"), P("bts"), P("brs"), F("birthParams"), n, n, n))
    ffi.call("sbo_birth_alloc", n, n, H["filState"], H["freeCount"], H["freeOffsets"], H["freeList"],
             H["acceptFlag"], H["rankOffsets"], H["reqCoord"], H["reqUVec"], H["reqYVec"], n, H["coord"], H["uVec"],
             H["yVec"], H["bts"], H["brs"], fp["birthParams"])
    for k in ("freeList",):   # only the first nFree entries of the free list are defined
        nf = int(H["freeOffsets"][n])
        dl = ctx.to_host(D[k])
        assert_same(dl[:nf], H[k][:nf], "alloc: freeList")
        H[k][nf:] = dl[nf:]
    check("allocate")
    assert H["rankOffsets"][n] > 0
    ctx.check(lib.sbx_growth_split_wire(ctx.h, P("rankOffsets"), P("freeList"), P("freeOffsets"), P("mc"), P("coord"),
                                        P("uVec"), P("e1s"), P("e1d"), P("e2s"), P("e2d"), P("seedNode"),
                                        F("splitParams"), n, n))
    ffi.call("sbo_split_wire", n, n, H["rankOffsets"], H["freeList"], H["freeOffsets"], H["mc"], H["coord"], H["uVec"],
             H["e1s"], H["e1d"], H["e2s"], H["e2d"], H["seedNode"], fp["splitParams"])
    check("splitWire")
    ctx.check(lib.sbx_aging_split_inherit_nuc(ctx.h, P("rankOffsets"), P("freeList"), P("freeOffsets"), P("nucFrac"),
                                              n, n))
    ffi.call("sbo_aging_split_inherit_nuc", n, n, H["rankOffsets"], H["freeList"], H["freeOffsets"], H["nucFrac"])
    ctx.check(lib.sbx_cofilin_fresh(ctx.h, P("rankOffsets"), P("freeList"), P("freeOffsets"), P("cofFrac"), n, n))
    ffi.call("sbo_nucleate_fresh_cofilin", n, n, H["rankOffsets"], H["freeList"], H["freeOffsets"], H["cofFrac"])
    check("splitInheritNuc + nucleateFreshCofilin")
    ctx.check(lib.sbx_growth_recompute_drag(ctx.h, P("mc"), P("segLength"), P("e1s"), P("e2s"), P("bTransGam"),
                                            P("bRotGam"), P("bTransDiff"), P("bRotDiff"), F("dragParams"), 1, n))
    ffi.call("sbo_recompute_drag", n, H["mc"], H["segLength"], H["e1s"], H["e2s"], H["bTransGam"], H["bRotGam"],
             H["bTransDiff"], H["bRotDiff"], fp["dragParams"], 1)
    check("recomputeDrag")


def _dev_filaments(ctx, fs):
    from softbox_amd.stores import DeviceFilaments

    return DeviceFilaments(ctx, fs)


def _compare_turnover(ctx, dt_dev, dfil, orc, what):
    from softbox_amd.turnover import TurnoverStore

    ts = dt_dev.pull(TurnoverStore(dt_dev.C))
    ot = orc.turn
    for k in ("filState", "monomerCount", "seedNode", "nucFrac", "cofFrac"):
        assert_same(getattr(ts, k), getattr(ot.ts, k), f"{what}: {k}")
    assert ts.pool[0] == ot.conc, f"{what}: pool {ts.pool[0]} vs {ot.conc}"
    assert tuple(ts.ledger) == ot.ledger, f"{what}: ledger"
    assert ts.stats[:6].tolist() == ot.last, f"{what}: stats {ts.stats[:6].tolist()} vs {ot.last}"


@pytest.mark.parametrize("rng", [RNG_PHILOX, RNG_WANG])
def test_turnover_cadence_matches_oracle(ctx, rng):
    from softbox_amd.turnover import DeviceTurnover

    t = _scene(rng=rng, seed=5)
    orc = ffi.OracleNet(t)
    fs = t.fs.copy()
    dfil = _dev_filaments(ctx, fs)
    dts = DeviceTurnover(ctx, t.turn, diff=(fs.bTransDiff, fs.bRotDiff))
    fstruct, tstruct, params = dfil.struct(), dts.struct(), t.turn.params()
    totals = np.zeros(6, np.int64)
    for k in range(30):
        step = 5 * k
        ctx.check(ctx.lib.sbx_turnover_cadence(ctx.h, C.byref(fstruct), C.byref(tstruct), C.byref(params), step, rng, 0))
        orc.cadence(step)
        totals += np.array(orc.turn.last)
        if k % 5 == 4:
            _compare_turnover(ctx, dts, dfil, orc, f"cadence {k}")
            got = fs.copy()
            dfil.pull(got, ("coord", "uVec", "yVec", "segLength", "bTransGam", "bRotGam", "brownTransScale",
                            "brownRotScale", "end1NbrSlot", "end1NbrSide", "end2NbrSlot", "end2NbrSide"))
            for key in ("coord", "uVec", "yVec", "segLength", "bTransGam", "bRotGam", "brownTransScale",
                        "brownRotScale", "end1NbrSlot", "end1NbrSide", "end2NbrSlot", "end2NbrSide"):
                assert_same(getattr(got, key), getattr(orc.fs, key), f"cadence {k}: {key}")
            assert_same(ctx.to_host(dts.t["bTransDiff"]), orc.fs.bTransDiff, f"cadence {k}: bTransDiff")
    assert (totals[[0, 1, 3, 4, 5]] > 0).all(), totals


def _compare_net(net, orc, sc, what, motors=False):
    from softbox_amd.turnover import TurnoverStore

    fs = sc.fs.copy()
    xl = sc.xl.copy() if sc.xl is not None else None
    net.pull(fs, xl, shape=True)
    for k in ("coord", "uVec", "yVec", "zVec", "end1", "end2", "forceSum", "torqueSum") + net.SHAPE:
        assert_same(getattr(fs, k), getattr(orc.fs, k), f"{what}: {k}")
    if xl is not None:
        for k in ("linkState", "linkFilA", "linkFilB", "strainHist", "strainPlace", "xlinkData", "activeLinkCount"):
            assert_same(getattr(xl, k), getattr(orc.xl, k), f"{what}: {k}")
    ts = net.pull_turnover(TurnoverStore(sc.fs.n), fs)
    for k in ("filState", "monomerCount", "seedNode", "nucFrac", "cofFrac"):
        assert_same(getattr(ts, k), getattr(orc.turn.ts, k), f"{what}: {k}")
    assert ts.pool[0] == orc.turn.conc and tuple(ts.ledger) == orc.turn.ledger, what
    if motors:
        mot = sc.mot.copy()
        net.pull_motors(mot, None)
        assert_same(mot.boundSeg, orc.mot.boundSeg, f"{what}: motor boundSeg")
        assert_same(mot.body.coord, orc.mot.body.coord, f"{what}: motor coord")


@pytest.mark.parametrize("periodic,rng,graph,formation", [(True, RNG_PHILOX, 0, False), (False, RNG_WANG, 1, False),
                                                          (True, RNG_WANG, 2, True), (False, RNG_PHILOX, 1, True)])
def test_network_with_turnover_matches_oracle(ctx, periodic, rng, graph, formation):
    from softbox_amd.network import Network

    sc = _scene(seed=7, periodic=periodic, rng=rng, formation=formation)
    orc = ffi.OracleNet(sc)
    net = Network(ctx, sc)
    if graph:
        net.use_graph(graph)
    t = 0
    for chunk in (3, 9, 1, 12, 7, 13, 25, 30):
        orc.step(t, chunk)
        net.step(t, chunk)
        t += chunk
        _compare_net(net, orc, sc, f"periodic={periodic} rng={rng} graph={graph} formation={formation} after {t}")
    assert orc.turn.ledger[0] > 0 and orc.turn.ledger[1] > 0
    net.close()


def test_actomyosin_with_turnover_matches_oracle(ctx):
    from softbox_amd.network import Network

    sc = scenes.contractile(n_fil=60, segs=10, box=3.0, n_mini=12, dimers_each_end=4, n_links=150, seed=13,
                            mini_near_fil=True, eta=1.0)
    sc = scenes.with_turnover(sc, spare=120, growth_rate=10.0, speedup=150.0, cofilin_ratio=0.35, interval=4)
    orc = ffi.OracleContractile(sc)
    net = Network(ctx, sc)
    t = 0
    for chunk in (5, 11, 16, 20):
        orc.step(t, chunk)
        net.step(t, chunk)
        t += chunk
        _compare_net(net, orc, sc, f"actomyosin after {t}", motors=True)
    net.close()


def test_frames_of_a_gpu_run_equal_the_oracles(ctx, tmp_path):
    """Frame output of the network with turnover, written at a cadence from the GPU's state, is byte-identical
    to the frames written from the oracle's state at the same steps."""
    from softbox_amd.frames import FrameWriter, write_network_frame
    from softbox_amd.network import Network

    sc = _scene(seed=9)
    orc = ffi.OracleNet(sc)
    net = Network(ctx, sc)
    wg = FrameWriter(str(tmp_path / "gpu"), sc.box, sc.box, sc.box)
    wo = FrameWriter(str(tmp_path / "orc"), sc.box, sc.box, sc.box)
    dt = float(sc.p["params"][0])
    t = 0
    for chunk in (10, 10, 15):
        net.step(t, chunk)
        orc.step(t, chunk)
        t += chunk
        a = write_network_frame(wg, net, sc, t * dt)
        b = wo.write_frame(orc.fs, t * dt, orc.turn.ts)
        assert open(a).read() == open(b).read(), f"frame at step {t}"
    net.close()
