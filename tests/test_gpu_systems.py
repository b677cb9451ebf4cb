"""GPU parity of every System entry point against the CPU oracle, same seeded inputs.

Integer/index outputs must be bit-exact.  Float outputs are compared for exact equality: the HIP
kernels use the reference's float/double types and operation order with -ffp-contract=off and IEEE
division/sqrt; the only non-IEEE-pinned steps are double log/sin/cos/exp, whose last-bit differences
between the device library and glibc are absorbed by the cast to float (checked here as exact; a
looser bound is reported in the message if they ever differ).
"""
import numpy as np
import pytest

from _common import assert_same, max_rel, rod_state, small_network
from oracle import ffi
from softbox_amd import scenes
from softbox_amd.systems import (BrownianForceSystem, ChainBendingForceSystem, ContainmentSystem,
                                 CrossBridgeSystem, CrosslinkerSystem, DerivedGeometrySystem, FilamentStoreKernels,
                                 MechanicsFusion, RigidRodLangevinIntegrationSystem, SpatialGrid, RNG_PHILOX,
                                 RNG_WANG)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from softbox_amd.context import Context

    c = Context(0)
    yield c
    c.close()


def dev(ctx, a):
    return ctx.to_device(a)


def host(ctx, t):
    ctx.sync()
    return ctx.to_host(t)


@pytest.mark.parametrize("rng", [RNG_WANG, RNG_PHILOX])
def test_brownian_force(ctx, rng):
    n = 6000
    fs = rod_state(n, 3).fs
    params = fs.params
    counts = np.array([n, 1234, 77, 1], np.int32)
    rf = np.zeros(3 * n, np.float32)
    rt = np.zeros(3 * n, np.float32)
    ffi.call("sbo_brownian_force", n, rf, rt, fs.bTransGam, fs.bRotGam, fs.brownTransScale, fs.brownRotScale,
             np.float32(params[1]), 1234, 77, rng, 5, 0)
    d = {k: dev(ctx, getattr(fs, k)) for k in ("bTransGam", "bRotGam", "brownTransScale", "brownRotScale")}
    grf, grt = dev(ctx, np.zeros(3 * n, np.float32)), dev(ctx, np.zeros(3 * n, np.float32))
    BrownianForceSystem.brownianForce(ctx, grf, grt, d["bTransGam"], d["bRotGam"], d["brownTransScale"],
                                      d["brownRotScale"], params, counts, rng=rng, stream=5)
    assert_same(host(ctx, grf), rf, "randForce")
    assert_same(host(ctx, grt), rt, "randTorque")


def test_force_build_zeroes_and_draws(ctx):
    n = 3000
    fs = rod_state(n, 4).fs
    counts = np.array([n, 9, 3, 1], np.int32)
    rf = np.zeros(3 * n, np.float32)
    rt = np.zeros(3 * n, np.float32)
    ffi.call("sbo_brownian_force", n, rf, rt, fs.bTransGam, fs.bRotGam, fs.brownTransScale, fs.brownRotScale,
             np.float32(fs.params[1]), 9, 3, RNG_PHILOX, 0, 0)
    t = {k: dev(ctx, getattr(fs, k)) for k in ("forceSum", "torqueSum", "randForce", "randTorque", "bTransGam",
                                               "bRotGam", "brownTransScale", "brownRotScale")}
    MechanicsFusion.forceBuild(ctx, t["forceSum"], t["torqueSum"], t["randForce"], t["randTorque"], t["bTransGam"],
                               t["bRotGam"], t["brownTransScale"], t["brownRotScale"], fs.params, counts)
    assert not host(ctx, t["forceSum"]).any() and not host(ctx, t["torqueSum"]).any()
    assert_same(host(ctx, t["randForce"]), rf, "randForce")
    assert_same(host(ctx, t["randTorque"]), rt, "randTorque")


def test_integrate_and_derive(ctx):
    n = 5000
    fs = rod_state(n, 5).fs
    o = fs.copy()
    ffi.call("sbo_integrate", n, o.coord, o.uVec, o.yVec, o.forceSum, o.torqueSum, o.randForce, o.randTorque,
             o.bTransGam, o.bRotGam, np.float32(o.params[0]))
    o_int = o.copy()
    ffi.call("sbo_derive", n, o.coord, o.uVec, o.yVec, o.zVec, o.end1, o.end2, o.segLength)
    t = {k: dev(ctx, getattr(fs, k)) for k in fs.VEC + fs.SCALAR_F}
    RigidRodLangevinIntegrationSystem.integrate(ctx, t["coord"], t["uVec"], t["yVec"], t["forceSum"],
                                                t["torqueSum"], t["randForce"], t["randTorque"], t["bTransGam"],
                                                t["bRotGam"], fs.params, fs.counts)
    for k in ("coord", "uVec", "yVec"):
        assert_same(host(ctx, t[k]), getattr(o_int, k), f"integrate {k}")
    DerivedGeometrySystem.derive(ctx, t["coord"], t["uVec"], t["yVec"], t["zVec"], t["end1"], t["end2"],
                                 t["segLength"], fs.counts)
    for k in ("yVec", "zVec", "end1", "end2"):
        assert_same(host(ctx, t[k]), getattr(o, k), f"derive {k}")


@pytest.mark.parametrize("confined", [False, True])
def test_integ_derive_fused_equals_sequential(ctx, confined):
    n = 4000
    sc = rod_state(n, 6)
    fs = sc.fs
    # push some rods through the walls so the confined form exercises the wall law
    box = 8.0
    boxParams = np.array([1.0e-4, box, box, box, 0.0035, 0.5, 1.0, fs.params[0]], np.float32)
    fs.plane("coord")[:] = ((fs.plane("coord") / 10.0) - 0.5) * (box + 1.0)
    scenes.derive_host(fs)
    o = fs.copy()
    if confined:
        ffi.call("sbo_confine", n, o.coord, o.uVec, o.segLength, o.bTransGam, o.forceSum, o.torqueSum, boxParams, 0)
    ffi.call("sbo_integrate", n, o.coord, o.uVec, o.yVec, o.forceSum, o.torqueSum, o.randForce, o.randTorque,
             o.bTransGam, o.bRotGam, np.float32(o.params[0]))
    ffi.call("sbo_derive", n, o.coord, o.uVec, o.yVec, o.zVec, o.end1, o.end2, o.segLength)
    t = {k: dev(ctx, getattr(fs, k)) for k in fs.VEC + fs.SCALAR_F}
    args = (t["coord"], t["uVec"], t["yVec"], t["zVec"], t["segLength"], t["end1"], t["end2"], t["forceSum"],
            t["torqueSum"], t["randForce"], t["randTorque"], t["bTransGam"], t["bRotGam"])
    if confined:
        MechanicsFusion.integDeriveConfined(ctx, *args, boxParams, np.array([n, 0, 0, 1], np.int32))
    else:
        MechanicsFusion.integDerive(ctx, *args, fs.params)
    for k in ("coord", "uVec", "yVec", "zVec", "end1", "end2", "forceSum", "torqueSum"):
        assert_same(host(ctx, t[k]), getattr(o, k), f"integDerive {k}")


def test_confine_standalone(ctx):
    n = 3000
    fs = rod_state(n, 7).fs
    box = 6.0
    fs.plane("coord")[:] = ((fs.plane("coord") / 10.0) - 0.5) * (box + 1.5)
    boxParams = np.array([1.0e-4, box, box, box, 0.0035, 0.5, 4.0, 1e-5], np.float32)
    for step in (8, 9):   # cadence step fires, off-cadence step is a no-op
        o = fs.copy()
        ffi.call("sbo_confine", n, o.coord, o.uVec, o.segLength, o.bTransGam, o.forceSum, o.torqueSum, boxParams,
                 step)
        t = {k: dev(ctx, getattr(fs, k)) for k in ("coord", "uVec", "segLength", "bTransGam", "forceSum",
                                                   "torqueSum")}
        ContainmentSystem.confine(ctx, t["coord"], t["uVec"], t["segLength"], t["bTransGam"], t["forceSum"],
                                  t["torqueSum"], boxParams, np.array([n, step, 0, 1], np.int32))
        assert_same(host(ctx, t["forceSum"]), o.forceSum, f"confine forceSum step {step}")
        assert_same(host(ctx, t["torqueSum"]), o.torqueSum, f"confine torqueSum step {step}")
        if step == 9:
            assert_same(o.forceSum, fs.forceSum, "off-cadence is a no-op")


@pytest.mark.parametrize("spring", [0.0, 1.0])
def test_chain_forces(ctx, spring):
    sc = small_network(seed=2, links=0)
    fs = sc.fs
    g = np.random.default_rng(9)
    # bend the chains so every joint has strain and torsion
    fs.plane("coord")[:] += (g.standard_normal((3, fs.n)) * 0.01).astype(np.float32)
    u = fs.plane("uVec") + (g.standard_normal((3, fs.n)) * 0.2)
    fs.plane("uVec")[:] = (u / np.linalg.norm(u, axis=0)).astype(np.float32)
    scenes.derive_host(fs)
    cp = fs.chainParams.copy()
    cp[4] = spring
    cp[5] = 2.7e-20
    fs.forceSum[:] = (g.standard_normal(3 * fs.n) * 1e-13).astype(np.float32)
    o = fs.copy()
    ffi.call("sbo_chain_forces", fs.n, o.coord, o.uVec, o.segLength, o.end2NbrSlot, o.end2NbrSide, o.end1NbrSlot,
             o.end1NbrSide, o.bTransGam, o.bRotGam, o.forceSum, o.torqueSum, cp)
    t = {k: dev(ctx, getattr(fs, k)) for k in fs.VEC + fs.SCALAR_F + fs.SCALAR_I}
    ChainBendingForceSystem.chainForces(ctx, t["coord"], t["uVec"], t["segLength"], t["end2NbrSlot"],
                                        t["end2NbrSide"], t["end1NbrSlot"], t["end1NbrSide"], t["bTransGam"],
                                        t["bRotGam"], t["forceSum"], t["torqueSum"], cp, fs.counts)
    gf, gt = host(ctx, t["forceSum"]), host(ctx, t["torqueSum"])
    assert max_rel(gf, o.forceSum, 1e-20) <= 1e-6 and max_rel(gt, o.torqueSum, 1e-30) <= 1e-6
    assert_same(gf, o.forceSum, "chain forceSum")
    assert_same(gt, o.torqueSum, "chain torqueSum")


@pytest.mark.parametrize("periodic", [False, True])
def test_grid_and_broad_phase(ctx, periodic):
    sc = small_network(seed=3, links=0, periodic=periodic, n_fil=300)
    fs, p = sc.fs, sc.p
    n = fs.n
    # reference view: center = coord, radius = L/2 + r
    view = np.zeros(3 * n, np.float32)
    rad = np.zeros(n, np.float32)
    ffi.call("sbo_publish_filaments", n, n, fs.coord, fs.segLength, np.float32(p["actinRadius"]), view, rad)
    gp = np.array(p["gridParams"], np.float32)
    gd = np.array(p["gridDims"], np.int32)
    bl = np.array(p["boxLen"], np.float32)
    cells = int(gd[3])
    cell = np.zeros(n, np.int32)
    ffi.call("sbo_body_cell", n, n, view, gp, gd, int(periodic), cell)
    cnt = np.zeros(cells, np.int32)
    off = np.zeros(cells + 1, np.int32)
    cont = np.zeros(n, np.int32)
    ffi.call("sbo_grid_build_serial", n, n, cell, gd, cnt, off, cont)
    mc = 256
    cand = np.full(n * mc, -1, np.int32)
    ccount = np.zeros(n, np.int32)
    ffi.call("sbo_broad_phase", n, n, view, rad, cell, off, cont, gd, gp, int(periodic), bl, mc, cand, ccount)

    counts = np.array([0, n, 0, 0], np.int32)
    d_coord, d_len = dev(ctx, fs.coord), dev(ctx, fs.segLength)
    d_view, d_rad = dev(ctx, np.zeros(3 * n, np.float32)), dev(ctx, np.zeros(n, np.float32))
    d_os, d_sl = dev(ctx, np.zeros(n, np.int32)), dev(ctx, np.zeros(n, np.int32))
    FilamentStoreKernels.publishToBodyView(ctx, d_coord, d_len, d_view, d_rad, d_os, d_sl,
                                           np.array([p["actinRadius"]], np.float32))
    assert_same(host(ctx, d_view), view, "publish center")
    assert_same(host(ctx, d_rad), rad, "publish radius")
    d_cell = dev(ctx, np.zeros(n, np.int32))
    SpatialGrid.bodyCell(ctx, d_view, gp, gd, counts, d_cell, periodic=int(periodic))
    assert_same(host(ctx, d_cell), cell, "bodyCell")
    d_off, d_cont = dev(ctx, np.zeros(cells + 1, np.int32)), dev(ctx, np.zeros(n, np.int32))
    SpatialGrid.gridBuild(ctx, d_cell, counts, gd, d_off, d_cont)
    assert_same(host(ctx, d_off), off, "gridCellOffsets")
    assert_same(host(ctx, d_cont), cont, "gridCellContents (within-cell order)")
    d_cand, d_cc = dev(ctx, np.full(n * mc, -1, np.int32)), dev(ctx, np.zeros(n, np.int32))
    SpatialGrid.broadPhase(ctx, d_view, d_rad, d_cell, d_off, d_cont, gd, gp, counts, d_cand, d_cc, mc,
                           periodic=int(periodic))
    assert_same(host(ctx, d_cc), ccount, "candCount")
    assert_same(host(ctx, d_cand), cand, "candPartner (stencil order)")
    assert ccount.sum() > n  # non-trivial: chain neighbours + bundles


def test_csr_build(ctx):
    g = np.random.default_rng(11)
    n_keys, n_items = 5000, 40000
    key = g.integers(-1, n_keys, n_items).astype(np.int32)
    cnt = np.zeros(n_keys, np.int32)
    off = np.zeros(n_keys + 1, np.int32)
    lst = np.zeros(n_items, np.int32)
    ffi.call("sbo_csr_build", n_keys, n_items, key, cnt, off, lst)
    d_off, d_lst = dev(ctx, np.zeros(n_keys + 1, np.int32)), dev(ctx, np.zeros(n_items, np.int32))
    CrossBridgeSystem.csrBuild(ctx, dev(ctx, key), n_keys, d_off, d_lst)
    assert_same(host(ctx, d_off), off, "csr offsets")
    m = int(off[-1])
    assert_same(host(ctx, d_lst)[:m], lst[:m], "csr list")


@pytest.mark.parametrize("sizes", [(5000, 33, 32, 1), (4096, 4097, 64, 31, 2), (20000,), (40, 33, 34, 35, 36)])
def test_csr_build_dense_clusters(ctx, sizes):
    """Keys shared by many items (a dense cluster in one grid cell, many motors on one segment): short
    segments are ranked per item, segments over 32 items sorted by a workgroup in LDS (up to 4096) or
    ranked over the whole segment beyond that.  Bit-exact against the serial counting sort, with the
    items of each key scattered through the index range and mixed with unkeyed items."""
    g = np.random.default_rng(len(sizes))
    n_keys = 300
    keys = np.concatenate([np.full(sz, 7 + 13 * i, np.int32) for i, sz in enumerate(sizes)]
                          + [g.integers(-2, n_keys, 3000).astype(np.int32)])
    key = keys[g.permutation(keys.size)]
    n_items = key.size
    cnt = np.zeros(n_keys, np.int32)
    off = np.zeros(n_keys + 1, np.int32)
    lst = np.zeros(n_items, np.int32)
    ffi.call("sbo_csr_build", n_keys, n_items, key, cnt, off, lst)
    for rep in range(2):  # a second sort on the same scratch starts from its cleaned state
        d_off, d_lst = dev(ctx, np.zeros(n_keys + 1, np.int32)), dev(ctx, np.full(n_items, -7, np.int32))
        CrossBridgeSystem.csrBuild(ctx, dev(ctx, key), n_keys, d_off, d_lst)
        assert_same(host(ctx, d_off), off, f"csr offsets (sort {rep})")
        m = int(off[-1])
        assert_same(host(ctx, d_lst)[:m], lst[:m], f"csr list (sort {rep})")
    # the grid form: bodies of a dense cluster binned into few cells
    gd = np.array([4, 4, 4, 64], np.int32)
    cell = np.where(key >= 0, key % 64, -1).astype(np.int32)
    counts = np.array([0, n_items, 0, 0], np.int32)
    gcnt = np.zeros(64, np.int32)
    goff = np.zeros(65, np.int32)
    gcont = np.zeros(n_items, np.int32)
    ffi.call("sbo_grid_build_serial", n_items, n_items, cell, gd, gcnt, goff, gcont)
    d_off, d_cont = dev(ctx, np.zeros(65, np.int32)), dev(ctx, np.zeros(n_items, np.int32))
    SpatialGrid.gridBuild(ctx, dev(ctx, cell), counts, gd, d_off, d_cont)
    assert_same(host(ctx, d_off), goff, "dense gridCellOffsets")
    assert_same(host(ctx, d_cont)[: int(goff[-1])], gcont[: int(goff[-1])], "dense gridCellContents")


@pytest.mark.parametrize("rng", [RNG_WANG, RNG_PHILOX])
def test_link_systems(ctx, rng):
    sc = small_network(seed=4, links=600)
    fs, xl = sc.fs, sc.xl
    g = np.random.default_rng(12)
    fs.plane("coord")[:] += (g.standard_normal((3, fs.n)) * 0.02).astype(np.float32)
    scenes.derive_host(fs)
    xl.strainHist[:] = (g.random(xl.strainHist.size) * 0.3).astype(np.float32)
    xl.torqueMagHist[:] = (g.random(xl.torqueMagHist.size) * 1e-19).astype(np.float32)
    xl.strainPlace[:] = g.integers(0, 10, xl.strainPlace.size)
    xl.torqueMagPlace[:] = g.integers(0, 5, xl.torqueMagPlace.size)
    xl.linkOrientSame[:] = g.integers(0, 2, xl.linkOrientSame.size)
    xl.linkState[g.random(xl.nLinks) < 0.1] = -1
    offp = np.array([0.0, 0.1, 4.26, 1e-5, 0.035], np.float32)
    offp_hot = np.array([50.0, 100.0, 4.26, 1e-3, 0.035], np.float32)  # large break probability: exercise deaths
    counts = np.array([xl.nLinks, 321, 8, fs.n], np.int32)
    L, n = xl.nLinks, fs.n
    o = xl.copy()
    ffi.call("sbo_link_unbind", n, L, fs.uVec, fs.end1, o.linkFilA, o.linkFilB, o.loc1, o.loc2, o.linkState,
             o.strainHist, o.strainPlace, offp_hot, 321, 8, rng, 2)
    ffi.call("sbo_count_active_links", L, n, o.linkState, o.linkFilA, o.linkFilB, o.activeLinkCount)
    ffi.call("sbo_link_forces", n, L, fs.coord, fs.uVec, fs.end1, fs.bTransGam, o.linkFilA, o.linkFilB, o.loc1,
             o.loc2, o.activeLinkCount, o.xlinkData, xl.xlParams)
    ffi.call("sbo_link_torsion", n, L, fs.uVec, o.linkFilA, o.linkFilB, o.linkState, o.linkOrientSame,
             o.torqueMagHist, o.torqueMagPlace, o.xlinkData, xl.torsionParams)
    F = np.zeros(3 * n, np.float32)
    T = np.zeros(3 * n, np.float32)
    cA, oA, iA = np.zeros(n, np.int32), np.zeros(n + 1, np.int32), np.zeros(L, np.int32)
    cB, oB, iB = np.zeros(n, np.int32), np.zeros(n + 1, np.int32), np.zeros(L, np.int32)
    ffi.call("sbo_csr_build", n, L, o.linkFilA, cA, oA, iA)
    ffi.call("sbo_seg_gather_a", n, oA, iA, o.xlinkData, o.linkState, F, T)
    ffi.call("sbo_csr_build", n, L, o.linkFilB, cB, oB, iB)
    ffi.call("sbo_seg_gather_b", n, oB, iB, o.xlinkData, o.linkState, F, T)
    assert (o.linkState < 0).sum() > (xl.linkState < 0).sum()   # some deaths happened

    f = {k: dev(ctx, getattr(fs, k)) for k in ("coord", "uVec", "end1", "bTransGam")}
    x = {k: dev(ctx, getattr(xl, k)) for k in ("linkFilA", "linkFilB", "loc1", "loc2", "linkState", "strainHist",
                                               "strainPlace", "linkOrientSame", "torqueMagHist", "torqueMagPlace",
                                               "xlinkData", "activeLinkCount")}
    CrosslinkerSystem.unbind(ctx, f["coord"], f["uVec"], f["end1"], x["linkFilA"], x["linkFilB"], x["loc1"],
                             x["loc2"], x["linkState"], x["strainHist"], x["strainPlace"], offp_hot, counts, rng=rng,
                             stream=2)
    assert_same(host(ctx, x["linkState"]), o.linkState, "unbind linkState (death events)")
    assert_same(host(ctx, x["strainHist"]), o.strainHist, "unbind strain ring")
    assert_same(host(ctx, x["strainPlace"]), o.strainPlace, "unbind strain place")
    CrosslinkerSystem.countActiveLinks(ctx, x["linkState"], x["linkFilA"], x["linkFilB"], x["activeLinkCount"])
    assert_same(host(ctx, x["activeLinkCount"]), o.activeLinkCount, "activeLinkCount")
    CrosslinkerSystem.linkForces(ctx, f["coord"], f["uVec"], f["end1"], f["bTransGam"], x["linkFilA"],
                                 x["linkFilB"], x["loc1"], x["loc2"], x["activeLinkCount"], x["xlinkData"],
                                 xl.xlParams)
    CrosslinkerSystem.linkTorsion(ctx, f["uVec"], x["linkFilA"], x["linkFilB"], x["linkState"],
                                  x["linkOrientSame"], x["torqueMagHist"], x["torqueMagPlace"], x["xlinkData"],
                                  xl.torsionParams)
    assert_same(host(ctx, x["xlinkData"]), o.xlinkData, "xlinkData payload")
    assert_same(host(ctx, x["torqueMagHist"]), o.torqueMagHist, "torque ring")
    dA, dAi = dev(ctx, np.zeros(n + 1, np.int32)), dev(ctx, np.zeros(L, np.int32))
    dB, dBi = dev(ctx, np.zeros(n + 1, np.int32)), dev(ctx, np.zeros(L, np.int32))
    CrossBridgeSystem.csrBuild(ctx, x["linkFilA"], n, dA, dAi)
    CrossBridgeSystem.csrBuild(ctx, x["linkFilB"], n, dB, dBi)
    dF, dT = dev(ctx, np.zeros(3 * n, np.float32)), dev(ctx, np.zeros(3 * n, np.float32))
    CrosslinkerSystem.segGatherA(ctx, dA, dAi, x["xlinkData"], x["linkState"], dF, dT)
    CrosslinkerSystem.segGatherB(ctx, dB, dBi, x["xlinkData"], x["linkState"], dF, dT)
    assert_same(host(ctx, dF), F, "gather forceSum")
    assert_same(host(ctx, dT), T, "gather torqueSum")
