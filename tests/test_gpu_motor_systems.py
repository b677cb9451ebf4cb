"""GPU parity of every myosin System entry point against the oracle, on the same inputs.

The inputs are a seeded actomyosin scene advanced a few steps by the oracle (so motors are bound,
cooling down and cycling).  Every System is then called once on the GPU and once in the oracle on
identical copies of that state; integer outputs must match exactly and floats bit-for-bit.
"""
import numpy as np
import pytest

from _common import assert_same
from oracle import ffi
from softbox_amd import scenes
from softbox_amd.systems import (BindingDetectionSystem, CrossBridgeSystem, DimerCouplingSystem, MiniFilamentSystem,
                                 MotorJointSystem, MotorStoreKernels, NucleotideCycleSystem, RNG_PHILOX, RNG_WANG)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from softbox_amd.context import Context

    c = Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def state():
    sc = scenes.contractile(n_fil=60, segs=10, box=2.5, n_mini=24, dimers_each_end=4, n_links=150, seed=11,
                            mini_near_fil=True)
    orc = ffi.OracleContractile(sc)
    for t in range(30):
        orc.step(t, 1)
    return orc


def D(ctx, a):
    return ctx.to_device(np.ascontiguousarray(a))


def H(ctx, t):
    ctx.sync()
    return ctx.to_host(t)


def test_publish_heads_and_view(ctx, state):
    o = state
    mot, nM, n = o.mot, o.mot.nM, o.fs.n
    b = mot.body
    head, hu, ru = (np.zeros(3 * nM, np.float32) for _ in range(3))
    ffi.call("sbo_publish_head_from_body", nM, b.coord, b.uVec, b.segLength, head, hu, ru)
    dh, dhu, dru = D(ctx, np.zeros_like(head)), D(ctx, np.zeros_like(hu)), D(ctx, np.zeros_like(ru))
    MotorStoreKernels.publishHeadFromBody(ctx, D(ctx, b.coord), D(ctx, b.uVec), D(ctx, b.segLength), dh, dhu, dru,
                                          np.array([nM, 0, 0, n], np.int32))
    assert_same(H(ctx, dh), head, "head")
    assert_same(H(ctx, dhu), hu, "head uVec")
    assert_same(H(ctx, dru), ru, "rod uVec")
    cap = n + nM
    center, br = np.zeros(3 * cap, np.float32), np.zeros(cap, np.float32)
    os_, sl = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
    ffi.call("sbo_publish_motors", nM, cap, n, head, mot.reach, center, br, os_, sl)
    dc, db, dos, dsl = D(ctx, np.zeros_like(center)), D(ctx, np.zeros_like(br)), D(ctx, np.zeros_like(os_)), \
        D(ctx, np.zeros_like(sl))
    MotorStoreKernels.publishToBodyView(ctx, dh, D(ctx, mot.reach), dc, db, dos, dsl, np.array([n], np.int32),
                                        np.array([nM], np.int32))
    assert_same(H(ctx, dc), center, "center")
    assert_same(H(ctx, db), br, "boundingRadius")
    assert_same(H(ctx, dos), os_, "ownerStore")
    assert_same(H(ctx, dsl), sl, "ownerSlot")


def test_grid_reachable_and_bind_nearest(ctx, state):
    o = state
    mot, nM, n = o.mot, o.mot.nM, o.fs.n
    mc = int(o.p["maxCand"])
    gp = np.array(o.p["gridParams"], np.float32)
    gd = np.array(o.p["gridDims"], np.int32)
    g = o.grid
    rs, rc = np.full(nM * mc, -1, np.int32), np.zeros(nM, np.int32)
    ffi.call("sbo_grid_reachable", nM, n, mot.head, mot.uVec, mot.rodUVec, o.fs.end1, o.fs.end2, gp, gd,
             g["gridCellOffsets"], g["gridCellContents"], o.mreach["ownerStore"], o.mreach["ownerSlot"], mc, rs, rc,
             mot.kinParams)
    assert rc.sum() > 0, "no motor in reach: the scene does not exercise binding"
    drs, drc = D(ctx, np.full_like(rs, -1)), D(ctx, np.zeros_like(rc))
    dhead, dhu, dru = D(ctx, mot.head), D(ctx, mot.uVec), D(ctx, mot.rodUVec)
    de1, de2 = D(ctx, o.fs.end1), D(ctx, o.fs.end2)
    BindingDetectionSystem.gridReachable(ctx, dhead, dhu, dru, de1, de2, gp, gd, D(ctx, g["gridCellOffsets"]),
                                         D(ctx, g["gridCellContents"]), D(ctx, o.mreach["ownerStore"]),
                                         D(ctx, o.mreach["ownerSlot"]), drs, drc, mot.kinParams,
                                         np.array([nM, 0, 0, n], np.int32))
    got_rc = H(ctx, drc)
    assert_same(got_rc, rc, "reachCount")
    got_rs = H(ctx, drs).reshape(nM, mc)
    for m in np.flatnonzero(rc):
        c = min(int(rc[m]), mc)
        assert_same(got_rs[m, :c], rs.reshape(nM, mc)[m, :c], f"reachSeg[{m}]")
    # every motor free: all motors with candidates bind to their nearest segment
    bs0 = np.full(nM, -1, np.int32)
    bs0[::7] = -2                                      # some in cooldown: must not bind
    arc0 = np.zeros(nM, np.float32)
    bs, arc = bs0.copy(), arc0.copy()
    ffi.call("sbo_bind_nearest", nM, n, mot.head, mot.uVec, mot.rodUVec, o.fs.end1, o.fs.end2, rs, rc, mc, bs, arc,
             mot.kinParams)
    dbs, darc = D(ctx, bs0), D(ctx, arc0)
    BindingDetectionSystem.bindNearest(ctx, dhead, dhu, dru, de1, de2, drs, drc, dbs, darc, mot.kinParams,
                                       np.array([nM, 0, 0, n], np.int32))
    assert (bs >= 0).sum() > 0
    assert_same(H(ctx, dbs), bs, "boundSeg")
    assert_same(H(ctx, darc), arc, "bindArc")


def test_brute_and_node_aware_binding(ctx, state):
    """bruteReachable (segment order) and the node-aware pair (BindingDetectionSystem.java:410-473) with a third of
    the in-reach segments held by a node: candidates and bindings equal the oracle's."""
    o = state
    mot, nM, n = o.mot, o.mot.nM, o.fs.n
    mc = int(o.p["maxCand"])
    counts = np.array([nM, 0, 0, n], np.int32)
    kp = mot.kinParams.copy()
    kp[7] = np.float32(0.06)                         # a 60 nm reach: many heads with several candidates
    dhead, dhu, dru = D(ctx, mot.head), D(ctx, mot.uVec), D(ctx, mot.rodUVec)
    de1, de2 = D(ctx, o.fs.end1), D(ctx, o.fs.end2)
    rs, rc = np.full(nM * mc, -1, np.int32), np.zeros(nM, np.int32)
    ffi.call("sbo_brute_reachable", nM, n, mot.head, mot.uVec, mot.rodUVec, o.fs.end1, o.fs.end2, mc, rs, rc,
             kp)
    drs, drc = D(ctx, np.full_like(rs, -1)), D(ctx, np.zeros_like(rc))
    BindingDetectionSystem.bruteReachable(ctx, dhead, dhu, dru, de1, de2, drs, drc, kp, counts)
    assert rc.sum() > 0
    assert_same(H(ctx, drc), rc, "brute reachCount")
    for m in np.flatnonzero(rc):
        c = min(int(rc[m]), mc)
        assert_same(H(ctx, drs).reshape(nM, mc)[m, :c], rs.reshape(nM, mc)[m, :c], f"brute reachSeg[{m}]")
    seedNode = np.full(n, -1, np.int32)
    held = np.unique(rs[rs >= 0])[::3]
    seedNode[held] = 0
    rs2, rc2 = np.full(nM * mc, -1, np.int32), np.zeros(nM, np.int32)
    ffi.call("sbo_brute_reachable_node_aware", nM, n, mot.head, mot.uVec, mot.rodUVec, o.fs.end1, o.fs.end2, mc, rs2,
             rc2, seedNode, kp)
    dsn = D(ctx, seedNode)
    drs2, drc2 = D(ctx, np.full_like(rs2, -1)), D(ctx, np.zeros_like(rc2))
    BindingDetectionSystem.bruteReachableNodeAware(ctx, dhead, dhu, dru, de1, de2, drs2, drc2, dsn, kp,
                                                   counts)
    assert rc2.sum() < rc.sum()
    assert_same(H(ctx, drc2), rc2, "node-aware reachCount")
    for m in np.flatnonzero(rc2):
        c = min(int(rc2[m]), mc)
        assert_same(H(ctx, drs2).reshape(nM, mc)[m, :c], rs2.reshape(nM, mc)[m, :c], f"node-aware reachSeg[{m}]")
    bs0 = np.full(nM, -1, np.int32)
    arc0 = np.zeros(nM, np.float32)
    bs, arc = bs0.copy(), arc0.copy()
    ffi.call("sbo_bind_nearest_node_aware", nM, n, mot.head, mot.uVec, mot.rodUVec, o.fs.end1, o.fs.end2, rs, rc, mc,
             bs, arc, seedNode, kp)
    dbs, darc = D(ctx, bs0), D(ctx, arc0)
    BindingDetectionSystem.bindNearestNodeAware(ctx, dhead, dhu, dru, de1, de2, D(ctx, rs), D(ctx, rc), dbs, darc,
                                                dsn, kp, counts)
    assert (bs >= 0).sum() > 0 and not np.isin(bs[bs >= 0], held).any()
    assert_same(H(ctx, dbs), bs, "node-aware boundSeg")
    assert_same(H(ctx, darc), arc, "node-aware bindArc")


@pytest.mark.parametrize("rng", [RNG_PHILOX, RNG_WANG])
def test_catch_slip_and_cycle(ctx, state, rng):
    o = state
    mot, nM = o.mot, o.mot.nM
    g = np.random.default_rng(5)
    bs0 = np.where(g.random(nM) < 0.6, g.integers(0, o.fs.n, nM), -1).astype(np.int32)
    bs0[g.random(nM) < 0.15] = -2
    fdot = (g.standard_normal(nM) * 4e-12).astype(np.float32)
    fmag = (np.abs(g.standard_normal(nM)) * 10e-12).astype(np.float32)   # some above the 12 pN cap
    cool0 = g.integers(0, 3, nM).astype(np.int32)
    kp = mot.kinParams.copy()
    kp[0] = 5.0e4                                            # fast off-rate so the Bell branch fires
    kp[13] = 0.5                                             # refractory block probability < 1: RFRA draw
    kp[10] = 2.0
    step, seed = 17, 1234
    bs, cool = bs0.copy(), cool0.copy()
    stats, cap = np.zeros(2 * nM, np.int32), np.zeros(nM, np.int32)
    ffi.call("sbo_catch_slip_release", nM, bs, fdot, fmag, cool, stats, cap, kp, step, seed, rng, 0)
    dbs, dcool, dstats, dcap = D(ctx, bs0), D(ctx, cool0), D(ctx, np.zeros_like(stats)), D(ctx, np.zeros_like(cap))
    NucleotideCycleSystem.catchSlipRelease(ctx, dbs, D(ctx, fdot), D(ctx, fmag), dcool, dstats, dcap, kp,
                                           np.array([nM, step, seed], np.int32), rng=rng)
    assert stats[1::2].sum() > 0 and cap.sum() > 0, "catch-slip and cap branches not both exercised"
    assert_same(H(ctx, dbs), bs, "boundSeg")
    assert_same(H(ctx, dcool), cool, "cooldown")
    assert_same(H(ctx, dstats), stats, "stats")
    assert_same(H(ctx, dcap), cap, "capStats")
    nuc0 = g.integers(0, 4, nM).astype(np.int32)
    hist = (g.standard_normal(10 * nM) * 1e-12).astype(np.float32)
    npar = mot.nucParams.copy()
    npar[1:] *= 20.0                                         # fast rates: every transition fires somewhere
    nuc = nuc0.copy()
    ffi.call("sbo_nucleotide_cycle", nM, nuc, bs, hist, npar, step, seed, rng, 0)
    dnuc = D(ctx, nuc0)
    NucleotideCycleSystem.cycle(ctx, dnuc, dbs, D(ctx, hist), npar, np.array([nM, step, seed], np.int32), rng=rng)
    assert (nuc != nuc0).sum() > 0
    assert_same(H(ctx, dnuc), nuc, "nucleotideState")


def test_bond_head_forcedot_gather(ctx, state):
    o = state
    mot, nM, n = o.mot, o.mot.nM, o.fs.n
    b = mot.body
    g = np.random.default_rng(9)
    bs = np.where(g.random(nM) < 0.5, g.integers(0, n, nM), -1).astype(np.int32)
    arc = (g.random(nM) * 0.1).astype(np.float32)
    nuc = g.integers(0, 4, nM).astype(np.int32)
    bond = np.zeros(13 * nM, np.float32)
    ffi.call("sbo_bond_forces", nM, n, b.coord, b.uVec, b.yVec, b.bRotGam, o.fs.coord, o.fs.uVec, o.fs.yVec,
             o.fs.bRotGam, o.fs.segLength, bs, arc, nuc, bond, mot.xbParams)
    dbond = D(ctx, np.zeros_like(bond))
    dbs = D(ctx, bs)
    CrossBridgeSystem.bondForces(ctx, D(ctx, b.coord), D(ctx, b.uVec), D(ctx, b.yVec), D(ctx, b.bRotGam),
                                 D(ctx, o.fs.coord), D(ctx, o.fs.uVec), D(ctx, o.fs.yVec), D(ctx, o.fs.bRotGam),
                                 D(ctx, o.fs.segLength), dbs, D(ctx, arc), D(ctx, nuc), dbond, mot.xbParams)
    assert np.abs(bond).max() > 0
    assert_same(H(ctx, dbond), bond, "bondData")
    F = (g.standard_normal(9 * nM) * 1e-12).astype(np.float32)
    T = (g.standard_normal(9 * nM) * 1e-19).astype(np.float32)
    Fo, To = F.copy(), T.copy()
    ffi.call("sbo_apply_head_force", nM, bond, Fo, To)
    dF, dT = D(ctx, F), D(ctx, T)
    CrossBridgeSystem.applyHeadForce(ctx, dbond, dF, dT)
    assert_same(H(ctx, dF), Fo, "head forceSum")
    assert_same(H(ctx, dT), To, "head torqueSum")
    fdf, fm = np.zeros(nM, np.float32), np.zeros(nM, np.float32)
    hist = (g.standard_normal(10 * nM) * 1e-12).astype(np.float32)
    place = g.integers(0, 10, nM).astype(np.int32)
    h2, p2 = hist.copy(), place.copy()
    ffi.call("sbo_register_force_dot", nM, bond, bs, fdf, fm, h2, p2)
    dfdf, dfm, dh, dp = D(ctx, np.zeros_like(fdf)), D(ctx, np.zeros_like(fm)), D(ctx, hist), D(ctx, place)
    CrossBridgeSystem.registerForceDot(ctx, dbond, dbs, dfdf, dfm, dh, dp)
    assert_same(H(ctx, dfdf), fdf, "forceDotFil")
    assert_same(H(ctx, dfm), fm, "forceMag")
    assert_same(H(ctx, dh), h2, "forceDotHist")
    assert_same(H(ctx, dp), p2, "forceDotPlace")
    cnt, off, lst = np.zeros(n, np.int32), np.zeros(n + 1, np.int32), np.zeros(nM, np.int32)
    ffi.call("sbo_csr_build", n, nM, bs, cnt, off, lst)
    doff, dlst = D(ctx, np.zeros_like(off)), D(ctx, np.zeros_like(lst))
    CrossBridgeSystem.csrBuild(ctx, dbs, n, doff, dlst)
    assert_same(H(ctx, doff), off, "segMotorOffsets")
    nb = int(off[-1])
    assert_same(H(ctx, dlst)[:nb], lst[:nb], "segMotorMyo")
    fF = (g.standard_normal(3 * n) * 1e-12).astype(np.float32)
    fT = (g.standard_normal(3 * n) * 1e-19).astype(np.float32)
    fFo, fTo = fF.copy(), fT.copy()
    ffi.call("sbo_motor_seg_gather", n, off, lst, bond, fFo, fTo)
    dfF, dfT = D(ctx, fF), D(ctx, fT)
    CrossBridgeSystem.segGather(ctx, doff, dlst, dbond, dfF, dfT)
    assert_same(H(ctx, dfF), fFo, "filament forceSum")
    assert_same(H(ctx, dfT), fTo, "filament torqueSum")


def test_joints_dimer_tether_backbone(ctx, state):
    o = state
    mot, dim, mini = o.mot, o.dim, o.mini
    nM, nB = mot.nM, 3 * mot.nM
    b = mot.body
    g = np.random.default_rng(21)
    nuc = g.integers(0, 4, nM).astype(np.int32)
    F, T = np.zeros(3 * nB, np.float32), np.zeros(3 * nB, np.float32)
    ffi.call("sbo_motor_joints", nB, b.coord, b.uVec, b.segLength, b.bTransGam, b.bRotGam, F, T, nuc, mot.jointParams)
    dC, dU, dL, dTG, dRG = (D(ctx, getattr(b, k)) for k in ("coord", "uVec", "segLength", "bTransGam", "bRotGam"))
    dF, dT = D(ctx, np.zeros_like(F)), D(ctx, np.zeros_like(T))
    MotorJointSystem.joints(ctx, dC, dU, dL, dTG, dRG, dF, dT, D(ctx, nuc), mot.jointParams)
    assert np.abs(F).max() > 0
    assert_same(H(ctx, dF), F, "joint forceSum")
    assert_same(H(ctx, dT), T, "joint torqueSum")
    bs = np.where(g.random(nM) < 0.5, 3, -1).astype(np.int32)
    par = dim.parallel.copy()
    par[::3] = 0                                           # antiparallel dimers too
    ffi.call("sbo_dimer_couple", nB, dim.nD, b.coord, b.uVec, b.segLength, b.bTransGam, b.bRotGam, F, T, dim.motorA,
             dim.motorB, par, dim.dimerParams, bs)
    DimerCouplingSystem.couple(ctx, dC, dU, dL, dTG, dRG, dF, dT, D(ctx, dim.motorA), D(ctx, dim.motorB), D(ctx, par),
                               dim.dimerParams, D(ctx, bs))
    assert_same(H(ctx, dF), F, "dimer forceSum")
    assert_same(H(ctx, dT), T, "dimer torqueSum")
    bb = mini.backbone
    md = np.zeros_like(mini.miniData)
    ffi.call("sbo_mini_tether", nB, mini.nBb, mini.nMD, b.coord, b.uVec, b.segLength, b.bTransGam, b.bRotGam, F, T,
             bb.coord, bb.uVec, mini.bbInvDragY, mini.headBackboneSlot, mini.motorA, mini.attachAxial, md,
             mini.miniParams)
    dmd = D(ctx, np.zeros_like(md))
    MiniFilamentSystem.tether(ctx, dC, dU, dL, dTG, dRG, dF, dT, D(ctx, bb.coord), D(ctx, bb.uVec),
                              D(ctx, mini.bbInvDragY), D(ctx, mini.headBackboneSlot), D(ctx, mini.motorA),
                              D(ctx, mini.attachAxial), dmd, mini.miniParams)
    assert np.abs(md).max() > 0
    assert_same(H(ctx, dmd), md, "miniData")
    assert_same(H(ctx, dF), F, "tether forceSum")
    assert_same(H(ctx, dT), T, "tether torqueSum")
    nBb, nMD = mini.nBb, mini.nMD
    cnt, off, lst = np.zeros(nBb, np.int32), np.zeros(nBb + 1, np.int32), np.zeros(nMD, np.int32)
    ffi.call("sbo_csr_build", nBb, nMD, mini.headBackboneSlot, cnt, off, lst)
    bF = (g.standard_normal(3 * nBb) * 1e-12).astype(np.float32)
    bT = (g.standard_normal(3 * nBb) * 1e-19).astype(np.float32)
    bFo, bTo = bF.copy(), bT.copy()
    ffi.call("sbo_backbone_gather", nBb, off, lst, md, bFo, bTo)
    dbF, dbT = D(ctx, bF), D(ctx, bT)
    MiniFilamentSystem.backboneGather(ctx, D(ctx, off), D(ctx, lst), dmd, dbF, dbT)
    assert_same(H(ctx, dbF), bFo, "backbone forceSum")
    assert_same(H(ctx, dbT), bTo, "backbone torqueSum")
