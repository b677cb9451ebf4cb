"""GPU parity of node nucleation (NodeNucleationSystem entry points) against the oracle.

Each entry point is called on the GPU and in the oracle on identical inputs; then the reference harness's
nucleation + dynamics step (NodeNucleationHarness.runCpu, NodeNucleationHarness.java:104-126: count, emit,
allocate, tag, newborn init, pool depletion, Brownian, seed tether, dissolve, integrate, derive) is composed
from the GPU entry points and from the oracle's functions and stepped side by side.  Integers must match
exactly and floats bit for bit.
"""
import numpy as np
import pytest

from _common import assert_same
from oracle import ffi
from softbox_amd.systems import (BrownianForceSystem, ChainBendingForceSystem, DerivedGeometrySystem,
                                 NodeNucleationSystem, RigidRodLangevinIntegrationSystem, RNG_PHILOX, RNG_WANG)
from test_oracle_nucleation import DT, SEED, NucScene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from softbox_amd.context import Context

    c = Context(0)
    yield c
    c.close()


class DevNuc:
    """The device arrays of a NucScene and the harness step composed from the library's entry points."""

    F_ARR = ("coord", "uVec", "yVec", "zVec", "end1", "end2", "segLength", "bTransGam", "bRotGam", "forceSum",
             "torqueSum", "randForce", "randTorque", "brownTransScale", "brownRotScale")

    def __init__(self, ctx, sc: NucScene):
        self.ctx, self.sc = ctx, sc
        D = ctx.to_device
        self.f = {k: D(getattr(sc.f, k)) for k in self.F_ARR}
        self.a = {k: D(getattr(sc, k)) for k in ("filState", "monomerCount", "freeCount", "freeOffsets", "freeList",
                                                  "accept", "rankOffsets", "reqCoord", "reqUVec", "reqYVec")}
        self.nodeCoord = D(sc.nd.node.coord)
        self.nodeInv = D(sc.nd.nodeInvTransY)
        self.seedNode = D(sc.nuc.seedNode)
        self.bound = D(sc.nuc.nodeBoundFil)
        self.nuc = sc.nuc.__class__.__new__(sc.nuc.__class__)
        self.nuc.__dict__.update({k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in sc.nuc.__dict__.items()})

    def step(self, t, brownian=False, pool_gate=False):
        ctx, sc, nuc, f, a = self.ctx, self.sc, self.nuc, self.f, self.a
        from softbox_amd.context import ptr

        nuc.set_counts(t, SEED)
        if pool_gate:
            nuc.refresh_pool_gate()
        else:
            nuc.nucCounts[3] = 1
        NodeNucleationSystem.countBoundFil(ctx, self.seedNode, self.bound, nuc.nucCounts)
        NodeNucleationSystem.emit(ctx, self.nodeCoord, self.bound, a["accept"], a["reqCoord"], a["reqUVec"],
                                  a["reqYVec"], nuc.nucParams, nuc.nucCounts, rng=sc.rng)
        bp = np.ascontiguousarray(sc.birthParams, np.float32)
        import ctypes as C

        ctx.check(ctx.lib.sbx_birth_alloc(ctx.h, ptr(a["filState"]), ptr(a["freeCount"]), ptr(a["freeOffsets"]),
                                          ptr(a["freeList"]), ptr(a["accept"]), ptr(a["rankOffsets"]),
                                          ptr(a["reqCoord"]), ptr(a["reqUVec"]), ptr(a["reqYVec"]), ptr(f["coord"]),
                                          ptr(f["uVec"]), ptr(f["yVec"]), ptr(f["brownTransScale"]),
                                          ptr(f["brownRotScale"]), bp.ctypes.data_as(C.POINTER(C.c_float)), sc.C,
                                          sc.K, sc.K))
        alloc = (sc.C, sc.K)
        NodeNucleationSystem.tagSeeds(ctx, a["rankOffsets"], a["freeList"], a["freeOffsets"], self.seedNode, alloc)
        NodeNucleationSystem.initNewborn(ctx, a["rankOffsets"], a["freeList"], a["freeOffsets"], a["monomerCount"],
                                         f["segLength"], nuc.seedParams, alloc)
        if pool_gate:
            ctx.sync()
            nuc.deplete_for_births(ctx.to_host(a["rankOffsets"]), sc.K, ctx.to_host(a["freeOffsets"]), sc.C)
        ChainBendingForceSystem.zeroAccumulators(ctx, f["forceSum"], f["torqueSum"])
        counts = np.array([sc.C, t, SEED, 1], np.int32)
        if brownian:
            BrownianForceSystem.brownianForce(ctx, f["randForce"], f["randTorque"], f["bTransGam"], f["bRotGam"],
                                              f["brownTransScale"], f["brownRotScale"], sc.f.params, counts,
                                              rng=sc.rng)
        else:
            f["randForce"].zero_()
            f["randTorque"].zero_()
        NodeNucleationSystem.seedTether(ctx, f["coord"], f["uVec"], f["segLength"], f["bTransGam"], f["forceSum"],
                                        f["torqueSum"], self.nodeCoord, self.nodeInv, self.seedNode, nuc.tetherParams)
        NodeNucleationSystem.dissolve(ctx, self.seedNode, nuc.dissolveParams, nuc.nucCounts, rng=sc.rng)
        RigidRodLangevinIntegrationSystem.integrate(ctx, f["coord"], f["uVec"], f["yVec"], f["forceSum"],
                                                    f["torqueSum"], f["randForce"], f["randTorque"], f["bTransGam"],
                                                    f["bRotGam"], sc.f.params, counts)
        DerivedGeometrySystem.derive(ctx, f["coord"], f["uVec"], f["yVec"], f["zVec"], f["end1"], f["end2"],
                                     f["segLength"], counts)

    def compare(self, what):
        ctx, sc = self.ctx, self.sc
        ctx.sync()
        assert_same(ctx.to_host(self.seedNode), sc.nuc.seedNode, f"{what}: seedNode")
        for k in ("filState", "monomerCount"):
            assert_same(ctx.to_host(self.a[k]), getattr(sc, k), f"{what}: {k}")
        for k in ("coord", "uVec", "yVec", "zVec", "end1", "end2", "segLength", "forceSum", "torqueSum",
                  "brownTransScale"):
            assert_same(ctx.to_host(self.f[k]), getattr(sc.f, k), f"{what}: {k}")
        assert self.nuc.taken == sc.nuc.taken and self.nuc.conc == sc.nuc.conc, what


@pytest.mark.parametrize("rng", [RNG_WANG, RNG_PHILOX])
def test_nucleation_harness_steps_match_oracle(ctx, rng):
    """300 steps of the harness step on 12 nodes with births, a formin cap of 3, dissolution, a pool that runs
    out part way and damped Brownian seeds: the GPU composition equals the oracle's at every checkpoint."""
    sc = NucScene(12, 64, formins=3, detach=300.0, kNuc=2000.0, born=1.0 / 30.0, rng=rng)
    sc.nuc.conc = 30 * 3 * sc.nuc.uMPerMonomer * (1 + 1e-12)   # 30 seeds' worth of actin
    dev = DevNuc(ctx, sc)
    for t in range(300):
        sc.step(t, brownian=True, pool_gate=True)
        dev.step(t, brownian=True, pool_gate=True)
        if t % 50 == 49:
            dev.compare(f"rng={rng} after {t + 1}")
    assert 0 < int((sc.filState >= 0).sum()) <= 30
    assert int((sc.nuc.seedNode >= 0).sum()) < int((sc.filState >= 0).sum())   # some bonds dissolved


def test_node_react_and_aim_match_oracle(ctx):
    sc = NucScene(5, 96, formins=20, kNuc=1.0e5, born=1.0 / 30.0)
    for t in range(8):
        sc.step(t, brownian=True)
    f, nd, nuc = sc.f, sc.nd, sc.nuc
    assert int((nuc.seedNode >= 0).sum()) >= 30
    g = np.random.default_rng(3)
    nF = (g.standard_normal(3 * nd.nNodes) * 1e-13).astype(np.float32)
    want = nF.copy()
    ffi.call("sbo_nuc_seed_tether_node_react", sc.C, nd.nNodes, f.coord, f.uVec, f.segLength, f.bTransGam, want,
             nd.node.coord, nd.nodeInvTransY, nuc.seedNode, nuc.tetherParams)
    D = ctx.to_device
    dF = D(nF)
    NodeNucleationSystem.seedTetherNodeReact(ctx, D(f.coord), D(f.uVec), D(f.segLength), D(f.bTransGam), dF,
                                             D(nd.node.coord), D(nd.nodeInvTransY), D(nuc.seedNode), nuc.tetherParams)
    ctx.sync()
    assert not np.array_equal(want, nF)
    assert_same(ctx.to_host(dF), want, "node forceSum")
    aim = (g.standard_normal(3 * sc.C)).astype(np.float32)
    aim.reshape(3, -1)[:, ::3] = 0.0
    T = (g.standard_normal(3 * sc.C) * 1e-20).astype(np.float32)
    Tw = T.copy()
    ap = np.array([0.3, DT], np.float32)
    ffi.call("sbo_nuc_seed_aim_torque", sc.C, f.uVec, f.bRotGam, Tw, aim, ap)
    dT = D(T)
    NodeNucleationSystem.seedAimTorque(ctx, D(f.uVec), D(f.bRotGam), dT, D(aim), ap)
    ctx.sync()
    assert_same(ctx.to_host(dT), Tw, "aim torque")
