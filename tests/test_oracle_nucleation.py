"""Pins of the oracle's node nucleation (CPU only): NodeNucleationSystem restated, checked against the gates
of the reference's NodeNucleationHarness (NodeNucleationHarness.java:186-342: nucleation rate, the elastic
tether against its double-precision formula and its relaxation, the dissolution rate, the actin-pool gate, the
no-op when off), the geometry of an emitted seed, the allocator's rank-to-slot rule and Newton's third law
between a seed and its node.
"""
import math

import numpy as np
import pytest

from oracle import ffi
from softbox_amd.physics import AETA_REF, ACTIN_MONO_RADIUS, ACTIN_RADIUS, KT_REF, brownian_force_mag, set_rod_drag
from softbox_amd.stores import FilamentStore, NodeNucleationStore, NodeStore

DT = 1.0e-5
SEED = 0x4E0DE
BOX_VOL = 8.0
SEED_LEN = (NodeNucleationStore.ACTIN_SEED + 1) * ACTIN_MONO_RADIUS
WANG, PHILOX = 0, 1


class NucScene:
    """NodeNucleationHarness.build (NodeNucleationHarness.java:70-97): nNodes fixed nodes 0.2 um apart on x,
    filCap FREE seed-length slots (inert), born seeds with Brownian scale bornScale, one request per node."""

    def __init__(self, nNodes, filCap, formins=4, detach=0.0, pool0=1.0e9, born=0.0, kNuc=None, rng=WANG):
        self.rng = rng
        nd = NodeStore(nNodes, 1)
        nd.node.plane("coord")[0] = 0.2 * np.arange(nNodes)
        nd.node.plane("uVec")[0] = 1.0
        nd.node.plane("yVec")[1] = 1.0
        nd.init_node_drag(AETA_REF, KT_REF)
        self.nd = nd
        f = FilamentStore(filCap)
        # DragTensorSystem.run (DragTensorSystem.java:70-95): a lone seed takes the drag of the at-end minimum
        # length stdSegLength monomer radii, and keeps its own length
        set_rod_drag(f, 32 * ACTIN_MONO_RADIUS, ACTIN_RADIUS, AETA_REF, KT_REF)
        f.segLength[:] = np.float32(SEED_LEN)
        f.set_params(DT, brownian_force_mag(DT, KT_REF))
        f.plane("uVec")[0] = 1.0
        f.plane("yVec")[1] = 1.0
        self.f = f
        self.filState = np.full(filCap, -1, np.int32)
        self.monomerCount = np.full(filCap, NodeNucleationStore.ACTIN_SEED, np.int32)
        self.birthParams = np.array([born, born], np.float32)
        C, K = filCap, nNodes
        self.C, self.K = C, K
        self.freeCount, self.freeOffsets = np.zeros(C, np.int32), np.zeros(C + 1, np.int32)
        self.freeList = np.zeros(max(1, C), np.int32)
        self.accept, self.rankOffsets = np.zeros(K, np.int32), np.zeros(K + 1, np.int32)
        self.reqCoord, self.reqUVec, self.reqYVec = (np.zeros(3 * K, np.float32) for _ in range(3))
        nuc = NodeNucleationStore(nNodes, filCap, pool0, BOX_VOL)
        nuc.set_nuc_params(kNuc if kNuc is not None else NodeNucleationStore.K_NODE_NUC, DT, SEED_LEN, formins)
        nuc.set_tether_params(NodeNucleationStore.FRAC_MOVE, DT)
        nuc.set_dissolve_params(detach, DT)
        self.nuc = nuc

    def step(self, t, brownian=False, pool_gate=False):
        """NodeNucleationHarness.runCpu's step (NodeNucleationHarness.java:104-126) on the oracle, with
        initNewborn after tagSeeds (FullSystemDemoHarness.java:690-691)."""
        f, nuc, nd = self.f, self.nuc, self.nd
        nuc.set_counts(t, SEED)
        if pool_gate:
            nuc.refresh_pool_gate()
        else:
            nuc.nucCounts[3] = 1
        ffi.call("sbo_nuc_count_bound_fil", self.C, nd.nNodes, nuc.seedNode, nuc.nodeBoundFil)
        ffi.call("sbo_nuc_emit", nd.nNodes, self.K, nd.node.coord, nuc.nodeBoundFil, self.accept, self.reqCoord,
                 self.reqUVec, self.reqYVec, nuc.nucParams, t, SEED, int(nuc.nucCounts[3]), self.rng, 0)
        ffi.call("sbo_birth_alloc", self.C, self.K, self.filState, self.freeCount, self.freeOffsets, self.freeList,
                 self.accept, self.rankOffsets, self.reqCoord, self.reqUVec, self.reqYVec, self.K, f.coord, f.uVec,
                 f.yVec, f.brownTransScale, f.brownRotScale, self.birthParams)
        ffi.call("sbo_nuc_tag_seeds", self.C, self.K, self.rankOffsets, self.freeList, self.freeOffsets, nuc.seedNode)
        ffi.call("sbo_nuc_init_newborn", self.C, self.K, self.rankOffsets, self.freeList, self.freeOffsets,
                 self.monomerCount, f.segLength, nuc.seedParams)
        if pool_gate:
            nuc.deplete_for_births(self.rankOffsets, self.K, self.freeOffsets, self.C)
        ffi.call("sbo_zero_accumulators", self.C, f.forceSum, f.torqueSum)
        if brownian:
            ffi.call("sbo_brownian_force", self.C, f.randForce, f.randTorque, f.bTransGam, f.bRotGam,
                     f.brownTransScale, f.brownRotScale, float(f.params[1]), t, SEED, self.rng, 0, 0)
        else:
            f.randForce[:] = 0.0
            f.randTorque[:] = 0.0
        ffi.call("sbo_nuc_seed_tether", self.C, nd.nNodes, f.coord, f.uVec, f.segLength, f.bTransGam, f.forceSum,
                 f.torqueSum, nd.node.coord, nd.nodeInvTransY, nuc.seedNode, nuc.tetherParams)
        ffi.call("sbo_nuc_dissolve", self.C, nuc.seedNode, nuc.dissolveParams, t, SEED, self.rng, 0)
        ffi.call("sbo_integrate", self.C, f.coord, f.uVec, f.yVec, f.forceSum, f.torqueSum, f.randForce, f.randTorque,
                 f.bTransGam, f.bRotGam, DT)
        ffi.call("sbo_derive", self.C, f.coord, f.uVec, f.yVec, f.zVec, f.end1, f.end2, f.segLength)


def test_event_words():
    """The first word of the four-word draw is the event draw; WANG keys node k at step t as the reference does:
    wang((k 1000003) ^ (t 999983) ^ (seed 7919) ^ "NUC1"), then the same base ^ 0x9e3779b9 / 0x85ebca6b /
    0xc2b2ae35 (NodeNucleationSystem.java:76-83)."""
    lib = ffi.load()
    import ctypes as C

    out = (C.c_uint32 * 4)()
    lib.sbo_event_bits4.argtypes = [C.c_int, C.c_uint32, C.c_int32, C.c_int32, C.c_uint32, C.c_uint32, C.c_void_p]
    lib.sbo_event_bits.argtypes = [C.c_int, C.c_uint32, C.c_int32, C.c_int32, C.c_uint32, C.c_uint32]
    lib.sbo_event_bits.restype = C.c_uint32
    lib.sbo_wang_hash.argtypes = [C.c_uint32]
    lib.sbo_wang_hash.restype = C.c_uint32
    M = 0xFFFFFFFF
    for rng in (WANG, PHILOX):
        for k, t, sd in ((0, 0, 0), (7, 123, SEED), (399, 99999, 0x310501)):
            lib.sbo_event_bits4(rng, k, t, sd, 5, 0x4E554331, out)
            assert out[0] == lib.sbo_event_bits(rng, k, t, sd, 5, 0x4E554331)
            if rng == WANG:
                base = ((k * 1000003) & M) ^ ((t * 999983) & M) ^ ((sd * 7919) & M) ^ 0x4E554331
                want = [lib.sbo_wang_hash(base ^ x) for x in (0, 0x9e3779b9, 0x85ebca6b, 0xc2b2ae35)]
                assert list(out) == want
            else:
                assert len(set(out)) == 4


@pytest.mark.parametrize("rng", [WANG, PHILOX])
def test_nucleation_rate(rng):
    """checkRate: the per-node-step request probability equals P_nuc (here 1e-3, 5 sigma over 1.28M draws)."""
    sc = NucScene(64, 8, formins=100, kNuc=100.0, rng=rng)
    nuc = sc.nuc
    n, steps = 64, 20000
    hits = 0
    for t in range(steps):
        ffi.call("sbo_nuc_emit", n, sc.K, sc.nd.node.coord, nuc.nodeBoundFil, sc.accept, sc.reqCoord, sc.reqUVec,
                 sc.reqYVec, nuc.nucParams, t, SEED, 1, rng, 0)
        hits += int(sc.accept.sum())
    p = float(nuc.nucParams[0])
    mean, sd = n * steps * p, math.sqrt(n * steps * p * (1 - p))
    assert abs(hits - mean) < 5 * sd, (hits, mean, sd)


def test_emitted_seed_pose_and_gates():
    """An emitted seed's barbed end2 = coord + (L/2) u sits on its node's centre, u and y are unit and
    orthogonal; nodes at their formin cap and a dry pool emit nothing."""
    sc = NucScene(16, 8, formins=2, kNuc=1.0e5)            # P_nuc = 1: every eligible node requests
    nuc = sc.nuc
    nuc.nodeBoundFil[:] = 0
    nuc.nodeBoundFil[::4] = 2
    ffi.call("sbo_nuc_emit", 16, sc.K, sc.nd.node.coord, nuc.nodeBoundFil, sc.accept, sc.reqCoord, sc.reqUVec,
             sc.reqYVec, nuc.nucParams, 3, SEED, 1, WANG, 0)
    assert np.array_equal(sc.accept, (nuc.nodeBoundFil < 2).astype(np.int32))
    c = sc.reqCoord.reshape(3, -1).astype(np.float64)
    u = sc.reqUVec.reshape(3, -1).astype(np.float64)
    y = sc.reqYVec.reshape(3, -1).astype(np.float64)
    half = float(nuc.nucParams[1])
    for k in np.flatnonzero(sc.accept):
        e2 = c[:, k] + half * u[:, k]
        node = sc.nd.node.plane("coord")[:, k].astype(np.float64)
        assert np.all(np.abs(e2 - node) <= 2.0e-7 * (1.0 + np.abs(node)))     # float32 round-off of the pose
        assert abs(np.linalg.norm(u[:, k]) - 1) < 1e-6 and abs(np.linalg.norm(y[:, k]) - 1) < 1e-6
        assert abs(u[:, k] @ y[:, k]) < 1e-6
    ffi.call("sbo_nuc_emit", 16, sc.K, sc.nd.node.coord, nuc.nodeBoundFil, sc.accept, sc.reqCoord, sc.reqUVec,
             sc.reqYVec, nuc.nucParams, 3, SEED, 0, WANG, 0)
    assert not sc.accept.any()


def test_born_slots_tagged_in_rank_order():
    """After the allocator, request r's born slot (the rank-th free slot) is held by node r and starts as a seed;
    with more requests than free slots only the first nFree are born."""
    sc = NucScene(6, 10, kNuc=1.0e5)
    sc.filState[:] = 0
    sc.filState[[2, 5, 9]] = -1                              # three free slots
    sc.monomerCount[:] = 40
    sc.f.segLength[:] = 0.5
    sc.accept[:] = [0, 1, 0, 1, 1, 1]                          # four requests
    ffi.call("sbo_birth_alloc", sc.C, sc.K, sc.filState, sc.freeCount, sc.freeOffsets, sc.freeList, sc.accept,
             sc.rankOffsets, sc.reqCoord, sc.reqUVec, sc.reqYVec, sc.K, sc.f.coord, sc.f.uVec, sc.f.yVec,
             sc.f.brownTransScale, sc.f.brownRotScale, sc.birthParams)
    ffi.call("sbo_nuc_tag_seeds", sc.C, sc.K, sc.rankOffsets, sc.freeList, sc.freeOffsets, sc.nuc.seedNode)
    ffi.call("sbo_nuc_init_newborn", sc.C, sc.K, sc.rankOffsets, sc.freeList, sc.freeOffsets, sc.monomerCount,
             sc.f.segLength, sc.nuc.seedParams)
    want = np.full(10, -1)
    want[[2, 5, 9]] = [1, 3, 4]
    assert np.array_equal(sc.nuc.seedNode, want)
    assert (sc.filState >= 0).all()
    assert np.array_equal(sc.monomerCount[[2, 5, 9]], [3, 3, 3]) and (np.delete(sc.monomerCount, [2, 5, 9]) == 40).all()
    assert np.allclose(sc.f.segLength[[2, 5, 9]], SEED_LEN)


def test_seed_tether_closed_form_and_node_reaction():
    """checkTether: F = fracMove 1e-6 |c - e2| / (dt (1/g_par + 1/g_node)) toward the node centre c, applied at
    end2 (torque (L/2 u 1e-6) x F); the node's reaction is -F."""
    sc = NucScene(2, 4)
    f, nuc, nd = sc.f, sc.nuc, sc.nd
    s, k = 1, 1
    sc.filState[s] = 0
    nuc.seedNode[s] = k
    half = 0.5 * SEED_LEN
    node = nd.node.plane("coord")[:, k].astype(np.float64)
    f.plane("coord")[:, s] = node + half * np.array([1.0, 0, 0]) + np.array([0.004, 0.003, -0.001])
    f.plane("uVec")[:, s] = (-1.0, 0.0, 0.0)
    ffi.call("sbo_zero_accumulators", sc.C, f.forceSum, f.torqueSum)
    ffi.call("sbo_nuc_seed_tether", sc.C, nd.nNodes, f.coord, f.uVec, f.segLength, f.bTransGam, f.forceSum,
             f.torqueSum, nd.node.coord, nd.nodeInvTransY, nuc.seedNode, nuc.tetherParams)
    c = f.plane("coord")[:, s].astype(np.float64)
    u = f.plane("uVec")[:, s].astype(np.float64)
    L = float(f.segLength[s])
    e2 = c + 0.5 * L * u
    d = nd.node.plane("coord")[:, k].astype(np.float64) - e2
    g_par, inv_node = float(f.bTransGam[s]), float(nd.nodeInvTransY[k])
    mag = 0.5 * 1e-6 * np.linalg.norm(d) / (DT * (1.0 / g_par + inv_node))
    F = f.plane("forceSum")[:, s].astype(np.float64)
    assert np.allclose(F, mag * d / np.linalg.norm(d), rtol=1e-6, atol=0)
    assert F[0] < 0                                           # toward the node (it sits at smaller x)
    T = f.plane("torqueSum")[:, s].astype(np.float64)
    assert np.allclose(T, np.cross(0.5 * L * u * 1e-6, F), rtol=1e-5, atol=1e-30)
    assert not np.delete(f.plane("forceSum"), s, axis=1).any()
    nF = np.zeros(3 * nd.nNodes, np.float32)
    ffi.call("sbo_nuc_seed_tether_node_react", sc.C, nd.nNodes, f.coord, f.uVec, f.segLength, f.bTransGam, nF,
             nd.node.coord, nd.nodeInvTransY, nuc.seedNode, nuc.tetherParams)
    assert np.allclose(nF.reshape(3, -1)[:, k], -F, rtol=1e-6, atol=0) and not nF.reshape(3, -1)[:, 0].any()


def test_tether_relaxes_a_displaced_seed():
    """checkTether's second half: with Brownian motion off the held seed's end2 returns to within 10 nm of the
    node centre and the strain shrinks."""
    sc = NucScene(1, 8, formins=1, kNuc=1.0e5)
    sc.step(0)
    s = int(np.flatnonzero(sc.nuc.seedNode >= 0)[0])
    sc.f.plane("coord")[:, s] += np.array([0.004, 0.003, 0.0], np.float32)

    def strain():
        c, u = sc.f.plane("coord")[:, s].astype(np.float64), sc.f.plane("uVec")[:, s].astype(np.float64)
        return np.linalg.norm(c + 0.5 * float(sc.f.segLength[s]) * u - sc.nd.node.plane("coord")[:, 0])

    s0 = strain()
    for t in range(1, 2001):
        sc.step(t)
    assert strain() < s0 and strain() < 0.01
    assert int((sc.nuc.seedNode >= 0).sum()) == 1             # the formin cap of 1 holds


@pytest.mark.parametrize("rng", [WANG, PHILOX])
def test_dissolution_rate(rng):
    """checkDissolution: held seeds detach at P_detach per step (5 sigma); a freed seed stays an ACTIVE slot."""
    N, M, p = 4000, 200, 0.02
    sc = NucScene(1, N, formins=0, detach=p / DT, rng=rng)
    sc.filState[:] = 0
    sc.nuc.seedNode[:] = 0
    detached = tethered = 0
    for t in range(M):
        before = int((sc.nuc.seedNode >= 0).sum())
        ffi.call("sbo_nuc_dissolve", N, sc.nuc.seedNode, sc.nuc.dissolveParams, t, SEED, rng, 0)
        detached += before - int((sc.nuc.seedNode >= 0).sum())
        tethered += before
    pd = float(sc.nuc.dissolveParams[0])
    assert abs(detached - tethered * pd) < 5 * math.sqrt(tethered * pd * (1 - pd))
    assert (sc.filState >= 0).all()


def test_pool_gate_limits_births():
    """checkActinPool: with a pool worth 5 seeds, at most 5 seeds are born, the pool runs dry and its integer
    ledger counts 3 monomers per birth."""
    sc = NucScene(8, 32, formins=4, kNuc=2000.0)
    budget = 5
    sc.nuc.conc = budget * 3 * sc.nuc.uMPerMonomer * (1 + 1e-12)
    for t in range(400):
        sc.step(t, pool_gate=True)
    births = int((sc.filState >= 0).sum())
    assert 0 < births <= budget and sc.nuc.taken == 3 * births
    assert not sc.nuc.available(3)


def test_no_op_when_off():
    """checkNoOpWhenOff: forminsPerNode 0 births nothing and moves nothing."""
    sc = NucScene(4, 8, formins=0, kNuc=1.0e5)
    c0 = sc.f.coord.copy()
    for t in range(50):
        sc.step(t)
    assert not (sc.filState >= 0).any() and not (sc.nuc.seedNode >= 0).any()
    assert np.array_equal(sc.f.coord, c0)


def test_node_reaction_sums_every_held_seed():
    """seedTetherNodeReact over several held seeds per node equals minus the sum of their tether forces."""
    sc = NucScene(3, 24, formins=6, kNuc=1.0e5, born=1.0 / 30.0)   # born seeds wander (damped Brownian)
    for t in range(6):
        sc.step(t, brownian=True)
    f, nd, nuc = sc.f, sc.nd, sc.nuc
    assert int((nuc.seedNode >= 0).sum()) >= 6
    ffi.call("sbo_zero_accumulators", sc.C, f.forceSum, f.torqueSum)
    ffi.call("sbo_nuc_seed_tether", sc.C, nd.nNodes, f.coord, f.uVec, f.segLength, f.bTransGam, f.forceSum,
             f.torqueSum, nd.node.coord, nd.nodeInvTransY, nuc.seedNode, nuc.tetherParams)
    nF = np.zeros(3 * nd.nNodes, np.float32)
    ffi.call("sbo_nuc_seed_tether_node_react", sc.C, nd.nNodes, f.coord, f.uVec, f.segLength, f.bTransGam, nF,
             nd.node.coord, nd.nodeInvTransY, nuc.seedNode, nuc.tetherParams)
    Fp = f.plane("forceSum").astype(np.float64)
    for k in range(nd.nNodes):
        want = -Fp[:, nuc.seedNode == k].sum(axis=1)
        assert np.allclose(nF.reshape(3, -1)[:, k], want, rtol=1e-6, atol=1e-30)
    assert np.abs(nF).max() > 0


def test_aim_torque_closed_form():
    """seedAimTorque: T = coeff (g_rot,y / dt) (u x aim) for segments with an aim, nothing for the others."""
    C = 3
    u = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float64)
    aim = np.array([[0, 1, 0], [0, 0, 0], [0.6, 0, 0.8]], np.float64)
    fu = np.ascontiguousarray(u.T.ravel(), np.float32)
    fa = np.ascontiguousarray(aim.T.ravel(), np.float32)
    gR = np.full(3 * C, 2.0e-22, np.float32)
    T = np.zeros(3 * C, np.float32)
    ffi.call("sbo_nuc_seed_aim_torque", C, fu, gR, T, fa, np.array([0.3, DT], np.float32))
    Tp = T.reshape(3, C)
    for s in range(C):
        want = 0.3 * float(gR[C + s]) / float(np.float32(DT)) * np.cross(u[s], aim[s]) if aim[s] @ aim[s] > 0.25 else 0
        assert np.allclose(Tp[:, s], want, rtol=1e-6, atol=1e-40)
