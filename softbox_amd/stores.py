"""Host-side SoA stores in the reference layout (planar x|y|z planes, plane stride n).

FilamentStore mirrors FilamentStore.java:29-87 / RigidRodBody.java:26-57 and CrosslinkerStore
mirrors CrosslinkerStore.java:31-179 (the fields the mechanics hot path reads and writes).
`DeviceFilaments` / `DeviceCrosslinks` hold the device mirrors (torch tensors used as plain
device memory) and produce the C-ABI structs of include/softbox_hip.h.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from .physics import KT_REF, ACTIN_MONO_RADIUS

SENTINEL_NO_NBR = -1
LINK_ACTIVE = 0
LINK_FREE = -1
XL_STRIDE = _lib.SBX_XL_STRIDE
STRAIN_WIN = _lib.SBX_STRAIN_WIN
TORQUE_WIN = _lib.SBX_TORQUE_WIN


class FilamentStore:
    """FilamentStore.java:29-199 (pose, derived, drag, accumulators, Brownian scales, chain topology)."""

    VEC = ("coord", "uVec", "yVec", "zVec", "end1", "end2", "bTransGam", "bRotGam", "bTransDiff", "bRotDiff",
           "forceSum", "torqueSum", "randForce", "randTorque")
    SCALAR_F = ("segLength", "brownTransScale", "brownRotScale")
    SCALAR_I = ("end1NbrSlot", "end1NbrSide", "end2NbrSlot", "end2NbrSide")

    def __init__(self, n: int):
        self.n = n
        for k in self.VEC:
            setattr(self, k, np.zeros(3 * n, np.float32))
        for k in self.SCALAR_F:
            setattr(self, k, np.zeros(n, np.float32))
        for k in self.SCALAR_I:
            setattr(self, k, np.full(n, SENTINEL_NO_NBR, np.int32))
        self.params = np.zeros(2, np.float32)        # [0]=dt [1]=sqrt(2kT/dt)
        self.counts = np.zeros(4, np.int32)          # [0]=n [1]=step [2]=seed [3]=loadOn
        self.counts[3] = 1
        self.chainParams = np.zeros(7, np.float32)

    # planar helpers (FilamentStore.java:240-260)
    def plane(self, name: str) -> np.ndarray:
        return getattr(self, name).reshape(3, self.n)

    def set_params(self, dt: float, brownian_force_mag: float) -> None:       # FilamentStore.java:330-333
        self.params[0] = np.float32(dt)
        self.params[1] = np.float32(brownian_force_mag)

    def set_counts(self, step: int, seed: int) -> None:                      # FilamentStore.java:334-338
        self.counts[0] = self.n
        self.counts[1] = step
        self.counts[2] = seed

    def set_chain_params(self, dt: float, fracMove=0.5, fracR=0.1, fracMoveTorq=0.265, springActive=0.0,
                         spring=1.0e-20, monoRadius=ACTIN_MONO_RADIUS) -> None:  # FilamentStore.java:345-353
        self.chainParams[:] = np.array([dt, fracMove, fracR, fracMoveTorq, springActive, spring, monoRadius],
                                       np.float32)

    def copy(self) -> "FilamentStore":
        o = FilamentStore(self.n)
        for k in self.VEC + self.SCALAR_F + self.SCALAR_I + ("params", "counts", "chainParams"):
            setattr(o, k, getattr(self, k).copy())
        return o


class CrosslinkerStore:
    """CrosslinkerStore.java:31-179 (topology, payload, lifecycle, strain and torque rings, params)."""

    def __init__(self, nLinks: int, nSeg: int):
        cap = max(1, nLinks)
        self.nLinks = nLinks
        self.nSeg = nSeg
        self.linkFilA = np.full(cap, -1, np.int32)
        self.linkFilB = np.full(cap, -1, np.int32)
        self.loc1 = np.zeros(cap, np.float32)
        self.loc2 = np.zeros(cap, np.float32)
        self.filLinkCt = np.zeros(nSeg, np.int32)
        self.xlinkData = np.zeros(cap * XL_STRIDE, np.float32)
        self.linkState = np.full(cap, LINK_FREE, np.int32)
        self.strainHist = np.zeros(cap * STRAIN_WIN, np.float32)
        self.strainPlace = np.zeros(cap, np.int32)
        self.linkOrientSame = np.zeros(cap, np.int32)
        self.torqueMagHist = np.zeros(cap * TORQUE_WIN, np.float32)
        self.torqueMagPlace = np.zeros(cap, np.int32)
        self.activeLinkCount = np.zeros(nSeg, np.int32)
        self.xlParams = np.zeros(3, np.float32)
        self.offParams = np.zeros(5, np.float32)
        self.torsionParams = np.zeros(2, np.float32)
        self.counts = np.array([nLinks, 0, 0, nSeg], np.int32)

    def set_link(self, k, filA, loc1, filB, loc2):                  # CrosslinkerStore.java:235-241
        self.linkFilA[k] = filA
        self.linkFilB[k] = filB
        self.loc1[k] = np.float32(loc1)
        self.loc2[k] = np.float32(loc2)
        self.linkState[k] = LINK_ACTIVE

    def set_params(self, restLength, fracMoveBase, dt):               # CrosslinkerStore.java:211-215
        self.xlParams[:] = np.array([restLength, fracMoveBase, dt], np.float32)

    def set_off_params(self, const, coeff, exp_, dt, restLength):   # CrosslinkerStore.java:219-225
        self.offParams[:] = np.array([const, coeff, exp_, dt, restLength], np.float32)

    def set_torsion_params(self, spring, active):                   # CrosslinkerStore.java:205-208
        self.torsionParams[:] = np.array([spring, 1.0 if active else 0.0], np.float32)

    def set_counts(self, step, seed):                               # CrosslinkerStore.java:228-231
        self.counts[1] = step
        self.counts[2] = seed

    def compute_fil_link_ct(self):                                  # CrosslinkerStore.java:249-257
        self.filLinkCt[:] = 0
        act = self.linkState[: self.nLinks] >= 0
        np.add.at(self.filLinkCt, self.linkFilA[: self.nLinks][act], 1)
        np.add.at(self.filLinkCt, self.linkFilB[: self.nLinks][act], 1)

    def copy(self) -> "CrosslinkerStore":
        o = CrosslinkerStore(self.nLinks, self.nSeg)
        for k, v in self.__dict__.items():
            setattr(o, k, v.copy() if isinstance(v, np.ndarray) else v)
        return o


class DeviceFilaments:
    """Device mirror of a FilamentStore; `struct()` gives the sbx_filaments of the C ABI."""

    def __init__(self, ctx, fs: FilamentStore, keep_forces: bool = True):
        self.n = fs.n
        self.ctx = ctx
        names = ("coord", "uVec", "yVec", "zVec", "end1", "end2", "segLength", "bTransGam", "bRotGam",
                 "brownTransScale", "brownRotScale", "end1NbrSlot", "end1NbrSide", "end2NbrSlot", "end2NbrSide")
        for k in names:
            setattr(self, k, ctx.to_device(getattr(fs, k)))
        self.forceSum = ctx.to_device(fs.forceSum) if keep_forces else None
        self.torqueSum = ctx.to_device(fs.torqueSum) if keep_forces else None

    def struct(self) -> _lib.SbxFilaments:
        from .context import ptr

        s = _lib.SbxFilaments()
        s.n = self.n
        for name, _ in _lib.SbxFilaments._fields_[1:]:
            setattr(s, name, ptr(getattr(self, name)).value)
        return s

    def pull(self, fs: FilamentStore, names=("coord", "uVec", "yVec", "zVec", "end1", "end2")) -> FilamentStore:
        for k in names:
            t = getattr(self, k)
            if t is not None:
                getattr(fs, k)[:] = self.ctx.to_host(t)
        return fs


class DeviceCrosslinks:
    NAMES = ("linkFilA", "linkFilB", "loc1", "loc2", "linkState", "strainHist", "strainPlace", "linkOrientSame",
             "torqueMagHist", "torqueMagPlace", "xlinkData", "activeLinkCount")

    def __init__(self, ctx, xl: CrosslinkerStore):
        self.ctx = ctx
        self.nLinks = xl.nLinks
        for k in self.NAMES:
            setattr(self, k, ctx.to_device(getattr(xl, k)))

    def struct(self) -> _lib.SbxCrosslinks:
        from .context import ptr

        s = _lib.SbxCrosslinks()
        s.nLinks = self.nLinks
        for name in self.NAMES:
            setattr(s, name, ptr(getattr(self, name)).value)
        return s

    def pull(self, xl: CrosslinkerStore, names=None) -> CrosslinkerStore:
        for k in names or self.NAMES:
            getattr(xl, k)[:] = self.ctx.to_host(getattr(self, k))
        return xl


NUC_NONE, NUC_ATP, NUC_ADPPI, NUC_ADP = 0, 1, 2, 3
FREE_BINDABLE, FREE_COOLDOWN = -1, -2
XB_STRIDE = 13
MINI_STRIDE = 6


class MotorStore:
    """MotorStore.java:32-263: a 3-sub-body myosin per motor (3m rod, 3m+1 lever, 3m+2 head) as a
    rigid-rod body of n = 3*nMotors, the published binding interface and the kinetic state."""

    ROD_LEN, LEVER_LEN, HEAD_LEN = 0.080, 0.008, 0.020      # um (MotorStore.java:62)
    ROD_R, LEVER_R, HEAD_R = 0.003, 0.002, 0.010

    def __init__(self, nMotors: int):
        self.nM = nMotors
        self.body = FilamentStore(3 * nMotors)               # the shared rigid-rod layout (no chain topology)
        self.head = np.zeros(3 * nMotors, np.float32)
        self.uVec = np.zeros(3 * nMotors, np.float32)
        self.rodUVec = np.zeros(3 * nMotors, np.float32)
        self.reach = np.zeros(nMotors, np.float32)
        self.boundSeg = np.full(nMotors, FREE_BINDABLE, np.int32)
        self.bindArc = np.zeros(nMotors, np.float32)
        self.cooldown = np.zeros(nMotors, np.int32)
        self.stats = np.zeros(2 * nMotors, np.int32)
        self.capStats = np.zeros(nMotors, np.int32)
        self.nucleotideState = np.full(nMotors, NUC_NONE, np.int32)
        self.forceDotFil = np.zeros(nMotors, np.float32)
        self.forceMag = np.zeros(nMotors, np.float32)
        self.forceDotHist = np.zeros(10 * nMotors, np.float32)
        self.forceDotPlace = np.zeros(nMotors, np.int32)
        self.bondData = np.zeros(XB_STRIDE * nMotors, np.float32)
        self.bodyParams = np.zeros(2, np.float32)
        self.jointParams = np.zeros(11, np.float32)
        self.kinParams = np.zeros(14, np.float32)
        self.nucParams = np.zeros(8, np.float32)
        self.xbParams = np.zeros(6, np.float32)

    def set_joint_params(self, dt, stall_pn=6.0):                       # MotorStore.java:244-250
        self.jointParams[:] = np.array([dt, 0.4, 0.4, 0.4, 0.0, 0.4, 0.4, 0.0, 96.0, 0.4, stall_pn], np.float32)

    def set_kin_params(self, myoColTol, alignTol, dt, kT):               # MotorStore.java:190-221
        self.kinParams[:] = np.array([100.0, 0.92, 0.08, 2.5e-9, 0.4e-9, kT, dt, myoColTol, alignTol, 0.0,
                                      np.ceil(1.0e-5 / dt), 12.0e-12, 0.0, 1.0], np.float32)

    def set_nuc_params(self, dt):                                       # MotorStore.java:255-263
        self.nucParams[:] = np.array([dt, 2.0e4, 100.0, 100.0, 1.0e4, 0.0, 1.0e3, 1.0e3], np.float32)

    def copy(self) -> "MotorStore":
        o = MotorStore.__new__(MotorStore)
        for k, v in self.__dict__.items():
            setattr(o, k, v.copy() if hasattr(v, "copy") else v)
        return o


class DimerStore:
    """DimerStore.java:26-66: disjoint 1:1 motor pairing (motorA = 2d, motorB = 2d + 1)."""

    def __init__(self, nDimers: int):
        self.nD = nDimers
        self.motorA = np.zeros(nDimers, np.int32)
        self.motorB = np.zeros(nDimers, np.int32)
        self.parallel = np.ones(nDimers, np.int32)
        self.dimerParams = np.zeros(5, np.float32)

    def set_dimer_params(self, dt):                                     # DimerStore.java:59-65
        self.dimerParams[:] = np.array([dt, 0.2, 0.4, 160.0, MotorStore.ROD_LEN], np.float32)

    def copy(self):
        o = DimerStore(self.nD)
        for k, v in self.__dict__.items():
            setattr(o, k, v.copy() if hasattr(v, "copy") else v)
        return o


class MiniFilamentStore:
    """MiniFilamentStore.java:29-135: backbone rigid rods + per-dimer tether data."""

    BACKBONE_LEN, BACKBONE_R, HEAD_ZONE = 0.180, 0.005, 0.05

    def __init__(self, nBackbones: int, nDimers: int):
        self.nBb = nBackbones
        self.nMD = nDimers
        self.backbone = FilamentStore(nBackbones)
        self.headBackboneSlot = np.zeros(nDimers, np.int32)
        self.motorA = np.zeros(nDimers, np.int32)
        self.attachAxial = np.zeros(nDimers, np.float32)
        self.miniData = np.zeros(MINI_STRIDE * nDimers, np.float32)
        self.bbInvDragY = np.zeros(2 * nBackbones, np.float32)
        self.miniParams = np.zeros(3, np.float32)
        self.bbBodyParams = np.zeros(2, np.float32)

    def copy(self):
        o = MiniFilamentStore.__new__(MiniFilamentStore)
        for k, v in self.__dict__.items():
            setattr(o, k, v.copy() if hasattr(v, "copy") else v)
        return o


class NodeStore:
    """NodeStore.java:36-153: protein nodes, rigid spheres with an orientation, each owning motors tethered to
    body-frame radial points on its surface.  The node bodies sit in the backbone slots of the motor step
    (the same Brownian / containment / integration / derive and the same CSR gather of the per-attachment
    reactions); the backbone-named properties below are that view."""

    NODE_RADIUS = 0.05        # um (NodeStore.java:40)
    ATTN_FORCE = 0.4          # NodeStore.java:41
    DIMER_FRACMOVE = 0.2      # DimerStore.ROD_FRAC_MOVE (NodeStore.java:42)

    def __init__(self, nNodes: int, nAttach: int, mode: int = 1):
        self.nNodes, self.nAttach = nNodes, nAttach
        self.node = FilamentStore(nNodes)
        self.attachNode = np.zeros(nAttach, np.int32)
        self.attachMotor = np.zeros(nAttach, np.int32)
        self.attachAtEnd1 = np.zeros(nAttach, np.int32)
        self.attachCoeff = np.zeros(nAttach, np.float32)
        self.radial = np.zeros(3 * nAttach, np.float32)
        self.attachCoeffK = np.zeros(nAttach, np.float32)
        self.nodeData = np.zeros(MINI_STRIDE * nAttach, np.float32)
        self.nodeInvTransY = np.zeros(nNodes, np.float32)
        self.nodeParams = np.zeros(2, np.float32)       # dt, boundCoeff (tetherBoundStiffen)
        self.nodeBodyParams = np.zeros(2, np.float32)   # dt, sqrt(2 kT / dt)
        self.mode = mode                                # 1 tether, 2 tetherBoundStiffen
        self._axial = np.zeros(nAttach, np.float32)

    @property
    def attachKey(self) -> np.ndarray:
        """The packed [node(nA) | motor(nA)] key of the per-System calls."""
        return np.concatenate([self.attachNode, self.attachMotor]).astype(np.int32)

    def attach(self, a, k, m, ru, ry, rz, coeff, at_end1):           # NodeStore.java:113-123
        self.attachNode[a], self.attachMotor[a] = k, m
        self.attachAtEnd1[a] = 1 if at_end1 else 0
        self.attachCoeff[a] = np.float32(coeff)
        n = self.nAttach
        self.radial[a], self.radial[n + a], self.radial[2 * n + a] = np.float32(ru), np.float32(ry), np.float32(rz)
        self.attachCoeffK[a] = np.float32(coeff if at_end1 else -coeff)

    def init_node_drag(self, aeta: float, kT: float) -> None:        # NodeStore.java:136-151
        from .physics import sphere_drag_si

        g = sphere_drag_si(self.NODE_RADIUS, aeta)
        b = self.node
        b.segLength[:] = 0.0
        for a in range(3):
            b.plane("bTransGam")[a][:] = np.float32(g[a])
            b.plane("bRotGam")[a][:] = np.float32(g[3 + a])
            b.plane("bTransDiff")[a][:] = np.float32(kT / g[a])
            b.plane("bRotDiff")[a][:] = np.float32(kT / g[3 + a])
        self.nodeInvTransY[:] = np.float32(1.0 / g[1])

    def set_params(self, dt: float, kT: float, bound_coeff: float = 0.0) -> None:   # NodeStore.java:125-133
        self.nodeParams[:] = np.array([dt, bound_coeff], np.float32)
        self.nodeBodyParams[:] = np.array([dt, math.sqrt(2.0 * kT / dt)], np.float32)

    # the backbone view of the motor step
    nBb = property(lambda self: self.nNodes)
    nMD = property(lambda self: self.nAttach)
    backbone = property(lambda self: self.node)
    headBackboneSlot = property(lambda self: self.attachNode)
    motorA = property(lambda self: self.attachMotor)
    attachAxial = property(lambda self: self._axial)
    miniData = property(lambda self: self.nodeData)
    bbInvDragY = property(lambda self: self.nodeInvTransY)
    bbBodyParams = property(lambda self: self.nodeBodyParams)
    miniParams = property(lambda self: np.array([self.nodeParams[0], self.nodeParams[1], 0.0], np.float32))

    def copy(self) -> "NodeStore":
        o = NodeStore.__new__(NodeStore)
        for k, v in self.__dict__.items():
            setattr(o, k, v.copy() if hasattr(v, "copy") else v)
        return o


class NodeNucleationStore:
    """NodeNucleationStore.java:35-110: the node bond of each filament slot (seedNode, -1 free), the per-node
    count of held seeds and the nucleation, tether and dissolution parameters; the actin pool (ActinPool.java)
    as a host scalar with its integer ledger."""

    ACTIN_SEED = 3            # Constants.actinSeed
    K_NODE_NUC = 10.0         # Constants.kNodeNuc, /node/s
    FRAC_MOVE = 0.5           # Constants.fracMove

    def __init__(self, nNodes: int, filCap: int, pool0: float, box_volume_um3: float):
        self.nNodes, self.filCap = nNodes, filCap
        self.seedNode = np.full(filCap, -1, np.int32)
        self.nodeBoundFil = np.zeros(nNodes, np.int32)
        self.nucParams = np.zeros(3, np.float32)
        self.tetherParams = np.zeros(2, np.float32)
        self.dissolveParams = np.zeros(1, np.float32)
        self.seedParams = np.array([self.ACTIN_SEED, 0.0], np.float32)
        self.nucCounts = np.array([nNodes, 0, 0, 0], np.int32)
        self.conc = float(pool0)
        self.uMPerMonomer = (1e5 * 1e5 * 1e5 * 1e6) / (box_volume_um3 * 6.022e23)
        self.taken = self.returned = 0

    def set_nuc_params(self, kNodeNuc, dt, seedLenUm, forminsPerNode):      # NodeNucleationStore.java:85-90
        self.nucParams[:] = np.array([kNodeNuc * dt, 0.5 * seedLenUm, forminsPerNode], np.float32)
        self.seedParams[1] = np.float32(seedLenUm)

    def set_tether_params(self, fracMove, dt):                              # :91-94
        self.tetherParams[:] = np.array([fracMove, dt], np.float32)

    def set_dissolve_params(self, detachRate, dt):                          # :95-97
        self.dissolveParams[0] = np.float32(detachRate * dt)

    def set_counts(self, step, seed):                                       # :98-101
        self.nucCounts[1], self.nucCounts[2] = step, seed

    def available(self, monomers: int) -> bool:                            # ActinPool.available
        return self.conc >= monomers * self.uMPerMonomer

    def take(self, monomers: int) -> None:                                  # ActinPool.take
        self.conc -= monomers * self.uMPerMonomer
        if self.conc < 0.0:
            self.conc = 0.0
        self.taken += monomers

    def refresh_pool_gate(self) -> None:                                    # :103
        self.nucCounts[3] = 1 if self.available(self.ACTIN_SEED) else 0

    def deplete_for_births(self, rankOffsets, K: int, freeOffsets, C: int) -> int:   # :107-114
        births = min(int(rankOffsets[K]), int(freeOffsets[C]))
        if births > 0:
            self.take(births * self.ACTIN_SEED)
        return births


class DeviceMotors:
    """Device mirror of a MotorStore + DimerStore + MiniFilamentStore; `struct()` gives the
    sbx_motors of the C ABI and `params()` the sbx_motor_params."""

    _BODY = ("coord", "uVec", "yVec", "zVec", "end1", "end2", "segLength", "bTransGam", "bRotGam",
             "brownTransScale", "brownRotScale", "forceSum", "torqueSum")
    _KIN = ("head", "rodUVec", "reach", "boundSeg", "cooldown", "stats", "capStats", "nucleotideState",
            "forceDotPlace", "bindArc", "forceDotFil", "forceMag", "forceDotHist", "bondData")
    _BB = (("coord", "bbCoord"), ("uVec", "bbUVec"), ("yVec", "bbYVec"), ("zVec", "bbZVec"), ("end1", "bbEnd1"),
           ("end2", "bbEnd2"), ("segLength", "bbSegLength"), ("bTransGam", "bbTransGam"), ("bRotGam", "bbRotGam"),
           ("brownTransScale", "bbBrownTransScale"), ("brownRotScale", "bbBrownRotScale"))

    def __init__(self, ctx, mot: MotorStore, dim: DimerStore, mini: MiniFilamentStore, max_cand: int,
                 seeds=(0, 0), keep_bb_forces: bool = True):
        self.ctx = ctx
        self.nM, self.nD, self.nBb, self.nMD = mot.nM, dim.nD, mini.nBb, mini.nMD
        self.max_cand = max_cand
        self.t = {}
        for k in self._BODY:
            self.t[k] = ctx.to_device(getattr(mot.body, k))
        self.t["hUVec"] = ctx.to_device(mot.uVec)
        for k in self._KIN:
            self.t[k] = ctx.to_device(getattr(mot, k))
        self.t["reachSeg"] = ctx.to_device(np.full(max(1, mot.nM * max_cand), -1, np.int32))
        self.t["reachCount"] = ctx.to_device(np.zeros(mot.nM, np.int32))
        for k in ("motorA", "motorB", "parallel"):
            self.t[k] = ctx.to_device(getattr(dim, k)) if dim.nD else None
        bb = mini.backbone
        for src, dst in self._BB:
            self.t[dst] = ctx.to_device(getattr(bb, src)) if mini.nBb else None
        self.t["bbInvDragY"] = ctx.to_device(mini.bbInvDragY) if mini.nBb else None
        self.t["bbForceSum"] = ctx.to_device(bb.forceSum) if (mini.nBb and keep_bb_forces) else None
        self.t["bbTorqueSum"] = ctx.to_device(bb.torqueSum) if (mini.nBb and keep_bb_forces) else None
        self.t["headBackboneSlot"] = ctx.to_device(mini.headBackboneSlot) if mini.nMD else None
        self.t["tetherMotor"] = ctx.to_device(mini.motorA) if mini.nMD else None
        self.t["attachAxial"] = ctx.to_device(mini.attachAxial) if mini.nMD else None
        self.t["miniData"] = ctx.to_device(mini.miniData) if mini.nMD else None
        node = isinstance(mini, NodeStore)
        self.t["radial"] = ctx.to_device(mini.radial) if (node and mini.nMD) else None
        self.t["attachCoeffK"] = ctx.to_device(mini.attachCoeffK) if (node and mini.nMD) else None
        p = _lib.SbxMotorParams()
        for k, src in (("bodyParams", mot.bodyParams), ("jointParams", mot.jointParams), ("kinParams", mot.kinParams),
                       ("nucParams", mot.nucParams), ("xbParams", mot.xbParams), ("dimerParams", dim.dimerParams),
                       ("miniParams", mini.miniParams), ("bbBodyParams", mini.bbBodyParams)):
            arr = getattr(p, k)
            for i, v in enumerate(src):
                arr[i] = float(v)
        p.seedMotor, p.seedBackbone = int(seeds[0]), int(seeds[1])
        p.nodeMode = int(mini.mode) if node else 0
        self._params = p

    def struct(self) -> _lib.SbxMotors:
        from .context import ptr

        s = _lib.SbxMotors()
        s.nM, s.nD, s.nBb, s.nMD = self.nM, self.nD, self.nBb, self.nMD
        for name, _ in _lib.SbxMotors._fields_:
            if name in self.t:
                v = self.t[name]
                setattr(s, name, ptr(v).value if v is not None else None)
        return s

    def params(self) -> _lib.SbxMotorParams:
        return self._params

    def pull(self, mot: MotorStore = None, mini: MiniFilamentStore = None) -> None:
        h = self.ctx.to_host
        if mot is not None:
            for k in self._BODY:
                getattr(mot.body, k)[:] = h(self.t[k])
            mot.uVec[:] = h(self.t["hUVec"])
            for k in self._KIN:
                getattr(mot, k)[:] = h(self.t[k])
        if mini is not None and self.nBb:
            for src, dst in self._BB:
                getattr(mini.backbone, src)[:] = h(self.t[dst])
            if self.t["bbForceSum"] is not None:
                mini.backbone.forceSum[:] = h(self.t["bbForceSum"])
                mini.backbone.torqueSum[:] = h(self.t["bbTorqueSum"])
            if self.nMD:
                mini.miniData[:] = h(self.t["miniData"])
