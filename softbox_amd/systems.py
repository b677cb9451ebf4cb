"""System-API mirror: the reference's static System methods, same names and argument order.

Each method forwards to the C-ABI entry point that replaces the TornadoVM task of the same name.
Big arrays are device tensors (planar SoA, stride n); the small `params` / `counts` / `*Params`
arrays are host arrays read at call time (the reference re-uploads them every execution).
Errors raise `SoftboxError` with the library's message; nothing falls back to the CPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .context import Context, ptr

RNG_WANG = _lib.SBX_RNG_WANG
RNG_PHILOX = _lib.SBX_RNG_PHILOX


def _f(a):
    arr = np.ascontiguousarray(a, dtype=np.float32)
    return arr, arr.ctypes.data_as(C.POINTER(C.c_float))


def _i(a):
    arr = np.ascontiguousarray(a, dtype=np.int32)
    return arr, arr.ctypes.data_as(C.POINTER(C.c_int))


def _n3(t) -> int:
    return t.numel() // 3


class ChainBendingForceSystem:
    """ChainBendingForceSystem.java:34-331"""

    @staticmethod
    def zeroAccumulators(ctx: Context, forceSum, torqueSum, counts=None):
        return ctx.check(ctx.lib.sbx_zero_accumulators(ctx.h, ptr(forceSum), ptr(torqueSum), _n3(forceSum)))

    @staticmethod
    def chainForces(ctx: Context, coord, uVec, segLength, end2NbrSlot, end2NbrSide, end1NbrSlot, end1NbrSide,
                    bTransGam, bRotGam, forceSum, torqueSum, chainParams, counts):
        cp, cpp = _f(chainParams)
        cn, cnp = _i(counts)
        return ctx.check(ctx.lib.sbx_chain_forces(ctx.h, ptr(coord), ptr(uVec), ptr(segLength), ptr(end2NbrSlot),
                                                  ptr(end2NbrSide), ptr(end1NbrSlot), ptr(end1NbrSide),
                                                  ptr(bTransGam), ptr(bRotGam), ptr(forceSum), ptr(torqueSum), cpp,
                                                  cnp, _n3(coord)))


class BrownianForceSystem:
    """BrownianForceSystem.java:44-106 (rng selects the reference Wang stream or Philox)."""

    @staticmethod
    def brownianForce(ctx: Context, randForce, randTorque, bTransGam, bRotGam, brownTransScale, brownRotScale,
                      params, counts, rng: int = RNG_PHILOX, stream: int = 0):
        pa, pp = _f(params)
        cn, cnp = _i(counts)
        return ctx.check(ctx.lib.sbx_brownian_force(ctx.h, ptr(randForce), ptr(randTorque), ptr(bTransGam),
                                                    ptr(bRotGam), ptr(brownTransScale), ptr(brownRotScale), pp, cnp,
                                                    _n3(randForce), rng, stream))


class RigidRodLangevinIntegrationSystem:
    """RigidRodLangevinIntegrationSystem.java:30-112"""

    @staticmethod
    def integrate(ctx: Context, coord, uVec, yVec, forceSum, torqueSum, randForce, randTorque, bTransGam, bRotGam,
                  params, counts):
        pa, pp = _f(params)
        cn, cnp = _i(counts)
        return ctx.check(ctx.lib.sbx_integrate(ctx.h, ptr(coord), ptr(uVec), ptr(yVec), ptr(forceSum),
                                               ptr(torqueSum), ptr(randForce), ptr(randTorque), ptr(bTransGam),
                                               ptr(bRotGam), pp, cnp, _n3(coord)))


class DerivedGeometrySystem:
    """DerivedGeometrySystem.java:27-73"""

    @staticmethod
    def derive(ctx: Context, coord, uVec, yVec, zVec, end1, end2, segLength, counts):
        cn, cnp = _i(counts)
        return ctx.check(ctx.lib.sbx_derive(ctx.h, ptr(coord), ptr(uVec), ptr(yVec), ptr(zVec), ptr(end1),
                                            ptr(end2), ptr(segLength), cnp, _n3(coord)))


class MechanicsFusion:
    """MechanicsFusion.java:33-338"""

    @staticmethod
    def forceBuild(ctx: Context, forceSum, torqueSum, randForce, randTorque, bTransGam, bRotGam, brownTransScale,
                   brownRotScale, params, counts, rng: int = RNG_PHILOX, stream: int = 0):
        pa, pp = _f(params)
        cn, cnp = _i(counts)
        return ctx.check(ctx.lib.sbx_force_build(ctx.h, ptr(forceSum), ptr(torqueSum), ptr(randForce),
                                                 ptr(randTorque), ptr(bTransGam), ptr(bRotGam), ptr(brownTransScale),
                                                 ptr(brownRotScale), pp, cnp, _n3(forceSum), rng, stream))

    @staticmethod
    def integDerive(ctx: Context, coord, uVec, yVec, zVec, segLength, end1, end2, forceSum, torqueSum, randForce,
                    randTorque, bTransGam, bRotGam, params):
        pa, pp = _f(params)
        return ctx.check(ctx.lib.sbx_integ_derive(ctx.h, ptr(coord), ptr(uVec), ptr(yVec), ptr(zVec),
                                                  ptr(segLength), ptr(end1), ptr(end2), ptr(forceSum),
                                                  ptr(torqueSum), ptr(randForce), ptr(randTorque), ptr(bTransGam),
                                                  ptr(bRotGam), pp, _n3(coord)))

    @staticmethod
    def integDeriveConfined(ctx: Context, coord, uVec, yVec, zVec, segLength, end1, end2, forceSum, torqueSum,
                            randForce, randTorque, bTransGam, bRotGam, boxParams, counts):
        bp, bpp = _f(boxParams)
        cn, cnp = _i(counts)
        return ctx.check(ctx.lib.sbx_integ_derive_confined(ctx.h, ptr(coord), ptr(uVec), ptr(yVec), ptr(zVec),
                                                           ptr(segLength), ptr(end1), ptr(end2), ptr(forceSum),
                                                           ptr(torqueSum), ptr(randForce), ptr(randTorque),
                                                           ptr(bTransGam), ptr(bRotGam), bpp, cnp, _n3(coord)))


class ContainmentSystem:
    """ContainmentSystem.java:55-137"""

    @staticmethod
    def confine(ctx: Context, coord, uVec, segLength, bTransGam, forceSum, torqueSum, boxParams, counts):
        bp, bpp = _f(boxParams)
        cn, cnp = _i(counts)
        return ctx.check(ctx.lib.sbx_confine(ctx.h, ptr(coord), ptr(uVec), ptr(segLength), ptr(bTransGam),
                                             ptr(forceSum), ptr(torqueSum), bpp, cnp, _n3(coord)))


class FilamentStoreKernels:
    """FilamentStore.publishToBodyView (FilamentStore.java:275-295)"""

    @staticmethod
    def publishToBodyView(ctx: Context, coord, segLength, center, boundingRadius, ownerStore, ownerSlot,
                          viewParams, counts=None):
        vp, vpp = _f(viewParams)
        return ctx.check(ctx.lib.sbx_publish_filaments(ctx.h, ptr(coord), ptr(segLength), ptr(center),
                                                       ptr(boundingRadius), ptr(ownerStore), ptr(ownerSlot), vpp,
                                                       _n3(coord), _n3(center)))


class SpatialGrid:
    """SpatialGrid.java:42-445 (gridBuild = the whole chunked counting sort + scan)."""

    MAX_CAND = _lib.SBX_MAX_CAND

    @staticmethod
    def bodyCell(ctx: Context, center, gridParams, gridDims, counts, bodyCell, periodic: int = 0):
        gp, gpp = _f(gridParams)
        gd, gdp = _i(gridDims)
        cn, cnp = _i(counts)
        return ctx.check(ctx.lib.sbx_body_cell(ctx.h, ptr(center), gpp, gdp, cnp, ptr(bodyCell), _n3(center),
                                               periodic))

    @staticmethod
    def gridBuild(ctx: Context, bodyCell, counts, gridDims, gridCellOffsets, gridCellContents):
        gd, gdp = _i(gridDims)
        cn, cnp = _i(counts)
        return ctx.check(ctx.lib.sbx_grid_build(ctx.h, ptr(bodyCell), cnp, gdp, ptr(gridCellOffsets),
                                                ptr(gridCellContents), gridCellContents.numel()))

    @staticmethod
    def broadPhase(ctx: Context, center, boundingRadius, bodyCell, gridCellOffsets, gridCellContents, gridDims,
                   gridParams, counts, candPartner, candCount, maxCand: int = _lib.SBX_MAX_CAND, periodic: int = 0):
        gp, gpp = _f(gridParams)
        gd, gdp = _i(gridDims)
        cn, cnp = _i(counts)
        return ctx.check(ctx.lib.sbx_broad_phase(ctx.h, ptr(center), ptr(boundingRadius), ptr(bodyCell),
                                                 ptr(gridCellOffsets), ptr(gridCellContents), gdp, gpp, cnp,
                                                 ptr(candPartner), ptr(candCount), _n3(center), maxCand, periodic))


class CrosslinkerSystem:
    """CrosslinkerSystem.java:47-808 (mechanics part)"""

    @staticmethod
    def countActiveLinks(ctx: Context, linkState, linkFilA, linkFilB, activeLinkCount):
        return ctx.check(ctx.lib.sbx_count_active_links(ctx.h, ptr(linkState), ptr(linkFilA), ptr(linkFilB),
                                                        ptr(activeLinkCount), linkState.numel(),
                                                        activeLinkCount.numel()))

    @staticmethod
    def unbind(ctx: Context, filCoord, filUVec, filEnd1, linkFilA, linkFilB, loc1, loc2, linkState, strainHist,
               strainPlace, offParams, counts, rng: int = RNG_PHILOX, stream: int = 0):
        op, opp = _f(offParams)
        cn, cnp = _i(counts)
        return ctx.check(ctx.lib.sbx_link_unbind(ctx.h, ptr(filCoord), ptr(filUVec), ptr(filEnd1), ptr(linkFilA),
                                                 ptr(linkFilB), ptr(loc1), ptr(loc2), ptr(linkState),
                                                 ptr(strainHist), ptr(strainPlace), opp, cnp, _n3(filUVec), rng,
                                                 stream))

    @staticmethod
    def linkForces(ctx: Context, filCoord, filUVec, filEnd1, filBTransGam, linkFilA, linkFilB, loc1, loc2,
                   filLinkCt, xlinkData, xlParams):
        xp, xpp = _f(xlParams)
        return ctx.check(ctx.lib.sbx_link_forces(ctx.h, ptr(filCoord), ptr(filUVec), ptr(filEnd1),
                                                 ptr(filBTransGam), ptr(linkFilA), ptr(linkFilB), ptr(loc1),
                                                 ptr(loc2), ptr(filLinkCt), ptr(xlinkData), xpp, _n3(filCoord),
                                                 linkFilA.numel()))

    @staticmethod
    def linkTorsion(ctx: Context, filUVec, linkFilA, linkFilB, linkState, linkOrientSame, torqueMagHist,
                    torqueMagPlace, xlinkData, torsionParams):
        tp, tpp = _f(torsionParams)
        return ctx.check(ctx.lib.sbx_link_torsion(ctx.h, ptr(filUVec), ptr(linkFilA), ptr(linkFilB), ptr(linkState),
                                                  ptr(linkOrientSame), ptr(torqueMagHist), ptr(torqueMagPlace),
                                                  ptr(xlinkData), tpp, _n3(filUVec), linkFilA.numel()))

    @staticmethod
    def segGatherA(ctx: Context, segOffsets, segLinkIdx, xlinkData, linkState, filForceSum, filTorqueSum,
                   counts=None):
        return ctx.check(ctx.lib.sbx_seg_gather(ctx.h, ptr(segOffsets), ptr(segLinkIdx), ptr(xlinkData),
                                                ptr(linkState), ptr(filForceSum), ptr(filTorqueSum),
                                                _n3(filForceSum), 0))

    @staticmethod
    def segGatherB(ctx: Context, segOffsets, segLinkIdx, xlinkData, linkState, filForceSum, filTorqueSum,
                   counts=None):
        return ctx.check(ctx.lib.sbx_seg_gather(ctx.h, ptr(segOffsets), ptr(segLinkIdx), ptr(xlinkData),
                                                ptr(linkState), ptr(filForceSum), ptr(filTorqueSum),
                                                _n3(filForceSum), 1))


class MotorStoreKernels:
    """MotorStore.java:300-347: the published binding interface of the motor bodies."""

    @staticmethod
    def publishHeadFromBody(ctx: Context, bodyCoord, bodyUVec, bodySegLength, head, uVec, rodUVec, counts):
        cn, cnp = _i(counts)
        return ctx.check(ctx.lib.sbx_publish_head_from_body(ctx.h, ptr(bodyCoord), ptr(bodyUVec),
                                                            ptr(bodySegLength), ptr(head), ptr(uVec), ptr(rodUVec),
                                                            cnp, int(cn[0])))

    @staticmethod
    def publishToBodyView(ctx: Context, head, reach, center, boundingRadius, ownerStore, ownerSlot, publishParams,
                          counts):
        pp, ppp = _i(publishParams)
        cn, cnp = _i(counts)
        return ctx.check(ctx.lib.sbx_publish_motors(ctx.h, ptr(head), ptr(reach), ptr(center), ptr(boundingRadius),
                                                    ptr(ownerStore), ptr(ownerSlot), ppp, cnp, int(cn[0]),
                                                    boundingRadius.numel()))


class BindingDetectionSystem:
    """BindingDetectionSystem.java:143-389: grid-filtered reach tests and nearest-segment binding."""

    @staticmethod
    def gridReachable(ctx: Context, head, uVec, rodUVec, segEnd1, segEnd2, gridParams, gridDims, gridCellOffsets,
                      gridCellContents, ownerStore, ownerSlot, motorReachSeg, motorReachCount, kinParams, counts):
        gp, gpp = _f(gridParams)
        gd, gdp = _i(gridDims)
        kp, kpp = _f(kinParams)
        cn, cnp = _i(counts)
        nM = int(cn[0])
        return ctx.check(ctx.lib.sbx_grid_reachable(ctx.h, ptr(head), ptr(uVec), ptr(rodUVec), ptr(segEnd1),
                                                    ptr(segEnd2), gpp, gdp, ptr(gridCellOffsets),
                                                    ptr(gridCellContents), ptr(ownerStore), ptr(ownerSlot),
                                                    ptr(motorReachSeg), ptr(motorReachCount), kpp, cnp,
                                                    _n3(segEnd1), motorReachSeg.numel() // max(1, nM)))

    @staticmethod
    def bindNearest(ctx: Context, head, uVec, rodUVec, segEnd1, segEnd2, motorCandSeg, motorCandCount, boundSeg,
                    bindArc, kinParams, counts):
        kp, kpp = _f(kinParams)
        cn, cnp = _i(counts)
        nM = int(cn[0])
        return ctx.check(ctx.lib.sbx_bind_nearest(ctx.h, ptr(head), ptr(uVec), ptr(rodUVec), ptr(segEnd1),
                                                  ptr(segEnd2), ptr(motorCandSeg), ptr(motorCandCount),
                                                  ptr(boundSeg), ptr(bindArc), kpp, cnp, _n3(segEnd1),
                                                  motorCandSeg.numel() // max(1, nM)))

    @staticmethod
    def bindNearestNodeAware(ctx: Context, head, uVec, rodUVec, segEnd1, segEnd2, motorCandSeg, motorCandCount,
                             boundSeg, bindArc, seedNode, kinParams, counts):
        kp, kpp = _f(kinParams)
        cn, cnp = _i(counts)
        nM = int(cn[0])
        return ctx.check(ctx.lib.sbx_bind_nearest_node_aware(ctx.h, ptr(head), ptr(uVec), ptr(rodUVec), ptr(segEnd1),
                                                             ptr(segEnd2), ptr(motorCandSeg), ptr(motorCandCount),
                                                             ptr(boundSeg), ptr(bindArc), ptr(seedNode), kpp, cnp,
                                                             _n3(segEnd1), motorCandSeg.numel() // max(1, nM)))

    @staticmethod
    def bruteReachable(ctx: Context, head, uVec, rodUVec, segEnd1, segEnd2, bruteReachSeg, bruteReachCount,
                       kinParams, counts):
        kp, kpp = _f(kinParams)
        cn, cnp = _i(counts)
        nM = int(cn[0])
        return ctx.check(ctx.lib.sbx_brute_reachable(ctx.h, ptr(head), ptr(uVec), ptr(rodUVec), ptr(segEnd1),
                                                     ptr(segEnd2), ptr(bruteReachSeg), ptr(bruteReachCount), kpp, cnp,
                                                     _n3(segEnd1), bruteReachSeg.numel() // max(1, nM)))

    @staticmethod
    def bruteReachableNodeAware(ctx: Context, head, uVec, rodUVec, segEnd1, segEnd2, bruteReachSeg, bruteReachCount,
                                seedNode, kinParams, counts):
        kp, kpp = _f(kinParams)
        cn, cnp = _i(counts)
        nM = int(cn[0])
        return ctx.check(ctx.lib.sbx_brute_reachable_node_aware(ctx.h, ptr(head), ptr(uVec), ptr(rodUVec),
                                                                ptr(segEnd1), ptr(segEnd2), ptr(bruteReachSeg),
                                                                ptr(bruteReachCount), ptr(seedNode), kpp, cnp,
                                                                _n3(segEnd1), bruteReachSeg.numel() // max(1, nM)))


class NucleotideCycleSystem:
    """NucleotideCycleSystem.java:33-135: catch-slip unbinding and the ATPase cycle."""

    @staticmethod
    def catchSlipRelease(ctx: Context, boundSeg, forceDotFil, forceMag, cooldown, stats, capStats, kinParams, counts,
                         rng: int = RNG_PHILOX, stream: int = 0):
        kp, kpp = _f(kinParams)
        cn, cnp = _i(counts)
        return ctx.check(ctx.lib.sbx_catch_slip_release(ctx.h, ptr(boundSeg), ptr(forceDotFil), ptr(forceMag),
                                                        ptr(cooldown), ptr(stats), ptr(capStats), kpp, cnp,
                                                        boundSeg.numel(), rng, stream))

    @staticmethod
    def cycle(ctx: Context, nucleotideState, boundSeg, forceDotHist, nucParams, counts, rng: int = RNG_PHILOX,
              stream: int = 0):
        npar, npp = _f(nucParams)
        cn, cnp = _i(counts)
        return ctx.check(ctx.lib.sbx_nucleotide_cycle(ctx.h, ptr(nucleotideState), ptr(boundSeg), ptr(forceDotHist),
                                                      npp, cnp, nucleotideState.numel(), rng, stream))


class CrossBridgeSystem:
    """CrossBridgeSystem.java:58-320: cross-bridge bond forces, head force, force-dot ring, segment CSR + gather.
    csrBuild (CrossBridgeSystem.java:189-298) is the CSR inverse: histogram + scan + scatter in one call."""

    @staticmethod
    def csrBuild(ctx: Context, key, nKeys: int, offsets, items):
        return ctx.check(ctx.lib.sbx_csr_build(ctx.h, ptr(key), key.numel(), nKeys, ptr(offsets), ptr(items)))

    @staticmethod
    def bondForces(ctx: Context, motorCoord, motorUVec, motorYVec, motorBRotGam, filCoord, filUVec, filYVec,
                   filBRotGam, filSegLength, boundSeg, bindArc, nucleotideState, bondData, xbParams):
        xp, xpp = _f(xbParams)
        return ctx.check(ctx.lib.sbx_bond_forces(ctx.h, ptr(motorCoord), ptr(motorUVec), ptr(motorYVec),
                                                 ptr(motorBRotGam), ptr(filCoord), ptr(filUVec), ptr(filYVec),
                                                 ptr(filBRotGam), ptr(filSegLength), ptr(boundSeg), ptr(bindArc),
                                                 ptr(nucleotideState), ptr(bondData), xpp, boundSeg.numel(),
                                                 filSegLength.numel()))

    @staticmethod
    def applyHeadForce(ctx: Context, bondData, bodyForceSum, bodyTorqueSum):
        return ctx.check(ctx.lib.sbx_apply_head_force(ctx.h, ptr(bondData), ptr(bodyForceSum), ptr(bodyTorqueSum),
                                                      _n3(bodyForceSum) // 3))

    @staticmethod
    def registerForceDot(ctx: Context, bondData, boundSeg, forceDotFil, forceMag, forceDotHist, forceDotPlace):
        return ctx.check(ctx.lib.sbx_register_force_dot(ctx.h, ptr(bondData), ptr(boundSeg), ptr(forceDotFil),
                                                        ptr(forceMag), ptr(forceDotHist), ptr(forceDotPlace),
                                                        boundSeg.numel()))

    @staticmethod
    def segGather(ctx: Context, segMotorOffsets, segMotorMyo, bondData, filForceSum, filTorqueSum):
        return ctx.check(ctx.lib.sbx_motor_seg_gather(ctx.h, ptr(segMotorOffsets), ptr(segMotorMyo), ptr(bondData),
                                                      ptr(filForceSum), ptr(filTorqueSum), _n3(filForceSum)))


class MotorJointSystem:
    """MotorJointSystem.java:61-198: rod-lever and lever-head joints."""

    @staticmethod
    def joints(ctx: Context, coord, uVec, segLength, bTransGam, bRotGam, forceSum, torqueSum, nucleotideState,
               jointParams):
        jp, jpp = _f(jointParams)
        return ctx.check(ctx.lib.sbx_motor_joints(ctx.h, ptr(coord), ptr(uVec), ptr(segLength), ptr(bTransGam),
                                                  ptr(bRotGam), ptr(forceSum), ptr(torqueSum), ptr(nucleotideState),
                                                  jpp, _n3(coord)))


class DimerCouplingSystem:
    """DimerCouplingSystem.java:83-184: rod-end coupling and lever torsion of a dimer."""

    @staticmethod
    def couple(ctx: Context, coord, uVec, segLength, bTransGam, bRotGam, forceSum, torqueSum, motorA, motorB,
               parallel, dimerParams, boundSeg):
        dp, dpp = _f(dimerParams)
        return ctx.check(ctx.lib.sbx_dimer_couple(ctx.h, ptr(coord), ptr(uVec), ptr(segLength), ptr(bTransGam),
                                                  ptr(bRotGam), ptr(forceSum), ptr(torqueSum), ptr(motorA),
                                                  ptr(motorB), ptr(parallel), dpp, ptr(boundSeg), _n3(coord),
                                                  motorA.numel()))


class MiniFilamentSystem:
    """MiniFilamentSystem.java:65-155: dimer-to-backbone tethers and the backbone gather."""

    @staticmethod
    def tether(ctx: Context, mCoord, mUVec, mSegLength, mBTransGam, mBRotGam, mForceSum, mTorqueSum, bbCoord, bbUVec,
               bbInvDragY, headBackboneSlot, motorA, attachAxial, miniData, miniParams):
        mp, mpp = _f(miniParams)
        return ctx.check(ctx.lib.sbx_mini_tether(ctx.h, ptr(mCoord), ptr(mUVec), ptr(mSegLength), ptr(mBTransGam),
                                                 ptr(mBRotGam), ptr(mForceSum), ptr(mTorqueSum), ptr(bbCoord),
                                                 ptr(bbUVec), ptr(bbInvDragY), ptr(headBackboneSlot), ptr(motorA),
                                                 ptr(attachAxial), ptr(miniData), mpp, _n3(mCoord), _n3(bbCoord),
                                                 headBackboneSlot.numel()))

    @staticmethod
    def backboneGather(ctx: Context, bbDimerOffsets, bbDimerList, miniData, bbForceSum, bbTorqueSum):
        return ctx.check(ctx.lib.sbx_backbone_gather(ctx.h, ptr(bbDimerOffsets), ptr(bbDimerList), ptr(miniData),
                                                     ptr(bbForceSum), ptr(bbTorqueSum), _n3(bbForceSum)))


class NodeSystem:
    """NodeSystem.java:48-231: the radial tether of motor rods to node spheres, its bound-stiffened variant and
    the cross-bridge load injection.  attachKey = [node(nA) | motor(nA)]; the per-node gather is
    MiniFilamentSystem.backboneGather over the CSR of attachNode."""

    @staticmethod
    def tether(ctx: Context, mCoord, mUVec, mSegLength, mBTransGam, mForceSum, mTorqueSum, nCoord, nUVec, nYVec,
               nInvTransY, attachKey, radial, attachCoeffK, nodeData, nodeParams):
        npar, npp = _f(nodeParams)
        return ctx.check(ctx.lib.sbx_node_tether(ctx.h, ptr(mCoord), ptr(mUVec), ptr(mSegLength), ptr(mBTransGam),
                                                 ptr(mForceSum), ptr(mTorqueSum), ptr(nCoord), ptr(nUVec),
                                                 ptr(nYVec), ptr(nInvTransY), ptr(attachKey), ptr(radial),
                                                 ptr(attachCoeffK), ptr(nodeData), npp, _n3(mCoord), _n3(nCoord),
                                                 attachCoeffK.numel()))

    @staticmethod
    def tetherBoundStiffen(ctx: Context, mCoord, mUVec, mSegLength, mBTransGam, mForceSum, mTorqueSum, nCoord, nUVec,
                           nYVec, nInvTransY, attachKey, radial, attachCoeffK, boundSeg, nodeData, nodeParams):
        npar, npp = _f(nodeParams)
        return ctx.check(ctx.lib.sbx_node_tether_bound_stiffen(
            ctx.h, ptr(mCoord), ptr(mUVec), ptr(mSegLength), ptr(mBTransGam), ptr(mForceSum), ptr(mTorqueSum),
            ptr(nCoord), ptr(nUVec), ptr(nYVec), ptr(nInvTransY), ptr(attachKey), ptr(radial), ptr(attachCoeffK),
            ptr(boundSeg), ptr(nodeData), npp, _n3(mCoord), _n3(nCoord), attachCoeffK.numel()))

    @staticmethod
    def xbridgeInject(ctx: Context, bondData, nCoord, nUVec, nYVec, attachKey, radial, attachCoeffK, boundSeg,
                      nodeData, nodeParams=None):
        return ctx.check(ctx.lib.sbx_node_xbridge_inject(ctx.h, ptr(bondData), ptr(nUVec), ptr(nYVec), ptr(attachKey),
                                                         ptr(radial), ptr(attachCoeffK), ptr(boundSeg),
                                                         ptr(nodeData), _n3(nCoord), attachCoeffK.numel()))


class NodeNucleationSystem:
    """NodeNucleationSystem.java:55-257: seeds nucleated at nodes, their node bond and tether, its dissolution.
    nucCounts (host) = [nNodes, step, seed, poolOK]; allocCounts (host) = [C, K] as FilamentBirthSystem's."""

    @staticmethod
    def countBoundFil(ctx: Context, seedNode, nodeBoundFil, nucCounts):
        return ctx.check(ctx.lib.sbx_nuc_count_bound_fil(ctx.h, ptr(seedNode), ptr(nodeBoundFil), seedNode.numel(),
                                                         int(nucCounts[0])))

    @staticmethod
    def emit(ctx: Context, nodeCoord, nodeBoundFil, acceptFlag, reqCoord, reqUVec, reqYVec, nucParams, nucCounts,
             rng: int = RNG_WANG, stream: int = 0):
        npar, npp = _f(nucParams)
        return ctx.check(ctx.lib.sbx_nuc_emit(ctx.h, ptr(nodeCoord), ptr(nodeBoundFil), ptr(acceptFlag), ptr(reqCoord),
                                              ptr(reqUVec), ptr(reqYVec), npp, int(nucCounts[1]), int(nucCounts[2]),
                                              int(nucCounts[3]), int(nucCounts[0]), _n3(reqCoord), rng, stream))

    @staticmethod
    def tagSeeds(ctx: Context, rankOffsets, freeList, freeOffsets, seedNode, allocCounts):
        return ctx.check(ctx.lib.sbx_nuc_tag_seeds(ctx.h, ptr(rankOffsets), ptr(freeList), ptr(freeOffsets),
                                                   ptr(seedNode), int(allocCounts[0]), int(allocCounts[1])))

    @staticmethod
    def initNewborn(ctx: Context, rankOffsets, freeList, freeOffsets, monomerCount, filSegLength, seedParams,
                    allocCounts):
        sp, spp = _f(seedParams)
        return ctx.check(ctx.lib.sbx_nuc_init_newborn(ctx.h, ptr(rankOffsets), ptr(freeList), ptr(freeOffsets),
                                                      ptr(monomerCount), ptr(filSegLength), spp, int(allocCounts[0]),
                                                      int(allocCounts[1])))

    @staticmethod
    def seedTether(ctx: Context, filCoord, filUVec, filSegLength, filBTransGam, filForceSum, filTorqueSum, nodeCoord,
                   nodeInvTransX, seedNode, tetherParams):
        tp, tpp = _f(tetherParams)
        return ctx.check(ctx.lib.sbx_nuc_seed_tether(ctx.h, ptr(filCoord), ptr(filUVec), ptr(filSegLength),
                                                     ptr(filBTransGam), ptr(filForceSum), ptr(filTorqueSum),
                                                     ptr(nodeCoord), ptr(nodeInvTransX), ptr(seedNode), tpp,
                                                     _n3(filCoord), _n3(nodeCoord)))

    @staticmethod
    def seedTetherNodeReact(ctx: Context, filCoord, filUVec, filSegLength, filBTransGam, nodeForceSum, nodeCoord,
                            nodeInvTransX, seedNode, tetherParams):
        tp, tpp = _f(tetherParams)
        return ctx.check(ctx.lib.sbx_nuc_seed_tether_node_react(ctx.h, ptr(filCoord), ptr(filUVec), ptr(filSegLength),
                                                                ptr(filBTransGam), ptr(nodeForceSum), ptr(nodeCoord),
                                                                ptr(nodeInvTransX), ptr(seedNode), tpp,
                                                                _n3(filCoord), _n3(nodeCoord)))

    @staticmethod
    def seedAimTorque(ctx: Context, filUVec, filBRotGam, filTorqueSum, aimDir, aimParams):
        ap, app = _f(aimParams)
        return ctx.check(ctx.lib.sbx_nuc_seed_aim_torque(ctx.h, ptr(filUVec), ptr(filBRotGam), ptr(filTorqueSum),
                                                         ptr(aimDir), app, _n3(filUVec)))

    @staticmethod
    def dissolve(ctx: Context, seedNode, dissolveParams, nucCounts, rng: int = RNG_WANG, stream: int = 0):
        dp, dpp = _f(dissolveParams)
        return ctx.check(ctx.lib.sbx_nuc_dissolve(ctx.h, ptr(seedNode), dpp, int(nucCounts[1]), int(nucCounts[2]),
                                                  seedNode.numel(), rng, stream))
