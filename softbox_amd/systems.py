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
    """BrownianForceSystem.java:170-245 (rng selects the reference Wang stream or Philox)."""

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
    """DerivedGeometrySystem.java:137-186"""

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


class CrossBridgeSystem:
    """CrossBridgeSystem.java:189-298 (CSR inverse: histogram + scan + scatter in one call)."""

    @staticmethod
    def csrBuild(ctx: Context, key, nKeys: int, offsets, items):
        return ctx.check(ctx.lib.sbx_csr_build(ctx.h, ptr(key), key.numel(), nKeys, ptr(offsets), ptr(items)))


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
