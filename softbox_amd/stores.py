"""Host-side SoA stores in the reference layout (planar x|y|z planes, plane stride n).

FilamentStore mirrors FilamentStore.java:29-87 / RigidRodBody.java:26-57 and CrosslinkerStore
mirrors CrosslinkerStore.java:31-179 (the fields the mechanics hot path reads and writes).
`DeviceFilaments` / `DeviceCrosslinks` hold the device mirrors (torch tensors used as plain
device memory) and produce the C-ABI structs of include/softbox_hip.h.
"""
from __future__ import annotations

import ctypes as C

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
