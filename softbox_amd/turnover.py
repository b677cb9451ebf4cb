"""Actin turnover state and its device mirror (the sbx_turnover / sbx_turnover_params of the C ABI).

TurnoverStore mirrors the per-slot lifecycle the reference spreads over FilamentStore (filState,
monomerCount: FilamentStore.java:50, 112-114), NodeNucleationStore.seedNode (the node holding a growing
barbed tip), AgingStore.nucFrac (ATP | ADP-Pi | ADP fractions), SeverStore.cofFrac and ActinPool (free
actin in uM with its integer ledger), plus the cadence parameters of GrowthStore / DepolyStore /
AgingStore / SeverStore (GrowthStore.java, DepolyStore.java, AgingStore.java, SeverStore.java).
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .physics import A_ORTHOG, A_PARALLEL, A_TURNING, ACTIN_MONO_RADIUS, ACTIN_RADIUS, KT_REF

FIL_ACTIVE, FIL_FREE = 0, -1
STD_SEG_LENGTH = 32          # Constants.stdSegLength (Constants.java:42)
ACTIN_SEED = 3               # Constants.actinSeed (Constants.java:72)
MIN_MONOMER_CT = 30          # Constants.minMonomerCt (Constants.java:81)
BIOCHEM_DT = 1.0e-3          # Constants.biochemDeltaT (Constants.java:79)
K_ATP_ON_FORMIN = 11.6       # Constants.kATPOn2WithFormin, /uM/s (Constants.java:80)
K_ATP_OFF1 = 0.8             # Constants.kATPOff1, /s (Constants.java:86)
K_ADP_OFF1 = 2.7             # Constants.kADPOff1, /s (Constants.java:87)
K_HYDROLYSIS = 0.3           # Constants.kHydrolysis, /s (Constants.java:92)
K_DISSOCIATION = 1.0         # Constants.kDissociation, /s (Constants.java:93)
COFILIN_RATE = 0.1           # Constants.cofilinRate, /uM/s (Constants.java:100)
COFILIN_CONC = 3.0           # Constants.cofilinConc, uM (Constants.java:101)
AVOGADRO = 6.022e23          # Constants.AvogadroNum (Constants.java:70)


def um_per_monomer(box_volume_um3: float) -> float:
    """ActinPool's uM per monomer (ActinPool.java): (1e5^3 * 1e6) / (volume * Avogadro)."""
    return (1e5 * 1e5 * 1e5 * 1e6) / (box_volume_um3 * AVOGADRO)


class TurnoverStore:
    """Host turnover state over C slots plus the cadence parameters (defaults: the reference's constants)."""

    def __init__(self, C: int):
        self.C = C
        self.filState = np.full(C, FIL_ACTIVE, np.int32)
        self.monomerCount = np.zeros(C, np.int32)
        self.seedNode = np.full(C, -1, np.int32)
        self.nucFrac = np.zeros(3 * C, np.float32)
        self.nucFrac[:C] = 1.0                      # fresh ATP (AgingStore.setATP)
        self.cofFrac = np.zeros(C, np.float32)
        self.pool = np.zeros(1, np.float64)          # free actin, uM
        self.ledger = np.zeros(2, np.int64)          # monomers taken, returned
        self.stats = np.zeros(8, np.int32)
        # cadence parameters
        self.interval = 100
        self.seed = 0
        self.growthOn = 1
        self.proxy = 1
        self.growParams = np.array([0.0, ACTIN_MONO_RADIUS, 2 * STD_SEG_LENGTH], np.float32)
        self.splitParams = np.array([ACTIN_MONO_RADIUS, 2 * STD_SEG_LENGTH, STD_SEG_LENGTH], np.float32)
        self.dragParams = np.zeros(9, np.float32)
        self.depolyParams = np.array([K_ATP_OFF1 * BIOCHEM_DT, ACTIN_MONO_RADIUS, ACTIN_SEED], np.float32)
        self.depolyRateParams = np.array([K_ATP_OFF1 * BIOCHEM_DT, K_ADP_OFF1 * BIOCHEM_DT], np.float32)
        self.agingParams = np.array([K_HYDROLYSIS * BIOCHEM_DT, K_DISSOCIATION * BIOCHEM_DT], np.float32)
        self.cofilinParams = np.array([COFILIN_CONC * COFILIN_RATE * BIOCHEM_DT, 1.0], np.float32)
        self.birthParams = np.array([1.0, 1.0], np.float32)
        self.onRate = K_ATP_ON_FORMIN
        self.biochemDeltaT = BIOCHEM_DT
        self.uMPerMonomer = 0.0
        self.set_drag(0.1, KT_REF)

    def set_drag(self, aeta: float, kT: float, radius: float = ACTIN_RADIUS) -> None:
        """GrowthStore.dragParams (GrowthStore.java:48-57): halfmono, aeta, radius, kT, the three slender-body
        end corrections, stdSegLength, minMonomerCt (float32, as the reference stores them)."""
        self.dragParams[:] = np.array([ACTIN_MONO_RADIUS, aeta, radius, kT, A_PARALLEL, A_ORTHOG, A_TURNING,
                                       STD_SEG_LENGTH, MIN_MONOMER_CT], np.float32)

    def copy(self) -> "TurnoverStore":
        o = TurnoverStore.__new__(TurnoverStore)
        for k, v in self.__dict__.items():
            setattr(o, k, v.copy() if isinstance(v, np.ndarray) else v)
        return o

    def params(self) -> _lib.SbxTurnoverParams:
        p = _lib.SbxTurnoverParams()
        p.interval, p.seed, p.growthOn, p.proxy = int(self.interval), int(self.seed), int(self.growthOn), int(self.proxy)
        for k in ("growParams", "splitParams", "dragParams", "depolyParams", "depolyRateParams", "agingParams",
                  "cofilinParams", "birthParams"):
            arr = getattr(p, k)
            for i, v in enumerate(getattr(self, k)):
                arr[i] = float(v)
        p.onRate, p.biochemDeltaT, p.uMPerMonomer = float(self.onRate), float(self.biochemDeltaT), float(self.uMPerMonomer)
        return p

    def active_monomers(self) -> int:
        return int(self.monomerCount[self.filState >= 0].astype(np.int64).sum())


class DeviceTurnover:
    """Device mirror of a TurnoverStore; `struct()` gives the sbx_turnover of the C ABI."""

    NAMES = ("filState", "monomerCount", "seedNode", "nucFrac", "cofFrac", "pool", "ledger", "stats")

    def __init__(self, ctx, ts: TurnoverStore, diff=None):
        self.ctx, self.C = ctx, ts.C
        self.t = {k: ctx.to_device(getattr(ts, k)) for k in self.NAMES}
        self.t["bTransDiff"] = ctx.to_device(diff[0]) if diff is not None else None
        self.t["bRotDiff"] = ctx.to_device(diff[1]) if diff is not None else None

    def struct(self) -> _lib.SbxTurnover:
        from .context import ptr

        s = _lib.SbxTurnover()
        s.C = self.C
        for name, _ in _lib.SbxTurnover._fields_[1:]:
            v = self.t[name]
            setattr(s, name, ptr(v).value if v is not None else None)
        return s

    def pull(self, ts: TurnoverStore) -> TurnoverStore:
        self.ctx.sync()
        for k in self.NAMES:
            getattr(ts, k)[:] = self.ctx.to_host(self.t[k])
        return ts


def seed_barbed_tips(fs, ts: TurnoverStore) -> None:
    """Mark every chain's barbed tip (end2 free, end1 bonded or a lone segment) as a formin-held growing tip:
    seedNode = the chain's index (GrowthSystem grows node-held tips; these chains stand as formin-capped)."""
    tips = np.flatnonzero((fs.end2NbrSlot < 0) & (ts.filState >= 0))
    ts.seedNode[tips] = np.arange(tips.size, dtype=np.int32)


def monomers_for(seg_len_um):
    """Monomer count whose (count + 1) * actinMonoRadius is nearest seg_len (DragTensorSystem.java:76-80);
    scalar or array."""
    return np.maximum(1, np.rint(np.asarray(seg_len_um, np.float64) / ACTIN_MONO_RADIUS).astype(np.int64) - 1)


def pool_for(conc_uM: float, box_um: float, ts: TurnoverStore) -> None:
    ts.pool[0] = conc_uM
    ts.uMPerMonomer = um_per_monomer(box_um ** 3)


def growth_conc_for(rate_subunits_per_s: float) -> float:
    """Free actin concentration giving a barbed-end growth rate (P = onRate conc biochemDeltaT per cadence)."""
    return rate_subunits_per_s / K_ATP_ON_FORMIN


__all__ = ["TurnoverStore", "DeviceTurnover", "seed_barbed_tips", "monomers_for", "pool_for", "growth_conc_for",
           "FIL_ACTIVE", "FIL_FREE"]
