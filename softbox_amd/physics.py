"""Physical constants and the host-side drag initialisation.

Constants follow Constants.java:25-67; drag follows DragTensorSystem.java:33-69 (the reference runs
it on the host once at scene build, as here).  The viscosity is an argument: the reference's
Constants.aeta is 0.1 Pa s, the workloads here use water (1e-3 Pa s).
"""
from __future__ import annotations

import math

import numpy as np

BOLTZ = 1.380662e-23                 # J/K      Constants.java:25
TEMP_K = 298.15                      # K        Constants.java:26
KT_REF = BOLTZ * TEMP_K              # J        Constants.java:27  (4.11644e-21)
AETA_REF = 0.1                       # Pa s     Constants.java:33
ACTIN_MONO_RADIUS = 0.0054 / 2.0     # um       Constants.java:37
ACTIN_RADIUS = 0.007 / 2.0           # um       Constants.java:39
A_PARALLEL = -0.20                   # Constants.java:49
A_ORTHOG = 0.84                      # Constants.java:50
A_TURNING = -0.662                   # Constants.java:51


def brownian_force_mag(dt: float, kT: float = KT_REF) -> float:
    """sqrt(2 kT / dt) (Constants.java:65-67)."""
    return math.sqrt(2.0 * kT / dt)


def rod_drag_si(length_um: float, radius_um: float, aeta: float = AETA_REF) -> tuple:
    """Slender-body rod drag (DragTensorSystem.rodDragSI, DragTensorSystem.java:33-45):
    (gT_par, gT_perp, gT_perp, gR_par, gR_perp, gR_perp) in SI."""
    LM = 1.0e-6 * length_um
    RM = radius_um * 1.0e-6
    logT = math.log(LM / (2 * RM))
    gtx = (2 * math.pi * aeta * LM) / (logT + A_PARALLEL)
    gty = (4 * math.pi * aeta * LM) / (logT + A_ORTHOG)
    grx = 4 * math.pi * aeta * RM * RM * LM
    gry = (math.pi * aeta * (LM * LM * LM)) / (3 * (logT + A_TURNING))
    return (gtx, gty, gty, grx, gry, gry)


def sphere_drag_si(radius_um: float, aeta: float = AETA_REF) -> tuple:
    """Isotropic sphere drag (DragTensorSystem.sphereDragSI, DragTensorSystem.java:52-58): 6 pi eta R
    translational and 8 pi eta R^3 rotational, R in metres."""
    RM = radius_um * 1.0e-6
    gt = 6 * math.pi * aeta * RM
    gr = 8 * math.pi * aeta * (RM * RM * RM)
    return (gt, gt, gt, gr, gr, gr)


def set_rod_drag(fs, length_um: float, radius_um: float = ACTIN_RADIUS, aeta: float = AETA_REF,
                 kT: float = KT_REF, idx=None) -> None:
    """Store drag + Einstein diffusion for rods of one length (DragTensorSystem.storeDrag,
    DragTensorSystem.java:61-69) into the planar arrays of `fs` (all rods, or the index array idx)."""
    g = rod_drag_si(length_um, radius_um, aeta)
    sel = slice(None) if idx is None else idx
    fs.segLength[sel] = np.float32(length_um)
    for a in range(3):
        fs.plane("bTransGam")[a][sel] = np.float32(g[a])
        fs.plane("bRotGam")[a][sel] = np.float32(g[3 + a])
        fs.plane("bTransDiff")[a][sel] = np.float32(kT / g[a])
        fs.plane("bRotDiff")[a][sel] = np.float32(kT / g[3 + a])
