"""Long-horizon accuracy of the FP32 restatement against the FP64 run.

The FP64 run is the same restatement compiled with every float of the state widened to double
(liboracle_softbox_f64.so, `wide=True` oracles): pose, drag, forces, link payloads and rings in double,
the random bits and every discrete rule unchanged.  The FP32 restatement is what the HIP step matches
bit-for-bit (the GPU tests), so its distance from the FP64 run is the GPU path's distance from it.

Gates (1000 steps, checked every 100):
  - positions within 1e-5 relative (|dx| / max(|x|, 1 um)), rod orientations u within 1e-5;
  - the body frame's spin vectors y, z within 5e-5: the spin about u is the fastest degree of freedom
    of a thin rod (its rotational drag about the axis scales as r^2 L, the tumbling one as L^3), each
    step turns y and z by tens of milliradians, and the float rounding of those turns accumulates as a
    random walk;
  - no discrete event differs (link states, strain rings' slots);
  - forces and torques teacher-forced: the FP32 state at the checkpoint is stepped once by both builds
    and the two force sums agree within 1e-5 (norm-relative).  Free-running forces are not gated: the
    stored FP32 pose cannot hold a chain joint's nanometre gap better than float rounding of a
    micrometre coordinate, which both the reference's FP32 stores and this path share; the
    free-running force difference is reported by the test (about 3e-5 for the crosslinked network).
The FP64 build carries no pins of its own beyond the FP32 restatement's (it is the same source); the
first test shows that it does compute in double.
"""
import dataclasses

import numpy as np
import pytest

from _common import small_network
from oracle import ffi
from softbox_amd import scenes

POS_TOL = 1.0e-5
ORI_TOL = 1.0e-5
SPIN_TOL = 5.0e-5
FORCE_TOL = 1.0e-5


def _pos_rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float((np.abs(a - b) / np.maximum(np.abs(b), 1.0)).max())


def _norm_rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_fp64_build_computes_in_double():
    sc = scenes.free_rods(n=256, seed=1, grid=False)
    o32 = ffi.OracleNet(sc)
    o64 = ffi.OracleNet(sc, wide=True)
    assert o64.fs.coord.dtype == np.float64 and o32.fs.coord.dtype == np.float32
    o32.step(0, 1)
    o64.step(0, 1)
    c64 = o64.fs.coord
    # one step apart only by float rounding ...
    assert _pos_rel(o32.fs.coord, c64) < 1e-6
    # ... and the FP64 run's values are not float32 numbers (its arithmetic is double)
    assert np.mean(c64 != c64.astype(np.float32).astype(np.float64)) > 0.9
    assert np.array_equal(o32.fs.brownTransScale, o64.fs.brownTransScale.astype(np.float32))


def _teacher_forced(sc, o32, step):
    """The FP32 run's state stepped once by the FP32 and the FP64 builds: force-sum agreement."""
    st = dataclasses.replace(sc, fs=o32.fs.copy(), xl=None if o32.xl is None else o32.xl.copy())
    a = ffi.OracleNet(st)
    b = ffi.OracleNet(st, wide=True)
    a.step(step, 1)
    b.step(step, 1)
    return _norm_rel(a.fs.forceSum, b.fs.forceSum), _norm_rel(a.fs.torqueSum, b.fs.torqueSum)


CASES = {
    # config-1 free Brownian rods at full size (1024 rods, 10 um periodic cube, water, Philox seed 0)
    "free_rods_1024": lambda: scenes.free_rods(n=1024, seed=0),
    # crosslinked chains at crowded-cytoplasm viscosity (0.1 Pa s): chain bending, links, Bell unbinding
    "network_150": lambda: small_network(seed=5, n_fil=150, links=400, eta=0.1),
    "network_150_wang_box": lambda: small_network(seed=6, n_fil=150, links=400, eta=0.1, periodic=False,
                                                  rng=scenes._lib.SBX_RNG_WANG),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_fp32_tracks_fp64_over_1000_steps(case):
    sc = CASES[case]()
    o32 = ffi.OracleNet(sc)
    o64 = ffi.OracleNet(sc, wide=True)
    worst = dict(pos=0.0, ori=0.0, spin=0.0, free_force=0.0, force=0.0, torque=0.0)
    for t0 in range(0, 1000, 100):
        o32.step(t0, 100)
        o64.step(t0, 100)
        worst["pos"] = max(worst["pos"], _pos_rel(o32.fs.coord, o64.fs.coord))
        worst["ori"] = max(worst["ori"], float(np.abs(o32.fs.uVec - o64.fs.uVec).max()))
        worst["spin"] = max(worst["spin"], *(float(np.abs(getattr(o32.fs, k) - getattr(o64.fs, k)).max())
                                             for k in ("yVec", "zVec")))
        worst["free_force"] = max(worst["free_force"], _norm_rel(o32.fs.forceSum, o64.fs.forceSum))
        if o32.xl is not None:
            assert np.array_equal(o32.xl.linkState, o64.xl.linkState), f"{case}: a link event differs by step {t0 + 100}"
            assert np.array_equal(o32.xl.strainPlace, o64.xl.strainPlace)
        fr, tr = _teacher_forced(sc, o32, t0 + 100)
        worst["force"] = max(worst["force"], fr)
        worst["torque"] = max(worst["torque"], tr)
    print(case, {k: f"{v:.2e}" for k, v in worst.items()})
    assert worst["pos"] <= POS_TOL, worst
    assert worst["ori"] <= ORI_TOL, worst
    assert worst["spin"] <= SPIN_TOL, worst
    assert worst["force"] <= FORCE_TOL, worst
    assert worst["torque"] <= FORCE_TOL, worst
    if o32.xl is not None:
        assert (o32.xl.linkState >= 0).sum() > 0.5 * sc.xl.nLinks, "links died: the network case tests nothing"
