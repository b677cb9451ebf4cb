"""Pins of the CPU oracle against what the reference's own records and the mathematics fix.

Each test is chosen so that a plausible restatement mistake (a dropped term, a wrong sign or
index, a transposed operand) fails it.  These run on the CPU (no GPU needed).
"""
import ctypes as C
import math

import numpy as np
import pytest

from _common import assert_same
from oracle import ffi
from softbox_amd import scenes
from softbox_amd.physics import ACTIN_MONO_RADIUS, KT_REF, rod_drag_si
from softbox_amd.stores import FilamentStore


def oracle_drag(length, radius, aeta):
    g = (C.c_double * 6)()
    ffi.load().sbo_rod_drag_si(length, radius, aeta, g)
    return list(g)


# ---------------------------------------------------------------- drag (DragTensorSystem.java:33-45)
def test_drag_matches_recorded_reference_values():
    # JOURNAL.md:2527-2532: L = 0.1755 um (monomerCt 64), aeta = 0.1 Pa s, r = 3.5 nm
    g = oracle_drag(0.1755, 0.0035, 0.1)
    assert abs(g[0] / 3.649e-8 - 1) < 2e-4
    assert abs(g[1] / 5.430e-8 - 1) < 2e-4
    assert abs(g[4] / 2.211e-22 - 1) < 3e-4
    # JOURNAL.md:2532 FDT prediction column: D_par 0.11280, D_perp 0.07581 um^2/s, D_rot 18.615 rad^2/s
    assert abs(KT_REF / g[0] * 1e12 / 0.11280 - 1) < 2e-4
    assert abs(KT_REF / g[1] * 1e12 / 0.07581 - 1) < 2e-4
    assert abs(KT_REF / g[4] / 18.615 - 1) < 2e-4
    # the host (python) restatement agrees with the C oracle exactly
    assert g == list(rod_drag_si(0.1755, 0.0035, 0.1))


def test_drag_free_rod_workload_closed_forms():
    # 1 um rods, d = 7 nm, water, kT = 4.11 pN nm: D_par, D_perp, D_rot_perp
    g = oracle_drag(1.0, 0.0035, 1.0e-3)
    kT = 4.11e-21
    assert abs(kT / g[0] * 1e12 - 3.11485) < 5e-5
    assert abs(kT / g[1] * 1e12 - 1.89757) < 5e-5
    assert abs(kT / g[4] - 16.8759) < 5e-4
    # slender-body limit: gamma_perp/gamma_par rises monotonically toward 2 as L/r grows
    ratios = [oracle_drag(L, 0.0035, 1.0e-3)[1] / oracle_drag(L, 0.0035, 1.0e-3)[0] for L in (0.1, 1, 1e2, 1e6)]
    assert all(a < b for a, b in zip(ratios, ratios[1:])) and 1.85 < ratios[-1] < 2.0


def test_sphere_drag_stokes():
    g = (C.c_double * 6)()
    ffi.load().sbo_sphere_drag_si(0.01, 1.0e-3, g)
    assert abs(g[0] - 6 * math.pi * 1e-3 * 1e-8) < 1e-20        # 6 pi eta r
    assert abs(g[3] - 8 * math.pi * 1e-3 * 1e-24) < 1e-36       # 8 pi eta r^3


# ---------------------------------------------------------------- random streams
def _philox(ctr, key):
    lib = ffi.load()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.sbo_philox4x32_10(c, k, o)
    return list(o)


def test_philox_known_answer_vectors():
    # Random123 kat_vectors, philox4x32_10
    assert _philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert _philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert _philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_wang_hash_java_int_semantics():
    lib = ffi.load()
    # hand walk-through of h(0): (0^61)^(0>>>16)=61; *9=549; 549^(549>>>4)=519; *0x27d4eb2d mod 2^32;
    # ^ (>>>15).  Exercises unsigned shifts and 32-bit wrap-around multiplication.
    x = 519 * 0x27D4EB2D % 2 ** 32
    assert lib.sbo_wang_hash(0) == x ^ (x >> 15)
    # negative Java ints: the hash of -1 (0xFFFFFFFF) uses logical (>>>) shifts
    s = 0xFFFFFFFF
    s = (s ^ 61) ^ (s >> 16)
    s = (s * 9) % 2 ** 32
    s ^= s >> 4
    s = (s * 0x27D4EB2D) % 2 ** 32
    s ^= s >> 15
    assert lib.sbo_wang_hash(0xFFFFFFFF) == s


def test_uniform_mapping_edges():
    lib = ffi.load()
    assert lib.sbo_uniform(0) == 0.0
    assert lib.sbo_uniform(1) == 0.0                      # >>>1 drops the low bit
    assert lib.sbo_uniform(0xFFFFFFFF) == 1.0             # float(2^31-1) rounds to 2^31 = float(2147483647)
    assert lib.sbo_uniform(2) == np.float32(1.0) / np.float32(2147483647.0)


@pytest.mark.parametrize("rng", [0, 1])
def test_brownian_draws_are_standard_normal_and_fdt_scaled(rng):
    """Box-Muller output: mean 0, variance 1, independent force/torque branches; amplitude
    tS*sqrt(2kT/dt)*sqrt(gamma) (BrownianForceSystem.java:95-104)."""
    n = 200000
    g = np.ones(3 * n, np.float32)
    scale = np.ones(n, np.float32)
    rf = np.zeros(3 * n, np.float32)
    rt = np.zeros(3 * n, np.float32)
    ffi.call("sbo_brownian_force", n, rf, rt, g, g, scale, scale, np.float32(1.0), 17, 3, rng, 0, 0)
    for a in (rf, rt):
        x = a.astype(np.float64)
        assert abs(x.mean()) < 5 * math.sqrt(1.0 / x.size)
        assert abs(x.var() - 1.0) < 5 * math.sqrt(2.0 / x.size)
        # kurtosis of a normal is 3
        assert abs((x ** 4).mean() / x.var() ** 2 - 3.0) < 0.05
    assert abs(np.corrcoef(rf[:n], rt[:n])[0, 1]) < 5 / math.sqrt(n)
    # amplitude: gamma = 4 and tS = 0.5 scale the same draws by 0.5 * 2
    rf2 = np.zeros_like(rf)
    rt2 = np.zeros_like(rt)
    ffi.call("sbo_brownian_force", n, rf2, rt2, np.full(3 * n, 4.0, np.float32), g, np.full(n, 0.5, np.float32),
             scale, np.float32(1.0), 17, 3, rng, 0, 0)
    assert np.allclose(rf2, rf, rtol=1e-6, atol=0)
    assert_same(rt2, rt, "torque branch independent of trans drag/scale")


def test_philox_and_wang_streams_differ_and_depend_on_every_key():
    lib = ffi.load()
    h = (C.c_uint32 * 6)()
    base = []
    for args in [(1, 5, 7, 9, 0), (1, 6, 7, 9, 0), (1, 5, 8, 9, 0), (1, 5, 7, 10, 0), (1, 5, 7, 9, 1),
                 (0, 5, 7, 9, 0)]:
        lib.sbo_brownian_bits(*args, h)
        base.append(tuple(h))
    assert len(set(base)) == len(base)


# ---------------------------------------------------------------- integrate / derive
def _one_rod(ux=(1, 0, 0), yx=(0, 1, 0), gam_t=(1e-9, 2e-9, 2e-9), gam_r=(1e-24, 2e-22, 2e-22)):
    fs = FilamentStore(1)
    fs.plane("uVec")[:, 0] = ux
    fs.plane("yVec")[:, 0] = yx
    fs.plane("bTransGam")[:, 0] = gam_t
    fs.plane("bRotGam")[:, 0] = gam_r
    fs.segLength[0] = 1.0
    return fs


def _integrate(fs, dt):
    ffi.call("sbo_integrate", fs.n, fs.coord, fs.uVec, fs.yVec, fs.forceSum, fs.torqueSum, fs.randForce,
             fs.randTorque, fs.bTransGam, fs.bRotGam, np.float32(dt))


def test_integrate_zero_force_is_stationary():
    sc = scenes.free_rods(n=500, seed=2, grid=False)
    fs = sc.fs
    before = fs.copy()
    _integrate(fs, 1e-5)
    assert_same(fs.coord, before.coord, "coord")
    assert np.abs(fs.uVec - before.uVec).max() < 2e-7
    assert np.abs(fs.yVec - before.yVec).max() < 2e-7


def test_integrate_translation_closed_form():
    # lab force along u (body x) moves the centre by dt * 1e6 * F / gamma_par along u; along z by
    # dt * 1e6 * F / gamma_perp.  A transposed frame (R vs R^T) or a dropped 1e6 fails this.
    u = np.array([0.6, 0.8, 0.0])
    y = np.array([-0.8, 0.6, 0.0])
    fs = _one_rod(tuple(u), tuple(y))
    fs.plane("forceSum")[:, 0] = (1e-12 * u + 2e-12 * np.array([0, 0, 1.0])).astype(np.float32)
    _integrate(fs, 1e-5)
    d = fs.plane("coord")[:, 0].astype(np.float64)
    want = 1e-5 * 1e6 * (1e-12 / 1e-9 * u + 2e-12 / 2e-9 * np.array([0, 0, 1.0]))
    assert np.allclose(d, want, rtol=2e-6, atol=1e-12)


def test_integrate_rotation_closed_form():
    # body torque about z (lab z for this frame) turns u toward y by atan(omega dt); y turns back
    # by the same angle (small-angle update of RigidRodLangevinIntegrationSystem.java:91-110)
    fs = _one_rod()
    fs.plane("torqueSum")[:, 0] = (0.0, 0.0, 2e-22 * 50.0)
    dt = 1e-4
    _integrate(fs, dt)
    w = 50.0 * dt
    ang = math.atan(w)
    assert np.allclose(fs.plane("uVec")[:, 0], [math.cos(ang), math.sin(ang), 0.0], atol=2e-7)
    assert np.allclose(fs.plane("yVec")[:, 0], [-math.sin(ang), math.cos(ang), 0.0], atol=2e-7)


def test_derive_invariants():
    sc = scenes.free_rods(n=2000, seed=4, grid=False)
    fs = sc.fs
    g = np.random.default_rng(1)
    fs.plane("yVec")[:] += (g.standard_normal((3, fs.n)) * 1e-3).astype(np.float32)  # drifted frame
    ffi.call("sbo_derive", fs.n, fs.coord, fs.uVec, fs.yVec, fs.zVec, fs.end1, fs.end2, fs.segLength)
    u, y, z = (fs.plane(k).astype(np.float64) for k in ("uVec", "yVec", "zVec"))
    assert np.abs((u * y).sum(0)).max() < 1e-6
    assert np.abs(np.linalg.norm(z, axis=0) - 1).max() < 1e-6
    assert np.abs(np.cross(u.T, y.T).T - z).max() < 1e-5
    c = fs.plane("coord").astype(np.float64)
    assert np.abs(fs.plane("end2") - (c + 0.5 * u)).max() < 1e-6
    assert np.abs(fs.plane("end1") - (c - 0.5 * u)).max() < 1e-6


# ---------------------------------------------------------------- chain: frozen deflection KAT
def _pin(fs, segs, ends, anchors):
    """DeflectionSupport.pinEndpoints (DeflectionSupport.java:40-59), float32 arithmetic."""
    n = fs.n
    f = np.float32
    for s, e, a in zip(segs, ends, anchors):
        c = fs.plane("coord")[:, s].copy()
        u = fs.plane("uVec")[:, s]
        half = f(0.5) * fs.segLength[s]
        ep = c - half * u if e == 1 else c + half * u
        fs.plane("coord")[:, s] = c + (np.asarray(a, np.float32) - ep)


@pytest.mark.parametrize("fracR,ratio_v2", [(0.1, 0.99831), (0.4, 2.18990)])
def test_chain_deflection_known_answer(fracR, ratio_v2):
    """fixtures/filament_characterization_v1.md:11-18 (DiffusionHarness.java:287-385): 11-segment pinned
    chain, Brownian off, midpoint load F = 48 EI frac / span^2, ratio = mean deflection / (frac*span)
    over the second half of 60000 steps at dt = 1e-4, aeta = 0.1.  Tolerance 0.05% (the fixture's)."""
    nSeg, monomerCt, dt, frac = 11, 32, 1e-4, 0.01
    segLen = (monomerCt + 1) * ACTIN_MONO_RADIUS
    span = nSeg * segLen
    EI = KT_REF * 15.0e-6
    forceN = 48.0 * EI * frac / (span * 1e-6) ** 2
    fs = FilamentStore(nSeg)
    x0 = -0.5 * span + 0.5 * segLen
    for k in range(nSeg):
        fs.plane("uVec")[:, k] = (1, 0, 0)
        fs.plane("yVec")[:, k] = (0, 1, 0)
        fs.plane("coord")[:, k] = (np.float32(x0 + k * segLen), 0, 0)
        if k < nSeg - 1:
            fs.end2NbrSlot[k], fs.end2NbrSide[k] = k + 1, 0
        if k > 0:
            fs.end1NbrSlot[k], fs.end1NbrSide[k] = k - 1, 1
    mono = np.full(nSeg, monomerCt, np.int32)
    ffi.call("sbo_drag_run_filament", nSeg, mono, fs.end1NbrSlot, fs.end2NbrSlot, fs.segLength, fs.bTransGam,
             fs.bRotGam, fs.bTransDiff, fs.bRotDiff, C.c_double(0.1), C.c_double(KT_REF))
    fs.set_chain_params(dt)
    cp = fs.chainParams.copy()
    cp[2] = np.float32(fracR)
    mid = nSeg // 2
    ext = np.zeros(3 * nSeg, np.float32)
    ext[nSeg + mid] = np.float32(-forceN)
    anchors = [(-0.5 * span, 0, 0), (0.5 * span, 0, 0)]
    loadM = 60000
    s = 0.0
    cnt = 0
    lib = ffi.load()
    P = ffi.P
    ptrs = [P(getattr(fs, k)) for k in ("coord", "uVec", "segLength", "end2NbrSlot", "end2NbrSide", "end1NbrSlot",
                                        "end1NbrSide", "bTransGam", "bRotGam", "forceSum", "torqueSum")]
    cpp = P(cp)
    for step in range(loadM):
        fs.forceSum[:] = ext                       # seedAccumulators with load on (counts[3] = 1)
        fs.torqueSum[:] = 0.0
        lib.sbo_chain_forces(nSeg, *ptrs, cpp)
        _integrate(fs, dt)
        _pin(fs, [0, nSeg - 1], [1, 2], anchors)
        ffi.call("sbo_derive", nSeg, fs.coord, fs.uVec, fs.yVec, fs.zVec, fs.end1, fs.end2, fs.segLength)
        if step + 1 > loadM // 2 and (step + 1) % 2 == 0:
            cy, cz = float(fs.plane("coord")[1, mid]), float(fs.plane("coord")[2, mid])
            s += math.sqrt(cy * cy + cz * cz)
            cnt += 1
    ratio = (s / cnt) / (frac * span)
    assert abs(ratio / ratio_v2 - 1) < 5e-4, ratio
    # tau_theo = N zeta_perp span^3 / (EI pi^4) = 0.05697 s (fixture: exact)
    tau = nSeg * float(fs.plane("bTransGam")[1, mid]) * (span * 1e-6) ** 3 / (EI * math.pi ** 4)
    assert abs(tau - 0.05697) < 5e-6


def test_chain_newton_third_law():
    """Every joint is evaluated from both sides with the owner sign (ChainBendingForceSystem.java:182-209):
    the chain's total translational force is zero up to float rounding."""
    sc = scenes.network(n_fil=50, segs=10, box=4.0, n_links=0, seed=8, grid=False)
    fs = sc.fs
    g = np.random.default_rng(2)
    fs.plane("coord")[:] += (g.standard_normal((3, fs.n)) * 0.01).astype(np.float32)
    ffi.call("sbo_chain_forces", fs.n, fs.coord, fs.uVec, fs.segLength, fs.end2NbrSlot, fs.end2NbrSide,
             fs.end1NbrSlot, fs.end1NbrSide, fs.bTransGam, fs.bRotGam, fs.forceSum, fs.torqueSum, fs.chainParams)
    F = fs.plane("forceSum").astype(np.float64).reshape(3, 50, 10)
    scale = np.abs(F).max()
    assert scale > 0
    assert np.abs(F.sum(2)).max() < 1e-5 * scale


# ---------------------------------------------------------------- accurate acos/asin
def test_accurate_acos_asin_match_libm():
    lib = ffi.load()
    xs = np.linspace(-1, 1, 2001)
    acos = np.array([lib.sbo_accurate_acos(float(x)) for x in xs])
    assert np.abs(acos - np.arccos(xs)).max() < 1e-7   # sqrt small-angle branch near |x|->1 is approximate
    mid = np.abs(xs) < 0.9
    assert np.abs(acos[mid] - np.arccos(xs[mid])).max() < 1e-14  # polynomial + 2 Newton steps: converged
    s = np.linspace(0, 0.7, 701)
    asin = np.array([lib.sbo_accurate_asin(float(v)) for v in s])
    assert np.abs(asin - np.arcsin(s)).max() < 1e-11   # Taylor seed + 2 Newton steps: ~2e-12 at s = 0.7
