"""Whole-step parity of the actomyosin network: sbx_net_step with motors attached against the
oracle's sbo_contractile_step (DenseContractileHarness.cpuStep, DenseContractileHarness.java:350-429).

Covers the binding stage (grid over filaments + heads, gridReachable, catchSlipRelease with the
break-force cap and the refractory cooldown, bindNearest, the nucleotide cycle), the motor joints,
dimer coupling, minifilament tethers, backbone gather + containment + integration, cross-bridge bond
forces, registerForceDot and the motor -> segment gather inside the filament block.  Integers must
match exactly and floats bit-for-bit.
"""
import numpy as np
import pytest

from _common import assert_same
from oracle import ffi
from softbox_amd import scenes
from softbox_amd.systems import RNG_PHILOX, RNG_WANG

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from softbox_amd.context import Context

    c = Context(0)
    yield c
    c.close()


def _scene(seed, rng, n_mini=12, box=3.0):
    return scenes.contractile(n_fil=60, segs=10, box=box, n_mini=n_mini, dimers_each_end=4, n_links=150, seed=seed,
                              mini_near_fil=True, rng=rng)


def _compare(net, orc, sc, what):
    fs = sc.fs.copy()
    xl = sc.xl.copy()
    net.pull(fs, xl)
    for k in ("coord", "uVec", "yVec", "zVec", "end1", "end2", "forceSum", "torqueSum"):
        assert_same(getattr(fs, k), getattr(orc.fs, k), f"{what}: filament {k}")
    for k in ("linkState", "strainHist", "xlinkData", "activeLinkCount"):
        assert_same(getattr(xl, k), getattr(orc.xl, k), f"{what}: {k}")
    mot = sc.mot.copy()
    mini = sc.mini.copy()
    net.pull_motors(mot, mini)
    for k in ("coord", "uVec", "yVec", "zVec", "end1", "end2", "forceSum", "torqueSum"):
        assert_same(getattr(mot.body, k), getattr(orc.mot.body, k), f"{what}: motor {k}")
    for k in ("boundSeg", "bindArc", "cooldown", "stats", "capStats", "nucleotideState", "forceDotFil", "forceMag",
              "forceDotHist", "forceDotPlace", "bondData", "head", "uVec", "rodUVec"):
        assert_same(getattr(mot, k), getattr(orc.mot, k), f"{what}: motor {k}")
    for k in ("coord", "uVec", "yVec", "zVec", "end1", "end2", "forceSum", "torqueSum"):
        assert_same(getattr(mini.backbone, k), getattr(orc.mini.backbone, k), f"{what}: backbone {k}")
    assert_same(mini.miniData, orc.mini.miniData, f"{what}: miniData")
    rc = net.ctx.to_host(net.mot.t["reachCount"])
    assert_same(rc, orc.mreach["reachCount"], f"{what}: reachCount")
    mc = int(sc.p["maxCand"])
    rs = net.ctx.to_host(net.mot.t["reachSeg"]).reshape(-1, mc)
    want = orc.mreach["reachSeg"].reshape(-1, mc)
    for m in np.flatnonzero(rc):
        c = min(int(rc[m]), mc)
        assert_same(rs[m, :c], want[m, :c], f"{what}: reachSeg[{m}]")


@pytest.mark.parametrize("rng,rod_threads", [(RNG_PHILOX, 256), (RNG_WANG, 256), (RNG_WANG, 512)])
def test_contractile_steps_match_oracle(ctx, rng, rod_threads):
    from softbox_amd.network import Network

    sc = _scene(seed=3, rng=rng)
    orc = ffi.OracleContractile(sc)
    net = Network(ctx, sc)
    net.set_shape(4, 8, rod_threads, 0)
    steps = 40
    for t in range(steps):
        orc.step(t, 1)
    net.step(0, steps)
    assert int((orc.mot.stats[1::2]).sum() + orc.mot.capStats.sum() + (orc.mot.boundSeg >= 0).sum()) > 0, \
        "scene produced no binding events"
    _compare(net, orc, sc, f"rng={rng} after {steps} steps")
    net.close()


def test_contractile_graph_replay_matches_eager(ctx):
    """Captured 2-step graph replays (with eager steps between) against the oracle."""
    from softbox_amd.network import Network

    sc = _scene(seed=4, rng=RNG_PHILOX)
    orc = ffi.OracleContractile(sc)
    net = Network(ctx, sc)
    net.use_graph(2)
    t = 0
    for chunk in (5, 6, 3, 4, 1, 7):
        net.step(t, chunk)
        for k in range(chunk):
            orc.step(t + k, 1)
        t += chunk
        _compare(net, orc, sc, f"graph after {t} steps")
    net.close()


def test_contractile_walls(ctx):
    """Containment walls act on filaments and backbones (box smaller than the scene)."""
    from softbox_amd.network import Network

    sc = _scene(seed=5, rng=RNG_PHILOX)
    b = 0.7 * sc.box
    sc.p["boxParams"] = [1.0e-4, b, b, b, 0.0035, 0.5, 1.0, sc.p["params"][0]]
    orc = ffi.OracleContractile(sc)
    net = Network(ctx, sc)
    for t in range(10):
        orc.step(t, 1)
    net.step(0, 10)
    _compare(net, orc, sc, "walls")
    net.close()


def test_contractile_workload_full_size_matches_oracle(ctx):
    """The actomyosin workload at full size (50k filaments x 10 rods, 2k minifilaments of 16 heads, 100k
    crosslinkers, 30 um box with walls), stepped through captured 2-step graphs: filaments, motors,
    backbones, crosslinks and the reach lists equal the oracle's bit for bit after 4 steps (motors
    bind within them)."""
    from softbox_amd.network import Network

    sc = scenes.contractile(seed=0)
    orc = ffi.OracleContractile(sc.copy())
    net = Network(ctx, sc)
    net.use_graph(2)
    steps = 4
    for t in range(steps):
        orc.step(t, 1)
    net.step(0, steps)
    assert int((orc.mot.boundSeg >= 0).sum()) > 0
    _compare(net, orc, sc, f"full-size contractile workload after {steps} steps")
    net.close()


def test_contractile_workload_2m_rods_matches_oracle(ctx):
    """The same recipe scaled to 2M rods (200k filaments, 8k minifilaments, 400k crosslinkers, 60 um
    box, 47M grid cells): 3 steps (a 2-step graph replay and an eager step) equal the oracle's."""
    from softbox_amd.network import Network

    sc = scenes.contractile(n_fil=200000, box=60.0, n_mini=8000, n_links=400000, seed=0)
    orc = ffi.OracleContractile(sc.copy())
    net = Network(ctx, sc)
    net.use_graph(2)
    steps = 3
    for t in range(steps):
        orc.step(t, 1)
    net.step(0, steps)
    assert int((orc.mot.boundSeg >= 0).sum()) > 0
    _compare(net, orc, sc, f"2M-rod contractile workload after {steps} steps")
    net.close()
