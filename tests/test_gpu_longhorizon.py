"""Long-horizon gates of the HIP step.

  - 1000 steps free-running against the FP32 oracle, compared bit-for-bit every 100 steps: crosslinked
    networks with formation (periodic Philox, boxed Wang) and the actomyosin network (both generators).
  - 1000 steps against the FP64 run (the oracle's FP64 build): positions within 1e-5 relative
    (|dx| / max(|x|, 1 um)), rod orientations within 1e-5, the spin vectors y, z within 5e-5 (the
    spin about a thin rod's axis is its fastest degree of freedom; its float rounding accumulates).
  - 32 seeds of the config-1 free Brownian rods (1024 rods each, one batched step, 10k steps) under the
    reference's diffusion protocols: D_par, D_perp and D_rot, the seed means each within 2 standard
    errors of their closed forms kT / gamma (slender-body drag).
  - 32 seeds of a small actomyosin network at the dense fixture's viscosity (1 Pa s,
    DenseContractileHarness.java:58): its contraction rate (the filament-network RMS extent's relative
    shrink per second, FullSystemDemoHarness.filRms, FullSystemDemoHarness.java:1047-1055) agrees with the
    FP64 run's within 2 standard errors.
"""
import numpy as np
import pytest

from _common import assert_same, small_network
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


def _compare(net, orc, sc, what, motors=False):
    fs = sc.fs.copy()
    xl = sc.xl.copy()
    net.pull(fs, xl)
    for k in ("coord", "uVec", "yVec", "zVec", "end1", "end2", "forceSum", "torqueSum"):
        assert_same(getattr(fs, k), getattr(orc.fs, k), f"{what}: {k}")
    for k in ("linkState", "linkFilA", "linkFilB", "loc1", "loc2", "strainHist", "strainPlace", "torqueMagHist",
              "torqueMagPlace", "xlinkData", "activeLinkCount"):
        assert_same(getattr(xl, k), getattr(orc.xl, k), f"{what}: {k}")
    if motors:
        mot = sc.mot.copy()
        mini = sc.mini.copy()
        net.pull_motors(mot, mini)
        for k in ("boundSeg", "cooldown", "stats", "nucleotideState"):
            assert_same(getattr(mot, k), getattr(orc.mot, k), f"{what}: motor {k}")
        assert_same(mot.body.coord, orc.mot.body.coord, f"{what}: motor coord")
        assert_same(mini.backbone.coord, orc.mini.backbone.coord, f"{what}: backbone coord")


def _network(periodic, rng, seed):
    sc = small_network(seed=seed, n_fil=150, links=400, periodic=periodic, rng=rng, eta=0.1)
    sc.p.update(scenes.formation_params(sc.box, 0.1, sc.fs.n, periodic, on_rate=400.0, interval=50))
    return sc


def _actomyosin(seed, rng=RNG_PHILOX):
    return scenes.contractile(n_fil=60, segs=10, box=3.0, n_mini=12, dimers_each_end=4, n_links=150, seed=seed,
                              mini_near_fil=True, eta=1.0, rng=rng)


@pytest.mark.parametrize("case", ["network_periodic_philox", "network_box_wang", "actomyosin_philox",
                                  "actomyosin_wang"])
def test_1000_steps_match_oracle(ctx, case):
    from softbox_amd.network import Network

    motors = case.startswith("actomyosin")
    if motors:
        sc = _actomyosin(11, RNG_WANG if case.endswith("wang") else RNG_PHILOX)
        orc = ffi.OracleContractile(sc)
    else:
        sc = _network(case == "network_periodic_philox", RNG_PHILOX if case.endswith("philox") else RNG_WANG, 21)
        orc = ffi.OracleNet(sc)
    net = Network(ctx, sc)
    for t0 in range(0, 1000, 100):
        orc.step(t0, 100)
        net.step(t0, 100)
        _compare(net, orc, sc, f"{case} at step {t0 + 100}", motors)
    if motors:
        events = int(orc.mot.stats[1::2].sum() + orc.mot.capStats.sum())
        assert events > 0, "no motor binding events in 1000 steps"
    else:
        assert net.form_stats()["formed_total"] > 0, "no link formed in 1000 steps"
    net.close()


@pytest.mark.parametrize("case", ["free_rods_1024", "network_150"])
def test_1000_steps_within_1e5_of_fp64_run(ctx, case):
    from softbox_amd.network import Network

    sc = scenes.free_rods(n=1024, seed=0) if case == "free_rods_1024" else small_network(seed=5, n_fil=150,
                                                                                           links=400, eta=0.1)
    o64 = ffi.OracleNet(sc, wide=True)
    net = Network(ctx, sc)
    fs = sc.fs.copy()
    worst = dict(pos=0.0, ori=0.0, spin=0.0)
    for t0 in range(0, 1000, 100):
        o64.step(t0, 100)
        net.step(t0, 100)
        net.pull(fs)
        c = fs.coord.astype(np.float64)
        worst["pos"] = max(worst["pos"], float((np.abs(c - o64.fs.coord) / np.maximum(np.abs(o64.fs.coord), 1.0)).max()))
        worst["ori"] = max(worst["ori"], float(np.abs(fs.uVec - o64.fs.uVec).max()))
        worst["spin"] = max(worst["spin"], float(np.abs(fs.yVec - o64.fs.yVec).max()),
                            float(np.abs(fs.zVec - o64.fs.zVec).max()))
    print(case, {k: f"{v:.2e}" for k, v in worst.items()})
    assert worst["pos"] <= 1e-5 and worst["ori"] <= 1e-5 and worst["spin"] <= 5e-5, worst
    net.close()


def _pull_planes(batch, fss, name):
    batch.publish()
    batch.ctx.sync()
    out = []
    for r, fs in enumerate(fss):
        batch.fil[r].pull(fs, (name,))
        out.append(fs.plane(name).astype(np.float64))
    return out


def _within_2sem(per_seed, want, what):
    x = np.asarray(per_seed, np.float64)
    mean, sem = x.mean(), x.std(ddof=1) / np.sqrt(x.size)
    print(f"{what}: mean {mean:.6g} closed form {want:.6g} sem {sem:.3g} z {(mean - want) / sem:+.2f}")
    assert abs(mean - want) <= 2.0 * sem, (what, mean, want, sem)


def _frozen_rods(seed, trans, rot):
    """Config-1 rods with every orientation +x (y = +y) and the given Brownian scales: the fixtures of
    DiffusionHarness.runTranslational / runRotational (DiffusionHarness.java:757-852)."""
    sc = scenes.free_rods(n=1024, seed=seed)
    fs = sc.fs
    fs.plane("uVec")[:] = np.array([[1.0], [0.0], [0.0]], np.float32)
    fs.plane("yVec")[:] = np.array([[0.0], [1.0], [0.0]], np.float32)
    fs.brownTransScale[:] = trans
    fs.brownRotScale[:] = rot
    scenes.derive_host(fs)
    return sc


def test_diffusion_32_seeds(ctx):
    """Config-1 free rods (1024 per seed, seeds 0..31, each protocol one batched run), the reference's
    protocols (DiffusionHarness.java:757-852):
      translational: rotational noise off, orientation frozen along +x, 10k steps; per axis the
        mean-square displacement every 200 steps, D = slope through the origin / 2 (x: D_par; y, z:
        D_perp);
      rotational: translational noise off, u starting along +x, C(t) = <u_x> every 20 steps over
        2000 steps (a fifth of the translational run), D_rot = -1/2 the least-squares slope of ln C
        over the samples with 0.2 < C < 0.95.
    Each seed mean within 2 standard errors of kT / gamma."""
    from softbox_amd.batch import Batch
    from softbox_amd.physics import rod_drag_si

    R = 32
    g = rod_drag_si(1.0, 0.0035, scenes.ETA_WATER)
    kT = scenes.KT_WORKLOAD
    Dpar, Dperp, Drot = kT / g[0] * 1e12, kT / g[1] * 1e12, kT / g[4]

    scs = [_frozen_rods(s, 1.0, 0.0) for s in range(R)]
    dt, box = float(scs[0].p["params"][0]), scs[0].box
    b = Batch(ctx, scs)
    fss = [sc.fs.copy() for sc in scs]
    start = _pull_planes(b, fss, "coord")
    prev = [c.copy() for c in start]
    unwrapped = [c.copy() for c in start]
    ts, msd = [], [[] for _ in range(R)]
    for t in range(0, 10000, 200):
        b.step(t, 200)
        cur = _pull_planes(b, fss, "coord")
        ts.append((t + 200) * dt)
        for r in range(R):
            d = cur[r] - prev[r]
            d -= box * np.rint(d / box)   # 200 steps move a rod ~0.11 um rms: the nearest image is the move
            unwrapped[r] += d
            msd[r].append(((unwrapped[r] - start[r]) ** 2).mean(1))
        prev = cur
    b.close()
    ts = np.array(ts)
    D = [(ts[:, None] * np.array(m)).sum(0) / (ts * ts).sum() / 2 for m in msd]
    _within_2sem([d[0] for d in D], Dpar, "D_par (um^2/s)")
    _within_2sem([0.5 * (d[1] + d[2]) for d in D], Dperp, "D_perp (um^2/s)")

    scs = [_frozen_rods(s, 0.0, 1.0) for s in range(R)]
    b = Batch(ctx, scs)
    fss = [sc.fs.copy() for sc in scs]
    ts, cs = [], [[] for _ in range(R)]
    for t in range(0, 2000, 20):
        b.step(t, 20)
        u = _pull_planes(b, fss, "uVec")
        ts.append((t + 20) * dt)
        for r in range(R):
            cs[r].append(u[r][0].mean())
    b.close()
    ts = np.array(ts)
    drot = []
    for r in range(R):
        c = np.array(cs[r])
        w = (c > 0.2) & (c < 0.95)
        assert w.sum() > 20
        drot.append(-np.polyfit(ts[w], np.log(c[w]), 1)[0] / 2)
    _within_2sem(drot, Drot, "D_rot (rad^2/s)")


def _filament_rms(coord):
    """FullSystemDemoHarness.filRms (FullSystemDemoHarness.java:1047-1055): RMS distance of the segment
    centres from their mean (every segment of these scenes is active)."""
    c = np.asarray(coord, np.float64).reshape(3, -1)
    return float(np.sqrt(((c - c.mean(1, keepdims=True)) ** 2).sum(0).mean()))


def test_contraction_rate_32_seeds_matches_fp64_run(ctx):
    """Contraction rate (rms(200) - rms(1000)) / (rms(200) * 800 dt) per seed, after the first 200 steps'
    relaxation of the placed links and heads; GPU seeds against the FP64 run's seeds."""
    from softbox_amd.network import Network

    R, t_a, t_b = 32, 200, 1000
    gpu, f64 = [], []
    for s in range(R):
        sc = _actomyosin(100 + s)
        dt = float(sc.p["params"][0])
        net = Network(ctx, sc)
        fs = sc.fs.copy()
        net.step(0, t_a)
        net.pull(fs)
        ra = _filament_rms(fs.coord)
        net.step(t_a, t_b - t_a)
        net.pull(fs)
        gpu.append((ra - _filament_rms(fs.coord)) / (ra * (t_b - t_a) * dt))
        net.close()
        o = ffi.OracleContractile(sc, wide=True)
        o.step(0, t_a)
        ra = _filament_rms(o.fs.coord)
        o.step(t_a, t_b - t_a)
        f64.append((ra - _filament_rms(o.fs.coord)) / (ra * (t_b - t_a) * dt))
    g, w = np.array(gpu), np.array(f64)
    sg, sw = g.std(ddof=1) / np.sqrt(R), w.std(ddof=1) / np.sqrt(R)
    se = np.hypot(sg, sw)
    print(f"contraction rate (1/s): GPU {g.mean():.4f} +- {sg:.4f}, FP64 run {w.mean():.4f} +- {sw:.4f}")
    assert g.mean() > 2 * sg, "the network does not contract: the statistic carries no signal"
    assert abs(g.mean() - w.mean()) <= 2 * se
