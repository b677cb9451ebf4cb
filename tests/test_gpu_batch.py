"""Batched step on the GPU: replicas of uneven size concatenated and advanced together.

Every replica's state after the batched steps -- pose, derived geometry, accumulated forces, crosslink
lifecycle and rings, grid CSR and candidate slices -- is bit-identical to stepping that replica alone
(eager and captured graphs, single stream and overlapped), and one replica is also checked against
the oracle directly.  Launch shapes and the autotuner change nothing but the time.
"""
import numpy as np
import pytest

from _common import assert_same
from oracle import ffi
from softbox_amd import scenes

pytestmark = pytest.mark.gpu

STATE = ("coord", "uVec", "yVec", "zVec", "end1", "end2", "forceSum", "torqueSum")
LINKS = ("linkState", "strainHist", "strainPlace", "torqueMagHist", "torqueMagPlace", "xlinkData", "activeLinkCount")


@pytest.fixture(scope="module")
def ctx():
    from softbox_amd.context import Context

    c = Context(0)
    yield c
    c.close()


def _replicas(seed=3, n_rep=4):
    # uneven sizes and densities; the last replica is box-clamped, one has no crosslinks
    reps = scenes.ensemble(n_rep - 1, seed=seed, fil_range=(60, 600), conc_range=(5.0, 30.0))
    reps[1] = scenes.network(n_fil=90, segs=10, box=reps[1].box, n_links=0, seed=seed * 1000 + 1)
    reps.append(scenes.network(n_fil=140, segs=10, box=2.0, n_links=300, seed=seed * 1000 + 9, periodic=False,
                               confine=False))
    return reps


def _alone(ctx, sc, steps, graph=False, overlap=True, rows=None):
    from softbox_amd.network import Network

    net = Network(ctx, sc.copy())
    net.set_overlap(overlap)
    if graph:
        net.use_graph(2)
    net.step(0, steps)
    fs = sc.fs.copy()
    xl = sc.xl.copy() if sc.xl is not None else None
    net.pull(fs, xl)
    cand = net.candidates(sc.fs.n, int(sc.p["maxCand"]), rows=rows)
    net.close()
    return fs, xl, cand


def _check_replica(bt, r, sc, want, what):
    fs_w, xl_w, cand_w = want
    fs = sc.fs.copy()
    xl = sc.xl.copy() if sc.xl is not None else None
    bt.pull(r, fs, xl)
    for k in STATE:
        assert_same(getattr(fs, k), getattr(fs_w, k), f"{what}: replica {r} {k}")
    if xl is not None:
        for k in LINKS:
            assert_same(getattr(xl, k), getattr(xl_w, k), f"{what}: replica {r} {k}")
    cand, cnt, off, cont, cell = bt.candidates(r)
    wc, wn, woff, wcont, wcell = cand_w
    mc = int(sc.p["maxCand"])
    assert_same(cnt, wn, f"{what}: replica {r} candCount")
    used = (np.arange(mc)[None, :] < np.minimum(wn, mc)[:, None]).ravel()
    assert_same(cand[used], wc[used], f"{what}: replica {r} candPartner")
    assert_same(off, woff, f"{what}: replica {r} gridCellOffsets")
    assert_same(cont, wcont, f"{what}: replica {r} gridCellContents")
    assert_same(cell, wcell, f"{what}: replica {r} bodyCell")


@pytest.mark.parametrize("graph,overlap,staged", [(False, False, 0), (False, True, 0), (True, True, 0), (True, True, 1),
                                                  (True, True, 2)])
def test_batch_replicas_equal_stepping_alone(ctx, graph, overlap, staged):
    from softbox_amd.batch import Batch

    reps = _replicas()
    steps = 12
    want = [_alone(ctx, sc, steps, graph=graph, overlap=overlap) for sc in reps]
    bt = Batch(ctx, [sc.copy() for sc in reps])
    bt.net.set_overlap(overlap)
    if staged:
        bt.net.set_shape(4, 8, 256, staged)
    if graph:
        bt.net.use_graph(2)
    bt.step(0, 5)
    bt.step(5, steps - 5)
    for r, sc in enumerate(reps):
        _check_replica(bt, r, sc, want[r], f"graph={graph} overlap={overlap} staged={staged}")
    bt.close()


def test_batch_replica_matches_oracle(ctx):
    from softbox_amd.batch import Batch

    reps = _replicas(seed=4)
    bt = Batch(ctx, [sc.copy() for sc in reps])
    steps = 8
    bt.step(0, steps)
    r = 2
    orc = ffi.OracleNet(reps[r].copy())
    for t in range(steps):
        orc.step(t, 1)
    fs, xl = reps[r].fs.copy(), reps[r].xl.copy()
    bt.pull(r, fs, xl)
    for k in STATE:
        assert_same(getattr(fs, k), getattr(orc.fs, k), f"oracle: {k}")
    for k in LINKS:
        assert_same(getattr(xl, k), getattr(orc.xl, k), f"oracle: {k}")
    cand, cnt, off, cont, cell = bt.candidates(r)
    assert_same(off, orc.grid["gridCellOffsets"], "oracle: gridCellOffsets")
    assert_same(cont, orc.grid["gridCellContents"], "oracle: gridCellContents")
    assert_same(cnt, orc.grid["candCount"], "oracle: candCount")
    mc = int(reps[r].p["maxCand"])
    used = (np.arange(mc)[None, :] < np.minimum(cnt, mc)[:, None]).ravel()
    assert_same(cand[used], orc.grid["candPartner"][used], "oracle: candPartner")
    bt.close()


def test_launch_shapes_and_tuner_change_only_time(ctx):
    """Every launch shape gives the same steps; the tuner's trial steps are steps of the run."""
    from softbox_amd.batch import Batch

    reps = _replicas(seed=5)
    bt = Batch(ctx, [sc.copy() for sc in reps])
    t = 0
    for shape in ((2, 4, 64, 0), (4, 16, 128, 1), (8, 8, 256, 0), (8, 4, 256, 1), (4, 8, 512, 0), (4, 8, 256, 0),
                  (4, 8, 128, 2), (8, 16, 256, 2)):
        bt.net.set_shape(*shape)
        bt.step(t, 3)
        t += 3
    res = bt.net.tune(t, trial_steps=1)  # (1 + 3 + 3) x 3 x 3 shapes, overlapped and not, 2 + 1 steps each
    assert res["steps"] == 126 * 3 and res["shape"][0] in (2, 4, 8) and res["ms_per_step"] > 0, res
    t += res["steps"]
    bt.step(t, 2)
    t += 2
    for r, sc in enumerate(reps):
        _check_replica(bt, r, sc, _alone(ctx, sc, t), f"shapes + tuner ({t} steps)")
    bt.close()


def test_ensemble_64_replicas_at_stated_size(ctx):
    """The batched ensemble as stated: 64 replicas of the crosslinked-network recipe, filament counts
    log-uniform in [2k, 40k], actin 5-30 uM (9.02M rods, 2.26M links), advanced together for 10 steps
    (captured two-step graphs, overlapped streams, the autotuned launch shape).  Every replica's pose,
    forces, link lifecycle, grid CSR and candidate counts equal stepping it alone, as do the candidate
    slices of its first 2048 bodies; the smallest replica also equals the oracle."""
    from softbox_amd.batch import Batch

    reps = scenes.ensemble(64, seed=0, eta=0.1)
    assert len(reps) == 64 and sum(sc.fs.n for sc in reps) > 8_000_000
    bt = Batch(ctx, [sc.copy() for sc in reps])
    bt.net.use_graph(2)
    tuned = bt.net.tune(0, trial_steps=1)
    t = tuned["steps"]
    bt.step(t, 10)
    t += 10
    rows = 2048
    for r, sc in enumerate(reps):
        fs_w, xl_w, cand_w = _alone(ctx, sc, t, graph=True, overlap=True, rows=(0, rows))
        fs, xl = sc.fs.copy(), sc.xl.copy()
        bt.pull(r, fs, xl)
        what = f"replica {r} ({sc.fs.n} rods)"
        for k in STATE:
            assert_same(getattr(fs, k), getattr(fs_w, k), f"{what}: {k}")
        for k in LINKS:
            assert_same(getattr(xl, k), getattr(xl_w, k), f"{what}: {k}")
        cand, cnt, off, cont, cell = bt.candidates(r, rows=(0, rows))
        wc, wn, woff, wcont, wcell = cand_w
        mc = int(sc.p["maxCand"])
        assert_same(cnt, wn, f"{what}: candCount")
        used = (np.arange(mc)[None, :] < np.minimum(wn[:rows], mc)[:, None]).ravel()
        assert_same(cand[used], wc[used], f"{what}: candPartner of the first {rows} bodies")
        assert_same(off, woff, f"{what}: gridCellOffsets")
        assert_same(cont, wcont, f"{what}: gridCellContents")
        assert_same(cell, wcell, f"{what}: bodyCell")
    small = int(np.argmin([sc.fs.n for sc in reps]))
    orc = ffi.OracleNet(reps[small].copy())
    orc.step(0, t)
    fs, xl = reps[small].fs.copy(), reps[small].xl.copy()
    bt.pull(small, fs, xl)
    for k in STATE:
        assert_same(getattr(fs, k), getattr(orc.fs, k), f"oracle, replica {small}: {k}")
    for k in LINKS:
        assert_same(getattr(xl, k), getattr(orc.xl, k), f"oracle, replica {small}: {k}")
    bt.close()


def test_cell_size_changes_only_slice_order_and_time(ctx):
    """Replicas on grids with 1x, 1.6x and 2.5x their cells: pose, forces and link lifecycle are
    bit-identical, and every body's candidate set (count and members) is the same; only the order
    within a slice follows the grid.  The autotuner picks one of the factors it timed."""
    from softbox_amd.batch import Batch, tune_cell_size

    reps = _replicas(seed=6)
    steps = 9
    runs = {}
    for f in (1.0, 1.6, 2.5):
        tiles = [scenes.retile(sc, f) for sc in reps]
        if f > 1.0:
            assert all(t.p["gridDims"][3] < sc.p["gridDims"][3] for t, sc in zip(tiles, reps) if sc.p["gridOn"])
        bt = Batch(ctx, tiles)
        bt.net.use_graph(2)
        bt.step(0, steps)
        out = []
        for r, sc in enumerate(tiles):
            fs, xl = sc.fs.copy(), sc.xl.copy() if sc.xl is not None else None
            bt.pull(r, fs, xl)
            cand, cnt, _, _, _ = bt.candidates(r)
            mc = int(sc.p["maxCand"])
            sets = [np.sort(cand.reshape(-1, mc)[i, :min(c, mc)]) for i, c in enumerate(cnt)]
            out.append((fs, xl, cnt, sets))
        bt.close()
        runs[f] = out
    for f in (1.6, 2.5):
        for r in range(len(reps)):
            fs_a, xl_a, cnt_a, sets_a = runs[1.0][r]
            fs_b, xl_b, cnt_b, sets_b = runs[f][r]
            for k in STATE:
                assert_same(getattr(fs_b, k), getattr(fs_a, k), f"cell x{f} replica {r}: {k}")
            if xl_a is not None:
                for k in LINKS:
                    assert_same(getattr(xl_b, k), getattr(xl_a, k), f"cell x{f} replica {r}: {k}")
            assert_same(cnt_b, cnt_a, f"cell x{f} replica {r}: candCount")
            for i in range(len(sets_a)):
                assert np.array_equal(sets_a[i], sets_b[i]), f"cell x{f} replica {r} body {i}: candidate set"
    res = tune_cell_size(ctx, reps, factors=(1.0, 2.0), trial_steps=2)
    assert res["factor"] in (1.0, 2.0) and all(v > 0 for v in res["ms_per_step"].values())
