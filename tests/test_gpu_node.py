"""GPU parity of the protein-node shell against the oracle.

The NodeSystem entry points (tether, tetherBoundStiffen, xbridgeInject) are called once on the GPU and once in
the oracle on identical inputs; the whole node-shell step (the 1x contractility standard's order,
V2OneXHarness.cpuStep, V2OneXHarness.java:353-423: binding, crosslinker formation and unbinding, motor joints,
the radial node tether, the node gather, containment and integration of the nodes, cross-bridges, crosslink
forces, filaments) runs on the GPU against sbo_contractile_step on the same scene.  Integers must match exactly
and floats bit for bit.
"""
import numpy as np
import pytest

from _common import assert_same
from oracle import ffi
from softbox_amd import scenes
from softbox_amd.systems import NodeSystem, RNG_PHILOX, RNG_WANG
from test_oracle_node import node_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from softbox_amd.context import Context

    c = Context(0)
    yield c
    c.close()


def D(ctx, a):
    return ctx.to_device(np.ascontiguousarray(a))


def H(ctx, t):
    ctx.sync()
    return ctx.to_host(t)


@pytest.mark.parametrize("stiffen", [False, True])
def test_node_tether_matches_oracle(ctx, stiffen):
    mot, nd = node_scene(n_nodes=3, n_sing=8, n_dim=5, seed=4)
    nd.nodeParams[1] = np.float32(0.3)
    b = mot.body
    nMB = b.n
    g = np.random.default_rng(5)
    F = (g.standard_normal(3 * nMB) * 1e-12).astype(np.float32)
    T = (g.standard_normal(3 * nMB) * 1e-19).astype(np.float32)
    bound = np.where(g.random(mot.nM) < 0.4, 3, -1).astype(np.int32)
    Fo, To, data = F.copy(), T.copy(), np.zeros_like(nd.nodeData)
    args = (nMB, nd.nNodes, nd.nAttach, b.coord, b.uVec, b.segLength, b.bTransGam, Fo, To, nd.node.coord,
            nd.node.uVec, nd.node.yVec, nd.nodeInvTransY, nd.attachKey, nd.radial, nd.attachCoeffK)
    if stiffen:
        ffi.call("sbo_node_tether_bound_stiffen", *args, bound, data, nd.nodeParams)
    else:
        ffi.call("sbo_node_tether", *args, data, nd.nodeParams)
    dF, dT, dd = D(ctx, F), D(ctx, T), D(ctx, np.zeros_like(data))
    dev = (D(ctx, b.coord), D(ctx, b.uVec), D(ctx, b.segLength), D(ctx, b.bTransGam), dF, dT, D(ctx, nd.node.coord),
           D(ctx, nd.node.uVec), D(ctx, nd.node.yVec), D(ctx, nd.nodeInvTransY), D(ctx, nd.attachKey),
           D(ctx, nd.radial), D(ctx, nd.attachCoeffK))
    if stiffen:
        NodeSystem.tetherBoundStiffen(ctx, *dev, D(ctx, bound), dd, nd.nodeParams)
    else:
        NodeSystem.tether(ctx, *dev, dd, nd.nodeParams)
    assert np.abs(data).max() > 0
    assert_same(H(ctx, dd), data, "nodeData")
    assert_same(H(ctx, dF), Fo, "rod forceSum")
    assert_same(H(ctx, dT), To, "rod torqueSum")


def test_node_xbridge_inject_matches_oracle(ctx):
    mot, nd = node_scene(n_nodes=2, n_sing=6, n_dim=6, seed=8)
    g = np.random.default_rng(9)
    nd.node.plane("uVec")[:, 1] = (0.0, 1.0, 0.0)
    nd.node.plane("yVec")[:, 1] = (0.0, 0.0, 1.0)
    bond = (g.standard_normal(13 * mot.nM) * 1e-12).astype(np.float32)
    bound = np.where(g.random(mot.nM) < 0.5, 2, -1).astype(np.int32)
    data = (g.standard_normal(6 * nd.nAttach) * 1e-13).astype(np.float32)
    want = data.copy()
    ffi.call("sbo_node_xbridge_inject", nd.nNodes, nd.nAttach, bond, nd.node.uVec, nd.node.yVec, nd.attachKey,
             nd.radial, nd.attachCoeffK, bound, want)
    dd = D(ctx, data)
    NodeSystem.xbridgeInject(ctx, D(ctx, bond), D(ctx, nd.node.coord), D(ctx, nd.node.uVec), D(ctx, nd.node.yVec),
                             D(ctx, nd.attachKey), D(ctx, nd.radial), D(ctx, nd.attachCoeffK), D(ctx, bound), dd)
    assert not np.array_equal(want, data)
    assert_same(H(ctx, dd), want, "nodeData")


def _compare(net, orc, sc, what):
    fs = sc.fs.copy()
    xl = sc.xl.copy()
    net.pull(fs, xl)
    for k in ("coord", "uVec", "yVec", "zVec", "end1", "end2", "forceSum", "torqueSum"):
        assert_same(getattr(fs, k), getattr(orc.fs, k), f"{what}: filament {k}")
    for k in ("linkState", "linkFilA", "linkFilB", "strainHist", "xlinkData", "activeLinkCount"):
        assert_same(getattr(xl, k), getattr(orc.xl, k), f"{what}: {k}")
    mot = sc.mot.copy()
    nd = sc.mini.copy()
    net.pull_motors(mot, nd)
    for k in ("coord", "uVec", "yVec", "zVec", "end1", "end2", "forceSum", "torqueSum"):
        assert_same(getattr(mot.body, k), getattr(orc.mot.body, k), f"{what}: motor {k}")
    for k in ("boundSeg", "bindArc", "cooldown", "stats", "capStats", "nucleotideState", "forceDotFil", "forceMag",
              "forceDotHist", "forceDotPlace", "bondData"):
        assert_same(getattr(mot, k), getattr(orc.mot, k), f"{what}: motor {k}")
    for k in ("coord", "uVec", "yVec", "zVec", "end1", "end2", "forceSum", "torqueSum"):
        assert_same(getattr(nd.node, k), getattr(orc.mini.node, k), f"{what}: node {k}")
    assert_same(nd.nodeData, orc.mini.nodeData, f"{what}: nodeData")


@pytest.mark.parametrize("rng,stiffen,graph,group", [(RNG_WANG, False, 0, 16), (RNG_PHILOX, True, 1, 64),
                                                      (RNG_WANG, True, 2, 32)])
def test_node_shell_step_matches_oracle(ctx, rng, stiffen, graph, group):
    from softbox_amd.network import Network

    sc = scenes.standard_1x(n_nodes=30, n_fil=150, box_xy=2.0, xlink_conc=50.0, rng=rng,
                            bound_coeff=0.25 if stiffen else None)
    orc = ffi.OracleContractile(sc)
    net = Network(ctx, sc)
    net.set_motor_group(group)
    if graph:
        net.use_graph(graph)
    t = 0
    for chunk in (7, 33, 60, 25):
        orc.step(t, chunk)
        net.step(t, chunk)
        t += chunk
        _compare(net, orc, sc, f"rng={rng} stiffen={stiffen} graph={graph} group={group} after {t}")
    assert int((orc.mot.boundSeg >= 0).sum()) > 0 and int((orc.xl.linkState >= 0).sum()) > 0
    assert np.abs(orc.mini.node.coord - sc.mini.node.coord).max() > 1e-4
    net.close()


def test_standard_scene_full_size_runs(ctx):
    """The full 1x standard scene (400 nodes x 24 myosins, 1000 filaments x 10 segments) takes 200 steps on the
    GPU and stays finite, binding heads; its first 3 steps equal the oracle's bit for bit."""
    from softbox_amd.network import Network

    sc = scenes.standard_1x()
    orc = ffi.OracleContractile(sc)
    net = Network(ctx, sc)
    orc.step(0, 3)
    net.step(0, 3)
    _compare(net, orc, sc, "standard 1x after 3")
    net.step(3, 197)
    fs = sc.fs.copy()
    net.pull(fs)
    mot = sc.mot.copy()
    nd = sc.mini.copy()
    net.pull_motors(mot, nd)
    assert np.isfinite(fs.coord).all() and np.isfinite(mot.body.coord).all() and np.isfinite(nd.node.coord).all()
    assert int((mot.boundSeg >= 0).sum()) > 0
    net.close()
