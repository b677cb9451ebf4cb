"""Pins of the oracle's protein-node shell (CPU only).

NodeSystem.tether is checked against the closed-form damping-limited spring and the geometry of a rotated
node frame, Newton's third law per attachment and per node, the node gather against a brute per-attachment
sum and the full-system momentum under cross-bridge load (ProteinNodeHarness gates #1a/#1b,
ProteinNodeHarness.java:275-411); tetherBoundStiffen against tether with the coefficient swapped;
xbridgeInject against the bound heads' forces and the world-frame torque arm.  The 1x standard scene is
checked against java.util.Random's published outputs and its own construction rules.
"""
import math

import numpy as np
import pytest

from oracle import ffi
from softbox_amd import scenes
from softbox_amd.physics import AETA_REF, KT_REF, sphere_drag_si
from softbox_amd.stores import MotorStore, NodeStore

DT = 1.0e-5


def planar(rows):
    a = np.asarray(rows, np.float64)
    return np.ascontiguousarray(a.T.ravel(), dtype=np.float32)


def one_attachment(node_c, node_u, node_y, radial, rod_c, rod_u, coeff, dimer, rod_len=0.08, g_rod=3.0e-9,
                   g_node=9.4e-8):
    """A single node and a single motor (rod, lever, head) with one attachment; returns the tether outputs."""
    nMB = 3
    mC = planar([rod_c, (0, 0, 0), (0, 0, 0)])
    mU = planar([rod_u, (1, 0, 0), (1, 0, 0)])
    mL = np.array([rod_len, 0.008, 0.02], np.float32)
    mG = np.full(3 * nMB, g_rod, np.float32)
    F, T = np.zeros(3 * nMB, np.float32), np.zeros(3 * nMB, np.float32)
    nInv = np.array([1.0 / g_node], np.float32)
    key = np.array([0, 0], np.int32)
    rad = planar([radial])
    ck = np.array([coeff if dimer else -coeff], np.float32)
    nd = np.zeros(6, np.float32)
    ffi.call("sbo_node_tether", nMB, 1, 1, mC, mU, mL, mG, F, T, planar([node_c]), planar([node_u]),
             planar([node_y]), nInv, key, rad, ck, nd, np.array([DT, 0.0], np.float32))
    return F.reshape(3, nMB)[:, 0], T.reshape(3, nMB)[:, 0], nd, dict(mG=mG, nInv=nInv, mL=mL)


def test_tether_closed_form_singlet():
    """Singlet: F = coeff 1e-6 |s - e1| / (dt (1/g_rod + 1/g_node)) toward the surface point s, at the rod centre
    (no torque); the node takes -F at its centre (no torque)."""
    c = np.array([0.3, -0.2, 0.1])
    rad = np.array([0.03, 0.0, 0.04])                  # |r| = 0.05 = NODE_RADIUS
    rod_u = np.array([0.6, 0.0, 0.8])
    rod_c = c + rad + np.array([0.004, -0.003, 0.002]) + 0.04 * rod_u
    coeff = 0.4 / 24
    F, T, nd, a = one_attachment(c, (1, 0, 0), (0, 1, 0), rad, rod_c, rod_u, coeff, dimer=False)
    rc32 = np.float32(rod_c).astype(np.float64)
    ru32 = np.float32(rod_u).astype(np.float64)
    e1 = rc32 - 0.5 * 0.08 * ru32
    s = np.float32(c).astype(np.float64) + np.float32(rad).astype(np.float64)
    d = s - e1
    strain = np.linalg.norm(d)
    g_rod, g_node = float(a["mG"][0]), 1.0 / float(a["nInv"][0])
    mag = np.float64(np.float32(coeff)) * 1e-6 * strain / (DT * (1.0 / g_rod + float(a["nInv"][0])))
    want = mag * d / strain
    assert np.allclose(F, want, rtol=2e-6, atol=0)
    assert mag > 1e-15                                   # a loaded spring, not a degenerate zero
    assert not T.any()
    assert np.array_equal(nd[:3], -F) and not nd[3:].any()
    del g_node


def test_tether_dimer_torques_and_body_frame_arm():
    """Dimer: the force acts at the rod's end1 (torque (-L/2 u 1e-6) x F on the rod); the node's torque uses
    the body-frame offset (ru, ry, rz) 1e-6 as its arm, even when the node frame is rotated."""
    c = np.array([0.0, 0.0, 0.0])
    nu, ny = np.array([0.0, 1.0, 0.0]), np.array([-1.0, 0.0, 0.0])   # z = u x y = (0, 0, 1)
    rad = np.array([0.05, 0.0, 0.0])
    s_world = c + rad[0] * nu                                           # the co-rotated surface point
    rod_u = np.array([0.0, 0.6, 0.8])
    rod_c = s_world + np.array([0.003, 0.001, -0.002]) + 0.04 * rod_u
    F, T, nd, a = one_attachment(c, nu, ny, rad, rod_c, rod_u, 0.08, dimer=True)
    e1 = np.float32(rod_c).astype(np.float64) - 0.04 * np.float32(rod_u).astype(np.float64)
    d = s_world - e1
    assert np.allclose(F / np.linalg.norm(F), d / np.linalg.norm(d), rtol=0, atol=1e-6)
    Rrod = -0.5 * 0.08 * np.float32(rod_u).astype(np.float64) * 1e-6
    assert np.allclose(T, np.cross(Rrod, F.astype(np.float64)), rtol=1e-5, atol=1e-30)
    assert np.array_equal(nd[:3], -F)
    assert np.allclose(nd[3:], np.cross(rad * 1e-6, -F.astype(np.float64)), rtol=1e-5, atol=1e-30)
    world = np.cross((s_world - c) * 1e-6, -F.astype(np.float64))
    assert not np.allclose(nd[3:], world, rtol=1e-3, atol=0)       # the world-frame arm would differ


def test_tether_relaxed_when_end1_on_the_rotated_surface_point():
    """A rod whose end1 sits exactly on the node-frame surface point c + ru u + ry y + rz (u x y) feels nothing."""
    c = np.array([0.5, 0.25, -0.125])
    nu, ny = np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0])     # z = u x y = (0, 1, 0)
    rad = np.array([0.0, 0.0, 0.0625])
    s = c + rad[2] * np.cross(nu, ny)
    rod_u = np.array([0.0, 1.0, 0.0])
    F, T, nd, _ = one_attachment(c, nu, ny, rad, s + 0.0625 * rod_u, rod_u, 0.4, dimer=True, rod_len=0.125)
    assert not F.any() and not T.any() and not nd.any()


def node_scene(n_nodes=2, n_sing=6, n_dim=6, seed=0):
    """ProteinNodeHarness.buildScene's node shells (identity frames, Fibonacci radial directions): singlets
    tethered at the rod centre and dimer arms tethered at end1, every other rod displaced off its point."""
    g = np.random.default_rng(seed)
    nChild = n_sing + n_dim
    per = n_sing + 2 * n_dim
    nM, nA = n_nodes * per, n_nodes * nChild
    mot = MotorStore(nM)
    nd = NodeStore(n_nodes, nA)
    b = mot.body
    R = NodeStore.NODE_RADIUS
    for k in range(n_nodes):
        cx, cy, cz = 0.0, -0.3 + 0.6 * k, 0.0
        nd.node.plane("coord")[:, k] = (cx, cy, cz)
        nd.node.plane("uVec")[:, k] = (1, 0, 0)
        nd.node.plane("yVec")[:, k] = (0, 1, 0)
        for c in range(nChild):
            yy = 1.0 - 2.0 * (c + 0.5) / nChild
            rr = math.sqrt(max(0.0, 1.0 - yy * yy))
            phi = c * 2.399963229728653
            u = (rr * math.cos(phi), yy, rr * math.sin(phi))
            m = k * per + (c if c < n_sing else n_sing + 2 * (c - n_sing))
            scenes._assemble_articulated(b, m, (cx + R * u[0], cy + R * u[1], cz + R * u[2]), u, 0.0)
            if c >= n_sing:
                scenes._assemble_articulated(b, m + 1, (cx + R * u[0], cy + R * u[1], cz + R * u[2]), u, 0.0)
            coeff = 0.4 / n_sing if c < n_sing else 0.4 * 0.2
            nd.attach(k * nChild + c, k, m, R * u[0], R * u[1], R * u[2], coeff, c >= n_sing)
    from softbox_amd.physics import set_rod_drag

    idx = np.arange(nM)
    for role, (L, Rr) in enumerate([(0.08, 0.003), (0.008, 0.002), (0.02, 0.01)]):
        set_rod_drag(b, L, Rr, AETA_REF, KT_REF, idx=3 * idx + role)
    nd.init_node_drag(AETA_REF, KT_REF)
    nd.set_params(DT, KT_REF)
    for a in range(0, nA, 2):                         # load every other tether (ProteinNodeHarness.java:280-283)
        rod = 3 * nd.attachMotor[a]
        b.plane("coord")[:, rod] += np.array([0.004, -0.003, 0.002], np.float32)
    b.plane("coord")[:] += (g.standard_normal((3, 3 * nM)) * 1e-4).astype(np.float32)
    scenes.derive_host(b)
    return mot, nd


def run_tether(mot, nd, boundSeg=None):
    b = mot.body
    nMB = b.n
    F, T = np.zeros(3 * nMB, np.float32), np.zeros(3 * nMB, np.float32)
    data = np.zeros_like(nd.nodeData)
    args = (nMB, nd.nNodes, nd.nAttach, b.coord, b.uVec, b.segLength, b.bTransGam, F, T, nd.node.coord,
            nd.node.uVec, nd.node.yVec, nd.nodeInvTransY, nd.attachKey, nd.radial, nd.attachCoeffK)
    if boundSeg is None:
        ffi.call("sbo_node_tether", *args, data, nd.nodeParams)
    else:
        ffi.call("sbo_node_tether_bound_stiffen", *args, boundSeg, data, nd.nodeParams)
    return F, T, data


def gather(nd, data):
    cnt, off, lst = (np.zeros(nd.nNodes, np.int32), np.zeros(nd.nNodes + 1, np.int32),
                     np.zeros(nd.nAttach, np.int32))
    ffi.call("sbo_csr_build", nd.nNodes, nd.nAttach, nd.attachNode, cnt, off, lst)
    nF, nT = np.zeros(3 * nd.nNodes, np.float32), np.zeros(3 * nd.nNodes, np.float32)
    ffi.call("sbo_backbone_gather", nd.nNodes, off, lst, data, nF, nT)
    return nF, nT


def test_gather_equals_brute_and_momentum():
    """ProteinNodeHarness #1a: the CSR gather of nodeData equals a brute per-attachment sum bit for bit, the
    nodes carry load, both attachment types take part, and each node's force cancels its rods' self-writes."""
    mot, nd = node_scene()
    F, T, data = run_tether(mot, nd)
    nF, nT = gather(nd, data)
    d = data.reshape(-1, 6).astype(np.float64)
    bF = np.zeros((3, nd.nNodes))
    bT = np.zeros((3, nd.nNodes))
    for a in range(nd.nAttach):
        bF[:, nd.attachNode[a]] += d[a, :3]
        bT[:, nd.attachNode[a]] += d[a, 3:]
    assert np.array_equal(nF, bF.astype(np.float32).ravel())
    assert np.array_equal(nT, bT.astype(np.float32).ravel())
    assert np.abs(nF).max() > 1e-15
    assert (nd.attachAtEnd1 == 0).any() and (nd.attachAtEnd1 == 1).any()
    Fp = F.reshape(3, -1).astype(np.float64)
    for k in range(nd.nNodes):
        rods = 3 * nd.attachMotor[nd.attachNode == k]
        assert np.abs(nF.reshape(3, -1)[:, k] + Fp[:, rods].sum(axis=1)).max() < 1e-18


def test_full_system_momentum_under_cross_bridge_load():
    """ProteinNodeHarness #1b: with outward heads bound to displaced segments, tether + node gather + bond
    forces + head force + segment gather leave the motor, node and filament totals summing to zero."""
    mot, nd = node_scene(n_nodes=1)
    b = mot.body
    nM = mot.nM
    heads = [m for m in range(nM) if b.plane("uVec")[0, 3 * m + 2] > 0.2]
    nSeg = len(heads)
    from softbox_amd.physics import ACTIN_RADIUS, set_rod_drag
    from softbox_amd.stores import FilamentStore

    fs = FilamentStore(nSeg)
    for s, m in enumerate(heads):
        h = 3 * m + 2
        tip = b.plane("coord")[:, h].astype(np.float64) + 0.01 * b.plane("uVec")[:, h]
        fs.plane("coord")[:, s] = tip + np.array([0.005, 0.003, 0.0])
        fs.plane("uVec")[:, s] = (0, 1, 0)
        fs.plane("yVec")[:, s] = (1, 0, 0)
        mot.boundSeg[m] = s
    set_rod_drag(fs, 32 * 0.0027, ACTIN_RADIUS)
    mot.bindArc[heads] = 0.5 * fs.segLength[0]
    mot.nucleotideState[:] = 2
    scenes.derive_host(fs)
    F, T, data = run_tether(mot, nd)
    nF, nT = gather(nd, data)
    xb = np.array([1.0e-9, 90.0, 0.4, DT, 0.02, 0.0], np.float32)
    bond = np.zeros(13 * nM, np.float32)
    ffi.call("sbo_bond_forces", nM, nSeg, b.coord, b.uVec, b.yVec, b.bRotGam, fs.coord, fs.uVec, fs.yVec, fs.bRotGam,
             fs.segLength, mot.boundSeg, mot.bindArc, mot.nucleotideState, bond, xb)
    ffi.call("sbo_apply_head_force", nM, bond, F, T)
    cnt, off, lst = np.zeros(nSeg, np.int32), np.zeros(nSeg + 1, np.int32), np.zeros(nM, np.int32)
    ffi.call("sbo_csr_build", nSeg, nM, mot.boundSeg, cnt, off, lst)
    fF, fT = np.zeros(3 * nSeg, np.float32), np.zeros(3 * nSeg, np.float32)
    ffi.call("sbo_motor_seg_gather", nSeg, off, lst, bond, fF, fT)
    assert np.abs(fF).max() > 1e-13
    tot = (F.reshape(3, -1).astype(np.float64).sum(1) + nF.reshape(3, -1).astype(np.float64).sum(1)
           + fF.reshape(3, -1).astype(np.float64).sum(1))
    assert np.abs(tot).max() < 1e-15 * max(1.0, np.abs(fF).max() / 1e-12)


def test_bound_stiffen_swaps_only_the_bound_coefficient():
    """tetherBoundStiffen equals tether for unbound motors and tether with |coeff| = boundCoeff (sign kept) for
    bound ones."""
    mot, nd = node_scene()
    nd.nodeParams[1] = np.float32(0.35)
    bound = np.full(mot.nM, -1, np.int32)
    bound[nd.attachMotor[::3]] = 7
    F, T, data = run_tether(mot, nd, boundSeg=bound)
    nd2 = nd.copy()
    sel = bound[nd.attachMotor] >= 0
    nd2.attachCoeffK[sel] = np.where(nd.attachCoeffK[sel] > 0, 0.35, -0.35).astype(np.float32)
    F2, T2, data2 = run_tether(mot, nd2)
    assert np.array_equal(F, F2) and np.array_equal(T, T2) and np.array_equal(data, data2)
    F0, _, data0 = run_tether(mot, nd)
    assert not np.array_equal(F0, F)
    unb = ~sel
    d, d0 = data.reshape(-1, 6), data0.reshape(-1, 6)
    assert np.array_equal(d[unb], d0[unb])


def test_xbridge_inject_adds_bound_head_forces_with_world_arm():
    """xbridgeInject adds each bound head's cross-bridge force (both heads of a dimer) to its attachment's
    record, with torque (R_world 1e-6) x f where R_world = ru u + ry y + rz (u x y)."""
    nN, nA = 1, 2
    nu, ny = np.array([0.0, 1.0, 0.0]), np.array([0.0, 0.0, 1.0])   # z = (1, 0, 0)
    rad = np.array([[0.01, 0.02, 0.04], [0.0, 0.03, -0.04]])
    key = np.array([0, 0, 0, 1], np.int32)             # attachment 0: singlet motor 0; 1: dimer motors 1, 2
    ck = np.array([-0.1, 0.08], np.float32)
    bond = np.zeros(13 * 3, np.float32)
    f = np.array([[1e-12, -2e-12, 3e-12], [4e-12, 5e-12, -6e-12], [-7e-12, 8e-12, 9e-12]], np.float32)
    for m in range(3):
        bond[13 * m: 13 * m + 3] = f[m]
    boundSeg = np.array([3, 4, 5], np.int32)
    data = np.zeros(6 * nA, np.float32)
    data[:] = np.arange(1, 13, dtype=np.float32) * 1e-13
    start = data.copy()
    ffi.call("sbo_node_xbridge_inject", nN, nA, bond, planar([nu]), planar([ny]), key, planar(rad), ck, boundSeg,
             data)
    nz = np.cross(nu, ny)
    for a, heads in ((0, [0]), (1, [1, 2])):
        fw = f[heads].astype(np.float64).sum(0)
        Rw = rad[a, 0] * nu + rad[a, 1] * ny + rad[a, 2] * nz
        assert np.allclose(data[6 * a: 6 * a + 3], start[6 * a: 6 * a + 3] + fw, rtol=1e-6, atol=0)
        assert np.allclose(data[6 * a + 3: 6 * a + 6], start[6 * a + 3: 6 * a + 6] + np.cross(Rw * 1e-6, fw),
                           rtol=1e-5, atol=1e-30)
    boundSeg[:] = -1
    d2 = start.copy()
    ffi.call("sbo_node_xbridge_inject", nN, nA, bond, planar([nu]), planar([ny]), key, planar(rad), ck, boundSeg, d2)
    assert np.array_equal(d2, start)


def test_java_random_published_outputs():
    """java.util.Random: new Random(42).nextDouble() = 0.7275636800328681, new Random(0).nextDouble() =
    0.730967787376657, new Random(42).nextInt() = -1170105035."""
    assert scenes.JavaRandom(42).nextDouble() == 0.7275636800328681
    assert scenes.JavaRandom(0).nextDouble() == 0.730967787376657
    v = scenes.JavaRandom(42).next(32)
    assert v - (1 << 32) == -1170105035


def test_standard_scene_construction():
    """The 1x standard scene (reduced sizes): every tethered rod starts with its end1 on its node's surface point
    (zero tether force at the initial condition), node drag is the sphere formula, containment every 1e-4 s,
    chains wired as the harness wires them, and the node-shell contractile step stays finite."""
    sc = scenes.standard_1x(n_nodes=12, n_fil=40)
    nd, mot = sc.mini, sc.mot
    F, T, data = run_tether(mot, nd)
    assert np.abs(F).max() < 1e-16 and np.abs(data).max() < 1e-16
    g = sphere_drag_si(0.05, AETA_REF)
    assert np.allclose(nd.node.bTransGam, np.float32(g[0])) and np.allclose(nd.node.bRotGam, np.float32(g[3]))
    assert np.allclose(nd.nodeInvTransY, np.float32(1.0 / g[1]))
    assert sc.p["boxParams"][6] == 10.0 and sc.p["boxParams"][3] == 0.5
    fs = sc.fs
    assert fs.end2NbrSlot[1] == 0 and fs.end2NbrSide[1] == 0 and fs.end1NbrSlot[0] == 1 and fs.end1NbrSide[0] == 1
    assert fs.end2NbrSlot[0] == -1 and fs.end1NbrSlot[9] == -1
    assert np.all(np.abs(nd.node.plane("coord")[2]) <= 0.25)
    o = ffi.OracleContractile(sc)
    o.step(0, 30)
    assert np.isfinite(o.fs.coord).all() and np.isfinite(o.mot.body.coord).all()
    assert np.isfinite(o.mini.node.coord).all()
    assert np.abs(o.mini.node.coord - sc.mini.node.coord).max() > 0     # the nodes are free bodies


@pytest.mark.parametrize("bound_coeff", [None, 0.2])
def test_node_step_wide_build_agrees(bound_coeff):
    """The FP64 build runs the node-shell step and stays within float32 round-off of the FP32 build over six
    steps (the same Brownian draws in both); later steps separate at the first nucleotide or binding decision
    that round-off flips, so this is a build check, not a long-horizon gate."""
    sc = scenes.standard_1x(n_nodes=10, n_fil=30, bound_coeff=bound_coeff)
    a = ffi.OracleContractile(sc)
    w = ffi.OracleContractile(sc, wide=True)
    a.step(0, 6)
    w.step(0, 6)
    assert np.allclose(a.mini.node.coord, w.mini.node.coord, rtol=0, atol=1e-6)
    assert np.allclose(a.mot.body.coord, w.mot.body.coord, rtol=0, atol=1e-6)
    assert np.abs(a.mini.node.coord - sc.mini.node.coord).max() > 1e-4
