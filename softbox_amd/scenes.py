"""Seeded synthetic workloads in the reference store layout.

free_rods      -- free Brownian rigid rods (the FDT workload of DiffusionHarness.java:742-851):
                  N rods of length L, centres uniform in a cube, orientations uniform on S^2.
network        -- crosslinked filament network (DenseContractileHarness.java:180-341 filament part):
                  straight chains of rigid rods with end2->next (side 0) / end1->prev (side 1)
                  topology (DiffusionHarness.java:546-547), crosslinks pre-placed at segment
                  midpoints between bundled or crossing filaments at the link rest length.

Both return a Scene: host stores + the scalar parameter block of the network step.  `periodic=True`
selects the periodic-cube extension (coordinates unwrapped, grid cells wrapped, minimum image in
the broad phase); `periodic=False` is the reference's clamped grid with the containment box
centred at the origin (ContainmentSystem.java:50-53).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import _lib
from .physics import ACTIN_MONO_RADIUS, ACTIN_RADIUS, brownian_force_mag, set_rod_drag
from .stores import CrosslinkerStore, FilamentStore

KT_WORKLOAD = 4.11e-21   # kT = 4.11 pN nm as the workloads state it (reference Constants.kT = 4.11644e-21)
ETA_WATER = 1.0e-3       # Pa s


@dataclasses.dataclass
class Scene:
    name: str
    fs: FilamentStore
    xl: CrosslinkerStore | None
    p: dict                     # sbx_net_params fields
    box: float
    seed: int
    n_fil: int = 0
    segs_per_fil: int = 1
    turn: object = None         # TurnoverStore (actin turnover at its cadence), or None

    def net_params(self) -> _lib.SbxNetParams:
        s = _lib.SbxNetParams()
        fields = {name for name, _ in _lib.SbxNetParams._fields_}
        for k, v in self.p.items():
            if k not in fields:
                continue
            f = getattr(s, k)
            if hasattr(f, "__len__"):
                for i, x in enumerate(v):
                    f[i] = x
            else:
                setattr(s, k, v)
        return s

    def copy(self) -> "Scene":
        return dataclasses.replace(self, fs=self.fs.copy(), xl=None if self.xl is None else self.xl.copy(),
                                   p={k: (list(v) if isinstance(v, (list, tuple, np.ndarray)) else v)
                                      for k, v in self.p.items()},
                                   turn=None if self.turn is None else self.turn.copy())


def _unit_vectors(rng: np.random.Generator, n: int) -> np.ndarray:
    z = 2.0 * rng.random(n) - 1.0
    phi = 2.0 * math.pi * rng.random(n)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.stack([r * np.cos(phi), r * np.sin(phi), z])


def _perp(u: np.ndarray) -> np.ndarray:
    """Unit vector perpendicular to u by Gram-Schmidt against z (or x when u is near z)
    (DiffusionHarness.java:692-702)."""
    a = np.zeros_like(u)
    near_z = np.abs(u[2]) >= 0.9
    a[2, ~near_z] = 1.0
    a[0, near_z] = 1.0
    dot = (a * u).sum(0)
    y = a - dot * u
    return y / np.linalg.norm(y, axis=0)


def grid_for(box: float, bound_r: float, cutoff: float, periodic: bool, min_cell: float = 0.0,
             margin: float = 1.0) -> dict:
    """Grid parameters.  Cell >= 2*maxBoundingRadius + cutoff makes the 27-cell stencil complete
    (SpatialGrid.java:19-27); periodic grids tile the box exactly."""
    need = max(2.0 * bound_r + cutoff, min_cell)
    if periodic:
        nb = max(3, int(math.floor(box / need)))
        cell = box / nb
        lo = 0.0
        box_len = [box, box, box]
    else:
        half = 0.5 * box + margin
        nb = 1 + int(math.ceil(2 * half / need))
        cell = need
        lo = -half
        box_len = [0.0, 0.0, 0.0]
    return dict(gridParams=[lo, lo, lo, cell, 1.0 / cell, cutoff], gridDims=[nb, nb, nb, nb ** 3],
                boxLen=box_len, periodic=1 if periodic else 0)


def retile(sc: "Scene", factor: float) -> "Scene":
    """A copy of the scene whose broad-phase grid has cells `factor` times the scene's (factor >= 1):
    the 27-cell stencil stays complete, so every candidate pair set is the same; the cell ids, CSR
    offsets and the order of each candidate slice are the coarser grid's.  Used by the cell-size
    autotuner; factor 1 is the reference's grid."""
    if factor < 1.0:
        raise ValueError("retile: cells can only grow (the stencil needs cell >= 2 r + cutoff)")
    out = sc.copy()
    if not sc.p.get("gridOn") or factor == 1.0:
        return out
    gp = sc.p["gridParams"]
    out.p.update(grid_for(sc.box, 0.0, float(gp[5]), bool(sc.p["periodic"]), min_cell=factor * float(gp[3])))
    return out


ACTIN_MONO_DIAM = 2.0 * ACTIN_MONO_RADIUS   # Constants.actinMonoDiam = 0.0054 um


def formation_params(box: float, seg_len: float, n_seg: int, periodic: bool, grab: float = 0.04,
                     max_angle: float = 0.6, min_sep: float = 5.0 * ACTIN_MONO_DIAM, max_links: int = 10,
                     on_rate: float = 40.0, conc: float = 1.0, dt: float = 1.0e-5, interval: int = 100,
                     jitter: float = 2.0 * ACTIN_MONO_DIAM, mode: int = 0, req_cap: int | None = None) -> dict:
    """Crosslinker formation at a cadence, with the dense preset of V2OneXHarness (V2OneXHarness.java:74-96:
    aperture 0.6 rad, spacing 5 monomer diameters, 10 links per segment, kon 40 /uM/s at 1 uM, jitter 2
    monomer diameters, every 100 steps) and the workloads' binding radius 40 nm.  P_form =
    1 - exp(-kon conc dt interval) (V2OneXHarness.pForm).  The formation grid is FormationGrid's
    (FormationGrid.java:63-76): cell 2 (L/2 + r) + grab, a cell of margin around the box (periodic: the
    box tiled exactly).  req_cap bounds the candidate list (the reference: n_seg * FORM_MAXC)."""
    cell = 2.0 * (0.5 * seg_len + ACTIN_RADIUS) + grab
    g = grid_for(box, 0.5 * seg_len + ACTIN_RADIUS, grab, periodic, margin=cell)
    p_form = 1.0 - math.exp(-on_rate * conc * dt * interval)
    if req_cap is None:
        req_cap = min(n_seg * 256, max(4096, 16 * n_seg))
    fp = np.array([max_angle, grab * grab, min_sep, max_links, p_form, jitter, grab], np.float32)
    return dict(formOn=1, formInt=int(interval), formParams=[float(v) for v in fp], formMode=int(mode),
                formGridParams=g["gridParams"], formGridDims=g["gridDims"], formReqCap=int(req_cap))


def _base_params(dt: float, kT: float) -> dict:
    return dict(params=[dt, brownian_force_mag(dt, kT)], chainParams=[dt, 0.5, 0.1, 0.265, 0.0, 1.0e-20,
                                                                       ACTIN_MONO_RADIUS],
                chainOn=0, xlParams=[0.0, 0.0, 1.0], offParams=[0.0] * 5, torsionParams=[0.0, 0.0], unbindOn=0,
                boxParams=[1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, dt], confineOn=0, gridOn=0, periodic=0,
                maxCand=_lib.SBX_MAX_CAND, gridParams=[0.0] * 6, gridDims=[1, 1, 1, 1], boxLen=[0.0] * 3,
                actinRadius=ACTIN_RADIUS, seed=0, rng=_lib.SBX_RNG_PHILOX, stream=0)


def free_rods(n: int = 1024, length: float = 1.0, box: float = 10.0, seed: int = 0, dt: float = 1.0e-5,
              periodic: bool = True, eta: float = ETA_WATER, kT: float = KT_WORKLOAD, grid: bool = True,
              cell: float = 1.0, rng: int = _lib.SBX_RNG_PHILOX) -> Scene:
    """Free Brownian rods, centres uniform in the cube, orientations uniform on S^2."""
    g = np.random.default_rng(seed)
    fs = FilamentStore(n)
    u = _unit_vectors(g, n)
    y = _perp(u)
    lo = 0.0 if periodic else -0.5 * box
    c = lo + box * g.random((3, n))
    fs.plane("coord")[:] = c
    fs.plane("uVec")[:] = u
    fs.plane("yVec")[:] = y
    set_rod_drag(fs, length, ACTIN_RADIUS, eta, kT)
    fs.brownTransScale[:] = 1.0
    fs.brownRotScale[:] = 1.0
    fs.set_params(dt, brownian_force_mag(dt, kT))
    fs.set_chain_params(dt)
    derive_host(fs)
    p = _base_params(dt, kT)
    p.update(seed=seed, rng=rng)
    if grid:
        p.update(grid_for(box, 0.5 * length + ACTIN_RADIUS, 0.0, periodic, min_cell=cell), gridOn=1)
    if not periodic:
        p.update(confineOn=1, boxParams=[1.0e-4, box, box, box, ACTIN_RADIUS, 0.5, 1.0, dt])
    return Scene("free_rods", fs, None, p, box, seed, n_fil=n, segs_per_fil=1)


def network(n_fil: int = 20000, segs: int = 10, seg_len: float = 0.1, box: float = 20.0, n_links: int = 50000,
            seed: int = 0, dt: float = 1.0e-5, periodic: bool = True, eta: float = ETA_WATER,
            kT: float = KT_WORKLOAD, rest: float = 0.035, grab: float = 0.04, bell=(0.0, 0.1, 4.26),
            torsion: float = 1.0e-19, unbind: bool = True, grid: bool = True, rng: int = _lib.SBX_RNG_PHILOX,
            confine: bool | None = None, seed_frac: float = 0.2) -> Scene:
    """Crosslinked network of straight `segs`-rod filaments.

    A fraction `seed_frac` of the filaments are placed uniformly (centre in the cube, direction
    uniform on S^2); every other filament attaches to a random earlier filament either parallel at
    distance `rest` (a bundle, several links between facing segments) or crossing it with its
    common perpendicular of length `rest` (one link).  Every link therefore starts at its rest length,
    so the Bell off-rate starts at k0 and the network persists (uniformly random filaments at this
    density have their nearest neighbour ~0.15 um away, which would start every link at strain ~4).
    Links sit at segment midpoints with orientSame = 1 (DenseContractileHarness.java:330-333).
    """
    g = np.random.default_rng(seed)
    n = n_fil * segs
    fs = FilamentStore(n)
    lo = 0.0 if periodic else -0.5 * box
    kk = np.arange(segs) - 0.5 * (segs - 1)
    fc = np.zeros((3, n_fil))
    fu = np.zeros((3, n_fil))
    n_seed = max(1, min(n_fil, int(round(seed_frac * n_fil))))
    fc[:, :n_seed] = lo + box * g.random((3, n_seed))
    fu[:, :n_seed] = _unit_vectors(g, n_seed)
    links = []
    attached = n_fil - n_seed
    r = n_links / attached if attached > 0 else 0.0
    k_par = int(min(segs, max(2, math.ceil(r) + 1)))
    p_par = min(1.0, max(0.0, (r - 1.0) / (k_par - 1.0))) if k_par > 1 else 1.0
    for f in range(n_seed, n_fil):
        par = int(g.integers(0, f))
        up = fu[:, par]
        if g.random() < p_par:
            w = _unit_vectors(g, 1)[:, 0]
            w -= (w @ up) * up
            w /= np.linalg.norm(w)
            fu[:, f] = up
            fc[:, f] = fc[:, par] + rest * w
            for k in g.choice(segs, size=k_par, replace=False):
                links.append((par * segs + int(k), f * segs + int(k)))
        else:
            uf = _unit_vectors(g, 1)[:, 0]
            w = np.cross(up, uf)
            if np.linalg.norm(w) < 1e-6:
                continue
            w /= np.linalg.norm(w)
            ks, kt = int(g.integers(0, segs)), int(g.integers(0, segs))
            mid = fc[:, par] + kk[ks] * seg_len * up
            fu[:, f] = uf
            fc[:, f] = mid + rest * w - kk[kt] * seg_len * uf
            links.append((par * segs + ks, f * segs + kt))
    fy = _perp(fu)
    cen = fc[:, :, None] + fu[:, :, None] * (kk[None, None, :] * seg_len)
    fs.plane("coord")[:] = cen.reshape(3, n)
    fs.plane("uVec")[:] = np.repeat(fu, segs, axis=1)
    fs.plane("yVec")[:] = np.repeat(fy, segs, axis=1)
    idx = np.arange(n)
    pos = idx % segs
    fs.end2NbrSlot[:] = np.where(pos < segs - 1, idx + 1, -1)
    fs.end2NbrSide[:] = np.where(pos < segs - 1, 0, -1)
    fs.end1NbrSlot[:] = np.where(pos > 0, idx - 1, -1)
    fs.end1NbrSide[:] = np.where(pos > 0, 1, -1)
    set_rod_drag(fs, seg_len, ACTIN_RADIUS, eta, kT)
    fs.brownTransScale[:] = 1.0
    fs.brownRotScale[:] = 1.0
    fs.set_params(dt, brownian_force_mag(dt, kT))
    fs.set_chain_params(dt)
    derive_host(fs)
    xl = None
    if n_links > 0:
        xl = CrosslinkerStore(n_links, n)
        order = g.permutation(len(links))[:n_links]
        for m, li in enumerate(order):
            a, b = links[li]
            xl.set_link(m, a, 0.5 * float(fs.segLength[a]), b, 0.5 * float(fs.segLength[b]))
            xl.linkOrientSame[m] = 1
        for m in range(len(order), n_links):     # pool slots beyond the placed links stay FREE
            xl.linkState[m] = -1
        xl.set_params(rest, 0.4, dt)
        xl.set_off_params(bell[0], bell[1], bell[2], dt, rest)
        xl.set_torsion_params(torsion, torsion > 0.0)
        xl.compute_fil_link_ct()
    p = _base_params(dt, kT)
    p.update(seed=seed, rng=rng, chainOn=1)
    if xl is not None:
        p.update(xlParams=list(xl.xlParams), offParams=list(xl.offParams), torsionParams=list(xl.torsionParams),
                 unbindOn=1 if unbind else 0)
    if grid:
        p.update(grid_for(box, 0.5 * seg_len + ACTIN_RADIUS, grab, periodic), gridOn=1)
    if confine if confine is not None else not periodic:
        p.update(confineOn=1, boxParams=[1.0e-4, box, box, box, ACTIN_RADIUS, 0.5, 1.0, dt])
    return Scene("network", fs, xl, p, box, seed, n_fil=n_fil, segs_per_fil=segs)


MONOMERS_PER_UM = 370.0   # actin subunits per micrometre of filament (2.7 nm rise)
AVOGADRO = 6.02214076e23


def ensemble(n_rep: int = 64, seed: int = 0, fil_range=(2000, 40000), conc_range=(5.0, 30.0), segs: int = 10,
             seg_len: float = 0.1, links_per_fil: float = 2.5, **kw) -> list[Scene]:
    """Replicas of the crosslinked-network recipe for a batched step: filament count log-uniform in
    fil_range, actin concentration uniform in conc_range (uM) setting the periodic box, links per
    filament as the 20k-filament network (50k links), replica r seeded with seed * 1000 + r."""
    g = np.random.default_rng(seed)
    out = []
    for r in range(n_rep):
        n_fil = int(round(math.exp(g.uniform(math.log(fil_range[0]), math.log(fil_range[1])))))
        conc = g.uniform(*conc_range) * 1.0e-6
        monomers = n_fil * segs * seg_len * MONOMERS_PER_UM
        box = (monomers / (conc * AVOGADRO) * 1.0e15) ** (1.0 / 3.0)  # litres -> um^3
        out.append(network(n_fil=n_fil, segs=segs, seg_len=seg_len, box=box,
                           n_links=int(round(links_per_fil * n_fil)), seed=seed * 1000 + r, **kw))
    return out


def derive_host(fs: FilamentStore) -> None:
    """Initial derived geometry in float32 with the reference's operation order
    (DerivedGeometrySystem.java:63-71), so the first step reads consistent end1/end2."""
    f = np.float32
    u = fs.plane("uVec")
    y = fs.plane("yVec")
    c = fs.plane("coord")
    ux, uy, uz = u[0].astype(f), u[1].astype(f), u[2].astype(f)
    yx, yy, yz = y[0].astype(f), y[1].astype(f), y[2].astype(f)
    zx = uy * yz - uz * yy
    zy = uz * yx - ux * yz
    zz = ux * yy - uy * yx
    m2 = zx * zx + zy * zy + zz * zz
    inv = np.where(m2 > 0, f(1.0) / np.sqrt(m2, dtype=f), f(1.0)).astype(f)
    zx, zy, zz = zx * inv, zy * inv, zz * inv
    fs.plane("zVec")[:] = np.stack([zx, zy, zz])
    fs.plane("yVec")[:] = np.stack([zy * uz - zz * uy, zz * ux - zx * uz, zx * uy - zy * ux])
    half = fs.segLength * f(0.5)
    fs.plane("end1")[:] = np.stack([c[0] - half * ux, c[1] - half * uy, c[2] - half * uz])
    fs.plane("end2")[:] = np.stack([c[0] + half * ux, c[1] + half * uy, c[2] + half * uz])


@dataclasses.dataclass
class ContractileScene(Scene):
    mot: object = None
    dim: object = None
    mini: object = None


def contractile(n_fil: int = 50000, segs: int = 10, seg_len: float = 0.1, box: float = 30.0,
                n_mini: int = 2000, dimers_each_end: int = 4, n_links: int = 100000, seed: int = 0,
                dt: float = 1.0e-5, eta: float = ETA_WATER, kT: float = KT_WORKLOAD, stall_pn: float = 4.0,
                rng: int = _lib.SBX_RNG_PHILOX, mini_near_fil: bool = False, unbind: bool = True,
                xlink_rest: float = 0.035, bell=(0.0, 0.1, 4.26)) -> ContractileScene:
    """Actomyosin network (DenseContractileHarness.buildScene, DenseContractileHarness.java:112-293):
    the crosslinked filament network above, plus minifilaments of 2*dimers_each_end parallel dimers
    (4 per end -> 16 heads) splayed around a 0.18 um backbone (ContractileAssayHarness.placeDimerAlong,
    ContractileAssayHarness.java:982-1017), in the reference's clamped box with containment walls.
    mini_near_fil=True places each backbone beside a random filament segment (small test scenes)."""
    from .physics import rod_drag_si
    from .stores import DimerStore, MiniFilamentStore, MotorStore

    base = network(n_fil=n_fil, segs=segs, seg_len=seg_len, box=box, n_links=n_links, seed=seed, dt=dt,
                   periodic=False, eta=eta, kT=kT, rest=xlink_rest, bell=bell, unbind=unbind, grid=False, rng=rng,
                   confine=True)
    g = np.random.default_rng(seed + 77)
    nD = n_mini * 2 * dimers_each_end
    nM = 2 * nD
    mot = MotorStore(nM)
    dim = DimerStore(nD)
    mini = MiniFilamentStore(n_mini, nD)
    MYO_COL_TOL, ALIGN_TOL = 0.025, -0.4
    SIN80, COS80 = math.sin(math.radians(80.0)), math.cos(math.radians(80.0))
    rl, ll, hl = MotorStore.ROD_LEN, MotorStore.LEVER_LEN, MotorStore.HEAD_LEN
    b = mot.body
    bb = mini.backbone
    fc = base.fs.plane("coord")
    fu = base.fs.plane("uVec")

    def place_arm(m, rc, d, p, e2, splay):
        rod, lever, head = 3 * m, 3 * m + 1, 3 * m + 2
        b.plane("coord")[:, rod] = rc
        b.plane("uVec")[:, rod] = d
        b.plane("yVec")[:, rod] = p
        lu = COS80 * d + splay * SIN80 * p
        nrm = np.cross(d, p)
        b.plane("coord")[:, lever] = e2 + 0.5 * ll * lu
        b.plane("uVec")[:, lever] = lu
        b.plane("yVec")[:, lever] = nrm
        le2 = e2 + ll * lu
        b.plane("coord")[:, head] = le2 + 0.5 * hl * lu
        b.plane("uVec")[:, head] = lu
        b.plane("yVec")[:, head] = nrm
        b.brownTransScale[[rod, head]] = 1.0
        b.brownRotScale[[rod, head]] = 0.3
        b.brownTransScale[lever] = 0.0
        b.brownRotScale[lever] = 0.0

    d_i = 0
    half = 0.5 * box
    for mfi in range(n_mini):
        if mini_near_fil:
            s = int(g.integers(0, base.fs.n))
            u = fu[:, s].astype(np.float64)
            side = _unit_vectors(g, 1)[:, 0]
            side -= (side @ u) * u
            side /= np.linalg.norm(side)
            c = fc[:, s].astype(np.float64) + 0.06 * side
        else:
            c = (g.random(3) - 0.5) * (box - 1.0)
            u = _unit_vectors(g, 1)[:, 0]
        y = _perp(u[:, None])[:, 0]
        bb.plane("coord")[:, mfi] = c
        bb.plane("uVec")[:, mfi] = u
        bb.plane("yVec")[:, mfi] = y
        bb.brownTransScale[mfi] = 1.0
        bb.brownRotScale[mfi] = 0.3
        for e in range(2):
            dirn = -1.0 if e == 0 else 1.0
            for j in range(dimers_each_end):
                mag = MiniFilamentStore.BACKBONE_LEN / 2.0 - (j + 0.5) / dimers_each_end * MiniFilamentStore.HEAD_ZONE
                ax = dirn * mag
                e1 = c + ax * u
                phi = (j + 0.5) / dimers_each_end * math.pi
                p = math.cos(phi) * y + math.sin(phi) * np.cross(u, y)
                d = dirn * u
                d = d / np.linalg.norm(d)
                p = p / np.linalg.norm(p)
                mA, mB = 2 * d_i, 2 * d_i + 1
                rc = e1 + 0.5 * rl * d
                e2 = e1 + rl * d
                place_arm(mA, rc, d, p, e2, +1)
                place_arm(mB, rc, d, p, e2, -1)
                dim.motorA[d_i], dim.motorB[d_i], dim.parallel[d_i] = mA, mB, 1
                mini.headBackboneSlot[d_i], mini.motorA[d_i], mini.attachAxial[d_i] = mfi, mA, np.float32(ax)
                mot.reach[mA] = mot.reach[mB] = MYO_COL_TOL
                d_i += 1
    # drag (DragTensorSystem.run(MotorStore), DragTensorSystem.java:105-122; MiniFilamentStore.initBackboneDrag)
    from .physics import set_rod_drag

    idx = np.arange(nM)
    for role, (L, R) in enumerate([(rl, MotorStore.ROD_R), (ll, MotorStore.LEVER_R)]):
        set_rod_drag(b, L, R, eta, kT, idx=3 * idx + role)
    gsph = (6 * math.pi * eta * MotorStore.HEAD_R * 1e-6, 8 * math.pi * eta * (MotorStore.HEAD_R * 1e-6) ** 3)
    hidx = 3 * idx + 2
    b.segLength[hidx] = np.float32(hl)
    for a in range(3):
        b.plane("bTransGam")[a][hidx] = np.float32(gsph[0])
        b.plane("bRotGam")[a][hidx] = np.float32(gsph[1])
    set_rod_drag(bb, MiniFilamentStore.BACKBONE_LEN, MiniFilamentStore.BACKBONE_R, eta, kT)
    gb = rod_drag_si(MiniFilamentStore.BACKBONE_LEN, MiniFilamentStore.BACKBONE_R, eta)
    mini.bbInvDragY[:n_mini] = np.float32(1.0 / gb[1])
    mini.bbInvDragY[n_mini:] = np.float32(1.0 / gb[4])
    derive_host(b)
    derive_host(bb)
    bfm = brownian_force_mag(dt, kT)
    mot.bodyParams[:] = np.array([dt, bfm], np.float32)
    mot.set_joint_params(dt, stall_pn)
    mot.set_kin_params(MYO_COL_TOL, ALIGN_TOL, dt, kT)
    mot.kinParams[12] = 1.0                        # v1 12 pN break-force cap on (DenseContractileHarness.java:206)
    mot.set_nuc_params(dt)
    mot.xbParams[:] = np.array([1.0e-9, 90.0, 0.4, dt, hl, 0.0], np.float32)
    dim.set_dimer_params(dt)
    mini.miniParams[:] = np.array([dt, 0.07, 0.01], np.float32)
    mini.bbBodyParams[:] = np.array([dt, bfm], np.float32)
    p = dict(base.p)
    seg_bound = 0.5 * seg_len + ACTIN_RADIUS
    cutoff = MYO_COL_TOL + 0.5 * rl
    p.update(grid_for(box, seg_bound, cutoff, False), gridOn=1)
    p.update(confineOn=1, boxParams=[1.0e-4, box, box, box, ACTIN_RADIUS, 0.5, 1.0, dt])
    p.update(seedMotor=seed + 0x0CDE5, seedBackbone=seed + 0x5C2F11)
    return ContractileScene("contractile", base.fs, base.xl, p, box, seed, n_fil=n_fil, segs_per_fil=segs,
                            mot=mot, dim=dim, mini=mini)


def with_turnover(sc: Scene, spare: int, growth_rate: float = 10.0, proxy: bool = True, cofilin_ratio: float = 1.0,
                  aging: bool = True, interval: int | None = None, seed: int | None = None, speedup: float = 1.0,
                  eta: float | None = None, kT: float = KT_WORKLOAD) -> Scene:
    """The scene with actin turnover: `spare` FREE slots appended to the filament store (parked at random
    positions, inert: Brownian scales 0, no chain links), every segment's monomer count from its length,
    every chain's barbed tip (end2 free) held as a growing tip, free actin set for `growth_rate` subunits/s at
    the barbed ends (GrowthSystem: P = kATPOn2WithFormin conc biochemDeltaT), pointed-end depolymerisation
    (nucleotide-dependent when `proxy`), aging, cofilin severing (off at cofilin_ratio >= 1) and the split
    allocator, every `interval` steps (the biochem cadence: biochemDeltaT / dt).  `speedup` multiplies every
    per-cadence probability (test scenes: many events in few steps)."""
    from .stores import CrosslinkerStore, FilamentStore
    from .turnover import (FIL_FREE, TurnoverStore, growth_conc_for, monomers_for, pool_for, seed_barbed_tips)

    n0 = sc.fs.n
    n = n0 + spare
    g = np.random.default_rng(sc.seed + 7777 if seed is None else seed)
    fs = FilamentStore(n)
    for k in FilamentStore.VEC:
        src = getattr(sc.fs, k).reshape(3, n0)
        dst = getattr(fs, k).reshape(3, n)
        dst[:, :n0] = src
        dst[:, n0:] = src[:, :1]
    for k in FilamentStore.SCALAR_F:
        a = getattr(fs, k)
        a[:n0] = getattr(sc.fs, k)
        a[n0:] = getattr(sc.fs, k)[0]
    for k in FilamentStore.SCALAR_I:
        getattr(fs, k)[:n0] = getattr(sc.fs, k)
    fs.params[:] = sc.fs.params
    fs.chainParams[:] = sc.fs.chainParams
    if spare:
        lo = 0.0 if sc.p.get("periodic") else -0.5 * sc.box
        c = fs.plane("coord")
        c[:, n0:] = lo + sc.box * g.random((3, spare))
        u = _unit_vectors(g, spare)
        fs.plane("uVec")[:, n0:] = u
        fs.plane("yVec")[:, n0:] = _perp(u)
        fs.brownTransScale[n0:] = 0.0
        fs.brownRotScale[n0:] = 0.0
        fs.plane("forceSum")[:, n0:] = 0.0
        fs.plane("torqueSum")[:, n0:] = 0.0
        derive_host(fs)
    xl = None
    if sc.xl is not None:
        xl = CrosslinkerStore(sc.xl.nLinks, n)
        for k, v in sc.xl.__dict__.items():
            if isinstance(v, np.ndarray) and k not in ("filLinkCt", "activeLinkCount", "counts"):
                setattr(xl, k, v.copy())
        xl.filLinkCt[:n0] = sc.xl.filLinkCt
        xl.activeLinkCount[:n0] = sc.xl.activeLinkCount
        xl.counts[:] = sc.xl.counts
        xl.counts[3] = n
    ts = TurnoverStore(n)
    ts.filState[n0:] = FIL_FREE
    ts.monomerCount[:n0] = monomers_for(sc.fs.segLength.astype(np.float64))
    seed_barbed_tips(fs, ts)
    dt = float(sc.p["params"][0])
    ts.interval = int(interval) if interval else max(1, int(round(ts.biochemDeltaT / dt)))
    ts.seed = int(sc.seed)
    ts.proxy = 1 if proxy else 0
    pool_for(growth_conc_for(growth_rate * speedup), sc.box, ts)
    ts.depolyParams[0] = np.float32(ts.depolyParams[0] * speedup)
    ts.depolyRateParams[:] = (ts.depolyRateParams.astype(np.float64) * speedup).astype(np.float32)
    if aging:
        ts.agingParams[:] = (ts.agingParams.astype(np.float64) * speedup).astype(np.float32)
    else:
        ts.agingParams[:] = 0.0
    ts.cofilinParams[0] = np.float32(ts.cofilinParams[0] * speedup)
    ts.cofilinParams[1] = np.float32(cofilin_ratio)
    ts.set_drag(float(eta) if eta is not None else _scene_eta(sc, kT), kT)
    out = dataclasses.replace(sc, fs=fs, xl=xl, p={k: (list(v) if isinstance(v, (list, tuple, np.ndarray)) else v)
                                                     for k, v in sc.p.items()}, turn=ts)
    if out.p.get("formOn"):
        out.p["formReqCap"] = int(min(n * 256, max(4096, 16 * n)))
    return out


def _scene_eta(sc: Scene, kT: float) -> float:
    """The viscosity the scene's drag was set from (inverts the parallel drag of its first rod)."""
    from .physics import rod_drag_si

    L = float(sc.fs.segLength[0])
    g1 = rod_drag_si(L, ACTIN_RADIUS, 1.0)[0]
    return float(f"{float(sc.fs.bTransGam[0]) / g1:.6g}")


class JavaRandom:
    """java.util.Random as its API specification defines it: a 48-bit linear congruential generator
    (multiplier 0x5DEECE66D, increment 0xB, the seed scrambled with the multiplier) whose nextDouble is
    ((next(26) << 27) + next(27)) * 2^-53.  The 1x standard scene draws its node centres and filament
    poses from it, so the scene here starts from the reference harness's initial condition."""

    MULT, ADD, MASK = 0x5DEECE66D, 0xB, (1 << 48) - 1

    def __init__(self, seed: int):
        self.s = (seed ^ self.MULT) & self.MASK

    def next(self, bits: int) -> int:
        self.s = (self.s * self.MULT + self.ADD) & self.MASK
        return self.s >> (48 - bits)

    def nextDouble(self) -> float:
        return ((self.next(26) << 27) + self.next(27)) * (1.0 / (1 << 53))


def _assemble_articulated(b, m, a, d, brown):
    """MotorStore.assembleArticulated (MotorStore.java:272-293) in float32 arithmetic: rod, lever and head
    laid out along d from the anchor a, y = unit(-dy, dx, 0)."""
    f = np.float32
    dx, dy, dz = f(d[0]), f(d[1]), f(d[2])
    px, py, pz = -dy, dx, f(0.0)
    pm = f(math.sqrt(float(px * px + py * py + pz * pz)))
    if pm < f(1.0e-4):
        px, py, pz, pm = f(1.0), f(0.0), f(0.0), f(1.0)
    px, py, pz = px / pm, py / pm, pz / pm
    RL, LL, HL = f(0.080), f(0.008), f(0.020)
    offs = (f(0.5) * RL, RL + f(0.5) * LL, f(0.080 + 0.008) + f(0.5) * HL)
    ax, ay, az = f(a[0]), f(a[1]), f(a[2])
    for role, off in enumerate(offs):
        i = 3 * m + role
        b.plane("coord")[:, i] = (ax + off * dx, ay + off * dy, az + off * dz)
        b.plane("uVec")[:, i] = (dx, dy, dz)
        b.plane("yVec")[:, i] = (px, py, pz)
    b.brownTransScale[3 * m] = b.brownRotScale[3 * m] = brown
    b.brownTransScale[3 * m + 1] = b.brownRotScale[3 * m + 1] = 0.0
    b.brownTransScale[3 * m + 2] = b.brownRotScale[3 * m + 2] = brown


def standard_1x(n_nodes: int = 400, n_sing: int = 24, n_fil: int = 1000, segs: int = 10, box_xy: float = 7.071,
                box_z: float = 0.5, dt: float = 1.0e-5, xlinks: bool = True, xlink_conc: float = 1.0,
                reach: float = 0.006, brown_off: bool = False, bound_coeff: float | None = None,
                rng: int = _lib.SBX_RNG_WANG) -> ContractileScene:
    """The 1x contractility standard (V2OneXHarness.build / buildShells / buildGrid / buildCrosslinkers,
    V2OneXHarness.java:153-349): n_nodes free protein nodes (radius 50 nm) in a box_xy x box_xy x box_z slab,
    each with n_sing singlet myosins tethered radially on a Fibonacci sphere; n_fil static filaments of `segs`
    64-monomer segments placed by the reference's two-point rule; crosslinkers formed every 100 steps
    (dense preset, kon 40 /uM/s); containment every 1e-4 s; viscosity 0.1 Pa s and the reference kT, no
    rescale; the reference's seeds (0x310501 motors/filaments/links, 0x315C2F nodes) and java.util.Random
    initial condition.  bound_coeff switches the node tether to tetherBoundStiffen.  brown_off zeroes every
    Brownian scale (the harness's -brownoff)."""
    from .physics import AETA_REF, KT_REF, brownian_force_mag, rod_drag_si, set_rod_drag
    from .stores import DimerStore, MotorStore, NodeStore

    GOLDEN = 2.399963229728653
    SEED, SEED_NODE = 0x310501, 0x315C2F
    ALIGN_TOL, BROWN_TRANS, BROWN_ROT, NODE_BROWN = -0.4, 1.0, 0.3, 0.03
    FIL_MONO = 64
    aeta, kT = AETA_REF, KT_REF
    half, half_z = 0.5 * box_xy, 0.5 * box_z
    jr = JavaRandom(0x31AC7E5 ^ ((n_nodes << 20) ^ n_fil))
    R = NodeStore.NODE_RADIUS
    rl, ll, hl = MotorStore.ROD_LEN, MotorStore.LEVER_LEN, MotorStore.HEAD_LEN
    marg = R + rl + (ll + hl) + 0.05
    mz = min(half_z - 0.02, marg)
    centers = []
    for _ in range(n_nodes):
        cx = (jr.nextDouble() - 0.5) * 2 * (half - marg)
        cy = (jr.nextDouble() - 0.5) * 2 * (half - marg)
        cz = (jr.nextDouble() - 0.5) * 2 * mz
        centers.append((cx, cy, cz))
    # ---- node shells (buildShells, singlets only) ----
    nM = n_nodes * n_sing
    mot = MotorStore(nM)
    dim = DimerStore(0)
    nd = NodeStore(n_nodes, nM, mode=2 if bound_coeff is not None else 1)
    b = mot.body
    for k, (cx, cy, cz) in enumerate(centers):
        nd.node.plane("coord")[:, k] = (cx, cy, cz)
        nd.node.plane("uVec")[:, k] = (1.0, 0.0, 0.0)
        nd.node.plane("yVec")[:, k] = (0.0, 1.0, 0.0)
        nd.node.brownTransScale[k] = nd.node.brownRotScale[k] = NODE_BROWN
        for c in range(n_sing):
            yy = 1.0 - 2.0 * (c + 0.5) / n_sing
            rr = math.sqrt(max(0.0, 1.0 - yy * yy))
            phi = c * GOLDEN
            ux, uy, uz = rr * math.cos(phi), yy, rr * math.sin(phi)
            m = k * n_sing + c
            _assemble_articulated(b, m, (cx + R * ux, cy + R * uy, cz + R * uz), (ux, uy, uz), BROWN_TRANS)
            b.brownRotScale[3 * m] = b.brownRotScale[3 * m + 2] = BROWN_ROT
            nd.attach(m, k, m, R * ux, R * uy, R * uz, NodeStore.ATTN_FORCE / n_sing, False)
    idx = np.arange(nM)
    for role, (L, Rr) in enumerate([(rl, MotorStore.ROD_R), (ll, MotorStore.LEVER_R)]):
        set_rod_drag(b, L, Rr, aeta, kT, idx=3 * idx + role)
    gsph = (6 * math.pi * aeta * MotorStore.HEAD_R * 1e-6, 8 * math.pi * aeta * (MotorStore.HEAD_R * 1e-6) ** 3)
    hidx = 3 * idx + 2
    b.segLength[hidx] = np.float32(hl)
    for a in range(3):
        b.plane("bTransGam")[a][hidx] = np.float32(gsph[0])
        b.plane("bRotGam")[a][hidx] = np.float32(gsph[1])
    nd.init_node_drag(aeta, kT)
    derive_host(b)
    derive_host(nd.node)
    bfm = brownian_force_mag(dt, kT)
    mot.bodyParams[:] = np.array([dt, bfm], np.float32)
    mot.set_joint_params(dt, 6.0)
    mot.set_kin_params(reach, ALIGN_TOL, dt, kT)
    mot.kinParams[12] = 1.0                       # setFaithfulRelease(true, 0): the 12 pN break-force cap
    mot.set_nuc_params(dt)
    mot.xbParams[:] = np.array([1.0e-9, 90.0, 0.4, dt, hl, 0.0], np.float32)
    nd.set_params(dt, kT, 0.0 if bound_coeff is None else bound_coeff)
    # ---- static filaments (two random points at contour separation) ----
    n = n_fil * segs
    fs = FilamentStore(n)
    L = (FIL_MONO + 1) * ACTIN_MONO_RADIUS
    span = segs * L
    for fi in range(n_fil):
        while True:
            ax_, ay_, az_ = ((jr.nextDouble() - 0.5) * 2 * half, (jr.nextDouble() - 0.5) * 2 * half,
                             (jr.nextDouble() - 0.5) * 2 * half_z)
            bx_, by_, bz_ = ((jr.nextDouble() - 0.5) * 2 * half, (jr.nextDouble() - 0.5) * 2 * half,
                             (jr.nextDouble() - 0.5) * 2 * half_z)
            dx, dy, dz = bx_ - ax_, by_ - ay_, bz_ - az_
            d = math.sqrt(dx * dx + dy * dy + dz * dz)
            if 0.90 * span <= d <= 1.10 * span:
                break
        ux, uy, uz = dx / d, dy / d, dz / d
        cx, cy, cz = 0.5 * (ax_ + bx_), 0.5 * (ay_ + by_), 0.5 * (az_ + bz_)
        yx, yy_, yz = -uy, ux, 0.0
        yn = math.sqrt(yx * yx + yy_ * yy_ + yz * yz)
        if yn < 1e-9:
            yx, yy_, yz = 0.0, -uz, uy
            yn = math.sqrt(yy_ * yy_ + yz * yz)
        yx, yy_, yz = yx / yn, yy_ / yn, yz / yn
        e1 = (cx - 0.5 * span * ux, cy - 0.5 * span * uy, cz - 0.5 * span * uz)
        for i in range(segs):
            sl = fi * segs + i
            fs.plane("coord")[:, sl] = (e1[0] + (0.5 + i) * L * ux, e1[1] + (0.5 + i) * L * uy,
                                        e1[2] + (0.5 + i) * L * uz)
            fs.plane("uVec")[:, sl] = (ux, uy, uz)
            fs.plane("yVec")[:, sl] = (yx, yy_, yz)
            fs.brownTransScale[sl] = 1.0
            fs.brownRotScale[sl] = 0.5 if (i == 0 or i == segs - 1) else 0.0
            # chain links exactly as the harness wires them (V2OneXHarness.java:214-215)
            if i > 0:
                fs.end2NbrSlot[sl], fs.end2NbrSide[sl] = sl - 1, 0
            if i < segs - 1:
                fs.end1NbrSlot[sl], fs.end1NbrSide[sl] = sl + 1, 1
    set_rod_drag(fs, L, ACTIN_RADIUS, aeta, kT)
    fs.set_params(dt, bfm)
    fs.set_chain_params(dt)
    derive_host(fs)
    if brown_off:
        for st in (b, nd.node, fs):
            st.brownTransScale[:] = 0.0
            st.brownRotScale[:] = 0.0
    p = _base_params(dt, kT)
    p.update(seed=SEED, rng=rng, chainOn=1, seedMotor=SEED, seedBackbone=SEED_NODE)
    # ---- binding grid over filaments + heads (buildGrid) ----
    seg_bound = 0.5 * L + ACTIN_RADIUS
    cutoff = reach + 0.5 * rl
    cell = 2.0 * seg_bound + cutoff
    gx, gz = half + 0.5, half_z + 0.3
    nX, nZ = 1 + int(math.ceil(2 * gx / cell)), 1 + int(math.ceil(2 * gz / cell))
    p.update(gridOn=1, periodic=0, boxLen=[0.0, 0.0, 0.0], gridParams=[-gx, -gx, -gz, cell, 1.0 / cell, cutoff],
             gridDims=[nX, nX, nZ, nX * nX * nZ])
    # ---- containment every 1e-4 s ----
    check = max(1, int(round(1.0e-4 / dt)))
    p.update(confineOn=1, boxParams=[1.0e-4, box_xy, box_xy, box_z, ACTIN_RADIUS, 0.5, float(check), dt])
    xl = None
    if xlinks:
        C = max(256, n * 4)
        xl = CrosslinkerStore(C, n)
        rest = 0.0125
        xl.set_params(rest, 0.4, dt)
        xl.set_off_params(1.0, 1.0, 2.0, dt, rest)
        xl.set_torsion_params(1.0e-19, True)
        p.update(xlParams=list(xl.xlParams), offParams=list(xl.offParams), torsionParams=list(xl.torsionParams),
                 unbindOn=1)
        grab = 2.0 * ACTIN_MONO_DIAM
        fcell = 2.0 * (0.5 * L + ACTIN_RADIUS) + grab
        fgx, fgz = half + fcell, half_z + fcell
        fnX, fnZ = 1 + int(math.ceil(2 * fgx / fcell)), 1 + int(math.ceil(2 * fgz / fcell))
        p_form = 1.0 - math.exp(-40.0 * xlink_conc * dt * 100)
        fp = np.array([0.6, grab * grab, 5.0 * ACTIN_MONO_DIAM, 10, p_form, 2.0 * ACTIN_MONO_DIAM, grab], np.float32)
        p.update(formOn=1, formInt=100, formParams=[float(v) for v in fp], formMode=0,
                 formGridParams=[-fgx, -fgx, -fgz, fcell, 1.0 / fcell, grab], formGridDims=[fnX, fnX, fnZ, fnX * fnX * fnZ],
                 formReqCap=int(min(n * 256, max(4096, 16 * n))))
    sc = ContractileScene("standard_1x", fs, xl, p, box_xy, SEED, n_fil=n_fil, segs_per_fil=segs, mot=mot, dim=dim,
                          mini=nd)
    return sc
