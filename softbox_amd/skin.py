"""Learned-skin broad phase: the skin model, its training data and its training loop.

No counterpart in the reference.  The per-step-rebuild network (SpatialGrid.broadPhase every
step) generates the trajectory; the library's kernels compute per-coarse-cell features and labels
(sbx_skin_*), and the MLP is trained with the library's forward/backward/Adam kernels (sbx_mlp_*).
The trained parameters are handed to sbx_net_set_skin(SBX_SKIN_LEARNED).  Whatever the model
predicts, the exact displacement guard keeps the candidate slices identical to the per-step
rebuild's; the model only sizes the neighbour lists.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from .context import Context, ptr

NPARAM = 5122 + 24
NTRAIN = 5122
N_FEATURES = 12


@dataclass
class SkinConfig:
    skin_default: float = 0.1     # um
    skin_min: float = 0.02
    skin_max: float = 0.3
    margin: float = 1.0           # learned skin = margin * exp(out0) (set by calibrate_margin)
    list_cap: int = 256           # neighbour slots per body (bundles are ~10x denser than average)
    horizon: int = 8              # steps a skin is trained to last; the run rebuilds that often
    window: int = 32              # steps tracked after a snapshot (censoring of the horizon label)
    extra_trigger: int = 64       # bodies outside their skins (lagged) that schedule a rebuild
    extra_cap: int = 8192         # capacity of the outside-skin list

    def params(self) -> _lib.SbxSkinParams:
        return _lib.SbxSkinParams(self.skin_default, self.skin_min, self.skin_max, self.margin, self.list_cap,
                                  self.horizon, self.extra_trigger, self.extra_cap)


@dataclass
class ReachSkinConfig:
    """Parameters of the reach skin (sbx_net_set_reach_skin): the binding grid's search from a coarse grid of
    the segments rebuilt between steps."""
    skin_default: float = 0.04    # um: fixed mode, and the first build of learned mode
    skin_min: float = 0.005
    skin_max: float = 0.3
    margin: float = 1.0           # learned skin = margin * exp(out0)
    cell_floor: float = 0.0       # smallest coarse cell, um (0: the binding grid's cell)
    horizon: int = 0              # steps between scheduled rebuilds (0: none)
    extra_trigger: int = 64       # outside-skin segments (lagged) that schedule a rebuild
    extra_cap: int = 4096         # capacity of the outside-skin list
    coverage: float = 0.999       # learned: the margin covers this quantile of the cells' largest displacement

    def params(self) -> _lib.SbxReachSkinParams:
        return _lib.SbxReachSkinParams(self.skin_default, self.skin_min, self.skin_max, self.margin,
                                       self.cell_floor, self.horizon, self.extra_trigger, self.extra_cap)


class SkinModel:
    """Device parameters of the 12 -> 64 -> 64 -> 2 GELU MLP (layout in softbox_amd/csrc/sbx_mlp.h)."""

    def __init__(self, ctx: Context, weights: np.ndarray):
        assert weights.size == NPARAM and ctx.lib.sbx_mlp_nparam() == NPARAM
        self.ctx = ctx
        self.w = ctx.to_device(np.ascontiguousarray(weights, np.float32))

    @staticmethod
    def initial(ctx: Context, mean: np.ndarray, std: np.ndarray, seed: int = 0) -> "SkinModel":
        rng = np.random.default_rng(seed)
        w = np.zeros(NPARAM, np.float32)
        w[0:768] = rng.standard_normal(768) * math.sqrt(2.0 / 12)
        w[832:832 + 4096] = rng.standard_normal(4096) * math.sqrt(2.0 / 64)
        w[4992:4992 + 128] = rng.standard_normal(128) * math.sqrt(1.0 / 64)
        w[NTRAIN:NTRAIN + 12] = mean
        w[NTRAIN + 12:] = std
        return SkinModel(ctx, w)

    def host(self) -> np.ndarray:
        return self.ctx.to_host(self.w)

    def forward(self, x):
        """x: device tensor [B, 12] -> device tensor [B, 2]."""
        import torch

        y = torch.empty((x.shape[0], 2), dtype=torch.float32, device=x.device)
        torch.cuda.synchronize()
        self.ctx.check(self.ctx.lib.sbx_mlp_forward(self.ctx.h, ptr(self.w), ptr(x), int(x.shape[0]), ptr(y)))
        self.ctx.sync()
        return y


def loss_grad(ctx: Context, model: SkinModel, x, t, out_weight=(1.0, 0.25), scratch=None):
    """(loss, grad) of the batch on the GPU; grad is a device tensor of NPARAM floats."""
    import torch

    B = int(x.shape[0])
    need = int(ctx.lib.sbx_mlp_scratch_floats(B))
    if scratch is None or scratch.numel() < need:
        scratch = torch.empty(need, dtype=torch.float32, device=x.device)
    grad = torch.zeros(NPARAM, dtype=torch.float32, device=x.device)
    loss = torch.zeros(1, dtype=torch.float32, device=x.device)
    ow = (C.c_float * 2)(*out_weight)
    torch.cuda.synchronize()
    ctx.check(ctx.lib.sbx_mlp_loss_grad(ctx.h, ptr(model.w), ptr(x), ptr(t), B, ow, ptr(scratch), ptr(grad),
                                        ptr(loss)))
    ctx.sync()
    return float(loss.item()), grad, scratch


def train(ctx: Context, model: SkinModel, x, t, iters: int = 400, batch: int = 4096, lr: float = 3e-3,
          out_weight=(1.0, 0.25), seed: int = 0, log_every: int = 0) -> list:
    """Adam on the GPU kernels; returns the full-data loss before training and after each epoch."""
    import torch

    S = int(x.shape[0])
    batch = min(batch, S)
    m = torch.zeros(NPARAM, dtype=torch.float32, device=x.device)
    v = torch.zeros(NPARAM, dtype=torch.float32, device=x.device)
    g = torch.Generator(device="cpu").manual_seed(seed)
    history = [loss_grad(ctx, model, x, t, out_weight)[0]]
    scratch = None
    per_epoch = max(1, S // batch)
    step = 0
    while step < iters:
        perm = torch.randperm(S, generator=g).to(x.device)
        xs, ts = x[perm].contiguous(), t[perm].contiguous()
        for e in range(per_epoch):
            if step >= iters:
                break
            xb, tb = xs[e * batch:(e + 1) * batch], ts[e * batch:(e + 1) * batch]
            _, grad, scratch = loss_grad(ctx, model, xb, tb, out_weight, scratch)
            step += 1
            ctx.check(ctx.lib.sbx_mlp_adam(ctx.h, ptr(model.w), ptr(grad), ptr(m), ptr(v), step, lr, 0.9, 0.999,
                                           1e-8))
        history.append(loss_grad(ctx, model, x, t, out_weight)[0])
        if log_every:
            print(f"[skin] step {step}: loss {history[-1]:.4f}", flush=True)
    ctx.sync()
    return history


class TrajectorySampler:
    """Features and labels of the skin model from the per-step-rebuild run of a network.

    One sample = one occupied coarse cell at a snapshot t0, members binned at t0 - h (h cycles over
    1..H):
      x = skin_cell_features of the members' motion t0 - h -> t0;
      y = (log rms over members of each one's largest displacement within H steps after t0 (the
           cell's displacement scale), log of the first step a member is 3 scales out (censored at
           `window`)).
    """

    SLACK = 1.0e-3  # um, as the library's list slack

    def __init__(self, ctx: Context, net, cfg: SkinConfig, cell_fn=None):
        """cell_fn(skin) -> the coarse cell a rebuild with that largest skin uses (default: the filament skin's
        rebuild grid; reach_cell_fn gives the reach skin's)."""
        import torch

        self.ctx, self.net, self.cfg = ctx, net, cfg
        sc = net.scene
        self.n = n = sc.fs.n
        p = sc.p
        self.periodic = int(p["periodic"])
        self.boxLen = (C.c_float * 3)(*(p["boxLen"] if self.periodic else [0.0, 0.0, 0.0]))
        gp, gd = p["gridParams"], p["gridDims"]
        self.fine_cell, self.cutoff, self.origin = float(gp[3]), float(gp[5]), [float(v) for v in gp[:3]]
        self.fine_dims = [int(v) for v in gd[:3]]
        self.box = float(p["boxLen"][0]) if self.periodic else None
        self.max_r = float((0.5 * sc.fs.segLength + p["actinRadius"]).max())
        if cell_fn is not None:
            self._cell = cell_fn
        # coarse cells span what the learned mode can build with: skins in [skin_min, skin_max]
        self.cell_lo = self._cell(cfg.skin_min)
        self.cell_hi = self._cell(cfg.skin_max)
        nmax = int(np.prod(self._dims(self.cell_lo)))
        dev = f"cuda:{ctx.device}"
        self.cell = torch.empty(n, dtype=torch.int32, device=dev)
        self.off = torch.empty(nmax + 1, dtype=torch.int32, device=dev)
        self.cont = torch.empty(n, dtype=torch.int32, device=dev)
        self.pPrev = torch.empty(3 * n, dtype=torch.float32, device=dev)
        self.uPrev = torch.empty(3 * n, dtype=torch.float32, device=dev)
        self.x0 = torch.empty(3 * n, dtype=torch.float32, device=dev)
        self.maxDisp = torch.empty(n, dtype=torch.float32, device=dev)
        self.thr = torch.empty(n, dtype=torch.float32, device=dev)
        self.breach = torch.empty(n, dtype=torch.int32, device=dev)
        self.feat = torch.empty(nmax * N_FEATURES, dtype=torch.float32, device=dev)
        self.sStar = torch.empty(nmax, dtype=torch.float32, device=dev)
        self.hStar = torch.empty(nmax, dtype=torch.float32, device=dev)
        self.mStar = torch.empty(nmax, dtype=torch.float32, device=dev)

    def _cell(self, skin):
        """The rebuild grid's cell for a largest skin (sbx_network.hip k_skin_geom)."""
        return max(self.fine_cell, 2.0 * self.max_r + self.cutoff + 2.0 * skin + 2.0 * self.SLACK)

    def _dims(self, cell):
        if self.periodic:
            nb = max(3, int(math.floor(self.box / cell)))
            return [nb, nb, nb]
        return [max(1, int(math.ceil(self.fine_cell * d / cell))) for d in self.fine_dims]

    def _grid(self, cell):
        dims = self._dims(cell)
        if self.periodic:
            cell = self.box / dims[0]
        gp = (C.c_float * 6)(self.origin[0], self.origin[1], self.origin[2], cell, 1.0 / cell, 0.0)
        gd = (C.c_int * 4)(dims[0], dims[1], dims[2], dims[0] * dims[1] * dims[2])
        return gp, gd, cell, dims[0] * dims[1] * dims[2]

    def _step(self, t):
        self.net.step(t, 1)
        self.net.publish()

    def _copy(self, dst, src):
        self.ctx.check(self.ctx.lib.sbx_copy(self.ctx.h, ptr(dst), ptr(src), dst.numel() * 4))

    def collect(self, t0: int, snapshots: int):
        """Run the network from step t0; returns device tensors x [S, 12], y [S, 2], the cells' largest
        member displacement m [S] (for the margin), and the next step."""
        import torch

        lib, h, fil = self.ctx.lib, self.ctx.h, self.net.fil
        cfg, n = self.cfg, self.n
        H, W = cfg.horizon, cfg.window
        links = self.net.xl.activeLinkCount if self.net.xl is not None else None
        xs, ys, ms = [], [], []
        t = t0
        counts = (C.c_int * 4)(n, n, 0, 0)
        for s_i in range(snapshots):
            # a rebuild-grid cell size between the smallest and the largest the learned mode uses, and a
            # feature look-back between 1 step and the horizon (at run time the features cover the
            # steps since the last rebuild, which the guard may cut short)
            u = (s_i + 0.5) / snapshots
            gp, gd, cell, nc = self._grid(self.cell_lo * (self.cell_hi / self.cell_lo) ** u)
            Hf = 1 + (s_i * 7) % H
            # t0 - H: positions and the member lists the features are taken over
            self.net.publish()
            self._copy(self.pPrev, fil.coord)
            self._copy(self.uPrev, fil.uVec)
            self.ctx.check(lib.sbx_body_cell(h, ptr(self.pPrev), gp, gd, counts, ptr(self.cell), n, self.periodic))
            self.ctx.check(lib.sbx_grid_build(h, ptr(self.cell), counts, gd, ptr(self.off), ptr(self.cont), n))
            for _ in range(Hf):
                self._step(t)
                t += 1
            self.ctx.check(lib.sbx_skin_features(h, ptr(self.off), ptr(self.cont), nc, n, ptr(fil.coord),
                                                 ptr(self.pPrev), ptr(fil.uVec), ptr(self.uPrev), ptr(links),
                                                 float(Hf), float(cell), self.periodic, self.boxLen, ptr(self.feat)))
            # labels: each member's largest displacement within H steps -> the cell's rms of it (the
            # displacement scale) and max (for the margin); then the first step a member is 3 scales out
            self._copy(self.x0, fil.coord)
            self.ctx.check(lib.sbx_fill_i32(h, ptr(self.maxDisp), 0, n))
            for k in range(1, W + 1):
                self._step(t)
                t += 1
                if k <= H:
                    self.ctx.check(lib.sbx_skin_track(h, ptr(fil.coord), ptr(self.x0), n, self.periodic,
                                                      self.boxLen, ptr(self.maxDisp), None, None, k))
                    if k == H:
                        self.ctx.check(lib.sbx_skin_cell_reduce(h, ptr(self.off), ptr(self.cont), nc,
                                                                ptr(self.maxDisp), None, 2, 0.0, ptr(self.sStar)))
                        self.ctx.check(lib.sbx_skin_cell_reduce(h, ptr(self.off), ptr(self.cont), nc,
                                                                ptr(self.maxDisp), None, 0, 0.0, ptr(self.mStar)))
                        self.ctx.check(lib.sbx_skin_cell_scatter(h, ptr(self.off), ptr(self.cont), nc,
                                                                 ptr(self.sStar), ptr(self.thr)))
                        self.ctx.sync()
                        self.thr *= 3.0
                        self.breach.copy_(torch.where(self.maxDisp > self.thr, H, -1).to(torch.int32))
                        torch.cuda.synchronize()
                else:
                    self.ctx.check(lib.sbx_skin_track(h, ptr(fil.coord), ptr(self.x0), n, self.periodic,
                                                      self.boxLen, None, ptr(self.thr), ptr(self.breach), k))
            self.ctx.check(lib.sbx_skin_cell_reduce(h, ptr(self.off), ptr(self.cont), nc, None, ptr(self.breach), 1,
                                                    float(W + 1), ptr(self.hStar)))
            self.ctx.sync()
            f = self.feat[: nc * N_FEATURES].view(nc, N_FEATURES)
            sS, hS, mS = self.sStar[:nc], self.hStar[:nc], self.mStar[:nc]
            occupied = (f[:, 11] > 0) & (sS > 0)
            xs.append(f[occupied].clone())
            ys.append(torch.stack([torch.log(sS[occupied]), torch.log(hS[occupied])], 1))
            ms.append(mS[occupied].clone())
        torch.cuda.synchronize()
        self.cells_per_snapshot = sum(int(m.shape[0]) for m in ms) / max(1, len(ms))
        return torch.cat(xs).contiguous(), torch.cat(ys).contiguous(), torch.cat(ms).contiguous(), t


def calibrate_margin(ctx: Context, model: SkinModel, x, m, cells_per_snapshot: float, coverage=None) -> float:
    """Skin margin from the training samples: the filament skin's guard rebuilds as soon as ANY body
    outruns its skin, so margin * exp(out0) must cover each cell's LARGEST member displacement m, over all
    cells of a rebuild: the (1 - 1/cells) quantile of m / exp(out0) (conformal calibration).  `coverage`
    replaces that quantile (the reach skin keeps the few segments that outrun their skins on a list)."""
    pred = ctx.to_host(model.forward(x))[:, 0]
    ratio = ctx.to_host(m) / np.exp(pred)
    q = coverage if coverage is not None else 1.0 - 1.0 / max(2.0, cells_per_snapshot)
    return float(max(1.0, np.quantile(ratio, q)))


def fit(ctx: Context, net, cfg: SkinConfig, t0: int = 0, snapshots: int = 16, iters: int = 400, seed: int = 0,
        log_every: int = 0, cell_fn=None, coverage=None):
    """Collect a trajectory of `net` (per-step rebuild) and train a skin model on it; sets
    cfg.margin by calibration; returns (model, loss history, samples, next step)."""
    import torch

    sampler = TrajectorySampler(ctx, net, cfg, cell_fn)
    x, y, m, t = sampler.collect(t0, snapshots)
    xh = ctx.to_host(x)
    mean = xh.mean(0)
    std = np.maximum(xh.std(0), 1e-3)
    model = SkinModel.initial(ctx, mean.astype(np.float32), std.astype(np.float32), seed)
    torch.cuda.synchronize()
    hist = train(ctx, model, x, y, iters=iters, seed=seed, log_every=log_every)
    cfg.margin = calibrate_margin(ctx, model, x, m, sampler.cells_per_snapshot, coverage)
    return model, hist, (x, y), t


def reach_cell_fn(sc, cfg: ReachSkinConfig):
    """The reach skin's coarse cell for a largest skin (sbx_network.hip k_rs_geom): the reach of the longest
    segment (0.5 L + myoColTol) + the skin + 2 slack, at least the floor (the binding grid's cell by default)."""
    r0 = float(0.5 * sc.fs.segLength.max()) + float(sc.mot.kinParams[7])
    floor = cfg.cell_floor if cfg.cell_floor > 0 else float(sc.p["gridParams"][3])
    return lambda skin: max(floor, r0 + skin + 2.0e-3)


def fit_reach(ctx: Context, net, cfg: ReachSkinConfig, t0: int = 0, snapshots: int = 12, iters: int = 400,
              seed: int = 0, horizon: int = 0, log_every: int = 0):
    """The skin model of the reach skin, trained like the filament skin's on a per-step-rebuild trajectory of
    the actomyosin network `net` (per-segment motion in the reach grid's coarse cells).  The model sizes skins
    to last `horizon` steps (or cfg.horizon, or 50); the run's schedule (cfg.horizon, the outside-skin
    trigger) is the caller's.  Sets cfg.margin; returns (model, loss history, samples, next step)."""
    H = horizon or cfg.horizon or 50
    scfg = SkinConfig(skin_default=cfg.skin_default, skin_min=cfg.skin_min, skin_max=cfg.skin_max, horizon=H,
                      window=4 * H)
    model, hist, samples, t = fit(ctx, net, scfg, t0=t0, snapshots=snapshots, iters=iters, seed=seed,
                                  log_every=log_every, cell_fn=reach_cell_fn(net.scene, cfg), coverage=cfg.coverage)
    cfg.margin = scfg.margin
    return model, hist, samples, t
