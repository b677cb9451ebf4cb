"""Batched network step: independent replicas of uneven size advanced together (sbx_batch_* of the C ABI).

No counterpart in the reference, whose harnesses step one network at a time.  The replicas are
concatenated in HBM (rods, links and grid cells of replica r start at its rod / link / cell base), so
every stage of the step is one launch over all of them, and the whole step can be one captured graph.
Each replica is binned in its own grid and draws under its own seed and own indices: its state after
any number of steps is bit-identical to stepping it alone with `Network`.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .context import Context
from .network import Network
from .stores import DeviceCrosslinks, DeviceFilaments


class _BatchNet(Network):
    """The network handle of the concatenated replicas (owned by the batch)."""

    def __init__(self, ctx: Context, h: C.c_void_p, params: _lib.SbxNetParams, n: int):
        self.ctx, self.h, self._p, self.n = ctx, h, params, n
        self.fil = self.xl = self.mot = None

    def close(self) -> None:  # the batch destroys it
        self.h = None


class Batch:
    def __init__(self, ctx: Context, scenes, keep_forces: bool = True):
        self.ctx = ctx
        self.scenes = list(scenes)
        R = len(self.scenes)
        self.fil = [DeviceFilaments(ctx, sc.fs, keep_forces=keep_forces) for sc in self.scenes]
        self.xl = [DeviceCrosslinks(ctx, sc.xl) if sc.xl is not None else None for sc in self.scenes]
        self._fa = (_lib.SbxFilaments * R)(*[f.struct() for f in self.fil])
        self._xa = (_lib.SbxCrosslinks * R)(*[x.struct() if x is not None else _lib.SbxCrosslinks() for x in self.xl])
        self._pa = (_lib.SbxNetParams * R)(*[sc.net_params() for sc in self.scenes])
        h = C.c_void_p()
        ctx.check(ctx.lib.sbx_batch_create(ctx.h, R, self._fa, self._xa, self._pa, C.byref(h)))
        self.h = h
        ctx.adopt(self)
        nh = C.c_void_p()
        ctx.check(ctx.lib.sbx_batch_net(h, C.byref(nh)))
        pc = self.scenes[0].net_params()
        pc.gridDims[3] = sum(int(sc.p["gridDims"][3]) for sc in self.scenes)
        self.n = sum(sc.fs.n for sc in self.scenes)
        self.net = _BatchNet(ctx, nh, pc, self.n)
        self.layout = []
        for r in range(R):
            rb, lb, cb = C.c_int(), C.c_int(), C.c_int()
            ctx.check(ctx.lib.sbx_batch_layout(h, r, C.byref(rb), C.byref(lb), C.byref(cb)))
            self.layout.append((rb.value, lb.value, cb.value))

    @property
    def rods(self) -> int:
        return self.n

    def step(self, step0: int, n: int = 1) -> None:
        self.ctx.check(self.ctx.lib.sbx_batch_step(self.h, step0, n))

    def publish(self) -> None:
        self.ctx.check(self.ctx.lib.sbx_batch_publish(self.h))

    def pull(self, r: int, fs=None, xl=None):
        """Replica r's state copied back into host stores (all replicas published first)."""
        self.publish()
        self.ctx.sync()
        if fs is not None:
            self.fil[r].pull(fs, ("coord", "uVec", "yVec", "zVec", "end1", "end2")
                             + (("forceSum", "torqueSum") if self.fil[r].forceSum is not None else ()))
        if xl is not None and self.xl[r] is not None:
            self.xl[r].pull(xl)
        return fs, xl

    def candidates(self, r: int, rows: tuple[int, int] | None = None):
        """Replica r's (candPartner, candCount, gridCellOffsets, gridCellContents, bodyCell) in its own
        rod and cell numbering (the batch's arrays, rebased); rows=(lo, hi) limits candPartner to the
        slices of the replica's bodies lo..hi-1."""
        sc = self.scenes[r]
        rb, _, cb = self.layout[r]
        n, mc, cells = sc.fs.n, int(sc.p["maxCand"]), int(sc.p["gridDims"][3])
        lo, hi = rows if rows is not None else (0, n)
        cand, cnt, off, cont, cell = self.net.candidates(self.n, mc, rows=(rb + lo, rb + hi))
        cand = cand.reshape(hi - lo, mc)
        cnt = cnt[rb:rb + n].copy()
        used = np.arange(mc)[None, :] < np.minimum(cnt[lo:hi], mc)[:, None]
        cand[used] -= rb
        return (cand.ravel(), cnt, off[cb:cb + cells + 1] - rb, cont[rb:rb + n] - rb, cell[rb:rb + n] - cb)

    def close(self) -> None:
        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            self.ctx.lib.sbx_batch_destroy(self.h)
        self.h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


def tune_cell_size(ctx: Context, scenes, factors=(1.0, 1.5, 2.0, 3.0), trial_steps: int = 10, step0: int = 0,
                   graph: bool = True) -> dict:
    """Cell-size autotuner: the replicas batched on grids with cells `factor` times their own (each
    replica's cell grown by the same factor, scenes.retile), trial_steps steps timed after two warm-up
    steps on copies of the scenes, the fastest factor kept.  The step's pose, forces and link events
    do not depend on the broad-phase grid, and every factor >= 1 keeps the stencil complete, so the
    candidate pair sets are the same for every factor; only cell ids, offsets and the order within
    each candidate slice follow the chosen grid.  Returns {"factor", "ms_per_step": {factor: ms}}."""
    import time

    from .scenes import retile

    times = {}
    for f in factors:
        b = Batch(ctx, [retile(sc, f) for sc in scenes], keep_forces=False)
        if graph:
            b.net.use_graph(True)
        b.step(step0, 2)
        ctx.sync()
        t0 = time.perf_counter()
        b.step(step0 + 2, trial_steps)
        ctx.sync()
        times[f] = 1e3 * (time.perf_counter() - t0) / trial_steps
        b.close()
    best = min(times, key=times.get)
    return {"factor": best, "ms_per_step": times}
