"""The whole network mechanics step on the GPU (sbx_net_* of the C ABI).

Stands in for a harness's `buildPlan` + per-step `plan.execute()` (DenseContractileHarness.java:
442-606): build once from device stores, then `step(step0, n)` advances n steps.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .context import Context, ptr
from .stores import DeviceCrosslinks, DeviceFilaments, DeviceMotors


class Network:
    def __init__(self, ctx: Context, scene, keep_forces: bool = True):
        self.ctx = ctx
        self.scene = scene
        self.fil = DeviceFilaments(ctx, scene.fs, keep_forces=keep_forces)
        self.xl = DeviceCrosslinks(ctx, scene.xl) if scene.xl is not None else None
        self._fs = self.fil.struct()
        self._xs = self.xl.struct() if self.xl is not None else None
        self._p = scene.net_params()
        h = C.c_void_p()
        ctx.check(ctx.lib.sbx_net_create(ctx.h, C.byref(self._fs), C.byref(self._xs) if self._xs else None,
                                         C.byref(self._p), C.byref(h)))
        self.h = h
        ctx.adopt(self)
        self.mot = None
        if getattr(scene, "mot", None) is not None:
            self.mot = DeviceMotors(ctx, scene.mot, scene.dim, scene.mini, int(scene.p["maxCand"]),
                                    seeds=(scene.p["seedMotor"], scene.p["seedBackbone"]))
            self._ms = self.mot.struct()
            self._mp = self.mot.params()
            ctx.check(ctx.lib.sbx_net_attach_motors(self.h, C.byref(self._ms), C.byref(self._mp)))
        self.turn = None
        if getattr(scene, "turn", None) is not None:
            from .turnover import DeviceTurnover

            self.turn = DeviceTurnover(ctx, scene.turn, diff=(scene.fs.bTransDiff, scene.fs.bRotDiff))
            self._ts = self.turn.struct()
            self._tp = scene.turn.params()
            ctx.check(ctx.lib.sbx_net_set_turnover(self.h, C.byref(self._ts), C.byref(self._tp)))

    def set_skin(self, mode: int, cfg=None, model=None) -> None:
        """Broad-phase mode: SBX_SKIN_REBUILD (the reference's per-step rebuild), SBX_SKIN_FIXED or
        SBX_SKIN_LEARNED (with a softbox_amd.skin.SkinModel)."""
        from .skin import SkinConfig

        cfg = cfg or SkinConfig()
        self._skin_p = cfg.params()
        w = ptr(model.w) if model is not None else None
        self.ctx.check(self.ctx.lib.sbx_net_set_skin(self.h, mode, C.byref(self._skin_p), w))

    def skin_stats(self) -> dict:
        out = (C.c_longlong * 8)()
        k = self.ctx.check(self.ctx.lib.sbx_net_skin_stats(self.h, out, 8))
        names = ("rebuilds", "outside_skin", "brute_steps", "block_bodies", "list_entries", "overflow_bodies",
                 "steps_since_rebuild", "skin_steps")
        return {names[i]: int(out[i]) for i in range(min(k, len(names)))}

    def set_reach_skin(self, mode: int, cfg=None, model=None) -> None:
        """Binding-grid search of the myosin stage: SBX_SKIN_REBUILD (the reference's per-step grid),
        SBX_SKIN_FIXED or SBX_SKIN_LEARNED (with a softbox_amd.skin.SkinModel) -- the reach skin."""
        from .skin import ReachSkinConfig

        cfg = cfg or ReachSkinConfig()
        self._reach_p = cfg.params()
        w = ptr(model.w) if model is not None else None
        self.ctx.check(self.ctx.lib.sbx_net_set_reach_skin(self.h, mode, C.byref(self._reach_p), w))
        self._reach_mode = (mode, cfg, model)

    def reach_mode(self) -> tuple:
        """(mode, config, model) of the last set_reach_skin (mode 0: the per-step binding grid)."""
        return getattr(self, "_reach_mode", (0, None, None))

    def reach_skin_stats(self) -> dict:
        out = (C.c_longlong * 9)()
        k = self.ctx.check(self.ctx.lib.sbx_net_reach_skin_stats(self.h, out, 9))
        names = ("rebuilds", "outside_skin", "brute_steps", "selected_heads", "cell_nm", "rebuilds_stale",
                 "rebuilds_horizon", "rebuilds_outside", "steps")
        return {names[i]: int(out[i]) for i in range(min(k, len(names)))}

    def form_stats(self) -> dict:
        """Formation counters: passes, candidates and links formed by the last pass, links formed in all."""
        out = (C.c_longlong * 4)()
        k = self.ctx.check(self.ctx.lib.sbx_net_form_stats(self.h, out, 4))
        names = ("passes", "candidates", "formed_last", "formed_total")
        return {names[i]: int(out[i]) for i in range(k)}

    def skin_state(self) -> dict:
        ptrs = [C.c_void_p() for _ in range(5)]
        nc = C.c_int()
        grid = (C.c_float * 8)()
        self.ctx.check(self.ctx.lib.sbx_net_skin_state(self.h, *[C.byref(p) for p in ptrs], C.byref(nc), grid))
        return {"refCoord": ptrs[0].value, "skin": ptrs[1].value, "coarseCell": ptrs[2].value,
                "coarseOffsets": ptrs[3].value, "coarseContents": ptrs[4].value, "nCells": nc.value,
                "grid": list(grid)}

    def step(self, step0: int, n: int = 1) -> None:
        self.ctx.check(self.ctx.lib.sbx_net_step(self.h, step0, n))

    def publish(self) -> None:
        self.ctx.check(self.ctx.lib.sbx_net_publish(self.h))

    def topology_changed(self) -> None:
        """After the caller edited linkFilA/linkFilB or linkState of the device store between steps
        (a formation pass): the crosslink CSR is rebuilt and the active links are recounted."""
        self.ctx.check(self.ctx.lib.sbx_net_topology_changed(self.h))

    def use_graph(self, on: bool | int = True) -> None:
        """True (or 2): replay captured steps (one per pose buffer for the per-step rebuild on the side
        streams, two at a time otherwise); False: eager."""
        self.ctx.check(self.ctx.lib.sbx_net_use_graph(self.h, int(on)))

    def set_overlap(self, on: bool = True) -> None:
        """Broad phase on a second stream, concurrent with the mechanics kernels (default on)."""
        self.ctx.check(self.ctx.lib.sbx_net_set_overlap(self.h, 1 if on else 0))

    def set_shape(self, bp_group: int = 4, scan_items: int = 8, rod_threads: int = 256, bp_staged: int = 0) -> None:
        """Launch shape of the step (results do not depend on it).  bp_staged: 0 global walk, 1 LDS-staged
        neighbourhoods, 2 stencil rows split over the body's lanes (1 and 2 take 4 or 8 lanes)."""
        self.ctx.check(self.ctx.lib.sbx_net_set_shape(self.h, bp_group, bp_staged, scan_items, rod_threads))

    def tune(self, step0: int, trial_steps: int = 20) -> dict:
        """Time every launch shape (with and without the concurrent streams) over real steps step0,
        step0+1, ... and keep the fastest; returns {"steps": steps taken, "shape": (bp_group,
        bp_staged, scan_items, rod_threads, overlap), "ms_per_step": best}."""
        adv = C.c_int()
        shape = (C.c_int * 5)()
        ms = C.c_float()
        self.ctx.check(self.ctx.lib.sbx_net_tune(self.h, step0, trial_steps, C.byref(adv), shape, C.byref(ms)))
        out = {"steps": adv.value, "shape": tuple(shape), "ms_per_step": ms.value}
        if self.mot is not None:
            out["motor_group"] = self.motor_group()
        return out

    def capture_log(self) -> np.ndarray:
        """The (op, stream, event) triples of the last captured graph (op 0 record, 1 wait; stream 0 the
        capturing stream, 1 broad phase, 2 sort), as sbx_capture_check takes them."""
        n = self.ctx.check(self.ctx.lib.sbx_net_capture_log(self.h, None, 0))
        out = np.zeros(max(1, n), np.int32)
        self.ctx.check(self.ctx.lib.sbx_net_capture_log(self.h, out.ctypes.data_as(C.c_void_p), n))
        return out[:n].reshape(-1, 3)

    def set_motor_group(self, lanes: int) -> None:
        """Lanes per motor of the fused reach search + binding (4, 8, 16, 32 or 64; the per-step grid walk takes at
        least 16; results do not depend on it)."""
        self.ctx.check(self.ctx.lib.sbx_net_set_motor_group(self.h, lanes))

    def motor_group(self) -> int:
        return self.ctx.check(self.ctx.lib.sbx_net_motor_group(self.h))

    def set_profiling(self, on: bool) -> None:
        self.ctx.check(self.ctx.lib.sbx_net_set_profiling(self.h, 1 if on else 0))

    def kernel_times(self) -> dict:
        cap = 32
        names = (C.c_char_p * cap)()
        ms = (C.c_double * cap)()
        cnt = (C.c_longlong * cap)()
        k = self.ctx.check(self.ctx.lib.sbx_net_kernel_times(self.h, names, ms, cnt, cap))
        return {names[i].decode(): (ms[i], cnt[i]) for i in range(k)}

    def time_broad_phase(self, reps: int = 20) -> float:
        """Mean microseconds per launch of the latest step's broad phase, re-run `reps` times back to
        back on its grid (the slices are rewritten with the same values)."""
        us = C.c_float()
        self.ctx.check(self.ctx.lib.sbx_net_time_broad_phase(self.h, reps, C.byref(us)))
        return us.value

    def candidates(self, n: int, max_cand: int, rows: tuple[int, int] | None = None):
        """Host copies of (candPartner[n*maxCand], candCount[n], offsets, contents, bodyCell); with
        rows=(lo, hi) only the candidate slices of bodies lo..hi-1 are copied (candPartner[(hi-lo)*maxCand])."""
        ptrs = [C.c_void_p() for _ in range(5)]
        self.ctx.check(self.ctx.lib.sbx_net_candidates(self.h, *[C.byref(p) for p in ptrs]))
        self.ctx.sync()
        cells = int(self._p.gridDims[3])
        lo, hi = rows if rows is not None else (0, n)
        sizes = [(hi - lo) * max_cand, n, cells + 1, n, n]
        skip = [lo * max_cand * 4, 0, 0, 0, 0]
        out = []
        for p, sz, sk in zip(ptrs, sizes, skip):
            if not p.value:
                out.append(None)
                continue
            a = np.empty(sz, np.int32)
            self.ctx.check(self.ctx.lib.sbx_pull(self.ctx.h, a.ctypes.data_as(C.c_void_p), C.c_void_p(p.value + sk),
                                                 a.nbytes))
            out.append(a)
        return out

    SHAPE = ("segLength", "bTransGam", "bRotGam", "end1NbrSlot", "end1NbrSide", "end2NbrSlot", "end2NbrSide")

    def pull(self, fs=None, xl=None, shape=False):
        """Synchronise and copy state back into host stores (pose published first); with shape, also the
        segment lengths, drag and chain links that turnover rewrites on the device."""
        self.publish()
        self.ctx.sync()
        if fs is not None:
            self.fil.pull(fs, ("coord", "uVec", "yVec", "zVec", "end1", "end2")
                          + (("forceSum", "torqueSum") if self.fil.forceSum is not None else ())
                          + (self.SHAPE if shape else ()))
        if xl is not None and self.xl is not None:
            self.xl.pull(xl)
        return fs, xl

    def pull_turnover(self, ts, fs=None):
        """The turnover state (and, with fs, the diffusion coefficients recomputeDrag writes) back into host
        stores."""
        self.turn.pull(ts)
        if fs is not None:
            fs.bTransDiff[:] = self.ctx.to_host(self.turn.t["bTransDiff"])
            fs.bRotDiff[:] = self.ctx.to_host(self.turn.t["bRotDiff"])
        return ts

    def pull_motors(self, mot=None, mini=None):
        """Synchronise and copy the motor and backbone state back into host stores."""
        self.ctx.sync()
        self.mot.pull(mot, mini)

    def close(self) -> None:
        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            self.ctx.lib.sbx_net_destroy(self.h)
        self.h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass
