"""Benchmark: rod-steps/s of the crosslinked-network mechanics step on one MI355X.

Workload (the one-GPU configuration the metric is quoted on): crosslinked actin network,
20k filaments x 10 rigid rods (0.1 um), 20 um periodic cube, 50k crosslinkers (rest 35 nm, Bell
k0 = 0.1/s, dx = 0.5 nm), water, dt = 1e-5 s, FP32 state, per-step grid rebuild + 27-cell broad phase
+ crosslink unbind/count/force/torsion + CSR gathers + chain + Brownian + integrate + derive.
Synthetic, seeded scene (softbox_amd.scenes.network); the whole state is resident in HBM before
the timed region starts.

Prints ONE JSON line (rank 0).  `roofline` is for the dominant kernel: algorithmic bytes per launch
divided by its mean duration measured with HIP events on the library's stream during the timed
region.  `cpu_baseline` times the oracle (single-threaded C restatement of the reference's
sequential CPU runner) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "rod-steps/sec + HBM GB/s % peak at 1 GPU; learned-skin vs per-step rebuild"


def kernel_bytes(sc, name: str) -> float | None:
    """Algorithmic (compulsory) HBM bytes of one launch of a network kernel on this workload:
    every input read once and every output written once, at its natural width."""
    n = sc.fs.n
    L = sc.xl.nLinks if sc.xl is not None else 0
    cells = int(sc.p["gridDims"][3])
    if name == "k_rod":
        # read pose 36 + segLength 4 + gamma 24 + Brownian scales 8 + topology 16 + CSR offsets 16,
        # write pose 36 + zVec/end1/end2 36 + active-link count 4 + next-step cell 4 -> 184 B/rod;
        # per link end: index 4 + state 4 + payload 24
        return 184.0 * n + 32.0 * 2 * L
    if name == "k_net_link":
        # ids 8 + locs 8 + state 4 + orient 4 + torque ring 20+4 + endpoints A/B (end1, u, coord, gamma) 2x40
        # + both segments' CSR bounds 2x16; write payload 48 + ring slot 4 + place 4
        return 224.0 * L
    if name == "k_net_unbind":
        return (8 + 8 + 4 + 40 + 4 + 2 * 24 + 40 + 4) * float(L)
    if name == "k_body_cell":
        return 16.0 * n
    if name == "grid_sort":
        # per body: count (key 4, slot 4, counts/tile atomics 8), scatter (key, offset, slot 12, tmp 4,
        # counts 4), place (key 4, offsets 8, segment 4, list 4); per cell: counts 4 + offsets 4
        return 56.0 * n + 8.0 * cells
    if name == "k_broad_phase":
        # per body: contents 4 + cell 4 + centre 12 + radius 4 + count 4; per cell: offsets 4
        return 28.0 * n + 4.0 * cells
    return None


def run_cpu_baseline(sc, budget_s: float = 15.0) -> dict:
    from oracle import ffi

    orc = ffi.OracleNet(sc.copy())
    orc.step(0, 1)  # warm (page-in)
    steps, t0 = 0, time.perf_counter()
    while True:
        orc.step(1 + steps, 1)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or steps >= 200:
            break
    return {"value": sc.fs.n * steps / el, "unit": "rod-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle C restatement (single thread, same seeded scene, {sc.fs.n} rods + "
                      f"{sc.xl.nLinks if sc.xl is not None else 0} links, full step incl. grid) "
                      f"for {steps} steps = {el:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--graph", type=int, default=0, help="1: replay the step as a captured hipGraph")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--n-fil", type=int, default=20000)
    ap.add_argument("--n-links", type=int, default=50000)
    ap.add_argument("--box", type=float, default=20.0)
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="before the warm-up, one device-to-device copy of this many MiB (a plain copy kernel "
                         "of known size that calibrates rocprof FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU design (spatial slabs + halo exchange) is documentation only", file=sys.stderr)
        return 2

    import torch

    from softbox_amd import scenes
    from softbox_amd.context import Context
    from softbox_amd.network import Network

    t_build = time.perf_counter()
    sc = scenes.network(n_fil=args.n_fil, segs=10, box=args.box, n_links=args.n_links, seed=0)
    n = sc.fs.n
    print(f"[bench] scene: {n} rods, {sc.xl.nLinks} links, grid {sc.p['gridDims'][:3]} "
          f"({time.perf_counter() - t_build:.1f} s)", file=sys.stderr, flush=True)
    ctx = Context(0)
    if args.calib_copy_mb > 0:
        nb = args.calib_copy_mb << 20
        src = torch.ones(nb // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        ctx.check(ctx.lib.sbx_copy(ctx.h, dst.data_ptr(), src.data_ptr(), nb))
        ctx.sync()
        del src, dst
        print(f"[bench] calibration copy of {args.calib_copy_mb} MiB done", file=sys.stderr, flush=True)
    net = Network(ctx, sc, keep_forces=False)
    if args.graph:
        net.use_graph(True)
    net.step(0, args.warmup)
    ctx.sync()
    torch.cuda.synchronize()
    print(f"[bench] warmup {args.warmup} steps done", file=sys.stderr, flush=True)

    # timed region: whole steps, per-kernel HIP events on the library stream (eager launches)
    prof = not args.graph
    if prof:
        net.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    net.step(args.warmup, args.steps)
    ctx.sync()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    times = net.kernel_times() if prof else {}
    if prof:
        net.set_profiling(False)
    ms_per_step = 1e3 * el / args.steps
    value = n * args.steps / el

    fs_end = sc.fs.copy()
    xl_end = sc.xl.copy()
    net.pull(fs_end, xl_end)
    assert np.isfinite(fs_end.coord).all(), "non-finite pose after the timed run"
    alive = int((xl_end.linkState[: xl_end.nLinks] >= 0).sum())

    roof = None
    per_kernel = {}
    if times:
        for k, (ms, cnt) in times.items():
            if cnt:
                per_kernel[k] = {"avg_us": 1e3 * ms / cnt, "launches": int(cnt)}
        dom = max((k for k in per_kernel if kernel_bytes(sc, k)), key=lambda k: per_kernel[k]["avg_us"])
        b = kernel_bytes(sc, dom)
        achieved = b / (per_kernel[dom]["avg_us"] * 1e-6) / 1e9
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_launch": b}
        for k, v in per_kernel.items():
            kb = kernel_bytes(sc, k)
            if kb:
                v["GBs"] = round(kb / (v["avg_us"] * 1e-6) / 1e9, 1)
    cpu = None if args.no_cpu else run_cpu_baseline(sc)

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "rod-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 state, f64 force laws", "data": "synthetic (seeded scene)",
        "config": {"workload": "crosslinked actin network: 20k filaments x 10 rods (0.1 um), 50k crosslinkers, "
                               "20 um periodic cube, dt 1e-5 s, per-step grid rebuild",
                   "rods": n, "links": sc.xl.nLinks, "links_alive_end": alive, "grid": list(sc.p["gridDims"][:3]),
                   "graph": bool(args.graph), "broad_phase": "per-step rebuild"},
        "roofline": roof, "cpu_baseline": cpu, "kernels": per_kernel,
    }
    print(json.dumps(out))
    net.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
