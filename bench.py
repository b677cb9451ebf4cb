"""Benchmark: rod-steps/s of the crosslinked-network mechanics step on one MI355X.

Workload (the one-GPU configuration the metric is quoted on): crosslinked actin network,
20k filaments x 10 rigid rods (0.1 um), 20 um periodic cube, 50k crosslinkers (rest 35 nm, Bell
k0 = 0.1/s, dx = 0.5 nm), viscosity 0.1 Pa s (the reference's Constants.aeta; the workload states
none), dt = 1e-5 s, FP32 state, broad phase (per-step grid rebuild + 27-cell stencil, or the learned
skin) + crosslink unbind/count/force/torsion + CSR gathers + chain + Brownian + integrate + derive.
Synthetic, seeded scene (softbox_amd.scenes.network); the whole state is resident in HBM before
the timed region starts.  The crosslinks stay alive through the timed steps (links_alive_start /
links_alive_end report how many).

Prints ONE JSON line (rank 0).  By default it runs the workload twice from the same initial state:
the reference's per-step grid rebuild, and the learned-skin broad phase (its skin model trained
first, outside the timed region, on a per-step-rebuild trajectory of the same scene).  The learned
skin keeps the per-step rebuild as its fallback and judge: `value` is the faster of the two (named
in config.broad_phase), `learned_vs_rebuild` is their ratio, and the two runs' candidate slices are
checked identical.  `roofline` is for the dominant kernel of the headline run: algorithmic bytes
per launch divided by its mean duration measured with HIP events on the library's stream.
`cpu_baseline` times the oracle (the C restatement of the reference's sequential CPU runner) on the
same steps from the same state the GPU's timed region starts from: the OpenMP build on all host
threads (`cpu_baseline`) and the single-thread build (`cpu_baseline_single_thread`).

Then `learned_skin_at_scale` (--at-scale-steps, default 300): the same recipe scaled to 2M rods with 8k
minifilaments (128k myosin heads), where the learned skin covers the binding grid of the myosin stage (the
reach skin): per-step rebuild and learned reach skin on the same steps, their ratio, and a check that the
reach lists, binding events and poses of the two runs are identical.
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


def kernel_bytes(sc, name: str, list_entries: int | None = None, cand_written: int | None = None) -> float | None:
    """Algorithmic (compulsory) HBM bytes of one launch of a network kernel on this workload:
    every input read once and every output written once, at its natural width.  cand_written: candidate
    indices the broad phase writes (the used part of every slice), counted when known."""
    n = sc.fs.n
    L = sc.xl.nLinks if sc.xl is not None else 0
    cells = int(sc.p["gridDims"][3])
    if name == "k_rod":
        # read pose 36 + segLength 4 + gamma 24 + Brownian scales 8 + topology 16 + CSR offsets 16 +
        # bounding radius 4, write pose 36 + zVec/end1/end2 36 + active-link count 4 + next-step cell 4
        # + histogram count/slot 8 + packed pose 16 + link pack 48 -> 260 B/rod;
        # per link end: index 4 + state 4 + payload 24
        return 260.0 * n + 32.0 * 2 * L
    if name == "k_net_link":
        # ids 8 + locs 8 + state 4 + orient 4 + torque ring 20+4 + endpoints A/B (end1, u, coord, gamma) 2x40
        # + both segments' active-link counts 2x4; write payload 48 + ring slot 4 + place 4
        return 192.0 * L
    if name == "k_net_unbind":
        return (8 + 8 + 4 + 40 + 4 + 2 * 24 + 40 + 4) * float(L)
    if name == "k_body_cell":
        return 16.0 * n
    if name == "grid_sort":
        # (the histogram is taken by k_rod) per body: scatter (key 4, offset 4, slot 4, count reset 4,
        # position 4), place (key 4, offsets 8, cell segment 4, contents 4, packed pose 16 + 16,
        # sorted cell 4); per cell: counts read twice (tile sums, scan) 8 + offsets 4
        return 76.0 * n + 12.0 * cells
    if name == "k_broad_phase":
        # per body: contents 4 + sorted pose 16 + sorted cell 4 + count 4; per cell: offsets 4; per candidate
        # written: its index 4
        return 28.0 * n + 4.0 * cells + 4.0 * (cand_written or 0)
    if name == "k_skin_pairs" and list_entries:
        # per body: order 4 + centre 12 + radius 4 + cell 4 + list length 4 + count 4; per list entry:
        # index 4 + centre 12 + radius 4 + cell 4
        return 32.0 * n + 24.0 * list_entries
    return None


def profiled_column(fname: str, name: str, col: int):
    """Column `col` of kernel `name`'s row in a committed rocprofv3 summary table under profiles/, as a
    float (a trailing ' %' dropped), or None when the file or the row is missing."""
    path = os.path.join(ROOT, "profiles", fname)
    if not os.path.exists(path):
        return None
    for line in open(path):
        cols = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cols) > col and cols[0] == name:
            try:
                return float(cols[col].rstrip("%").strip())
            except ValueError:
                return None
    return None


def kernel_traffic(name: str):
    """HBM bytes per launch of `name` from the rocprofv3 FETCH_SIZE/WRITE_SIZE summary of the 200k-rod
    workload (profiles/network_step_200k_rods_counters.md)."""
    return profiled_column("network_step_200k_rods_counters.md", name, 5)


def kernel_valu_issue(name: str):
    """Lower bound on the fraction of `name`'s solo duration spent issuing VALU instructions, from the
    rocprofv3 SQ summary of the 200k-rod workload (profiles/network_step_200k_rods_instructions.md)."""
    v = profiled_column("network_step_200k_rods_instructions.md", name, 7)
    return None if v is None else v / 100.0


def run_cpu_baseline(sc, step0: int, budget_s: float = 12.0, openmp: bool = False) -> dict:
    """The oracle stepping the workload from state `sc` (the GPU's state at the start of its timed
    region) over the same step numbers step0, step0 + 1, ... for about budget_s seconds."""
    from oracle import ffi

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)) if openmp else 1
    orc = ffi.OracleNet(sc.copy(), ffi.load_omp() if openmp else None)
    steps, t0 = 0, time.perf_counter()
    while True:
        orc.step(step0 + steps, 1)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or steps >= 400:
            break
    alive = int((orc.xl.linkState[: orc.xl.nLinks] >= 0).sum()) if orc.xl is not None else 0
    form = f"OpenMP build, {threads} threads" if openmp else "single thread"
    return {"value": sc.fs.n * steps / el, "unit": "rod-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle C restatement ({form}), the GPU's state at the start of its timed region "
                      f"({sc.fs.n} rods, {sc.xl.nLinks if sc.xl is not None else 0} link slots), steps "
                      f"{step0}..{step0 + steps - 1}, full step incl. grid: {steps} steps in {el:.1f} s; "
                      f"{alive} links alive at the end"}


def shape_of(args, tuned=None) -> dict:
    """The launch shape a run used: the tuner's (group, staged, items, threads, overlap) or --shape."""
    if tuned is not None:
        g, st, it, th, ov = tuned["shape"]
    else:
        v = [int(x) for x in args.shape.split(",")] + [0]
        g, it, th, st, ov = v[0], v[1], v[2], v[3], int(args.overlap)
    return {"bp_group": g, "bp_staged": st, "scan_items": it, "rod_threads": th, "overlap": ov}


def timed(net, ctx, step0: int, steps: int) -> float:
    """Seconds for `steps` whole steps, bracketed by device synchronisation."""
    import torch

    ctx.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    net.step(step0, steps)
    ctx.sync()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def profile(net, step0: int, steps: int) -> dict:
    """Per-kernel mean durations (HIP events on the library's stream) over `steps` eager steps."""
    net.use_graph(False)
    net.set_profiling(True)
    net.step(step0, steps)
    times = net.kernel_times()
    net.set_profiling(False)
    return {k: {"avg_us": 1e3 * ms / cnt, "launches": int(cnt)} for k, (ms, cnt) in times.items() if cnt}


def slices_digest(net, sc):
    """(candCount, order-sensitive digest of the used part of every candidate slice)."""
    n, mc = sc.fs.n, int(sc.p["maxCand"])
    cand, cnt, _, _, _ = net.candidates(n, mc)
    keep = (np.arange(mc)[None, :] < np.minimum(cnt, mc)[:, None])
    used = np.where(keep, cand.reshape(n, mc), -1).astype(np.int64)
    w = (np.arange(mc, dtype=np.int64) * 2654435761) % 1000003 + 1
    return cnt, int(((used + 2) * w[None, :]).sum() % (1 << 61))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--skin", choices=("compare", "rebuild", "fixed", "learned", "ablation"), default="compare",
                    help="compare: per-step rebuild and learned skin on identical trajectories (network: the filament "
                         "skin; actomyosin: the reach skin of the binding grid); ablation (actomyosin): rebuild, "
                         "fixed and learned reach skins")
    ap.add_argument("--at-scale-steps", type=int, default=300,
                    help="network workload: then the learned-skin comparison at scale (2M rods, 128k heads, the reach "
                         "skin of the binding grid) over this many steps (0: skip)")
    ap.add_argument("--reach-skin", type=float, default=0.04,
                    help="actomyosin: the fixed reach skin, um (and the learned mode's first build)")
    ap.add_argument("--reach-horizon", type=int, default=0,
                    help="actomyosin: steps between scheduled reach-skin rebuilds (0: only on stale state and the "
                         "outside-skin count; the learned model sizes its skins for 50 steps)")
    ap.add_argument("--graph", type=int, default=1, help="1: replay the step as a captured hipGraph")
    ap.add_argument("--overlap", type=int, default=1,
                    help="1: broad phase on a second stream, concurrent with the mechanics kernels")
    ap.add_argument("--profile-steps", type=int, default=40, help="eager steps timed per kernel (0: none)")
    ap.add_argument("--train-snapshots", type=int, default=12)
    ap.add_argument("--train-iters", type=int, default=400)
    ap.add_argument("--horizon", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--n-fil", type=int, default=None, help="filaments (network 20000, actomyosin 50000)")
    ap.add_argument("--n-links", type=int, default=None, help="crosslinker slots (network 50000, actomyosin 100000)")
    ap.add_argument("--box", type=float, default=None, help="box edge, um (network 20, actomyosin 30)")
    ap.add_argument("--workload", choices=("network", "actomyosin", "standard1x"), default="network",
                    help="network: the crosslinked network (the headline); actomyosin: the full actomyosin "
                         "network (50k filaments, 2k minifilaments of 16 heads, 100k crosslinkers, 30 um box); "
                         "standard1x: the reference's 1x contractility standard (400 protein nodes x 24 myosins, "
                         "1000 filaments x 10 segments, 7.071 x 7.071 x 0.5 um slab)")
    ap.add_argument("--eta", type=float, default=None,
                    help="viscosity, Pa s (default: the reference's Constants.aeta = 0.1 for the network, the "
                         "dense actomyosin fixture's 1.0 for the actomyosin network)")
    ap.add_argument("--formation", type=int, default=100,
                    help="crosslinker formation every this many steps (binding radius 40 nm; 0: off)")
    ap.add_argument("--tune-steps", type=int, default=8,
                    help="autotuner: steps timed per launch shape after the warm-up (0: default shape)")
    ap.add_argument("--shape", type=str, default="",
                    help="bp_group,scan_items,rod_threads[,bp_staged]: fixed launch shape instead of the autotuner")
    ap.add_argument("--batch", type=int, default=0,
                    help="N > 0: the batched ensemble instead (N replicas, 2k-40k filaments log-uniform, 5-30 uM)")
    ap.add_argument("--motor-group", type=int, default=0,
                    help="motor workloads: lanes per motor in the reach search (16, 32, 64; 0: the tuner's or 16)")
    ap.add_argument("--turnover", type=int, default=1,
                    help="actomyosin workload: actin turnover at its biochem cadence (1) or none (0)")
    ap.add_argument("--cell-factors", type=str, default="1,1.5,2,3",
                    help="batched ensemble: grid cell multiples the cell-size autotuner times ('' or '1': the "
                         "reference grid)")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="before the warm-up, one device-to-device copy of this many MiB (a plain copy kernel "
                         "of known size that calibrates rocprof FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the step runs and is measured on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU design (spatial slabs + halo exchange) is documentation only", file=sys.stderr)
        return 2

    import torch

    from softbox_amd import _lib, scenes
    from softbox_amd.context import Context
    from softbox_amd.network import Network
    from softbox_amd.skin import SkinConfig, fit

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    acto = args.workload == "actomyosin"
    if args.eta is None:
        args.eta = 1.0 if acto else 0.1
    if args.n_fil is None:
        args.n_fil = 50000 if acto else 20000
    if args.n_links is None:
        args.n_links = 100000 if acto else 50000
    if args.box is None:
        args.box = 30.0 if acto else 20.0
    if args.batch:
        return run_batch(args, log)
    if args.workload == "actomyosin":
        return run_actomyosin(args, log)
    if args.workload == "standard1x":
        return run_standard(args, log)
    t_build = time.perf_counter()
    sc = scenes.network(n_fil=args.n_fil, segs=10, box=args.box, n_links=args.n_links, seed=0, eta=args.eta)
    if args.formation > 0:  # crosslinker formation at its cadence: the workload's 40 nm binding radius
        sc.p.update(scenes.formation_params(args.box, 0.1, sc.fs.n, True, grab=0.04, interval=args.formation))
    n = sc.fs.n
    log(f"scene: {n} rods, {sc.xl.nLinks} links, grid {sc.p['gridDims'][:3]} ({time.perf_counter() - t_build:.1f} s)")
    ctx = Context(0)
    if args.calib_copy_mb > 0:
        nb = args.calib_copy_mb << 20
        src = torch.ones(nb // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        ctx.check(ctx.lib.sbx_copy(ctx.h, dst.data_ptr(), src.data_ptr(), nb))
        ctx.sync()
        del src, dst
        log(f"calibration copy of {args.calib_copy_mb} MiB done")

    cfg = SkinConfig(horizon=args.horizon, window=4 * args.horizon)
    model, train_info = None, None
    if args.skin == "learned" or args.skin == "compare":
        t_tr = time.perf_counter()
        tnet = Network(ctx, sc.copy(), keep_forces=False)
        model, hist, (x, _), _ = fit(ctx, tnet, cfg, snapshots=args.train_snapshots, iters=args.train_iters)
        tnet.close()
        train_info = {"samples": int(x.shape[0]), "loss_start": round(hist[0], 4), "loss_end": round(hist[-1], 4),
                      "margin": round(cfg.margin, 3), "horizon": cfg.horizon,
                      "trajectory_steps": args.train_snapshots * (cfg.horizon + cfg.window),
                      "seconds": round(time.perf_counter() - t_tr, 1)}
        log(f"skin model trained: {train_info}")

    modes = {"compare": ["rebuild", "learned"], "rebuild": ["rebuild"], "fixed": ["fixed"], "learned": ["learned"]}
    runs = {}
    extra = 0  # steps the tuner took in the rebuild run: the other runs warm up as long (same end step)
    for mode in modes[args.skin]:
        net = Network(ctx, sc.copy(), keep_forces=False)
        if mode != "rebuild":
            net.set_skin(_lib.SBX_SKIN_LEARNED if mode == "learned" else _lib.SBX_SKIN_FIXED, cfg,
                         model if mode == "learned" else None)
        net.set_overlap(bool(args.overlap))
        if args.graph:
            net.use_graph(True)
        net.step(0, args.warmup)
        t0 = args.warmup
        if args.shape:
            net.set_shape(*[int(v) for v in args.shape.split(",")])
        elif args.tune_steps and mode == "rebuild":
            tuned = net.tune(t0, args.tune_steps)
            t0 += tuned["steps"]
            extra = tuned["steps"]
            log(f"tuned launch shape {tuned}")
        elif extra:
            net.step(t0, extra)
            t0 += extra
        fs0, xl0 = net.pull(sc.fs.copy(), sc.xl.copy())
        links_start = int((xl0.linkState[: xl0.nLinks] >= 0).sum())
        if mode == modes[args.skin][0]:
            start_state = (sc.copy(), t0)
            start_state[0].fs, start_state[0].xl = fs0, xl0
        formed0 = net.form_stats().get("formed_total", 0) if args.formation > 0 else 0
        el = timed(net, ctx, t0, args.steps)
        r = {"value": n * args.steps / el, "ms_per_step": 1e3 * el / args.steps, "timed_steps": [t0, t0 + args.steps - 1],
             "links_alive_start": links_start}
        if args.formation > 0:
            fst = net.form_stats()
            r["links_formed_timed"] = fst["formed_total"] - formed0
            r["formation"] = fst
        if args.shape or (args.tune_steps and mode == "rebuild"):
            r["shape"] = shape_of(args, None if args.shape else tuned)
        r["digest"] = slices_digest(net, sc)
        r["cand_written"] = int(np.minimum(r["digest"][0], int(sc.p["maxCand"])).sum())
        if mode != "rebuild":
            r["skin"] = net.skin_stats()
        if args.profile_steps:
            r["kernels"] = profile(net, t0 + args.steps, args.profile_steps)
            if mode == "rebuild":  # the broad phase of the last profiled step, 20 launches back to back
                r["broad_phase_back_to_back_us"] = net.time_broad_phase(20)
        xl_end = sc.xl.copy()
        fs_end = sc.fs.copy()
        net.pull(fs_end, xl_end)
        assert np.isfinite(fs_end.coord).all(), "non-finite pose after the timed run"
        r["links_alive_end"] = int((xl_end.linkState[: xl_end.nLinks] >= 0).sum())
        net.close()
        runs[mode] = r
        log(f"{mode}: {r['value']:.4g} rod-steps/s, {r['ms_per_step']:.4f} ms/step")

    # the learned-skin form keeps the per-step rebuild as its fallback and judge: run whichever of
    # the two measured faster on this workload (both rates are reported)
    if "rebuild" in runs and "learned" in runs:
        choice = "learned" if runs["learned"]["value"] >= runs["rebuild"]["value"] else "rebuild"
        head = runs[choice]
    else:
        choice = modes[args.skin][0]
        head = runs[choice]
    if "rebuild" in runs and "learned" in runs:
        same = (np.array_equal(runs["rebuild"]["digest"][0], runs["learned"]["digest"][0])
                and runs["rebuild"]["digest"][1] == runs["learned"]["digest"][1])
        assert same, "learned-skin candidate slices differ from the per-step rebuild"
    roof = roof_rod = None
    per_kernel = head.get("kernels", {})
    entries = head.get("skin", {}).get("list_entries")
    cw = head.get("cand_written")
    if per_kernel:
        dom = max((k for k in per_kernel if kernel_bytes(sc, k, entries, cw)), key=lambda k: per_kernel[k]["avg_us"])
        b = kernel_bytes(sc, dom, entries, cw)
        # launch duration: back-to-back launches between two HIP events for the broad phase (a pair of
        # events per launch adds its dispatch and completion signalling), the per-launch timers else
        dur_us, timing = per_kernel[dom]["avg_us"], "HIP events around each launch of 40 eager steps"
        if dom == "k_broad_phase" and head.get("broad_phase_back_to_back_us"):
            dur_us, timing = head["broad_phase_back_to_back_us"], "HIP events around 20 back-to-back launches"
        achieved = b / (dur_us * 1e-6) / 1e9
        # the committed counter summaries are of the default workload (200k rods, 50k links, 20 um, water)
        profiled = (args.n_fil, args.n_links, args.box, args.eta, args.formation) == (20000, 50000, 20.0, 0.1, 100)
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": kernel_traffic(dom) if profiled else None, "bytes_per_launch": b,
                "duration_us": round(dur_us, 2), "timing": timing,
                "valu_issue_frac": kernel_valu_issue(dom) if profiled else None}
        if "k_rod" in per_kernel:  # the gather/integrate kernel (chain and link gathers + integrate)
            kr = kernel_bytes(sc, "k_rod") / (per_kernel["k_rod"]["avg_us"] * 1e-6) / 1e9
            roof_rod = {"kernel": "k_rod", "achieved": round(kr, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(kr / HBM_PEAK_GBS, 4), "bytes_per_launch": kernel_bytes(sc, "k_rod"),
                        "duration_us": round(per_kernel["k_rod"]["avg_us"], 2),
                        "traffic": kernel_traffic("k_rod") if profiled else None,
                        "note": "duration: HIP events around each launch, which add dispatch and completion"}
        for rr in runs.values():
            for k, v in rr.get("kernels", {}).items():
                kb = kernel_bytes(sc, k, rr.get("skin", {}).get("list_entries"), rr.get("cand_written"))
                if kb:
                    v["GBs"] = round(kb / (v["avg_us"] * 1e-6) / 1e9, 1)
    cpu = cpu1 = None
    if not args.no_cpu:
        cpu = run_cpu_baseline(start_state[0], start_state[1], openmp=True)
        cpu1 = run_cpu_baseline(start_state[0], start_state[1], openmp=False)

    bp = {"compare": f"learned-skin policy (faster of learned skin and per-step rebuild: {choice})",
          "rebuild": "per-step rebuild", "fixed": "fixed skin", "learned": "learned skin"}[args.skin]
    out = {
        "metric": METRIC, "value": round(head["value"], 1), "unit": "rod-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(head["ms_per_step"], 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 state, f64 force laws", "data": "synthetic (seeded scene)",
        "config": {"workload": f"crosslinked actin network: {args.n_fil // 1000}k filaments x 10 rods (0.1 um), "
                               f"{args.n_links // 1000}k crosslinker slots, {args.box:g} um periodic cube, dt 1e-5 s, "
                               f"viscosity {args.eta:g} Pa s"
                               + (f", formation every {args.formation} steps (binding radius 40 nm)"
                                  if args.formation > 0 else ""),
                   "rods": n, "links": sc.xl.nLinks, "grid": list(sc.p["gridDims"][:3]), "graph": bool(args.graph),
                   "broad_phase": bp},
        "roofline": roof, "roofline_gather_integrate": roof_rod, "cpu_baseline": cpu,
        "cpu_baseline_single_thread": cpu1,
        "links_alive_start": head["links_alive_start"], "links_alive_end": head["links_alive_end"],
    }
    if "rebuild" in runs and "learned" in runs:
        out["learned_vs_rebuild"] = round(runs["learned"]["value"] / runs["rebuild"]["value"], 3)
        out["pair_sets_identical"] = True
    out["runs"] = {m: {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items() if k != "digest"}
                   for m, r in runs.items()}
    if train_info:
        out["skin_training"] = train_info
    ctx.close()
    if args.at_scale_steps > 0:
        out["learned_skin_at_scale"] = learned_skin_at_scale(args, log)
    print(json.dumps(out))
    return 0


def learned_skin_at_scale(args, log) -> dict:
    """The learned-skin comparison at scale: the crosslinked-network recipe scaled to 200k filaments (2M rods) in
    a 60 um box with 400k crosslinker slots and 8k minifilaments (128k heads), 1 Pa s, where the skin covers the
    binding grid of the myosin stage (the reach skin).  Per-step rebuild and learned reach skin on the same steps;
    their reach lists, binding events and poses at the end are identical (checked)."""
    import copy

    a2 = copy.copy(args)
    a2.workload, a2.n_fil, a2.n_links, a2.box, a2.eta, a2.turnover = "actomyosin", 200000, 400000, 60.0, 1.0, 0
    a2.skin, a2.steps, a2.warmup, a2.no_cpu, a2.profile_steps = "compare", args.at_scale_steps, 10, True, 0
    sub = run_actomyosin(a2, log, emit=False)
    runs = sub["runs"]
    return {"workload": sub["config"]["workload"], "rods": sub["config"]["rods"],
            "motor_heads": sub["config"]["motor_heads"], "steps": a2.steps,
            "rebuild_ms_per_step": runs["rebuild"]["ms_per_step"], "learned_ms_per_step": runs["learned"]["ms_per_step"],
            "learned_rod_steps_per_s": round(runs["learned"]["value"], 1),
            "learned_vs_rebuild": sub["learned_vs_rebuild"],
            "reach_lists_and_binding_identical": sub["reach_lists_and_binding_identical"],
            "reach_skin": runs["learned"].get("reach_skin"), "training": sub.get("reach_skin_training")}


def run_cpu_baseline_actomyosin(sc, step0: int, budget_s: float = 12.0, openmp: bool = False) -> dict:
    """The oracle's actomyosin step (DenseContractileHarness.cpuStep order) from the GPU's state at the
    start of its timed region, over the same step numbers, for about budget_s seconds."""
    from oracle import ffi

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)) if openmp else 1
    orc = ffi.OracleContractile(sc, ffi.load_omp() if openmp else None)
    steps, t0 = 0, time.perf_counter()
    while True:
        orc.step(step0 + steps, 1)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or steps >= 200:
            break
    form = f"OpenMP build, {threads} threads" if openmp else "single thread"
    rods = int((sc.turn.filState[: sc.fs.n] >= 0).sum()) if getattr(sc, "turn", None) is not None else sc.fs.n
    return {"value": rods * steps / el, "unit": "rod-steps/s (filament rods)", "cores": threads, "kind": "port",
            "sample": f"oracle C restatement ({form}), the GPU's state at the start of its timed region "
                      f"({rods} filament rods, {sc.mot.nM} motor heads, {sc.xl.nLinks} link slots), steps "
                      f"{step0}..{step0 + steps - 1}, full actomyosin step incl. grid"
                      + (" and turnover" if getattr(sc, "turn", None) is not None else "")
                      + f": {steps} steps in {el:.1f} s"}


def actomyosin_turnover(sc, scenes):
    """The workload's turnover: barbed-end growth at 10 subunits/s at every chain's barbed tip, pointed-end
    depolymerisation (nucleotide-dependent off-rates, death below 3 monomers) with dead-slot reuse, aging,
    cofilin severing (dissolve ratio 0.8), splits at 64 monomers into 10% spare slots, every 100 steps.
    Ragged ends, as in a network that has been turning over: barbed tips hold 33-63 monomers and pointed
    tips 3-36, uniformly, so splits and deaths fall inside the timed steps."""
    t = scenes.with_turnover(sc, spare=sc.fs.n // 10, growth_rate=10.0, cofilin_ratio=0.8)
    ts = t.turn
    g = np.random.default_rng(2024)
    tips = np.flatnonzero(ts.seedNode >= 0)
    ts.monomerCount[tips] = g.integers(33, 64, tips.size)
    pointed = np.flatnonzero((t.fs.end1NbrSlot < 0) & (ts.filState >= 0) & (ts.seedNode < 0))
    ts.monomerCount[pointed] = g.integers(3, 37, pointed.size)
    return t


def reach_digest(net, sc) -> dict:
    """Arrays that must be identical between the per-step rebuild and the reach skin: the reach lists (count
    and the first maxCand entries), the binding state and the filament and motor poses."""
    mot, mini = sc.mot.copy(), sc.mini.copy()
    net.pull_motors(mot, mini)
    fs = sc.fs.copy()
    net.pull(fs, sc.xl.copy())
    mc = int(sc.p["maxCand"])
    rc = net.ctx.to_host(net.mot.t["reachCount"]).copy()
    rs = net.ctx.to_host(net.mot.t["reachSeg"]).reshape(-1, mc)
    lists = np.concatenate([rs[m, : min(int(rc[m]), mc)] for m in np.flatnonzero(rc)] or [np.zeros(0, np.int32)])
    return {"reachCount": rc, "reachSeg": lists, "boundSeg": mot.boundSeg.copy(), "bindArc": mot.bindArc.copy(),
            "stats": mot.stats.copy(), "coord": fs.coord.copy(), "motorCoord": mot.body.coord.copy()}


def run_actomyosin(args, log, emit: bool = True):
    """The full actomyosin network: 50k filaments x 10 rods (0.1 um) in the reference's clamped 30 um box
    with containment walls, 2k minifilaments (4 dimers each end -> 16 heads, stall 4 pN), 100k
    crosslinkers; motor reach grid + binding + catch-slip + cycle + bond forces + joints + dimer
    coupling + tethers + backbones on top of the network step (DenseContractileHarness.cpuStep order).
    Viscosity 1 Pa s, the dense fixture's crowded cytoplasm (DenseContractileHarness.java:58): at water
    viscosity the placed heads and links fling segments micrometres per step.

    --skin: the binding grid's search per step (rebuild: the reference's grid sorted every step), from the
    reach skin (fixed, learned), or both on the same trajectory (compare: rebuild and learned; ablation:
    rebuild, fixed and learned).  Every run starts from the same scene and takes the same steps; the runs'
    reach lists, binding state and poses at the end must be identical."""
    import torch  # noqa: F401  (device runtime first, as in the network workload)

    from softbox_amd import _lib, scenes
    from softbox_amd.context import Context
    from softbox_amd.network import Network
    from softbox_amd.skin import ReachSkinConfig, fit_reach

    n_fil, n_links, box = args.n_fil, args.n_links, args.box
    t_build = time.perf_counter()
    sc = scenes.contractile(n_fil=n_fil, segs=10, box=box, n_mini=max(1, n_fil // 25), n_links=n_links, seed=0,
                            eta=args.eta)
    if args.formation > 0:
        sc.p.update(scenes.formation_params(box, 0.1, sc.fs.n, False, grab=0.04, interval=args.formation))
    n_rods = sc.fs.n
    if args.turnover:
        sc = actomyosin_turnover(sc, scenes)
    n = n_rods  # the metric counts the network's filament rods (spare slots excluded)
    log(f"actomyosin scene: {n} filament rods (+{sc.fs.n - n_rods} free slots), {sc.mot.nM} heads, {sc.mini.nBb} "
        f"backbones, {sc.xl.nLinks} links ({time.perf_counter() - t_build:.1f} s)")
    ctx = Context(0)
    modes = {"compare": ["rebuild", "learned"], "rebuild": ["rebuild"], "fixed": ["fixed"], "learned": ["learned"],
             "ablation": ["rebuild", "fixed", "learned"]}[args.skin]
    fcfg = ReachSkinConfig(skin_default=args.reach_skin, horizon=args.reach_horizon)
    rcfg = ReachSkinConfig(skin_default=args.reach_skin, horizon=args.reach_horizon)
    model, train_info = None, None
    if "learned" in modes:
        t_tr = time.perf_counter()
        tnet = Network(ctx, sc.copy(), keep_forces=False)
        model, hist, (x, _), _ = fit_reach(ctx, tnet, rcfg, snapshots=args.train_snapshots, iters=args.train_iters,
                                           horizon=50)
        tnet.close()
        train_info = {"samples": int(x.shape[0]), "loss_start": round(hist[0], 4), "loss_end": round(hist[-1], 4),
                      "margin": round(rcfg.margin, 3), "trained_horizon": 50, "schedule_horizon": rcfg.horizon,
                      "seconds": round(time.perf_counter() - t_tr, 1)}
        log(f"reach skin model trained: {train_info}")
    runs, digests = {}, {}
    tuned, t0_all = None, None
    for mode in modes:
        net = Network(ctx, sc.copy(), keep_forces=False)
        net.set_overlap(bool(args.overlap))
        if mode != "rebuild":
            if mode == "learned":
                net.set_reach_skin(_lib.SBX_SKIN_LEARNED, rcfg, model)
            else:
                net.set_reach_skin(_lib.SBX_SKIN_FIXED, fcfg)
        if args.graph:
            net.use_graph(True)
        net.step(0, args.warmup)
        t0 = args.warmup
        if args.shape:
            net.set_shape(*[int(v) for v in args.shape.split(",")])
        elif args.tune_steps:
            # launch shape and lanes per head over real steps of the run; every network takes the same number
            # of tuning steps, so every run times the same step numbers
            tn = net.tune(t0, args.tune_steps)
            t0 += tn["steps"]
            if tuned is None:
                tuned = tn
            log(f"tuned launch shape {tn} ({mode})")
        if args.motor_group:
            net.set_motor_group(args.motor_group)
        if t0_all is None:
            t0_all = t0
            start = sc.copy()
            start.mot, start.dim, start.mini = sc.mot.copy(), sc.dim.copy(), sc.mini.copy()
            net.pull(start.fs, start.xl)
            net.pull_motors(start.mot, start.mini)
            if start.turn is not None:
                net.pull_turnover(start.turn, start.fs)
        links_start = int((start.xl.linkState[: start.xl.nLinks] >= 0).sum())
        bound_start = int((start.mot.boundSeg >= 0).sum())
        formed0 = net.form_stats().get("formed_total", 0) if args.formation > 0 else 0
        turn0 = net.pull_turnover(start.turn.copy()) if start.turn is not None else None
        el = timed(net, ctx, t0, args.steps)
        assert t0 == t0_all, "every run times the same steps"
        r = {"value": n * args.steps / el, "ms_per_step": 1e3 * el / args.steps,
             "timed_steps": [t0, t0 + args.steps - 1], "motor_group": net.motor_group()}
        if args.formation > 0:
            r["links_formed_timed"] = net.form_stats()["formed_total"] - formed0
        if turn0 is not None:
            tend = net.pull_turnover(start.turn.copy())
            r["turnover"] = {"cadences_timed": int(tend.stats[6] - turn0.stats[6]),
                             "monomers_grown": int(tend.ledger[0] - turn0.ledger[0]),
                             "monomers_returned": int(tend.ledger[1] - turn0.ledger[1]),
                             "segments_active_start": int((turn0.filState >= 0).sum()),
                             "segments_active_end": int((tend.filState >= 0).sum()),
                             "free_actin_uM_end": round(float(tend.pool[0]), 4)}
        if mode != "rebuild":
            r["reach_skin"] = net.reach_skin_stats()
        digests[mode] = reach_digest(net, sc)
        r["heads_bound_end"] = int((digests[mode]["boundSeg"] >= 0).sum())
        r["reach_hits_end"] = int(digests[mode]["reachCount"].sum())
        if args.profile_steps:
            r["kernels"] = profile(net, t0 + args.steps, args.profile_steps)
        end = sc.copy()
        end.mot, end.mini = sc.mot.copy(), sc.mini.copy()
        net.pull(end.fs, end.xl)
        net.pull_motors(end.mot, end.mini)
        assert np.isfinite(end.fs.coord).all() and np.isfinite(end.mot.body.coord).all(), "non-finite pose"
        r["links_alive_start"] = links_start
        r["links_alive_end"] = int((end.xl.linkState[: end.xl.nLinks] >= 0).sum())
        r["heads_bound_start"] = bound_start
        net.close()
        runs[mode] = r
        log(f"{mode}: {r['value']:.4g} rod-steps/s, {r['ms_per_step']:.4f} ms/step")
    # every form took the same steps: the reach lists, binding events and poses must be identical
    base = modes[0]
    for mode in modes[1:]:
        for k, v in digests[base].items():
            assert np.array_equal(v, digests[mode][k]), f"{mode} run differs from {base} in {k}"
    head_mode = "learned" if "learned" in runs else modes[0]
    head = runs[head_mode]
    kernels = head.get("kernels", {})
    roof = roof_rod = None
    if kernels:
        known = {k: kernel_bytes(sc, k) for k in kernels}
        dom = max((k for k in kernels if known[k]), key=lambda k: kernels[k]["avg_us"])
        b = known[dom]
        achieved = b / (kernels[dom]["avg_us"] * 1e-6) / 1e9
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "bytes_per_launch": b,
                "duration_us": round(kernels[dom]["avg_us"], 2),
                "timing": f"HIP events around each launch of {args.profile_steps} eager steps"}
        for rr in runs.values():
            for k, v in rr.get("kernels", {}).items():
                if known.get(k):
                    v["GBs"] = round(known[k] / (v["avg_us"] * 1e-6) / 1e9, 1)
        if known.get("k_rod"):  # the gather/integrate kernel (chain, link and motor gathers + integrate)
            kr = known["k_rod"] / (kernels["k_rod"]["avg_us"] * 1e-6) / 1e9
            roof_rod = {"kernel": "k_rod", "achieved": round(kr, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(kr / HBM_PEAK_GBS, 4), "bytes_per_launch": known["k_rod"],
                        "duration_us": round(kernels["k_rod"]["avg_us"], 2),
                        "note": "bytes: the filament formula (the motor gather's reads left out); duration: HIP "
                                "events around each launch, which add dispatch and completion"}
    cpu = cpu1 = None
    if not args.no_cpu:
        cpu = run_cpu_baseline_actomyosin(start, t0_all, openmp=True)
        cpu1 = run_cpu_baseline_actomyosin(start, t0_all, openmp=False)
    bp = {"rebuild": "per-step binding grid (the reference's gridReachable)", "fixed": "reach skin, fixed",
          "learned": "reach skin, learned"}[head_mode]
    out = {
        "metric": "rod-steps/sec (filament rods) of the actomyosin step at 1 GPU", "value": round(head["value"], 1),
        "unit": "rod-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 state, f64 force laws", "data": "synthetic (seeded scene)",
        "config": {"workload": f"full actomyosin network: {n_fil // 1000}k filaments x 10 rods (0.1 um), "
                               f"{sc.mini.nBb} minifilaments x 16 heads (stall 4 pN), {n_links // 1000}k crosslinker "
                               f"slots, {box:g} um clamped box with walls, dt 1e-5 s, viscosity {args.eta:g} Pa s"
                               + (f", formation every {args.formation} steps" if args.formation > 0 else ""),
                   "rods": n, "motor_heads": sc.mot.nM, "backbones": sc.mini.nBb, "links": sc.xl.nLinks,
                   "grid": list(sc.p["gridDims"][:3]), "graph": bool(args.graph),
                   "shape": shape_of(args, tuned) if (args.shape or tuned) else None,
                   "motor_group": head["motor_group"], "binding_search": bp},
        "roofline": roof, "roofline_gather_integrate": roof_rod, "cpu_baseline": cpu,
        "cpu_baseline_single_thread": cpu1,
        "links_alive_start": head["links_alive_start"], "links_alive_end": head["links_alive_end"],
        "heads_bound_start": head["heads_bound_start"], "heads_bound_end": head["heads_bound_end"],
        "timed_steps": head["timed_steps"],
    }
    if "rebuild" in runs and len(runs) > 1:
        for mode in runs:
            if mode != "rebuild":
                out[f"{mode}_vs_rebuild"] = round(runs[mode]["value"] / runs["rebuild"]["value"], 3)
        out["reach_lists_and_binding_identical"] = True
    if "links_formed_timed" in head:
        out["links_formed_timed"] = head["links_formed_timed"]
    if "turnover" in head:
        out["turnover"] = head["turnover"]
        out["config"]["workload"] += (", turnover every 100 steps (barbed growth 10 subunits/s, pointed-end "
                                      "depolymerisation, aging, cofilin severing, splits into 10% free slots)")
    out["runs"] = {m: {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()} for m, r in runs.items()}
    if train_info:
        out["reach_skin_training"] = train_info
    ctx.close()
    if not emit:
        return out
    print(json.dumps(out))
    return 0


STANDARD_1X_PUBLISHED = {"v1-CPU": 30.4, "v1-GPU": 21.4, "v2-CPU": 15.3, "v2-GPU": 49.5}   # steps/s


def run_standard(args, log) -> int:
    """The reference's 1x contractility standard (V2OneXHarness.java:14-47): 400 free protein nodes with 24
    singlet myosins each, 1000 static filaments x 10 segments (64 monomers), crosslinker formation every 100
    steps (dense preset, 1 uM), containment every 1e-4 s, viscosity 0.1 Pa s, dt 1e-5 s, the harness's seeds
    and java.util.Random initial condition.  Reported in the harness's own unit, steps/s, beside its
    published four-way figures (v1-CPU 30.4, v1-GPU 21.4, v2-CPU 15.3, v2-GPU 49.5 steps/s on the reference's
    hardware)."""
    import torch  # noqa: F401

    from softbox_amd import scenes
    from softbox_amd.context import Context
    from softbox_amd.network import Network

    t_build = time.perf_counter()
    sc = scenes.standard_1x()
    n = sc.fs.n
    log(f"1x standard scene: {n} segments, {sc.mot.nM} myosins on {sc.mini.nNodes} nodes, {sc.xl.nLinks} link "
        f"slots ({time.perf_counter() - t_build:.1f} s)")
    ctx = Context(0)
    net = Network(ctx, sc.copy(), keep_forces=False)
    net.set_overlap(bool(args.overlap))
    reach = args.skin == "fixed"  # the reach skin of the binding grid (the default here: the per-step grid)
    if reach:
        from softbox_amd import _lib
        from softbox_amd.skin import ReachSkinConfig

        net.set_reach_skin(_lib.SBX_SKIN_FIXED, ReachSkinConfig(skin_default=args.reach_skin,
                                                                horizon=args.reach_horizon))
    if args.graph:
        net.use_graph(True)
    net.step(0, args.warmup)
    t0 = args.warmup
    tuned = None
    if args.shape:
        net.set_shape(*[int(v) for v in args.shape.split(",")])
    elif args.tune_steps:
        tuned = net.tune(t0, args.tune_steps)
        t0 += tuned["steps"]
        log(f"tuned launch shape {tuned}")
    if args.motor_group:
        net.set_motor_group(args.motor_group)
    start = sc.copy()
    start.mot, start.dim, start.mini = sc.mot.copy(), sc.dim.copy(), sc.mini.copy()
    net.pull(start.fs, start.xl)
    net.pull_motors(start.mot, start.mini)
    formed0 = net.form_stats().get("formed_total", 0)
    el = timed(net, ctx, t0, args.steps)
    sps = args.steps / el
    formed = net.form_stats()["formed_total"] - formed0
    kernels = profile(net, t0 + args.steps, args.profile_steps) if args.profile_steps else {}
    end = sc.copy()
    end.mot, end.mini = sc.mot.copy(), sc.mini.copy()
    net.pull(end.fs, end.xl)
    net.pull_motors(end.mot, end.mini)
    assert np.isfinite(end.fs.coord).all() and np.isfinite(end.mot.body.coord).all(), "non-finite pose"
    assert np.isfinite(end.mini.node.coord).all(), "non-finite node pose"
    mgroup = net.motor_group()
    reach_stats = net.reach_skin_stats() if reach else None
    net.close()
    roof = None
    if kernels:
        dom = max(kernels, key=lambda k: kernels[k]["avg_us"])
        b = kernel_bytes(sc, dom)
        roof = {"bound": "launch", "kernel": dom, "duration_us": round(kernels[dom]["avg_us"], 2),
                "note": "at 10k segments and 29k motor bodies every kernel is a few waves deep: the step is "
                        "bound by launch and dependency latency, not by HBM"}
        if b:
            a = b / (kernels[dom]["avg_us"] * 1e-6) / 1e9
            roof.update(achieved=round(a, 1), peak=HBM_PEAK_GBS, unit="GB/s", frac=round(a / HBM_PEAK_GBS, 4),
                        bytes_per_launch=b)
    cpu = cpu1 = None
    if not args.no_cpu:
        for openmp in (True, False):
            c = run_cpu_baseline_actomyosin(start, t0, budget_s=10.0, openmp=openmp)
            c["steps_per_s"] = round(c["value"] / n, 2)
            if openmp:
                cpu = c
            else:
                cpu1 = c
    out = {
        "metric": "steps/s of the 1x contractility standard at 1 GPU", "value": round(sps, 1), "unit": "steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * el / args.steps, 4),
        "higher_is_better": True, "scaling": "weak",
        "vs_baseline": round(sps / STANDARD_1X_PUBLISHED["v2-GPU"], 1),
        "vs_published": {k: round(sps / v, 1) for k, v in STANDARD_1X_PUBLISHED.items()},
        "rod_steps_per_s": round(n * sps, 1),
        "dtype": "f32 state, f64 force laws", "data": "synthetic (the harness's seeded initial condition)",
        "config": {"workload": "1x contractility standard (V2OneXHarness): 400 free nodes x 24 singlet myosins, "
                               "1000 static filaments x 10 segments (64 monomers), crosslinker formation every 100 "
                               "steps (1 uM, kon 40 /uM/s), 7.071 x 7.071 x 0.5 um slab with containment, dt 1e-5 s, "
                               "viscosity 0.1 Pa s, Wang-hash RNG with the harness's seeds",
                   "segments": n, "myosins": sc.mot.nM, "nodes": sc.mini.nNodes, "link_slots": sc.xl.nLinks,
                   "grid": list(sc.p["gridDims"][:3]), "graph": bool(args.graph),
                   "shape": shape_of(args, tuned) if (args.shape or tuned) else None, "motor_group": mgroup,
                   "binding_search": "reach skin, fixed" if reach else "per-step binding grid (the reference's)"},
        "reach_skin": reach_stats,
        "roofline": roof, "cpu_baseline": cpu, "cpu_baseline_single_thread": cpu1,
        "links_formed_timed": formed, "links_alive_end": int((end.xl.linkState[: end.xl.nLinks] >= 0).sum()),
        "heads_bound_start": int((start.mot.boundSeg >= 0).sum()), "heads_bound_end": int((end.mot.boundSeg >= 0).sum()),
        "timed_steps": [t0, t0 + args.steps - 1],
    }
    if kernels:
        out["kernels"] = {k: {kk: (round(vv, 2) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                          for k, v in kernels.items()}
    print(json.dumps(out))
    ctx.close()
    return 0


def run_batch(args, log) -> int:
    """The batched ensemble: args.batch replicas of the network recipe (filament count log-uniform in
    [2k, 40k], actin 5-30 uM setting each periodic box, 2.5 links per filament), concatenated and
    advanced together; rod-steps/s over all replicas and the p50 / p99 of the batch step time."""
    import torch

    from softbox_amd import scenes
    from softbox_amd.batch import Batch, tune_cell_size
    from softbox_amd.context import Context

    t_build = time.perf_counter()
    reps = scenes.ensemble(args.batch, seed=0, eta=args.eta)
    rods = [sc.fs.n for sc in reps]
    n = sum(rods)
    log(f"ensemble: {len(reps)} replicas, {n} rods (per replica {min(rods)}..{max(rods)}), "
        f"{sum(sc.xl.nLinks for sc in reps)} links ({time.perf_counter() - t_build:.1f} s)")
    ctx = Context(0)
    factors = tuple(float(v) for v in args.cell_factors.split(",") if v.strip()) if args.cell_factors else (1.0,)
    cell = {"factor": 1.0, "ms_per_step": None}
    if len(factors) > 1:
        cell = tune_cell_size(ctx, reps, factors=factors, trial_steps=10)
        log(f"cell-size tuner: {cell}")
        if cell["factor"] != 1.0:
            reps = [scenes.retile(sc, cell["factor"]) for sc in reps]
    bt = Batch(ctx, reps, keep_forces=False)
    bt.net.set_overlap(bool(args.overlap))
    if args.graph:
        bt.net.use_graph(True)
    bt.step(0, args.warmup)
    t0 = args.warmup
    shape = None
    if args.shape:
        bt.net.set_shape(*[int(v) for v in args.shape.split(",")])
        shape = shape_of(args)
    elif args.tune_steps:
        tuned = bt.net.tune(t0, args.tune_steps)
        t0 += tuned["steps"]
        shape = shape_of(args, tuned)
        log(f"tuned launch shape {tuned}")
    el = timed(bt, ctx, t0, args.steps)
    t0 += args.steps
    # batch step times: pairs of steps (one captured graph) between device synchronisations
    per = []
    for _ in range(max(8, min(64, args.steps // 2))):
        torch.cuda.synchronize()
        ts = time.perf_counter()
        bt.step(t0, 2)
        t0 += 2
        ctx.sync()
        torch.cuda.synchronize()
        per.append(1e3 * (time.perf_counter() - ts) / 2.0)
    per = np.array(per)
    kernels = profile(bt.net, t0, args.profile_steps) if args.profile_steps else None
    fs = reps[0].fs.copy()
    bt.pull(0, fs)
    assert np.isfinite(fs.coord).all(), "non-finite pose after the timed run"
    out = {
        "metric": "rod-steps/sec, batched ensemble at 1 GPU", "value": round(n * args.steps / el, 1),
        "unit": "rod-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * el / args.steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 state, f64 force laws", "data": "synthetic (seeded ensemble)",
        "config": {"workload": f"{len(reps)} replicas of the crosslinked network recipe, 2k-40k filaments x 10 rods "
                               "(log-uniform), actin 5-30 uM (periodic box), 2.5 links per filament, dt 1e-5 s",
                   "rods": n, "replicas": len(reps), "rods_per_replica": [min(rods), int(np.median(rods)), max(rods)],
                   "graph": bool(args.graph), "shape": shape, "cell_factor": cell["factor"],
                   "cell_tuning_ms_per_step": cell["ms_per_step"]},
        "step_ms": {"p50": round(float(np.percentile(per, 50)), 4), "p99": round(float(np.percentile(per, 99)), 4)},
    }
    if kernels:
        out["kernels"] = {k: {"avg_us": round(v["avg_us"], 1), "launches": v["launches"]} for k, v in kernels.items()}
    print(json.dumps(out))
    bt.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
