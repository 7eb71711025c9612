"""Benchmark of the PBF step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-batched] [--no-adjoint]
                    [--no-late]

Workload: the 1M-particle dam break (h = 0.02, box 2.027 x 2.027 x 1.0135,
102 x 102 x 51 = 530,604 cells, 4 solver iterations, dt = 1/120), seeded
with PCG32 jitter; state stays resident in HBM for the whole timed region.

One JSON line on stdout:
  value        algorithmic HBM GB/s of the whole step, B_step / t_step with
               B_step = (100 + 44 I) N + (16 + 8 I) C bytes (the compulsory
               traffic of every stage: vec3 = 12 B, u32/f32 = 4 B)
  ms_per_step  wall time of K steps between two device synchronisations / K
  roofline     the dominant kernel: its algorithmic bytes per launch over its
               mean duration from HIP events recorded on the launch stream
               during the timed steps; traffic = HBM bytes per launch from
               rocprofv3 FETCH_SIZE / WRITE_SIZE (profiles/), else null
  cpu_baseline the oracle's FP32 step on this host's cores (OpenMP), timed
               on a bounded sample of the same workload
  late_steps   ms per step in windows of 20 over the first 120 steps of the
               same dam break (the step slows as dense corners form)
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
N_PARTICLES = 1_000_000
SCENE_DAM = 4
DT = 1.0 / 120.0


def stage_bytes(name, n, c):
    """Algorithmic (compulsory) bytes of one launch of each stage."""
    return {
        "predict_hash_count": 28 * n + 4 * c,   # R x,v (24N), W key (4N), counts (4C)
        "exclusive_scan": 8 * c,                 # R counts + W cell_start
        "stable_scatter_gather": 24 * n + 4 * c, # R key, cell_start; W index, slot map, sorted x
        "density_lambda": 16 * n + 4 * c,        # R sorted x (12N), cell_start; W lambda
        "apply_delta": 28 * n + 4 * c,           # R sorted x + lambda (16N), cell_start; W x'
        # R slot map, final iterate, x, neighbour count, lambda; W x, v, prev, foam, lambda
        "scatter_finalize": 100 * n,
        "count_neighbors": 16 * n + 4 * c,
    }.get(name, 0)


def step_bytes(n, c, iters):
    return (152 + 44 * iters) * n + (16 + 8 * iters) * c


def read_traffic(kernel):
    """HBM bytes per launch of `kernel` from profiles/hbm_traffic.txt
    (lines: `<kernel> <workload> fetch_bytes=<B> write_bytes=<B>`)."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 4 and parts[0] == kernel and parts[1] == "dam_break_1m":
                kv = dict(p.split("=", 1) for p in parts[2:] if "=" in p)
                try:
                    return float(kv["fetch_bytes"]) + float(kv["write_bytes"])
                except (KeyError, ValueError):
                    return None
    return None


def cpu_baseline(x, v, params, iters_budget_s=20.0):
    """The oracle's FP32 step on the host (all cores), on whole 1M steps."""
    from oracle import build as obuild
    from oracle import oracle as O
    if not os.path.exists(O.LIB_PATH):
        obuild.build(force=True)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    O.set_threads(threads)
    op = O.params(params)
    g = O.grid(op)
    O.step(x, v, op, DT, prec=32)  # warm-up
    times = []
    t_start = time.perf_counter()
    while len(times) < 3 and time.perf_counter() - t_start < iters_budget_s:
        t0 = time.perf_counter()
        O.step(x, v, op, DT, prec=32)
        times.append(time.perf_counter() - t0)
    t = statistics.median(times)
    b = step_bytes(len(x), g.total_cells, op.iterations)
    return {"value": round(b / t / 1e9, 3), "unit": "GB/s", "cores": O.get_threads(), "kind": "port",
            "ms_per_step": round(t * 1e3, 1),
            "sample": f"{len(times)} whole FP32 oracle steps (after 1 warm-up) of the same 1M dam break, "
                      f"median; OpenMP over particles, serial scan/scatter"}


def _time_steps(eng, steps, stream, torch):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.step(DT, stream)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def batched_leg(steps, warmup, stream, torch):
    """256 independent cube drops (N_s log-uniform in [2048, 65536], per-scene
    rest density and gravity, 3 iterations) in one batch, against one cube
    drop of the same total particle count stepped alone."""
    from mtlphys_amd import workloads as W
    import mtlphys_amd as mtlp
    b = W.make_ensemble()
    n = b.particle_count
    block = b.autotune(DT, os.path.join(ROOT, "mtlphys_amd", "autotune_cache.txt"))
    b.set_graph(True)
    for _ in range(warmup):
        b.step(DT, stream)
    t_batch = _time_steps(b, steps, stream, torch)
    _, co = b.offsets()
    iters = b.scene_params[0].iterations
    bytes_batch = sum(W.step_bytes(c, int(co[s + 1] - co[s]), iters) for s, c in enumerate(b.counts))
    flags = b.error_flags()
    b.close()
    p, x, v = W.cube_drop(n)
    single = mtlp.Engine(0, p)
    single.reset(n, mtlp.capi.SCENE_EMPTY)
    single.upload_state(x, v)
    single.autotune(DT, os.path.join(ROOT, "mtlphys_amd", "autotune_cache.txt"))
    single.set_graph(True)
    for _ in range(warmup):
        single.step(DT, stream)
    t_single = _time_steps(single, steps, stream, torch)
    bytes_single = W.step_bytes(n, single.total_cells, p.iterations)
    single.close()
    gbs_b, gbs_s = bytes_batch / t_batch / 1e9, bytes_single / t_single / 1e9
    return {"scenes": 256, "particles": n, "iterations": iters, "stencil_block": block,
            "ms_per_step": round(t_batch * 1e3, 4), "value": round(gbs_b, 2), "unit": "GB/s",
            "single_scene_ms_per_step": round(t_single * 1e3, 4), "single_scene_gbs": round(gbs_s, 2),
            "ratio_to_single_scene": round(gbs_b / gbs_s, 3), "error_flags": flags,
            "note": "one launch per stage for all scenes, HIP-graph replay; the single scene is one "
                    "cube drop with the same total particle count"}


def adjoint_leg(torch, reps=3):
    """The differentiable workload (65,536 particles, h = 0.04, 3 iterations,
    40 steps, checkpoints every 8): forward rollout + adjoint against 40
    plain steps from the same state, median of `reps`."""
    from mtlphys_amd import workloads as W
    import mtlphys_amd as mtlp
    p, x, v, target = W.differentiable()
    n, steps, every = len(x), W.ADJ_STEPS, W.ADJ_CKPT
    eng = mtlp.Engine(0, p)
    eng.reset(n, mtlp.capi.SCENE_EMPTY)
    dev = torch.device("cuda")
    stream = torch.cuda.current_stream()
    xbar = torch.zeros((n, 4), device=dev)
    xbar[:, 0] = 1.0 / n  # the centroid term's shape: any fixed adjoint input costs the same
    xb0, vb0, gb = (torch.empty((n, 4), device=dev), torch.empty((n, 4), device=dev), torch.empty(3, device=dev))
    t_plain, t_fwd, t_adj = [], [], []
    for rep in range(reps + 1):  # rep 0 warms up (allocations, code load)
        eng.upload_state(x, v)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.step_n(DT, steps, stream)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        eng.upload_state(x, v)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        eng.rollout(DT, steps, every, stream)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        eng.adjoint_rollout(DT, steps, every, xbar, None, xb0, vb0, gb, stream)
        torch.cuda.synchronize()
        t4 = time.perf_counter()
        if rep:
            t_plain.append(t1 - t0)
            t_fwd.append(t3 - t2)
            t_adj.append(t4 - t3)
    flags = eng.error_flags()
    finite = bool(torch.isfinite(vb0).all().item() and torch.isfinite(gb).all().item())
    eng.close()
    tp, tf, ta = (statistics.median(t) for t in (t_plain, t_fwd, t_adj))
    return {"workload": "differentiable_64k", "particles": n, "steps": steps, "ckpt_every": every,
            "iterations": p.iterations, "plain_ms": round(tp * 1e3, 3), "rollout_ms": round(tf * 1e3, 3),
            "adjoint_ms": round(ta * 1e3, 3), "fwd_plus_adjoint_over_plain": round((tf + ta) / tp, 3),
            "finite": finite, "error_flags": flags}


def step_curve(steps, torch, window=5):
    """ms per step of the 1M dam break in windows of `window` steps from the
    initial state (the step slows as the column collapses into dense
    corners), the slots with more than 96 neighbours at the end, and the
    per-stage times (HIP events) of one more step."""
    import mtlphys_amd as mtlp
    params = mtlp.scene_params(SCENE_DAM, N_PARTICLES)
    x, v = mtlp.generate_scene(SCENE_DAM, N_PARTICLES, 1, params)
    eng = mtlp.Engine(0, params)
    eng.reset(N_PARTICLES, mtlp.capi.SCENE_EMPTY)
    eng.upload_state(x, v)
    eng.set_stencil_block(64)
    stream = torch.cuda.current_stream()
    out = []
    for _ in range(max(1, steps // window)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(window):
            eng.step(DT, stream)
        torch.cuda.synchronize()
        out.append(round((time.perf_counter() - t0) / window * 1e3, 3))
    counts = eng.neighbor_counts_buffer().read()
    # one more step with the engine's per-kernel events: us per stage
    eng.set_profiling(True)
    eng.step(DT, stream)
    stages = {}
    for name, ms in eng.stage_times():
        stages[name] = round(stages.get(name, 0.0) + ms * 1e3, 1)
    eng.set_profiling(False)
    eng.close()
    return {"workload": "dam_break_1m", "window": window, "ms_per_step": out,
            "slots_over_96_neighbours": int((counts > 96).sum()), "max_neighbours": int(counts.max()),
            "next_step_stage_us": stages}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the forward+adjoint leg")
    ap.add_argument("--no-late", action="store_true",
                    help="skip the late-steps leg (ms per step over 120 steps from the initial state)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only run the forward+adjoint leg and print its JSON (for profiling)")
    ap.add_argument("--step-curve", type=int, default=0, metavar="STEPS",
                    help="only time STEPS steps of the 1M dam break from its initial state in windows of "
                         "5 (one-wave stencil workgroups) and print ms per step per window")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 device copies of MIB MiB (reference kernel of known "
                         "traffic for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (got --gpus {args.gpus}); "
              "a multi-GPU figure is not part of it", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.adjoint_only:
        print(json.dumps(adjoint_leg(torch, reps=2)), flush=True)
        return
    if args.step_curve:
        print(json.dumps(step_curve(args.step_curve, torch)), flush=True)
        return
    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": n * 4, "bytes_written": n * 4}), flush=True)
        return
    import mtlphys_amd as mtlp

    params = mtlp.scene_params(SCENE_DAM, N_PARTICLES)
    x, v = mtlp.generate_scene(SCENE_DAM, N_PARTICLES, 1, params)
    iters = params.iterations
    eng = mtlp.Engine(0, params)
    eng.reset(N_PARTICLES, mtlp.capi.SCENE_EMPTY)
    eng.upload_state(x, v)
    cells = eng.total_cells
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        eng.step(DT, stream)
    torch.cuda.synchronize()
    # workgroup size of the stencil kernels for this workload, tuned on the
    # warmed-up state (the engine's autotuner times a few steps per size and
    # restores the state)
    block = eng.autotune(DT, os.path.join(ROOT, "mtlphys_amd", "autotune_cache.txt"), stream)

    # timed region: K steps, HIP events after every kernel on the launch stream
    eng.set_profiling(True)
    per_stage = {}
    step_ms = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.step(DT, stream)
        st = eng.stage_times()  # waits for this step's last event
        step_ms.append(sum(ms for _, ms in st))
        for name, ms in st:
            per_stage.setdefault(name, []).append(ms)
    torch.cuda.synchronize()
    t_prof = (time.perf_counter() - t0) / args.steps
    eng.set_profiling(False)

    # the same K steps (same initial state and warm-up) without events or
    # host waits between steps
    eng.upload_state(x, v)
    for _ in range(args.warmup):
        eng.step(DT, stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.step(DT, stream)
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / args.steps
    flags = eng.error_flags()
    xs = eng.positions_buffer().read()
    eng.close()
    if flags or not np.isfinite(xs).all():
        print(f"bench.py: device error flags {flags:#x} / non-finite state", file=sys.stderr)
        sys.exit(1)

    b_step = step_bytes(N_PARTICLES, cells, iters)
    kernels = {}
    for name, ms in per_stage.items():
        mean_ms = statistics.fmean(ms)
        b = stage_bytes(name, N_PARTICLES, cells)
        kernels[name] = {"launches_per_step": len(ms) // args.steps, "mean_us": round(mean_ms * 1e3, 2),
                         "alg_bytes": b, "gbs": round(b / (mean_ms * 1e-3) / 1e9, 1) if mean_ms > 0 else None}
    dom = max(kernels, key=lambda k: kernels[k]["mean_us"] * kernels[k]["launches_per_step"])
    dk = kernels[dom]
    achieved = dk["gbs"]
    roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": read_traffic(dom),
                "alg_bytes_per_launch": dk["alg_bytes"], "mean_us": dk["mean_us"]}

    out = {
        "metric": "PBF step ms + HBM GB/s at 1M particles (dam break, 4 iterations)",
        "value": round(b_step / t_step / 1e9, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_step * 1e3, 4),
        "ms_per_step_p50": round(statistics.median(step_ms), 4),
        "ms_per_step_p95": round(float(np.percentile(step_ms, 95)), 4),
        "ms_per_step_with_events": round(t_prof * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: PCG32-jittered lattice dam break, seed 1",
        "config": {"workload": "dam_break_1m", "particles": N_PARTICLES, "h": params.h, "cells": cells,
                   "grid": list(mtlp.grid_dims(params)), "iterations": iters, "dt": DT, "stencil_block": block,
                   "alg_bytes_per_step": b_step},
        "roofline": roofline,
        "kernels": kernels,
        "batched": None if args.no_batched else batched_leg(args.steps, args.warmup, stream, torch),
        "adjoint": None if args.no_adjoint else adjoint_leg(torch),
        # the same dam break run on to step 120: the step slows as the column
        # piles into dense corners (per-lane imbalance), which the timed window
        # above is before
        "late_steps": None if args.no_late else step_curve(120, torch, window=20),
        "cpu_baseline": None if args.no_cpu else cpu_baseline(x, v, params),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
