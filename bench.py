"""Benchmark of the PBF step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-batched] [--no-adjoint]
                    [--no-late] [--no-hash]

Workload: the 1M-particle dam break (h = 0.02, box 2.027 x 2.027 x 1.0135,
102 x 102 x 51 = 530,604 cells, 4 solver iterations, dt = 1/120), seeded
with PCG32 jitter; state stays resident in HBM for the whole timed region.

One JSON line on stdout:
  value         algorithmic HBM GB/s of the whole step, B_step / t_step with
                B_step = (100 + 44 I) N + (16 + 8 I) C bytes (the compulsory
                traffic of every stage: vec3 = 12 B, u32/f32 = 4 B), over the
                K timed steps after W warm-up steps (steps W+1 .. W+K)
  ms_per_step   wall time of those K steps (each replayed from the engine's
                HIP graph, mtlp_set_graph) between two device
                synchronisations / K
  ms_per_step_late, value_late
                the same workload over steps 81-120 (after the column hits
                the floor the step slows: dense corners)
  roofline      the dominant kernel: its algorithmic bytes per launch over its
                mean duration from HIP events recorded on the launch stream
                during the timed steps; traffic = HBM bytes per launch from
                rocprofv3 FETCH_SIZE / WRITE_SIZE (profiles/hbm_traffic.txt)
  roofline_valu the same kernel against the vector-ALU issue limit: VALU
                instructions per launch (rocprofv3 SQ_INSTS_VALU,
                profiles/valu_dam_break_1m.txt) over its mean duration, against
                1,024 SIMDs x 2.4 GHz / 4 cycles per wave64 instruction
  hash_scan     hash + count + scan + stable scatter + gather on 4,194,304
                points in a 128^3 grid (uniform, clustered, and 95% of them in
                200 cells of ~19,900: the ranking of huge cells), the scan
                alone at 2^20 .. 2^24 elements, and buildSpatialHash on the
                dam break
  parity        per-output errors against the FP64 oracle at this workload,
                as recorded by tests/test_gpu_parity_1m.py
                (profiles/parity_dam_break_1m.json)
  cpu_baseline  the oracle's step on this host's cores (OpenMP over
                particles): FP32 and FP64, one thread and all cores, on the
                32^3 cube drop and bounded samples of the 1M dam break
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
# wave64 VALU instructions per second: 256 CUs x 4 SIMDs, a wave64
# instruction every 4 cycles per SIMD, 2.4 GHz peak engine clock
PEAK_VALU_GIPS = 1024 * 2.4 / 4.0
N_PARTICLES = 1_000_000
SCENE_DAM = 4
DT = 1.0 / 120.0
LATE_FIRST, LATE_LAST = 81, 120


def stage_bytes(name, n, c):
    """Compulsory bytes of one launch of each stage (vec3 = 12 B)."""
    return {
        "predict_hash_count": 28 * n + 4 * c,    # R x, v (24N), W key (4N), counts (4C)
        "exclusive_scan": 8 * c,                  # R counts + W cell_start
        "stable_scatter_gather": 20 * n + 4 * c,  # R key, cell_start; W index, sorted x
        "density_lambda": 16 * n + 4 * c,         # R sorted x (12N), cell_start; W lambda
        "apply_delta": 28 * n + 4 * c,            # R sorted x + lambda (16N), cell_start; W x'
        "scatter_finalize": 52 * n,               # R sorted x, index, prev / x; W x, v
        "count_neighbors": 16 * n + 4 * c,
    }.get(name, 0)


def step_bytes(n, c, iters):
    return (100 + 44 * iters) * n + (16 + 8 * iters) * c


def step_bytes_stored(n, c, iters):
    return (152 + 44 * iters) * n + (16 + 8 * iters) * c


def read_profile(fname, kernel, workload="dam_break_1m"):
    """key=value fields of `kernel` `workload` in profiles/<fname>, or None."""
    path = os.path.join(ROOT, "profiles", fname)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 3 and parts[0] == kernel and parts[1] == workload:
                try:
                    return {k: float(v) for k, v in (p.split("=", 1) for p in parts[2:] if "=" in p)}
                except ValueError:
                    return None
    return None


def read_traffic(kernel):
    kv = read_profile("hbm_traffic.txt", kernel)
    if not kv or "fetch_bytes" not in kv or "write_bytes" not in kv:
        return None
    return kv["fetch_bytes"] + kv["write_bytes"]


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(x_dam, v_dam, p_dam):
    """The oracle's step on the host: the 32^3 cube drop (FP32 and FP64, all
    cores for 100 steps and one thread for 10, after 3 warm-up steps each)
    and the 1M dam break (FP32 and FP64 on all cores, 5 steps after a
    warm-up; FP64 on one thread, 3 steps after a warm-up).  Each run reports
    the p50 of its per-step times."""
    from mtlphys_amd import workloads as W
    from oracle import build as obuild
    from oracle import oracle as O
    if not os.path.exists(O.LIB_PATH):
        obuild.build(force=True)
    all_threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1

    def run(workload, params, x, v, prec, threads, warm, steps):
        O.set_threads(threads)
        op = O.params(params)
        xs, vs = x, v
        for _ in range(warm):
            xs, vs = O.step(xs, vs, op, DT, prec=prec)
        times = []
        for _ in range(steps):
            t0 = time.perf_counter()
            xs, vs = O.step(xs, vs, op, DT, prec=prec)
            times.append(time.perf_counter() - t0)
        g = O.grid(op)
        t = statistics.median(times)
        return {"workload": workload, "precision": f"f{prec}", "threads": O.get_threads(), "warmup": warm,
                "steps": steps, "ms_per_step_p50": round(t * 1e3, 2),
                "gbs": round(step_bytes(len(x), g.total_cells, op.iterations) / t / 1e9, 3)}

    pc, xc, vc = W.cube_drop(32_768)
    runs = []
    for prec in (32, 64):
        runs.append(run("cube_drop_32k", pc, xc, vc, prec, all_threads, 3, 100))
        runs.append(run("cube_drop_32k", pc, xc, vc, prec, 1, 3, 10))
    for prec in (32, 64):
        runs.append(run("dam_break_1m", p_dam, x_dam, v_dam, prec, all_threads, 1, 5))
    runs.append(run("dam_break_1m", p_dam, x_dam, v_dam, 64, 1, 1, 3))
    head = runs[4]  # 1M dam break, FP32, all cores
    return {"value": head["gbs"], "unit": "GB/s", "cores": head["threads"], "kind": "port",
            "ms_per_step": head["ms_per_step_p50"],
            "sample": "5 FP32 oracle steps of the same 1M dam break after 1 warm-up (p50), OpenMP over "
                      "particles, serial scan/scatter; more runs in `runs`",
            "host": {"nproc": os.cpu_count(), "cpu_model": cpu_model()},
            "runs": runs}


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
    _, pull, packed = b.launch_config()
    b.set_graph(True)
    for _ in range(warmup):
        b.step(DT, stream)
    t_batch = _time_steps(b, steps, stream, torch)
    # one more step, eager, with per-kernel events: us per stage
    b.set_graph(False)
    b.set_profiling(True)
    b.step(DT, stream)
    stage_us = {}
    for name, ms in b.stage_times():
        stage_us[name] = round(stage_us.get(name, 0.0) + ms * 1e3, 1)
    b.set_profiling(False)
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
    return {"scenes": 256, "particles": n, "iterations": iters, "stencil_block": block, "tile_pull": pull, "packed_positions": packed,
            "ms_per_step": round(t_batch * 1e3, 4), "value": round(gbs_b, 2), "unit": "GB/s",
            "single_scene_ms_per_step": round(t_single * 1e3, 4), "single_scene_gbs": round(gbs_s, 2),
            "ratio_to_single_scene": round(gbs_b / gbs_s, 3), "error_flags": flags,
            "next_step_stage_us": stage_us,
            "note": "one launch per stage for all scenes, HIP-graph replay; the single scene is one "
                    "cube drop with the same total particle count; GB/s of the compulsory bytes"}


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
            "finite": finite, "error_flags": flags,
            "gradient_check": "tests/test_gpu_adjoint.py::test_differentiable_workload_full_length "
                              "(FP64 adjoint on the GPU's tapes, all 40 steps)"}


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
    eng.set_graph(True)
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
    eng.set_graph(False)
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


def _hash_stages(eng, reps, stream, warm=3):
    """Mean us per stage of `reps` buildSpatialHash calls (HIP events)."""
    for _ in range(warm):
        eng.build_spatial_hash(stream)
    eng.set_profiling(True)
    per = {}
    for _ in range(reps):
        eng.build_spatial_hash(stream)
        for name, ms in eng.stage_times():
            per.setdefault(name, []).append(ms * 1e3)
    eng.set_profiling(False)
    return {k: statistics.fmean(v) for k, v in per.items()}


def _hash_step_stages(eng, reps, stream, warm=3):
    """Mean us per stage of `reps` steps with no solver iteration (the hash
    of the current positions without countNeighbors) (HIP events)."""
    for _ in range(warm):
        eng.step(DT, stream)
    eng.set_profiling(True)
    per = {}
    for _ in range(reps):
        eng.step(DT, stream)
        for name, ms in eng.stage_times():
            per.setdefault(name, []).append(ms * 1e3)
    eng.set_profiling(False)
    return {k: statistics.fmean(v) for k, v in per.items()}


def hash_scan_leg(torch, dam_state=None, reps=20):
    """The isolated spatial hash (4,194,304 points, 128^3 cells, uniform and
    clustered), the scan alone at 2^20 .. 2^24 elements, and buildSpatialHash
    (with countNeighbors) on the 1M dam break."""
    from mtlphys_amd import workloads as W
    import mtlphys_amd as mtlp
    stream = torch.cuda.current_stream()
    out = {"hash_128cubed": {}, "scan_only": [], "dam_break_1m_build_spatial_hash": None}
    p = W.unit_grid_params()
    hash_stages = ("predict_hash_count", "exclusive_scan", "stable_scatter_gather")
    gen = {"uniform": W.uniform_points, "clustered": W.clustered_points, "huge_cells": W.huge_cell_points}
    for variant in ("uniform", "clustered", "huge_cells"):
        pts = gen[variant](W.HASH_POINTS, 2)
        q = p.copy()
        if variant == "huge_cells":  # no countNeighbors: its walk is quadratic in a cell's size
            q.iterations = 0
            for ax in range(3):
                q.gravity[ax] = 0.0
        eng = mtlp.Engine(0, q)
        eng.reset(W.HASH_POINTS, mtlp.capi.SCENE_EMPTY)
        eng.upload_state(pts, np.zeros_like(pts))
        st = _hash_stages(eng, reps, stream) if variant != "huge_cells" else _hash_step_stages(eng, reps, stream)
        flags = eng.error_flags()
        cells = eng.total_cells
        eng.close()
        t_us = sum(st[k] for k in hash_stages)
        b = W.hash_bytes(W.HASH_POINTS, cells)
        out["hash_128cubed"][variant] = {
            "points": W.HASH_POINTS, "cells": cells, "us": round(t_us, 2), "alg_bytes": b,
            "gbs": round(b / (t_us * 1e-6) / 1e9, 1), "frac": round(b / (t_us * 1e-6) / 1e9 / PEAK_HBM_GBS, 4),
            "stage_us": {k: round(v, 2) for k, v in st.items()},
            "count_neighbors_us": round(st.get("count_neighbors", 0.0), 2), "error_flags": flags,
            "note": "us = hash + count + scan + stable scatter + gather (28 N + 16 C bytes); countNeighbors "
                    "separately"}
    lib = mtlp.capi.lib()
    for k in range(20, 25):
        n = 1 << k
        g = torch.Generator(device="cuda")
        g.manual_seed(k)
        d_in = torch.randint(0, 8, (n,), dtype=torch.int32, device="cuda", generator=g)
        d_out = torch.empty_like(d_in)
        scratch = torch.empty(mtlp.Engine.scan_scratch_bytes(n), dtype=torch.uint8, device="cuda")
        s = stream.cuda_stream

        def alloc_scan():
            mtlp.capi.check(lib.mtlp_exclusive_scan(None, s, d_in.data_ptr(), d_out.data_ptr(),
                                                    scratch.data_ptr(), n), "mtlp_exclusive_scan")

        def noalloc_scan():
            mtlp.capi.check(lib.mtlp_exclusive_scan_noalloc(s, d_in.data_ptr(), d_out.data_ptr(), n),
                            "mtlp_exclusive_scan_noalloc")

        res = {"elements": n, "alg_bytes": 8 * n}
        for name, fn in (("us", alloc_scan), ("noalloc_us", noalloc_scan)):
            for _ in range(5):
                fn()
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ev0.record(stream)
            for _ in range(reps):
                fn()
            ev1.record(stream)
            ev1.synchronize()
            res[name] = round(ev0.elapsed_time(ev1) * 1e3 / reps, 2)
        ref = torch.cumsum(d_in, 0) - d_in
        res["exact"] = bool(torch.equal(d_out, ref.to(torch.int32)))
        res["gbs"] = round(8 * n / (res["us"] * 1e-6) / 1e9, 1)
        res["note"] = "per call: the scratch-clearing fill kernel + the single-pass scan"
        out["scan_only"].append(res)
        del d_in, d_out, scratch, ref
    if dam_state is not None:
        pd, xd, vd = dam_state
        eng = mtlp.Engine(0, pd)
        eng.reset(len(xd), mtlp.capi.SCENE_EMPTY)
        eng.upload_state(xd, vd)
        st = _hash_stages(eng, reps, stream)
        cells = eng.total_cells
        eng.close()
        total = sum(st.values())
        t_us = sum(st[k] for k in hash_stages)
        b = W.hash_bytes(len(xd), cells)
        out["dam_break_1m_build_spatial_hash"] = {
            "points": len(xd), "cells": cells, "us_with_count_neighbors": round(total, 2), "us": round(t_us, 2),
            "gbs": round(b / (t_us * 1e-6) / 1e9, 1), "stage_us": {k: round(v, 2) for k, v in st.items()},
            "state": "after the benchmark's warm-up steps",
            "reference_m4_pro_ms": 2.4,
            "note": "reference figure (README.md:55) is hash + count + scan + scatter + gather + neighbour count "
                    "on Apple M4 Pro, context only"}
    torch.cuda.synchronize()
    return out


def cube_drop_gpu_leg(torch):
    """The 32^3 cube drop (h = 0.1, 3 iterations) for 100 steps on the GPU."""
    from mtlphys_amd import workloads as W
    import mtlphys_amd as mtlp
    p, x, v = W.cube_drop(32_768)
    eng = mtlp.Engine(0, p)
    eng.reset(len(x), mtlp.capi.SCENE_EMPTY)
    eng.upload_state(x, v)
    stream = torch.cuda.current_stream()
    for _ in range(3):
        eng.step(DT, stream)
    eng.upload_state(x, v)
    t = _time_steps(eng, 100, stream, torch)
    cells = eng.total_cells
    eng.close()
    return {"workload": "cube_drop_32k", "particles": len(x), "cells": cells, "iterations": p.iterations,
            "steps": 100, "ms_per_step": round(t * 1e3, 4),
            "gbs": round(step_bytes(len(x), cells, p.iterations) / t / 1e9, 2)}


def parity_record():
    """Per-output errors at the 1M dam break from the committed record of
    tests/test_gpu_parity_1m.py, worst over iterations."""
    path = os.path.join(ROOT, "profiles", "parity_dam_break_1m.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    out = {"source": "profiles/parity_dam_break_1m.json (tests/test_gpu_parity_1m.py, FP64 oracle)"}
    for case, c in sorted(data.get("cases", {}).items()):
        its = [r for r in c["kernel_level"] if r["stage"] == "iteration"]
        fin = [r for r in c["kernel_level"] if r["stage"] == "finalize"][0]
        out[case] = {
            "kernel_level": {
                "lambda_rel_floor_max": max(r["lambda_rel_floor_max"] for r in its),
                "lambda_rel_floor_p99": max(r["lambda_rel_floor_p99"] for r in its),
                "x_over_h_max": max(r["x_over_h_max"] for r in its + [fin]),
                "x_over_h_p99": max(r["x_over_h_p99"] for r in its + [fin]),
                "v_over_h_dt_max": fin["v_over_h_dt_max"], "v_over_h_dt_p99": fin["v_over_h_dt_p99"],
                "iterations_checked": len(its)},
            "free_running": c["free_running"],
            "slots_over_96_neighbours": c.get("slots_over_96_neighbours")}
    return out


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
    ap.add_argument("--no-hash", action="store_true", help="skip the hash / scan legs")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only run the forward+adjoint leg and print its JSON (for profiling)")
    ap.add_argument("--hash-only", action="store_true",
                    help="only run the hash / scan legs and print their JSON (for profiling)")
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
    if args.hash_only:
        print(json.dumps(hash_scan_leg(torch, reps=5)), flush=True)
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
    _, pull, packed = eng.launch_config()
    warm_state = (params, eng.positions_buffer().read(), eng.velocities_buffer().read())

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
    # host waits between steps, each replayed from the engine's HIP graph
    # (captured during the warm-up)
    eng.upload_state(x, v)
    eng.set_graph(True)
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
    valu = read_profile("valu_dam_break_1m.txt", dom)
    roofline_valu = None
    if valu and "valu_insts" in valu:
        gips = valu["valu_insts"] / (dk["mean_us"] * 1e-6) / 1e9
        roofline_valu = {"bound": "valu", "kernel": dom, "achieved": round(gips, 1), "peak": PEAK_VALU_GIPS,
                         "unit": "G wave64 VALU instructions/s", "frac": round(gips / PEAK_VALU_GIPS, 4),
                         "valu_insts_per_launch": valu["valu_insts"],
                         "issue_floor_us": round(valu["valu_insts"] / (PEAK_VALU_GIPS * 1e3), 1),
                         "source": "profiles/valu_dam_break_1m.txt (rocprofv3 --pmc SQ_INSTS_VALU)"}

    late = None if args.no_late else step_curve(LATE_LAST, torch, window=20)
    late_ms = None
    if late:
        w = late["window"]
        wins = [late["ms_per_step"][k] for k in range(len(late["ms_per_step"]))
                if k * w + 1 >= LATE_FIRST and (k + 1) * w <= LATE_LAST]
        late_ms = statistics.fmean(wins) if wins else None

    out = {
        "metric": "PBF step ms + HBM GB/s at 1M particles (dam break, 4 iterations)",
        "value": round(b_step / t_step / 1e9, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_step * 1e3, 4),
        "timed_steps": [args.warmup + 1, args.warmup + args.steps],
        "ms_per_step_late": round(late_ms, 4) if late_ms else None,
        "value_late": round(b_step / (late_ms * 1e-3) / 1e9, 2) if late_ms else None,
        "late_steps_window": [LATE_FIRST, LATE_LAST],
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
                   "tile_pull": pull, "packed_positions": packed, "alg_bytes_per_step": b_step,
                   "stored_bytes_per_step": step_bytes_stored(N_PARTICLES, cells, iters)},
        "roofline": roofline,
        "roofline_valu": roofline_valu,
        "kernels": kernels,
        "parity": parity_record(),
        "hash_scan": None if args.no_hash else hash_scan_leg(torch, warm_state),
        "cube_drop_gpu": cube_drop_gpu_leg(torch),
        "batched": None if args.no_batched else batched_leg(args.steps, args.warmup, stream, torch),
        "adjoint": None if args.no_adjoint else adjoint_leg(torch),
        # the same dam break run on to step 120: the step slows as the column
        # piles into dense corners (per-lane imbalance)
        "late_steps": late,
        "cpu_baseline": None if args.no_cpu else cpu_baseline(x, v, params),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
