"""Shared helpers of the parity tests (comparison metrics, GPU runs and the
kernel-level check against the FP64 oracle)."""
import json
import os

import numpy as np

LAMBDA_TOL = 1e-4   # |dl| / max(|lambda_ref|, 1e-3 max|lambda_ref|)
X_TOL_H = 1e-5      # |dx| / h, every component
V_TOL = 1e-5        # |dv| / (h / dt), every component


def lambda_errors(lam_gpu, lam_ref):
    """Relative lambda error two ways: against max|lambda_ref| and against
    max(|lambda_ref|, 1e-3 max|lambda_ref|) (lambda ~ 0 near rest)."""
    lam_ref = np.asarray(lam_ref, np.float64)
    d = np.abs(np.asarray(lam_gpu, np.float64) - lam_ref)
    mx = max(np.abs(lam_ref).max(), 1e-30)
    floor = np.maximum(np.abs(lam_ref), 1e-3 * mx)
    return dict(rel_max=float(d.max() / mx), rel_floor=float((d / floor).max()),
                p99_floor=float(np.percentile(d / floor, 99)))


def position_error_h(x_gpu, x_ref, h):
    d = np.abs(np.asarray(x_gpu, np.float64)[:, :3] - np.asarray(x_ref, np.float64)[:, :3]).max(1) / h
    return dict(max=float(d.max()), p99=float(np.percentile(d, 99)), p50=float(np.median(d)))


def velocity_error(v_gpu, v_ref, h, dt):
    """Largest component error per particle in units of h / dt."""
    d = np.abs(np.asarray(v_gpu, np.float64)[:, :3] - np.asarray(v_ref, np.float64)[:, :3]).max(1) / (h / dt)
    return dict(max=float(d.max()), p99=float(np.percentile(d, 99)), p50=float(np.median(d)))


def run_gpu_step(mtlp, params, x, v, dt, iterations=None):
    """One mtlp_step from (x, v); returns every buffer the step leaves."""
    p = params.copy()
    if iterations is not None:
        p.iterations = iterations
    with mtlp.Engine(0, p) as e:
        e.reset(len(x), mtlp.capi.SCENE_EMPTY)
        e.upload_state(x, v)
        e.step(dt)
        out = dict(
            x=e.positions_buffer().read(), v=e.velocities_buffer().read(),
            prev=e.prev_positions_buffer().read(), keys=e.cell_hashes_buffer().read(),
            cell_start=e.cell_start_buffer().read(), sorted_index=e.sorted_index_buffer().read(),
            nbr=e.neighbor_counts_buffer().read(), lam=e.lambdas_buffer().read(),
            sorted=e.sorted_positions_buffer().read())
        flags = e.error_flags()
    assert flags == 0, f"device error flags {flags:#x}"
    return out


def kernel_level_errors(mtlp, orc, p, x, v, dt):
    """Runs the GPU step with 0..I iterations from (x, v) and feeds each
    iteration's GPU input to the FP64 oracle's density/lambda and delta
    kernels, and the last iterate to its finalize.  Returns one record per
    kernel: lambda (rel_floor max / p99) and positions (x/h max / p99) per
    iteration, then finalize's positions and velocities (v in h/dt).  Also
    returns the GPU runs."""
    op = orc.params(p)
    g = orc.grid(op)
    h = op.h
    runs = [run_gpu_step(mtlp, p, x, v, dt, iterations=k) for k in range(0, p.iterations + 1)]
    cs, sidx = runs[0]["cell_start"], runs[0]["sorted_index"]
    xs = orc.predict(x, v, op, dt)
    np.testing.assert_array_equal(runs[0]["sorted"].view(np.uint32)[:, :3], xs[sidx].view(np.uint32)[:, :3])
    report = []
    for k in range(p.iterations):
        inp = runs[k]["sorted"].astype(np.float64)
        _, lam_o, _ = orc.density_lambda(inp, cs, sidx, g, op, 64)
        le = lambda_errors(runs[k + 1]["lam"], lam_o)
        nxt = orc.apply_delta(inp, runs[k + 1]["lam"][sidx].astype(np.float64), cs, g, op, 64)
        xe = position_error_h(runs[k + 1]["sorted"], nxt, h)
        report.append(dict(stage="iteration", iteration=k, lambda_rel_floor_max=le["rel_floor"],
                           lambda_rel_floor_p99=le["p99_floor"], x_over_h_max=xe["max"], x_over_h_p99=xe["p99"]))
    final = runs[-1]
    xo = np.empty((len(x), 4), np.float64)
    xo[sidx] = final["sorted"].astype(np.float64)
    xr, vr = orc.finalize(xo, v, x, op, dt, 64)
    xe = position_error_h(final["x"], xr, h)
    ve = velocity_error(final["v"], vr, h, dt)
    report.append(dict(stage="finalize", x_over_h_max=xe["max"], x_over_h_p99=xe["p99"],
                       v_over_h_dt_max=ve["max"], v_over_h_dt_p99=ve["p99"]))
    return report, runs


def assert_kernel_level(report):
    for r in report:
        if r["stage"] == "iteration":
            assert r["lambda_rel_floor_max"] <= LAMBDA_TOL, r
            assert r["x_over_h_max"] <= X_TOL_H, r
        else:
            assert r["x_over_h_max"] <= X_TOL_H and r["v_over_h_dt_max"] <= V_TOL, r


def record_parity(case, entry):
    """Adds `entry` under cases[case] of the JSON file named by the
    MTLP_PARITY_RECORD environment variable (nothing without it)."""
    path = os.environ.get("MTLP_PARITY_RECORD")
    if not path:
        return
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    data.setdefault("cases", {})[case] = entry
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
