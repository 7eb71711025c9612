"""Shared helpers of the parity tests (comparison metrics and GPU runs)."""
import numpy as np


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
