"""Parity at the headline workload: the 1M-particle dam break (h = 0.02,
102 x 102 x 51 = 530,604 cells, 4 iterations, dt = 1/120) against the FP64
oracle, from two states:

  initial         the seeded column (PCG32 jitter, seed 1);
  after 80 steps  the state the GPU reaches after 80 steps from it: the
                  column has hit the floor, tens of thousands of slots have
                  more than 96 neighbours, and the heavy-slot queues, the
                  measured tile order and the chunk-scratch walks all run.

Kernel level: every iteration's density/lambda and delta kernels and the
finalize, each fed the GPU's own inputs (Constraints.metal:82-199, 305-353;
orchestration Engine.cpp:729-796), within
    lambda  1e-4 of max(|lambda_ref|, 1e-3 max|lambda_ref|)
    x       1e-5 h (every component)
    v       1e-5 h / dt (every component).
Free running: one whole step from the same state on the GPU and on the FP64
oracle, next to the oracle's own FP32 instantiation, with the stencil-centre
flips (a slot whose current cell differs between the GPU's and the FP64
iterate) counted per iteration.  The GPU's free-running error is bounded by 4x
the FP32 oracle's plus 1e-5 h (FP32 position stores are amplified by the
Jacobi iterations; no FP32 iterate can hold 1e-5 h after four of them).

Each output's error (max and p99) is printed and, with MTLP_PARITY_RECORD
set to a path, written there as JSON (one entry per state).
"""
import numpy as np
import pytest

from helpers import (LAMBDA_TOL, X_TOL_H, assert_kernel_level, kernel_level_errors, lambda_errors,
                     position_error_h, record_parity, velocity_error)

pytestmark = pytest.mark.gpu

N = 1_000_000
DT = 1.0 / 120.0
STATES = {"initial": 0, "after_80_steps": 80}


def _state(mtlp, steps):
    p = mtlp.scene_params(4, N)
    assert mtlp.grid_dims(p) == (102, 102, 51)
    x, v = mtlp.generate_scene(4, N, 1, p)
    if steps:
        with mtlp.Engine(0, p) as e:
            e.reset(N, mtlp.capi.SCENE_EMPTY)
            e.upload_state(x, v)
            for _ in range(steps):
                e.step(DT)
            x, v = e.positions_buffer().read(), e.velocities_buffer().read()
            assert e.error_flags() == 0
        assert np.isfinite(x).all() and np.isfinite(v).all()
    return p, x, v


def _cells(sx, op):
    o = np.asarray(op.bounds_min, np.float64)
    dims = np.array([102, 102, 51])
    c = np.floor((np.asarray(sx, np.float64)[:, :3] - o) * (1.0 / op.cell_size)).astype(np.int64)
    return np.clip(c, 0, dims - 1)


@pytest.mark.parametrize("state", list(STATES))
def test_dam_break_1m_all_iterations(mtlp, orc, state):
    steps = STATES[state]
    print(f"dam break 1M parity, state {state}: {steps} GPU steps first", flush=True)
    p, x, v = _state(mtlp, steps)
    op = orc.params(p)
    h = op.h

    report, runs = kernel_level_errors(mtlp, orc, p, x, v, DT)
    nbr = runs[1]["nbr"]
    for r in report:
        print("kernel level", r, flush=True)

    # free running: GPU full step vs FP64 oracle, and the FP32 oracle beside it
    full = runs[-1]
    x64, v64, t64 = orc.step(x, v, op, DT, prec=64, trace=True)
    x32, v32, t32 = orc.step(x, v, op, DT, prec=32, trace=True)
    flips = [int((_cells(runs[k]["sorted"], op) != _cells(t64["sx_iter"][k], op)).any(1).sum())
             for k in range(p.iterations)]
    flips32 = [int((_cells(t32["sx_iter"][k], op) != _cells(t64["sx_iter"][k], op)).any(1).sum())
               for k in range(p.iterations)]

    def free(xg, vg, lg):
        xe = position_error_h(xg, x64, h)
        ve = velocity_error(vg, v64, h, DT)
        le = lambda_errors(lg, t64["lambda_orig"])
        return dict(x_over_h_max=xe["max"], x_over_h_p99=xe["p99"], x_over_h_p50=xe["p50"],
                    v_over_h_dt_max=ve["max"], v_over_h_dt_p99=ve["p99"],
                    lambda_rel_floor_max=le["rel_floor"], lambda_rel_floor_p99=le["p99_floor"],
                    lambda_rel_max=le["rel_max"])

    fr_gpu = free(full["x"], full["v"], full["lam"])
    fr_f32 = free(x32, v32, t32["lambda_orig"])
    print("free running gpu", fr_gpu, flush=True)
    print("free running oracle fp32", fr_f32, flush=True)
    print("stencil-centre flips gpu", flips, "oracle fp32", flips32, flush=True)

    record_parity(state, dict(
        steps_before=steps, particles=N, iterations=p.iterations, dt=DT,
        slots_over_96_neighbours=int((nbr > 96).sum()), max_neighbours=int(nbr.max()),
        tolerances=dict(lambda_rel_floor=LAMBDA_TOL, x_over_h=X_TOL_H, v_over_h_dt=1e-5),
        kernel_level=report,
        free_running=dict(gpu=fr_gpu, oracle_fp32=fr_f32, stencil_centre_flips_gpu=flips,
                          stencil_centre_flips_oracle_fp32=flips32)))

    assert_kernel_level(report)
    assert fr_gpu["x_over_h_max"] <= 4 * fr_f32["x_over_h_max"] + X_TOL_H, (fr_gpu, fr_f32)
    assert fr_gpu["lambda_rel_max"] <= max(LAMBDA_TOL, 4 * fr_f32["lambda_rel_max"]), (fr_gpu, fr_f32)
    if steps:
        # the post-impact state exercises the dense-slot paths
        assert (nbr > 96).sum() > 1000, int((nbr > 96).sum())
