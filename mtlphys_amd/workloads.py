"""The benchmark workloads as parameter sets and seeded initial states.

dam_break(n)      1M-particle dam break (scene 4): h = 0.02, spacing h/2,
                  box 2.027 cbrt(n / 1e6) x same x half, 4 iterations
cube_drop(n)      jittered cube drop (scene 3): h = 0.1, a (cbrt n)^3 lattice in
                  [0.2 L, 0.6 L]^3 of the box [0, L]^3, 3 iterations
differentiable(n) the adjoint workload: the dam break at h = 0.04 (spacing
                  0.02) in a box of 2 x 2.027 cbrt(n / 1e6) (1.634 at 65,536
                  particles, grid 41 x 41 x 21) with the reference engine's
                  constants (rest density 8000 (0.1/h)^3, epsilon 800, s_corr
                  k 1e-4), 3 iterations, 40 steps with checkpoints every 8; loss
                  |centroid(x_T) - target|^2 + 0.1 mean((rho/rho0 - 1)^2) with
                  target = initial centroid + (0.2, 0, 0)
unit_grid_params, uniform_points, clustered_points, huge_cell_points
                  the isolated spatial hash: 4,194,304 points in the unit
                  box's 128^3 grid, uniform, with 90% of them in 5% of the
                  cells, or (a stress case) 95% of them in 200 cells
ensemble(...)     independent cube drops with N_s = round(exp(U[ln 2048,
                  ln 65536])), rest density U[5000, 7000] (times (0.1/h)^3),
                  gravity U[-12, -8], 3 iterations, drawn from numpy's PCG64
                  with the given seed; scene s is seeded with seed + s
"""
import numpy as np

from . import _capi as C
from .engine import Batch, Engine, generate_scene

DT = 1.0 / 120.0


def dam_break(n=1_000_000, seed=1):
    p = C.scene_params(C.SCENE_JITTER_DAM, n)
    x, v = generate_scene(C.SCENE_JITTER_DAM, n, seed, p)
    return p, x, v


def cube_drop(n=32_768, seed=1):
    p = C.scene_params(C.SCENE_JITTER_CUBE, n)
    x, v = generate_scene(C.SCENE_JITTER_CUBE, n, seed, p)
    return p, x, v


ADJ_STEPS = 40
ADJ_CKPT = 8
ADJ_SHIFT = (0.2, 0.0, 0.0)


def differentiable(n=65_536, seed=1):
    """(params, x, v, target) of the adjoint workload.  Constants are the
    reference engine's own (Engine.cpp:23-39: rest density 8000 at h = 0.1,
    scaled by (0.1/h)^3 for the same particle mass ratio; epsilon 800;
    s_corr k 1e-4; g = -9.81).  With the benchmark constants (s_corr k 0.1)
    the rollout is so chaotic that dL/dv0 grows ~1e8 per 5 steps and
    overflows FP32 long before 40 steps."""
    p = C.scene_params(C.SCENE_JITTER_DAM, n)
    d = C.default_params()
    h = 0.04
    S = 2.0 * 2.027 * float(np.cbrt(n / 1.0e6))
    p.h = h
    p.cell_size = h
    p.bounds_max[0] = S
    p.bounds_max[1] = S
    p.bounds_max[2] = 0.5 * S
    p.rest_density = d.rest_density * (0.1 / h) ** 3
    p.epsilon = d.epsilon
    p.scorr_k = d.scorr_k
    p.gravity[1] = d.gravity[1]
    p.iterations = 3
    x, v = generate_scene(C.SCENE_JITTER_DAM, n, seed, p)
    target = x[:, :3].astype(np.float64).mean(axis=0) + np.asarray(ADJ_SHIFT)
    return p, x, v, target


def ensemble_config(n_scenes=256, seed=3, n_min=2048, n_max=65536):
    """Per-scene particle counts and parameters of the batched workload."""
    rng = np.random.default_rng(seed)
    counts = np.rint(np.exp(rng.uniform(np.log(n_min), np.log(n_max), n_scenes))).astype(np.int64)
    rho0 = rng.uniform(5000.0, 7000.0, n_scenes)
    g = rng.uniform(-12.0, -8.0, n_scenes)
    params = []
    for s in range(n_scenes):
        p = C.scene_params(C.SCENE_JITTER_CUBE, int(counts[s]))
        p.rest_density = float(rho0[s])
        p.gravity[1] = float(g[s])
        params.append(p)
    return [int(c) for c in counts], params


def make_ensemble(n_scenes=256, seed=3, device=0):
    """A seeded Batch of the ensemble workload."""
    counts, params = ensemble_config(n_scenes, seed)
    b = Batch(counts, params, device)
    b.seed(C.SCENE_JITTER_CUBE, seed)
    return b


def step_bytes(n, cells, iters):
    """Compulsory bytes of one step, B_step = (100 + 44 I) N + (16 + 8 I) C
    (a vec3 counts 12 bytes, a u32 / f32 4): predict + hash + count 28 N +
    4 C, scan 8 C, stable scatter + gather 20 N + 4 C, per iteration the
    density pass 16 N + 4 C and the delta pass 28 N + 4 C, scatter back +
    finalize 52 N."""
    return (100 + 44 * iters) * n + (16 + 8 * iters) * cells


def step_bytes_stored(n, cells, iters):
    """Bytes of one step counting what this implementation stores as well:
    float4 (16-byte) positions and velocities, the slot map, foam and lambda
    outputs: (152 + 44 I) N + (16 + 8 I) C.  Reported as an extra, never as
    the headline."""
    return (152 + 44 * iters) * n + (16 + 8 * iters) * cells


# ---- the isolated spatial hash: 4M points in a 128^3 grid ------------------

HASH_POINTS = 4 * 1024 * 1024
HASH_CELLS = 128


def unit_grid_params(cells=HASH_CELLS):
    """The unit box [0, 1]^3 with cell = h = 1/cells (1/128 is exact)."""
    p = C.default_params()
    for a in range(3):
        p.bounds_min[a] = 0.0
        p.bounds_max[a] = 1.0
    p.cell_size = 1.0 / cells
    p.h = 1.0 / cells
    return p


def uniform_points(n=HASH_POINTS, seed=2):
    """n points U[0, 1)^3 (FP32, w = 1)."""
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.random((n, 3), dtype=np.float32), np.ones((n, 1), np.float32)], 1)


def clustered_points(n=HASH_POINTS, seed=2, cells=HASH_CELLS, frac_cells=0.05, frac_pts=0.9):
    """n points of which frac_pts fall in frac_cells of the cells (chosen by
    a seeded shuffle), the rest uniform; FP32, w = 1, in shuffled order."""
    rng = np.random.default_rng(seed)
    total = cells ** 3
    hot = rng.permutation(total)[: int(round(frac_cells * total))]
    n_hot = int(round(frac_pts * n))
    pick = hot[rng.integers(0, len(hot), n_hot)]
    cx, cy, cz = pick % cells, (pick // cells) % cells, pick // (cells * cells)
    off = rng.random((n_hot, 3), dtype=np.float32)
    hot_pts = (np.stack([cx, cy, cz], 1) + off * np.float32(0.999)) / np.float32(cells)
    cold_pts = rng.random((n - n_hot, 3), dtype=np.float32)
    pts = np.concatenate([hot_pts.astype(np.float32), cold_pts])[rng.permutation(n)]
    return np.concatenate([pts, np.ones((n, 1), np.float32)], 1).astype(np.float32)


def huge_cell_points(n=HASH_POINTS, seed=2, cells=HASH_CELLS, n_hot=200, frac_pts=0.95):
    """n points of which frac_pts fall in n_hot cells (about 19,900 each at
    the defaults: every one above the 16,384 the LDS sort of a cell holds),
    the rest uniform; FP32, w = 1, in shuffled order.  A stress case for the
    hash's ranking of huge cells."""
    rng = np.random.default_rng(seed)
    hot = rng.permutation(cells ** 3)[:n_hot]
    k = int(round(frac_pts * n))
    pick = hot[rng.integers(0, n_hot, k)]
    cx, cy, cz = pick % cells, (pick // cells) % cells, pick // (cells * cells)
    off = rng.random((k, 3), dtype=np.float32)
    hot_pts = (np.stack([cx, cy, cz], 1) + off * np.float32(0.999)) / np.float32(cells)
    cold_pts = rng.random((n - k, 3), dtype=np.float32)
    pts = np.concatenate([hot_pts.astype(np.float32), cold_pts])[rng.permutation(n)]
    return np.concatenate([pts, np.ones((n, 1), np.float32)], 1).astype(np.float32)


def hash_bytes(n, cells):
    """Compulsory bytes of hash + count + scan + stable scatter + gather:
    28 N + 16 C."""
    return 28 * n + 16 * cells


__all__ = ["DT", "ADJ_STEPS", "ADJ_CKPT", "ADJ_SHIFT", "differentiable", "dam_break", "cube_drop", "ensemble_config",
           "make_ensemble", "step_bytes", "step_bytes_stored", "HASH_POINTS", "HASH_CELLS", "unit_grid_params",
           "uniform_points", "clustered_points", "huge_cell_points", "hash_bytes", "Engine", "Batch"]
