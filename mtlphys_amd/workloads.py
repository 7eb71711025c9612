"""The benchmark workloads as parameter sets and seeded initial states.

dam_break(n)      1M-particle dam break (scene 4): h = 0.02, spacing h/2,
                  box 2.027 cbrt(n / 1e6) x same x half, 4 iterations
cube_drop(n)      jittered cube drop (scene 3): h = 0.1, a (cbrt n)^3 lattice in
                  [0.2 L, 0.6 L]^3 of the box [0, L]^3, 3 iterations
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
    """Compulsory bytes of one step: (100 + 44 I) N + (16 + 8 I) C."""
    return (100 + 44 * iters) * n + (16 + 8 * iters) * cells


__all__ = ["DT", "dam_break", "cube_drop", "ensemble_config", "make_ensemble", "step_bytes",
           "Engine", "Batch"]
