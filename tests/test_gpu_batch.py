"""GPU: batched ensembles, HIP-graph replay and stencil workgroup sizes.

A batch runs every stage as one launch over all scenes; each scene must come
out bit-identical to the same scene stepped alone (same kernels, same
per-slot summation order), and its integer stages must equal the oracle's.
Graph replays and every stencil workgroup size must reproduce eager steps
bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DT = 1.0 / 120.0


def _mixed_scenes(mtlp):
    counts = [300, 2048, 0, 5000, 1, 777]
    params = []
    for k, c in enumerate(counts):
        p = mtlp.scene_params(3, max(c, 8))
        p.rest_density = 5000.0 + 300.0 * k
        p.gravity[1] = -8.0 - 0.7 * k
        params.append(p)
    return counts, params


def _read_all(e):
    return dict(x=e.positions_buffer().read(), v=e.velocities_buffer().read(),
                prev=e.prev_positions_buffer().read(), keys=e.cell_hashes_buffer().read(),
                cs=e.cell_start_buffer().read(), sidx=e.sorted_index_buffer().read(),
                nbr=e.neighbor_counts_buffer().read(), lam=e.lambdas_buffer().read(),
                sorted=e.sorted_positions_buffer().read())


def test_batch_matches_scenes_stepped_alone(mtlp, orc):
    counts, params = _mixed_scenes(mtlp)
    b = mtlp.Batch(counts, params)
    b.seed(3, 11)
    po, co = b.offsets()
    assert po.tolist() == list(np.concatenate([[0], np.cumsum(counts)]))
    for _ in range(3):
        b.step(DT)
    got = _read_all(b)
    assert b.error_flags() == 0
    assert got["cs"][-1] == sum(counts)
    for s, (c, p) in enumerate(zip(counts, params)):
        if c == 0:
            assert po[s] == po[s + 1]
            continue
        x0, v0 = mtlp.generate_scene(3, c, 11 + s, p)
        with mtlp.Engine(0, p) as e:
            e.reset(c, mtlp.capi.SCENE_EMPTY)
            e.upload_state(x0, v0)
            for _ in range(3):
                e.step(DT)
            ref = _read_all(e)
        sl = slice(po[s], po[s + 1])
        for k in ("x", "v", "prev", "nbr", "lam"):
            np.testing.assert_array_equal(got[k][sl].view(np.uint32), ref[k].view(np.uint32), err_msg=(s, k))
        np.testing.assert_array_equal(got["keys"][sl], ref["keys"] + co[s])
        np.testing.assert_array_equal(got["sidx"][sl], ref["sidx"] + po[s])
        np.testing.assert_array_equal(got["sorted"][sl].view(np.uint32), ref["sorted"].view(np.uint32))
        np.testing.assert_array_equal(got["cs"][co[s]:co[s + 1]], ref["cs"][:-1] + po[s])


def test_batch_integer_stages_vs_oracle(mtlp, orc):
    counts, params = _mixed_scenes(mtlp)
    b = mtlp.Batch(counts, params)
    b.seed(3, 5)
    x0 = b.positions_buffer().read()
    v0 = b.velocities_buffer().read()
    b.step(DT)
    got = _read_all(b)
    po, co = b.offsets()
    for s, (c, p) in enumerate(zip(counts, params)):
        if c == 0:
            continue
        sl = slice(po[s], po[s + 1])
        op = orc.params(p)
        _, _, tr = orc.step(x0[sl], v0[sl], op, DT, prec=32, trace=True)
        np.testing.assert_array_equal(got["keys"][sl], tr["keys"] + co[s])
        np.testing.assert_array_equal(got["sidx"][sl], tr["sorted_index"] + po[s])
        np.testing.assert_array_equal(got["nbr"][sl], tr["nbr_count"])


def test_graph_replay_matches_eager(mtlp):
    p = mtlp.scene_params(3, 8000)
    x, v = mtlp.generate_scene(3, 8000, 2, p)
    with mtlp.Engine(0, p) as eager, mtlp.Engine(0, p) as graphed:
        for e in (eager, graphed):
            e.reset(8000, mtlp.capi.SCENE_EMPTY)
            e.upload_state(x, v)
        graphed.set_graph(True)
        for k in range(6):
            eager.step(DT)
            graphed.step(DT)
            # eager work and a synchronisation between replays
            a = eager.positions_buffer().read()
            g = graphed.positions_buffer().read()
            np.testing.assert_array_equal(a.view(np.uint32), g.view(np.uint32), err_msg=f"replay {k}")
            np.testing.assert_array_equal(eager.lambdas_buffer().read().view(np.uint32),
                                          graphed.lambdas_buffer().read().view(np.uint32))
        graphed.step(DT / 2)  # new dt: recaptured
        eager.step(DT / 2)
        np.testing.assert_array_equal(eager.velocities_buffer().read().view(np.uint32),
                                      graphed.velocities_buffer().read().view(np.uint32))
        graphed.set_graph(False)
        graphed.step(DT)
        eager.step(DT)
        np.testing.assert_array_equal(eager.positions_buffer().read().view(np.uint32),
                                      graphed.positions_buffer().read().view(np.uint32))
        assert eager.error_flags() == 0 and graphed.error_flags() == 0


def test_graph_replay_after_parameter_changes(mtlp):
    """New parameters between replays (gravity and rest density, then the
    iteration count, which changes the captured kernel sequence) drop the
    graph: the replayed steps stay bit-identical to eager steps."""
    p = mtlp.scene_params(3, 4000)
    x, v = mtlp.generate_scene(3, 4000, 4, p)
    with mtlp.Engine(0, p) as eager, mtlp.Engine(0, p) as graphed:
        for e in (eager, graphed):
            e.reset(4000, mtlp.capi.SCENE_EMPTY)
            e.upload_state(x, v)
        graphed.set_graph(True)
        q = mtlp.scene_params(3, 4000)
        q.gravity[0], q.gravity[1] = 0.5, -4.0
        q.rest_density = p.rest_density * 1.1
        r = mtlp.scene_params(3, 4000)
        r.iterations = 2
        for k, params in enumerate((None, q, None, r, None)):
            if params is not None:
                for e in (eager, graphed):
                    e.params = params
            eager.step(DT)
            graphed.step(DT)
            for name in ("positions_buffer", "velocities_buffer", "lambdas_buffer"):
                a = getattr(eager, name)().read()
                g = getattr(graphed, name)().read()
                np.testing.assert_array_equal(a.view(np.uint32), g.view(np.uint32), err_msg=f"step {k} {name}")
        assert eager.error_flags() == 0 and graphed.error_flags() == 0


def test_stencil_block_sizes_bitwise_and_autotune(mtlp, tmp_path):
    p, x, v = None, None, None
    p = mtlp.scene_params(3, 20000)
    x, v = mtlp.generate_scene(3, 20000, 4, p)
    outs = {}
    for blk in (64, 128, 256, 512):
        with mtlp.Engine(0, p) as e:
            e.reset(20000, mtlp.capi.SCENE_EMPTY)
            e.upload_state(x, v)
            e.set_stencil_block(blk)
            for _ in range(2):
                e.step(DT)
            outs[blk] = e.positions_buffer().read()
    np.testing.assert_array_equal(outs[64].view(np.uint32), outs[256].view(np.uint32))
    np.testing.assert_array_equal(outs[128].view(np.uint32), outs[256].view(np.uint32))
    np.testing.assert_array_equal(outs[512].view(np.uint32), outs[256].view(np.uint32))
    cache = tmp_path / "tune.txt"
    with mtlp.Engine(0, p) as e:
        e.reset(20000, mtlp.capi.SCENE_EMPTY)
        e.upload_state(x, v)
        e.step(DT)
        before = _read_all(e)
        before["foam"] = e.foam_intensity_buffer().read()
        blk = e.autotune(DT, cache)
        assert blk in (64, 128, 256, 512)
        cfg = e.launch_config()
        assert cfg[0] == blk and cfg[1] in (0, 1, 4) and cfg[2] in (True, False)
        after = _read_all(e)
        after["foam"] = e.foam_intensity_buffer().read()
        for k in before:  # every readable buffer as before the call
            np.testing.assert_array_equal(after[k].view(np.uint32), before[k].view(np.uint32), err_msg=k)
        assert e.autotune(DT, cache) == blk  # second call reads the cache
        assert e.launch_config() == cfg
        assert len(cache.read_text().splitlines()) == 1
        e.step(DT)  # the tuned configuration steps as the default one does
        tuned = e.positions_buffer().read()
    with mtlp.Engine(0, p) as e:
        e.reset(20000, mtlp.capi.SCENE_EMPTY)
        e.upload_state(x, v)
        e.step(DT)
        e.step(DT)
        np.testing.assert_array_equal(e.positions_buffer().read().view(np.uint32), tuned.view(np.uint32))
        with pytest.raises(mtlp.MtlpError):
            e.set_stencil_block(100)


def test_neighbor_lists_bitwise(mtlp):
    """Delta passes walking the density pass's neighbour lists give the same
    bits as walking the stencil, including slots with more than 64
    neighbours (a dense blob: they fall back to the stencil), on a single
    scene and on a batch."""
    p = mtlp.scene_params(3, 20000)
    x, v = mtlp.generate_scene(3, 20000, 4, p)
    rng = np.random.default_rng(6)
    blob = x[:400, :3].mean(axis=0) + rng.uniform(-0.3, 0.3, (400, 3)) * p.h
    x[:400, :3] = blob.astype(np.float32)
    outs = {}
    for lists in (False, True):
        with mtlp.Engine(0, p) as e:
            e.reset(20000, mtlp.capi.SCENE_EMPTY)
            e.upload_state(x, v)
            e.set_neighbor_lists(lists)
            e.step(DT)
            nbr = e.neighbor_counts_buffer().read()
            e.step(DT)
            outs[lists] = (e.positions_buffer().read(), e.velocities_buffer().read(), e.lambdas_buffer().read())
            assert e.error_flags() == 0
    assert (nbr > 64).sum() >= 100  # the fallback path ran
    for a, b in zip(outs[False], outs[True]):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    counts, params = _mixed_scenes(mtlp)
    got = {}
    for lists in (False, True):
        b = mtlp.Batch(counts, params)
        b.seed(3, 11)
        b.set_neighbor_lists(lists)
        for _ in range(2):
            b.step(DT)
        got[lists] = b.positions_buffer().read()
        b.close()
    np.testing.assert_array_equal(got[False].view(np.uint32), got[True].view(np.uint32))


def test_ensemble_256_scenes(mtlp):
    from mtlphys_amd import workloads as W
    b = W.make_ensemble()
    n = b.particle_count
    assert len(b.counts) == 256 and n == sum(b.counts)
    b.set_graph(True)
    for _ in range(3):
        b.step(W.DT)
    x = b.positions_buffer().read()
    assert b.error_flags() == 0 and np.isfinite(x).all()
    po, _ = b.offsets()
    for s in range(0, 256, 17):
        p = b.scene_params[s]
        xs = x[po[s]:po[s + 1], :3]
        assert (xs >= np.array(p.bounds_min) - 1e-6).all() and (xs <= np.array(p.bounds_max) + 1e-6).all()
    b.close()


def test_ensemble_scenes_kernel_level_vs_fp64(mtlp, orc):
    """Three scenes of distinct rest density and gravity taken from inside
    the 256-scene ensemble (N_s log-uniform in [2048, 65536], 4.58M particles
    in all, 3 iterations): every iteration's density/lambda and delta kernels
    and the finalize, run in the batch, against the FP64 oracle fed the
    batch's own inputs for that scene, within the single-scene tolerances
    (lambda 1e-4 of max(|lambda|, 1e-3 max|lambda|), x 1e-5 h, v 1e-5 h/dt)."""
    from helpers import LAMBDA_TOL, V_TOL, X_TOL_H, lambda_errors, position_error_h, velocity_error
    from mtlphys_amd import workloads as W
    counts, params = W.ensemble_config()
    rho = np.array([p.rest_density for p in params])
    grav = np.array([p.gravity[1] for p in params])
    pick = sorted({int(np.argmin(rho)), int(np.argmax(rho)), int(np.argmin(grav))})
    assert len(pick) == 3
    iters = params[0].iterations
    print("ensemble scenes", [(s, counts[s], float(rho[s]), float(grav[s])) for s in pick], flush=True)
    runs = []
    for k in range(iters + 1):
        pk = [p.copy() for p in params]
        for p in pk:
            p.iterations = k
        b = mtlp.Batch(counts, pk)
        b.seed(3, 3)
        po, co = b.offsets()
        if k == 0:
            x0, v0 = b.positions_buffer().read(), b.velocities_buffer().read()
        b.step(DT)
        got = _read_all(b)
        assert b.error_flags() == 0
        b.close()
        runs.append({s: dict(sorted=got["sorted"][po[s]:po[s + 1]], lam=got["lam"][po[s]:po[s + 1]],
                             x=got["x"][po[s]:po[s + 1]], v=got["v"][po[s]:po[s + 1]],
                             sidx=got["sidx"][po[s]:po[s + 1]] - po[s],
                             cs=got["cs"][co[s]:co[s + 1] + 1] - po[s]) for s in pick})
        del got
    for s in pick:
        sl = slice(po[s], po[s + 1])
        op = orc.params(params[s])
        g = orc.grid(op)
        h = op.h
        cs, sidx = runs[0][s]["cs"], runs[0][s]["sidx"]
        for k in range(iters):
            inp = runs[k][s]["sorted"].astype(np.float64)
            _, lam_o, _ = orc.density_lambda(inp, cs, sidx, g, op, 64)
            le = lambda_errors(runs[k + 1][s]["lam"], lam_o)
            nxt = orc.apply_delta(inp, runs[k + 1][s]["lam"][sidx].astype(np.float64), cs, g, op, 64)
            xe = position_error_h(runs[k + 1][s]["sorted"], nxt, h)
            print(f"scene {s} iteration {k}: lambda {le['rel_floor']:.2e} x {xe['max']:.2e} h", flush=True)
            assert le["rel_floor"] <= LAMBDA_TOL, (s, k, le)
            assert xe["max"] <= X_TOL_H, (s, k, xe)
        fin = runs[-1][s]
        xo = np.empty((len(sidx), 4), np.float64)
        xo[sidx] = fin["sorted"].astype(np.float64)
        xr, vr = orc.finalize(xo, v0[sl], x0[sl], op, DT, 64)
        xe, ve = position_error_h(fin["x"], xr, h), velocity_error(fin["v"], vr, h, DT)
        print(f"scene {s} finalize: x {xe['max']:.2e} h v {ve['max']:.2e} h/dt", flush=True)
        assert xe["max"] <= X_TOL_H and ve["max"] <= V_TOL, (s, xe, ve)


def test_batch_huge_cells_stable(mtlp):
    """Three scenes in one batch, two of them with cells far above the 16,384
    particles an LDS sort holds, whose members are ranked per block of the
    batch's particle indices: the permutation must be the stable argsort of
    the (global) cell keys, so every scene keeps its particles in index
    order within a cell.  Steps with no solver iteration and zero gravity
    and velocity: the hash of the positions as they are."""
    from mtlphys_amd import workloads as W
    p = W.unit_grid_params(64)
    p.iterations = 0
    for a in range(3):
        p.gravity[a] = 0.0
    rng = np.random.default_rng(31)

    def scene(n, hot):
        pts = [rng.random((n - sum(hot), 3), dtype=np.float32)]
        for m in hot:
            c = rng.integers(0, 64, 3)
            pts.append(((c + rng.random((m, 3)) * 0.999) / 64.0).astype(np.float32))
        x = np.concatenate(pts)[rng.permutation(n)]
        return np.concatenate([x, np.ones((n, 1), np.float32)], 1)

    counts = [40000, 8000, 30000]
    xs = [scene(40000, [17500, 17500]), scene(8000, []), scene(30000, [20000])]
    x = np.concatenate(xs).astype(np.float32)
    b = mtlp.Batch(counts, [p, p, p])
    try:
        b.upload_state(x, np.zeros_like(x))
        b.step(DT)
        keys = b.cell_hashes_buffer().read()
        sidx = b.sorted_index_buffer().read()
        s4 = b.sorted_positions_buffer().read()
        assert b.error_flags() == 0
    finally:
        b.close()
    assert (np.bincount(keys) > 16384).sum() == 3
    np.testing.assert_array_equal(sidx, np.argsort(keys, kind="stable").astype(np.uint32))
    np.testing.assert_array_equal(s4.view(np.uint32), x[sidx].view(np.uint32))
