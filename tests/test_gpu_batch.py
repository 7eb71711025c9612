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
        before = e.positions_buffer().read()
        blk = e.autotune(DT, cache)
        assert blk in (64, 128, 256, 512)
        np.testing.assert_array_equal(e.positions_buffer().read(), before)  # state restored
        assert e.autotune(DT, cache) == blk  # second call reads the cache
        assert len(cache.read_text().splitlines()) == 1
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
