"""GPU parity: the HIP kernels of libmtlp.so against the CPU oracle.

Integer stages (cell keys, counts/offsets, permutation, neighbour counts) are
compared bit for bit.  Floating-point stages are compared kernel by kernel:
each GPU kernel's input is read back and fed to the FP64 oracle kernel, and
the outputs must agree within
    lambda: |dl| <= 1e-4 * max(|lambda_ref|, 1e-3 max|lambda_ref|)
    x:      |dx| <= 1e-5 * h   (every component)
    v:      |dv| <= 1e-5 * h / dt
Free-running steps (GPU from the same (x, v) as the FP64 oracle) are
reported, and bounded by the error of the oracle's own FP32 instantiation.
"""
import numpy as np
import pytest

from helpers import (LAMBDA_TOL, X_TOL_H, assert_kernel_level, kernel_level_errors, lambda_errors,
                     position_error_h, run_gpu_step)

pytestmark = pytest.mark.gpu


# ---------------------------------------------------------------- scan ----
def _gpu_scan(mtlp, arr):
    import torch
    n = len(arr)
    d_in = torch.from_numpy(arr.astype(np.uint32).view(np.int32)).cuda()
    d_out = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    scratch = torch.empty(mtlp.Engine.scan_scratch_bytes(n), dtype=torch.uint8, device="cuda")
    st = torch.cuda.current_stream()
    mtlp.capi.check(mtlp.capi.lib().mtlp_exclusive_scan(
        None, st.cuda_stream, d_in.data_ptr(), d_out.data_ptr(), scratch.data_ptr(), n),
        "mtlp_exclusive_scan")
    torch.cuda.synchronize()
    return d_out.cpu().numpy().view(np.uint32)


def _ref_scan(a):
    out = np.zeros(len(a), np.uint64)
    out[1:] = np.cumsum(a[:-1].astype(np.uint64))
    return (out & 0xFFFFFFFF).astype(np.uint32)


# The reference's scan cases (tests/test_scan.cpp:59-110), inputs from this
# project's RNG (std::uniform_int_distribution is implementation-defined),
# plus tile edges and lengths past the reference's 1024^2 cap.
SCAN_CASES = [("zeros", 100, 0, 0), ("ones", 500, 1, 1), ("rand", 1024, 0xA, 7), ("rand", 700, 0xB, 7),
              ("rand", 2048, 0xC, 7), ("rand", 635626, 0xD, 3), ("rand", 1, 1, 7), ("rand", 4095, 2, 7),
              ("rand", 4096, 3, 7), ("rand", 4097, 4, 7), ("rand", 530605, 5, 7),
              ("rand", (1 << 20) + 1, 6, 7), ("rand", 2097153, 7, 7), ("rand", (1 << 24) + 1, 8, 7),
              ("big", 100000, 9, 0xFFFFFFFF),
              # tile edges (16,384 elements; 32,768 from 2^22 on) and the switch at 2^22
              ("rand", 16383, 10, 7), ("rand", 16384, 11, 7), ("rand", 16385, 12, 7),
              ("rand", (1 << 22) - 1, 13, 7), ("rand", 1 << 22, 14, 7), ("rand", (1 << 22) + 32769, 15, 7)]


@pytest.mark.parametrize("kind,n,seed,maxv", SCAN_CASES)
def test_exclusive_scan(mtlp, kind, n, seed, maxv):
    rng = np.random.default_rng(seed)
    if kind == "zeros":
        a = np.zeros(n, np.uint32)
    elif kind == "ones":
        a = np.ones(n, np.uint32)
    else:
        a = rng.integers(0, maxv, n, dtype=np.uint64, endpoint=True).astype(np.uint32)
    got = _gpu_scan(mtlp, a)
    np.testing.assert_array_equal(got, _ref_scan(a))


def test_scan_repeated_launches_epochs(mtlp):
    """The engine's scan reuses its look-back words across launches."""
    p = mtlp.scene_params(3, 32768)
    x, v = mtlp.generate_scene(3, 32768, 1, p)
    with mtlp.Engine(0, p) as e:
        e.reset(32768, mtlp.capi.SCENE_EMPTY)
        e.upload_state(x, v)
        first = None
        for _ in range(5):
            e.build_spatial_hash()
            cs = e.cell_start_buffer().read()
            if first is None:
                first = cs
            np.testing.assert_array_equal(cs, first)
        assert e.error_flags() == 0


# ------------------------------------------------------- spatial hash ----
def _hash_case(mtlp, orc, params, pts):
    pts = np.ascontiguousarray(pts, np.float32)
    n = len(pts)
    with mtlp.Engine(0, params) as e:
        e.reset(n, mtlp.capi.SCENE_EMPTY)
        e.upload_state(pts, np.zeros_like(pts))
        e.build_spatial_hash()
        got = dict(keys=e.cell_hashes_buffer().read(), cell_start=e.cell_start_buffer().read(),
                   sorted_index=e.sorted_index_buffer().read(),
                   sorted4=e.sorted_positions_buffer().read(), nbr=e.neighbor_counts_buffer().read())
        assert e.error_flags() == 0
        assert e.total_cells == int(np.prod(mtlp.grid_dims(params)))
    op = orc.params(params)
    ref = orc.build_spatial_hash(pts, orc.grid(op), np.float32(op.h) * np.float32(op.h))
    return got, ref


def _assert_hash_equal(got, ref):
    for k in ("keys", "cell_start", "sorted_index", "nbr"):
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    np.testing.assert_array_equal(got["sorted4"].view(np.uint32), ref["sorted4"].view(np.uint32))


def test_hash_reference_known_answers(mtlp, orc):
    """tests/test_spatial.cpp:52-70 (reference grid: [-3,3]x[-3,6]x[-3,3], cell 0.1)."""
    pts = np.array([[0.0, 0.0, 0.0, 1], [2.5, 4.0, -1.5, 1], [-2.9, 5.5, 2.5, 1], [1.0, -2.0, -2.5, 1]],
                   np.float32)
    got, ref = _hash_case(mtlp, orc, mtlp.default_params(), pts)
    # independent FP32 restatement of cpuCellHash (tests/TestUtils.cpp:30-42)
    o = np.float32(-3.0)
    inv = np.float32(1.0) / np.float32(0.1)
    dims = np.array([60, 90, 60])
    c = np.floor((pts[:, :3] - o) * inv).astype(np.int64)
    c = np.clip(c, 0, dims - 1)
    expect = c[:, 0] + c[:, 1] * 60 + c[:, 2] * 60 * 90
    np.testing.assert_array_equal(got["keys"], expect)
    _assert_hash_equal(got, ref)


def test_hash_grouping_reference_case(mtlp, orc):
    """tests/test_spatial.cpp:72-100: cell ranges hold exactly their particles,
    here in increasing index order."""
    pts = np.array([[0, 0, 0, 1], [0.01, 0.02, 0.01, 1], [0.03, 0.01, 0.02, 1], [2, 2, 2, 1],
                    [2.01, 2.02, 2.03, 1], [-2.5, -2.5, -2.5, 1]], np.float32)
    got, ref = _hash_case(mtlp, orc, mtlp.default_params(), pts)
    _assert_hash_equal(got, ref)
    assert got["cell_start"][-1] == len(pts)
    for key in np.unique(got["keys"]):
        s, e = got["cell_start"][key], got["cell_start"][key + 1]
        np.testing.assert_array_equal(got["sorted_index"][s:e], np.nonzero(got["keys"] == key)[0])


@pytest.mark.parametrize("case", ["isolated", "pair", "lattice"])
def test_neighbor_counts_reference_cases(mtlp, orc, case):
    """tests/test_spatial.cpp:102-145."""
    cs = 0.1
    if case == "isolated":
        pts = [[0, 0, 0, 1], [1.5, 1.5, 1.5, 1], [-1.5, -1.5, -1.5, 1]]
        expect = {0: 0, 1: 0, 2: 0}
    elif case == "pair":
        pts = [[0, 0, 0, 1], [cs * 0.5, 0, 0, 1]]
        expect = {0: 1, 1: 1}
    else:
        s = np.float32(cs) * np.float32(0.99)
        pts = [[x * s, y * s, z * s, 1] for z in (-1, 0, 1) for y in (-1, 0, 1) for x in (-1, 0, 1)]
        expect = {13: 6}
    got, ref = _hash_case(mtlp, orc, mtlp.default_params(), np.array(pts, np.float32))
    _assert_hash_equal(got, ref)
    for i, cnt in expect.items():
        assert got["nbr"][i] == cnt
    if case == "lattice":
        assert got["nbr"][0] < got["nbr"][13] and got["nbr"][26] < got["nbr"][13]


def _unit_grid_params(mtlp, cells=128):
    from mtlphys_amd import workloads as W
    return W.unit_grid_params(cells)


@pytest.mark.parametrize("variant", ["uniform", "clustered"])
def test_hash_4m_128cubed_bitexact(mtlp, orc, variant):
    """4M points in a 128^3 grid (2,097,152 cells; scan length 2,097,153)."""
    from mtlphys_amd import workloads as W
    n = W.HASH_POINTS
    pts = W.uniform_points(n, 2) if variant == "uniform" else W.clustered_points(n, 2)
    p = _unit_grid_params(mtlp)
    with mtlp.Engine(0, p) as e:
        e.reset(n, mtlp.capi.SCENE_EMPTY)
        e.upload_state(pts, np.zeros_like(pts))
        e.build_spatial_hash()
        keys = e.cell_hashes_buffer().read()
        cs = e.cell_start_buffer().read()
        sidx = e.sorted_index_buffer().read()
        s4 = e.sorted_positions_buffer().read()
        assert e.error_flags() == 0
    op = orc.params(p)
    g = orc.grid(op)
    ref_keys = orc.hash_keys(pts, g)
    np.testing.assert_array_equal(keys, ref_keys)
    counts = np.bincount(ref_keys, minlength=g.total_cells + 1).astype(np.uint32)
    np.testing.assert_array_equal(cs, _ref_scan(counts))
    np.testing.assert_array_equal(sidx, np.argsort(ref_keys, kind="stable").astype(np.uint32))
    np.testing.assert_array_equal(s4.view(np.uint32), pts[sidx].view(np.uint32))


@pytest.mark.parametrize("n_in_cell", [300, 5000, 16384, 16385, 70000])
def test_hash_big_cells_stable(mtlp, orc, n_in_cell):
    """Cells above the in-place ranking limit: LDS sort up to 16384 particles,
    ranks per index block beyond; mixed with ordinary cells."""
    rng = np.random.default_rng(n_in_cell)
    n_rest = 20000
    hot = np.tile(np.array([[0.5, 0.5, 0.5]], np.float32), (n_in_cell, 1))
    hot += rng.random((n_in_cell, 3), dtype=np.float32) * np.float32(0.5 / 128)
    rest = rng.random((n_rest, 3), dtype=np.float32)
    pts = np.concatenate([hot, rest])[rng.permutation(n_in_cell + n_rest)]
    pts = np.concatenate([pts, np.ones((len(pts), 1), np.float32)], 1).astype(np.float32)
    p = _unit_grid_params(mtlp)
    with mtlp.Engine(0, p) as e:
        e.reset(len(pts), mtlp.capi.SCENE_EMPTY)
        e.upload_state(pts, np.zeros_like(pts))
        e.build_spatial_hash()
        keys = e.cell_hashes_buffer().read()
        sidx = e.sorted_index_buffer().read()
        s4 = e.sorted_positions_buffer().read()
        assert e.error_flags() == 0
    np.testing.assert_array_equal(sidx, np.argsort(keys, kind="stable").astype(np.uint32))
    np.testing.assert_array_equal(s4.view(np.uint32), pts[sidx].view(np.uint32))


def _hash_only_step(mtlp, p, pts):
    """One step with no solver iteration, no gravity and zero velocities: the
    spatial hash of pts exactly (x* = x), without countNeighbors, whose
    stencil walk is quadratic in a cell's size.  Returns (keys, sorted
    index, sorted positions)."""
    q = p.copy()
    q.iterations = 0
    for a in range(3):
        q.gravity[a] = 0.0
    with mtlp.Engine(0, q) as e:
        e.reset(len(pts), mtlp.capi.SCENE_EMPTY)
        e.upload_state(pts, np.zeros_like(pts))
        e.step(1.0 / 120.0)
        out = (e.cell_hashes_buffer().read(), e.sorted_index_buffer().read(), e.sorted_positions_buffer().read())
        assert e.error_flags() == 0
    return out


@pytest.mark.parametrize("n,hot", [(400000, 20), (4194304, 200)])
def test_hash_many_huge_cells_stable(mtlp, n, hot):
    """95% of the points in `hot` cells of ~19,000 each (all above the
    16,384 the LDS sort holds), the rest uniform: every huge cell is ranked
    per index block, together in one pass over the indices, and the
    permutation must be the stable argsort of the keys."""
    from mtlphys_amd import workloads as W
    pts = W.huge_cell_points(n, 7, n_hot=hot)
    keys, sidx, s4 = _hash_only_step(mtlp, _unit_grid_params(mtlp), pts)
    counts = np.bincount(keys)
    assert (counts > 16384).sum() == hot, np.sort(counts)[-hot - 1:]
    np.testing.assert_array_equal(sidx, np.argsort(keys, kind="stable").astype(np.uint32))
    np.testing.assert_array_equal(s4.view(np.uint32), pts[sidx].view(np.uint32))


# ---------------------------------------------------------------- step ----
def _cube_case(mtlp, n=32768, seed=1):
    p = mtlp.scene_params(3, n)
    x, v = mtlp.generate_scene(3, n, seed, p)
    return p, x, v


def test_step_integer_stages_bitexact(mtlp, orc):
    p, x, v = _cube_case(mtlp)
    dt = 1.0 / 120.0
    got = run_gpu_step(mtlp, p, x, v, dt)
    op = orc.params(p)
    _, _, tr = orc.step(x, v, op, dt, prec=32, trace=True)
    for k in ("keys", "cell_start", "sorted_index"):
        np.testing.assert_array_equal(got[k], tr[k], err_msg=k)
    # iteration-0 neighbour count == countNeighbors on the predicted positions
    np.testing.assert_array_equal(got["nbr"], tr["nbr_count"])
    np.testing.assert_array_equal(got["prev"].view(np.uint32), x.view(np.uint32))


def _kernel_level(mtlp, orc, p, x, v, dt):
    """Every lambda/delta kernel and finalize against the FP64 oracle fed
    the GPU's own inputs (helpers.kernel_level_errors)."""
    report, _ = kernel_level_errors(mtlp, orc, p, x, v, dt)
    assert_kernel_level(report)
    return report


def test_step_kernel_level_cube(mtlp, orc):
    p, x, v = _cube_case(mtlp)
    _kernel_level(mtlp, orc, p, x, v, 1.0 / 120.0)


def test_step_kernel_level_after_settling(mtlp, orc):
    """Same check on a state 20 GPU steps into the cube drop."""
    p, x, v = _cube_case(mtlp, n=8000, seed=3)
    with mtlp.Engine(0, p) as e:
        e.reset(len(x), mtlp.capi.SCENE_EMPTY)
        e.upload_state(x, v)
        for _ in range(20):
            e.step(1.0 / 120.0)
        x1, v1 = e.positions_buffer().read(), e.velocities_buffer().read()
    assert np.isfinite(x1).all() and np.isfinite(v1).all()
    _kernel_level(mtlp, orc, p, x1, v1, 1.0 / 120.0)


def test_step_particles_on_and_past_the_walls(mtlp, orc):
    """Particles exactly on the box faces and up to h/2 beyond them (binned
    into the edge cells by the clamp of cellCoordOf, pulled back by the
    delta pass's box clamp): integer stages bit-exact, every kernel against
    the FP64 oracle."""
    p, x, v = _cube_case(mtlp, n=4000, seed=9)
    lo, hi, h = np.array(p.bounds_min), np.array(p.bounds_max), p.h
    rng = np.random.default_rng(9)
    k = len(x) // 10
    idx = rng.permutation(len(x))
    past = idx[:k]              # beyond the +x face by up to h/2
    x[past, 0] = hi[0] + rng.uniform(0.0, 0.5 * h, k).astype(np.float32)
    on = idx[k:2 * k]           # exactly on the +x face
    x[on, 0] = np.float32(hi[0])
    floor = idx[2 * k:3 * k]    # exactly on the floor
    x[floor, 1] = np.float32(lo[1])
    dt = 1.0 / 120.0
    got = run_gpu_step(mtlp, p, x, v, dt)
    _, _, tr = orc.step(x, v, orc.params(p), dt, prec=32, trace=True)
    for name in ("keys", "cell_start", "sorted_index"):
        np.testing.assert_array_equal(got[name], tr[name], err_msg=name)
    np.testing.assert_array_equal(got["nbr"], tr["nbr_count"])
    _kernel_level(mtlp, orc, p, x, v, dt)


@pytest.mark.parametrize("n", [1, 2, 63, 65, 257, 4097])
def test_step_ragged_sizes(mtlp, orc, n):
    """Partial waves, workgroups and tiles: 1 and 2 particles, one short of
    and one past a wave, one past a workgroup and one past a 4,096-slot
    tile.  Integer stages bit-exact against the oracle, then every
    density/lambda and delta kernel and finalize against the FP64 oracle."""
    p, x, v = _cube_case(mtlp, n=n, seed=5)
    dt = 1.0 / 120.0
    got = run_gpu_step(mtlp, p, x, v, dt)
    _, _, tr = orc.step(x, v, orc.params(p), dt, prec=32, trace=True)
    for k in ("keys", "cell_start", "sorted_index"):
        np.testing.assert_array_equal(got[k], tr[k], err_msg=k)
    np.testing.assert_array_equal(got["nbr"], tr["nbr_count"])
    _kernel_level(mtlp, orc, p, x, v, dt)


def _chunks_and_counts(orc, params, x, v, dt):
    """Per particle at iteration 0: the stencil chunks the density pass keeps
    (chunks of <= 32 candidate slots of a stencil row, as it cuts them, that
    hold a neighbour) and the neighbour count.  The chunks are counted from
    the float64 neighbour pairs of a k-d tree: a coverage count (a pair at
    |r| = h within rounding may fall either way), not a parity check."""
    from scipy.spatial import cKDTree
    op = orc.params(params)
    g = orc.grid(op)
    xs = orc.predict(x, v, op, dt)
    bh = orc.build_spatial_hash(xs, g, np.float32(op.h) * np.float32(op.h))
    cs = bh["cell_start"].astype(np.int64)
    s4 = bh["sorted4"][:, :3]
    dims = np.array(g.dim, np.int64)
    c = np.clip(np.floor((s4 - np.array(op.bounds_min, np.float32)) / np.float32(op.cell_size)).astype(np.int64),
                0, dims - 1)
    pairs = cKDTree(s4.astype(np.float64)).query_pairs(float(op.h), output_type="ndarray")
    i = np.concatenate([pairs[:, 0], pairs[:, 1]])
    j = np.concatenate([pairs[:, 1], pairs[:, 0]])
    row = (c[j, 2] - c[i, 2] + 1) * 3 + (c[j, 1] - c[i, 1] + 1)
    first = cs[np.maximum(c[i, 0] - 1, 0) + dims[0] * (c[j, 1] + dims[1] * c[j, 2])]  # the row's first slot
    chunk = (i * 9 + row) * 4096 + (j - first) // 32
    K = np.bincount(np.unique(chunk) // (9 * 4096), minlength=len(x))
    return K, bh["nbr"]


def _blob_case(mtlp, n=20000):
    """A cube-drop lattice with three blobs of particles packed 1.2 h, 4 h
    and 6 h wide: slots with few, many and very many stencil chunks, and
    tiles of very different cost."""
    p = mtlp.scene_params(3, n)
    x, v = mtlp.generate_scene(3, n, 5, p)
    rng = np.random.default_rng(8)
    centre = x[:, :3].mean(axis=0)
    h = p.h
    for lo, hi, side, off in ((0, 600, 1.2, (-2.5, 0.0, 0.0)), (600, 3600, 4.0, (2.5, 0.0, 0.0)),
                              (3600, 7000, 6.0, (0.0, 0.0, 3.0))):
        x[lo:hi, :3] = (centre + np.array(off) * h + rng.uniform(-0.5, 0.5, (hi - lo, 3)) * side * h).astype(
            np.float32)
    return p, x, v


def test_dense_regions_chunk_paths(mtlp, orc):
    """Blobs of three densities drive the density pass through every way it
    keeps a slot's stencil chunks: all in LDS (<= 15), the global scratch
    (16..32, with and without a neighbour list of <= 96 entries), and the
    second stencil walk (> 32).  Lists on and off must agree bit for bit
    over two steps (the second runs its tiles in the order the first left),
    and each iteration's kernels must match the FP64 oracle."""
    n = 20000
    p, x, v = _blob_case(mtlp, n)
    dt = 1.0 / 120.0
    K, cnt = _chunks_and_counts(orc, p, x, v, dt)
    assert (K <= 15).sum() > 1000
    assert ((K > 15) & (K <= 32) & (cnt <= 96)).sum() > 50
    assert ((K > 15) & (K <= 32) & (cnt > 96)).sum() > 50
    assert (K > 32).sum() > 50
    outs = {}
    for lists in (False, True):
        with mtlp.Engine(0, p) as e:
            e.reset(n, mtlp.capi.SCENE_EMPTY)
            e.upload_state(x, v)
            e.set_neighbor_lists(lists)
            outs[lists] = []
            for _ in range(2):
                e.step(dt)
                outs[lists] += [e.positions_buffer().read(), e.velocities_buffer().read(), e.lambdas_buffer().read()]
            assert e.error_flags() == 0
    for a, b in zip(outs[False], outs[True]):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    _kernel_level(mtlp, orc, p, x, v, dt)


@pytest.mark.parametrize("n,m", [(40000, 32000), (200000, 100000)])
def test_heavy_queue_overflow(mtlp, orc, n, m):
    """Two dense cubes, m / 4 particles at 150 per h^3 (about 600
    neighbours) and 3 m / 4 at 36 per h^3 (about 150), give thousands of
    slots with over 32 kept stencil chunks (a wave each from the delta
    pass's heavy queue) and more slots with over 96 neighbours in at most
    32 kept chunks than its heavy-lane queue holds (n / 8, rounded up to
    4,096): at n = 40,000 that queue (8,192) is short enough for the heavy
    waves to take its slots; at n = 200,000 (28,672) its slots are walked
    one per lane in the heavy-lane workgroups, and both queues overflow into
    the ordinary waves.
    Lists on and off must agree bit for bit over two steps, and each
    iteration's kernels must match the FP64 oracle."""
    p = mtlp.scene_params(3, n)
    x, v = mtlp.generate_scene(3, n, 5, p)
    rng = np.random.default_rng(11)
    centre = x[:, :3].mean(axis=0)
    m1, m2 = m // 4, m - m // 4
    s1, s2 = (m1 / 150.0) ** (1.0 / 3.0), (m2 / 36.0) ** (1.0 / 3.0)  # sides in h
    x[:m1, :3] = (centre + np.array([-0.5 * s2 - 1.0, 0.0, 0.0]) * p.h
                  + rng.uniform(-0.5, 0.5, (m1, 3)) * s1 * p.h).astype(np.float32)
    x[m1:m, :3] = (centre + np.array([0.5 * s1 + 1.0, 0.0, 0.0]) * p.h
                   + rng.uniform(-0.5, 0.5, (m2, 3)) * s2 * p.h).astype(np.float32)
    dt = 1.0 / 120.0
    K, cnt = _chunks_and_counts(orc, p, x, v, dt)
    lanes_cap = (n // 8 + 4095) // 4096 * 4096
    assert (K > 32).sum() > 2000, (K > 32).sum()
    assert ((K <= 32) & (cnt > 96)).sum() > lanes_cap + 2000, ((K <= 32) & (cnt > 96)).sum()
    outs = {}
    for lists in (False, True):
        with mtlp.Engine(0, p) as e:
            e.reset(n, mtlp.capi.SCENE_EMPTY)
            e.upload_state(x, v)
            e.set_neighbor_lists(lists)
            outs[lists] = []
            for _ in range(2):
                e.step(dt)
                outs[lists] += [e.positions_buffer().read(), e.lambdas_buffer().read()]
            assert e.error_flags() == 0
    for a, b in zip(outs[False], outs[True]):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    _kernel_level(mtlp, orc, p, x, v, dt)


@pytest.mark.parametrize("iterations", [4, 3])
def test_tile_order_does_not_change_results(mtlp, iterations):
    """The stencil passes take their tiles heaviest first, in an order the
    previous iteration measured; with odd or even iteration counts, with the
    stencil workgroup size changing between steps (which invalidates the
    order) and with lists off (no order at all), four steps must give the
    same bits."""
    n = 20000
    p, x, v = _blob_case(mtlp, n)
    p.iterations = iterations
    dt = 1.0 / 120.0
    runs = {"lists off": (False, [64, 64, 64, 64]), "lists, 64": (True, [64, 64, 64, 64]),
            "lists, sizes change": (True, [64, 128, 128, 64])}
    outs = {}
    for name, (lists, blocks) in runs.items():
        with mtlp.Engine(0, p) as e:
            e.reset(n, mtlp.capi.SCENE_EMPTY)
            e.upload_state(x, v)
            e.set_neighbor_lists(lists)
            outs[name] = []
            for b in blocks:
                e.set_stencil_block(b)
                e.step(dt)
                outs[name] += [e.positions_buffer().read(), e.lambdas_buffer().read()]
            assert e.error_flags() == 0
    for name in runs:
        for a, b in zip(outs["lists off"], outs[name]):
            np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32), err_msg=name)


@pytest.mark.parametrize("case", ["tiny_300", "blobs_20k", "dam_break_1m"])
def test_persistent_launches_bit_identical(mtlp, case):
    """Persistent stencil launches (resident workgroups taking 1 or 4 tiles
    at a time from one counter per XCD) give the bits of one workgroup per
    tile, also when the mode changes between steps; on the blob scene (tiles
    of very different cost, heavy queues, two workgroup sizes) and on the 1M
    dam break (15,625 tiles: each persistent workgroup takes several)."""
    if case == "tiny_300":  # fewer tiles than queues
        n = 300
        p, x, v = _cube_case(mtlp, n=n)
        blocks, steps = (64, 512), 4
    elif case == "blobs_20k":
        n = 20000
        p, x, v = _blob_case(mtlp, n)
        blocks, steps = (64, 256), 4
    else:
        n = 1_000_000
        p = mtlp.scene_params(4, n)
        x, v = mtlp.generate_scene(4, n, 1, p)
        blocks, steps = (64,), 2
    dt = 1.0 / 120.0
    for block in blocks:
        outs = {}
        for name, pulls in {"dispatch": [0] * steps, "pull 1": [1] * steps, "pull 4": [4] * steps,
                            "switching": ([0, 4, 1, 0] if steps == 4 else [4, 0])}.items():
            with mtlp.Engine(0, p) as e:
                e.reset(n, mtlp.capi.SCENE_EMPTY)
                e.upload_state(x, v)
                e.set_stencil_block(block)
                outs[name] = []
                for t in pulls:
                    e.set_tile_pull(t)
                    e.step(dt)
                    outs[name] += [e.positions_buffer().read(), e.lambdas_buffer().read()]
                assert e.launch_config()[:2] == (block, pulls[-1])
                assert e.error_flags() == 0
        for name in outs:
            for a, b in zip(outs["dispatch"], outs[name]):
                np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32), err_msg=(case, block, name))


@pytest.mark.parametrize("case", ["blobs_20k", "dam_break_1m"])
def test_packed_positions_bit_identical(mtlp, case):
    """The density pass's stencil filter reading the packed (x, y, z) copy of
    each iterate gives the bits of the filter reading the float4 iterate:
    positions, lambdas and neighbour counts, with the copy on, off, switched
    between steps, and in persistent launches; on the blob scene (heavy
    queues, chunk-scratch walks) and on the 1M dam break."""
    if case == "blobs_20k":
        n = 20000
        p, x, v = _blob_case(mtlp, n)
        steps = 4
    else:
        n = 1_000_000
        p = mtlp.scene_params(4, n)
        x, v = mtlp.generate_scene(4, n, 1, p)
        steps = 2
    dt = 1.0 / 120.0
    runs = {"off": [(False, 0)] * steps, "on": [(True, 0)] * steps, "on, pull 4": [(True, 4)] * steps,
            "switching": [(True, 0), (False, 0), (True, 1), (False, 4)][:steps]}
    outs = {}
    for name, cfgs in runs.items():
        with mtlp.Engine(0, p) as e:
            e.reset(n, mtlp.capi.SCENE_EMPTY)
            e.upload_state(x, v)
            outs[name] = []
            for packed, pull in cfgs:
                e.set_packed_positions(packed)
                e.set_tile_pull(pull)
                e.step(dt)
                outs[name] += [e.positions_buffer().read(), e.lambdas_buffer().read(),
                               e.neighbor_counts_buffer().read()]
            assert e.launch_config()[1:] == (cfgs[-1][1], cfgs[-1][0])
            assert e.error_flags() == 0
    for name in outs:
        for a, b in zip(outs["off"], outs[name]):
            np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32), err_msg=(case, name))


def test_step_free_running_vs_fp64(mtlp, orc):
    """One full step from the same state: the GPU may differ from the FP64
    oracle by no more than a few times the oracle's own FP32 instantiation
    (the Jacobi iterations amplify FP32 rounding in a few particles)."""
    p, x, v = _cube_case(mtlp)
    dt = 1.0 / 120.0
    op = orc.params(p)
    got = run_gpu_step(mtlp, p, x, v, dt)
    x64, v64, tr = orc.step(x, v, op, dt, prec=64, trace=True)
    x32, v32, t32 = orc.step(x, v, op, dt, prec=32, trace=True)
    e_gpu = position_error_h(got["x"], x64, op.h)
    e_f32 = position_error_h(x32, x64, op.h)
    l_gpu = lambda_errors(got["lam"], tr["lambda_orig"])
    l_f32 = lambda_errors(t32["lambda_orig"], tr["lambda_orig"])
    print("free-running x err/h gpu", e_gpu, "oracle-f32", e_f32, "lambda gpu", l_gpu, "f32", l_f32)
    assert e_gpu["p50"] <= X_TOL_H
    assert e_gpu["max"] <= 4 * e_f32["max"] + X_TOL_H
    assert l_gpu["rel_max"] <= max(LAMBDA_TOL, 4 * l_f32["rel_max"])


def test_dam_break_1m_integer_stages(mtlp, orc):
    """The 1M-particle dam break (h = 0.02, 530,604 cells): cell keys,
    offsets, permutation and the iteration-0 neighbour counts bit-exact.
    Every iteration's floating-point kernels at this size are checked in
    test_gpu_parity_1m.py."""
    n = 1_000_000
    p = mtlp.scene_params(4, n)
    assert mtlp.grid_dims(p) == (102, 102, 51)
    x, v = mtlp.generate_scene(4, n, 1, p)
    dt = 1.0 / 120.0
    op = orc.params(p)
    got1 = run_gpu_step(mtlp, p, x, v, dt, iterations=1)
    _, _, tr = orc.step(x, v, orc.params(p, iterations=1), dt, prec=32, trace=True)
    for k in ("keys", "cell_start", "sorted_index"):
        np.testing.assert_array_equal(got1[k], tr[k], err_msg=k)
    # countNeighbors is FP32 arithmetic (Spatial.metal:235-236): compare with
    # the oracle's FP32 count on the predicted positions
    xs = orc.predict(x, v, op, dt)
    ref_nbr = orc.build_spatial_hash(xs, orc.grid(op), np.float32(op.h) * np.float32(op.h))["nbr"]
    np.testing.assert_array_equal(got1["nbr"], ref_nbr)


# -------------------------------------------- reference pbf test cases ----
def test_isolated_particle_free_fall(mtlp):
    """tests/test_pbf.cpp:47-65: y = 1 + (-9.81 * 0.01) * 0.01 after dt = 0.01."""
    with mtlp.Engine(0) as e:
        e.reset(1)
        pos = e.positions_buffer()
        pos.write(np.array([[0, 1, 0, 1]], np.float32))
        e.velocities_buffer().write(np.zeros((1, 4), np.float32))
        e.step(0.01)
        y = pos.read()[0, 1]
        vy = e.velocities_buffer().read()[0, 1]
    assert abs(y - 0.99902) <= 1e-3 * 0.99902
    assert y == np.float32(1.0) + (np.float32(0.0) + np.float32(-9.81) * np.float32(0.01)) * np.float32(0.01)
    assert vy < 0


def test_lambda_sign_and_isolated_closed_form(mtlp):
    """tests/test_pbf.cpp:67-107 with the engine's own epsilon (800)."""
    h = 0.1
    s = np.float32(h * 0.4)
    pts = [[x * s, y * s, z * s, 1] for z in (-1, 0, 1) for y in (-1, 0, 1) for x in (-1, 0, 1)]
    pts.append([2, 2, 2, 1])
    pts = np.array(pts, np.float32)
    p = mtlp.default_params()
    with mtlp.Engine(0, p) as e:
        e.reset(len(pts))
        e.upload_state(pts, np.zeros_like(pts))
        e.step(1e-6)
        lam = e.lambdas_buffer().read()
    assert lam[13] < 0
    w0 = 315.0 / (64.0 * np.pi * h ** 3)
    expect = -(w0 / 8000.0 - 1.0) / 800.0
    assert abs(lam[-1] - expect) <= 1e-5 * abs(expect)


def test_fifty_steps_stable(mtlp):
    """tests/test_pbf.cpp:109-130: 1024 particles, 50 steps at dt = 1/240."""
    with mtlp.Engine(0) as e:
        e.reset(1024)
        for _ in range(50):
            e.step(1.0 / 240.0)
        x = e.positions_buffer().read()
    assert np.isfinite(x).all()
    lo, hi = np.array([-3.01, -3.01, -3.01]), np.array([3.01, 6.01, 3.01])
    assert ((x[:, :3] >= lo) & (x[:, :3] <= hi)).all()


def test_dt_nonpositive_is_noop(mtlp):
    with mtlp.Engine(0) as e:
        e.reset(1000)
        x0 = e.positions_buffer().read()
        e.step(0.0)
        e.step(-1.0)
        np.testing.assert_array_equal(e.positions_buffer().read(), x0)


def test_zero_particles(mtlp):
    with mtlp.Engine(0) as e:
        e.reset(0)
        e.step(0.01)
        e.build_spatial_hash()
        assert e.particle_count == 0
        assert e.cell_start_buffer().read()[-1] == 0
