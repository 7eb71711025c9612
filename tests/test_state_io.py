"""State file: numpy round trip on the CPU; on the GPU, engine save/load
round trips (single scene, batch, sphere) and numpy-written files."""
import numpy as np
import pytest

DT = 1.0 / 120.0


def test_numpy_round_trip(tmp_path, mtlp):
    p = mtlp.scene_params(3, 1000)
    x, v = mtlp.generate_scene(3, 1000, 2, p)
    path = tmp_path / "s.bin"
    mtlp.write_state(path, [p], [1000], x, v, scene=3, sphere=((1, 2, 3), (0, -1, 0), 0.4))
    st = mtlp.read_state(path)
    assert st["counts"] == [1000] and st["scene"] == 3 and not st["batch"]
    assert bytes(st["params"][0]) == bytes(p)
    np.testing.assert_array_equal(st["x"], x)
    np.testing.assert_array_equal(st["v"], v)
    assert st["sphere"] == ((1.0, 2.0, 3.0), (0.0, -1.0, 0.0), pytest.approx(0.4))
    raw = path.read_bytes()
    assert raw[:8] == b"MTLPSTAT" and len(raw) == 60 + 96 + 32 * 1000
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"NOTSTATE" + raw[8:])
    with pytest.raises(ValueError):
        mtlp.read_state(bad)


def _bits(a):
    return a.view(np.uint32)


@pytest.mark.gpu
def test_engine_save_load_continue(tmp_path, mtlp):
    """Save mid-run, load into a fresh engine, continue: bit-identical."""
    p = mtlp.scene_params(4, 20000)
    x, v = mtlp.generate_scene(4, 20000, 1, p)
    path = tmp_path / "mid.bin"
    with mtlp.Engine(0, p) as a:
        a.reset(20000, mtlp.capi.SCENE_EMPTY)
        a.upload_state(x, v)
        for _ in range(3):
            a.step(DT)
        a.save_state(path)
        for _ in range(3):
            a.step(DT)
        ref = a.positions_buffer().read(), a.velocities_buffer().read()
    st = mtlp.read_state(path)
    assert st["counts"] == [20000] and bytes(st["params"][0]) == bytes(p)
    with mtlp.Engine(0) as b:  # default params and no particles: the file sets both
        b.load_state(path)
        assert b.particle_count == 20000 and bytes(b.params) == bytes(p)
        for _ in range(3):
            b.step(DT)
        np.testing.assert_array_equal(_bits(b.positions_buffer().read()), _bits(ref[0]))
        np.testing.assert_array_equal(_bits(b.velocities_buffer().read()), _bits(ref[1]))


@pytest.mark.gpu
def test_batch_and_sphere_state(tmp_path, mtlp):
    counts = [500, 1200, 64]
    params = [mtlp.scene_params(3, c) for c in counts]
    b = mtlp.Batch(counts, params)
    b.seed(3, 5)
    b.step(DT)
    path = tmp_path / "batch.bin"
    b.save_state(path)
    x0 = b.positions_buffer().read()
    st = mtlp.read_state(path)
    assert st["batch"] and st["counts"] == counts
    np.testing.assert_array_equal(_bits(st["x"]), _bits(x0))
    b2 = mtlp.Batch(counts, params)
    b2.load_state(path)
    np.testing.assert_array_equal(_bits(b2.positions_buffer().read()), _bits(x0))
    b3 = mtlp.Batch([500, 1264], params[:2])
    with pytest.raises(mtlp.MtlpError):
        b3.load_state(path)
    for e in (b, b2, b3):
        e.close()
    with mtlp.Engine(0) as e:
        e.reset(4000, mtlp.capi.SCENE_SPHERE_DROP)
        for _ in range(5):
            e.step(DT)
        e.save_state(tmp_path / "sph.bin")
        sph = e.sphere()
    with mtlp.Engine(0) as f:
        f.load_state(tmp_path / "sph.bin")
        assert f.sphere() == sph
