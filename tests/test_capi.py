"""The C ABI of libmtlp.so without a GPU: every symbol include/mtlp.h declares
is exported, the host-only entry points work, and errors come back as status
codes (the reference aborts instead, Engine.cpp:69-85, 407-413, 849-854)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, "include", "mtlp.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mtlp_[a-z_0-9]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(mtlp):
    lib = mtlp.capi.lib()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(mtlp.capi.SIGNATURES), set(names) ^ set(mtlp.capi.SIGNATURES)


def test_params_struct_layout(mtlp):
    # 9 vec3 floats + 12 floats + 2 u32 = 23 * 4 bytes, no padding
    assert ctypes.sizeof(mtlp.Params) == 92


def test_default_params_match_reference(mtlp):
    p = mtlp.default_params()
    assert list(p.bounds_min) == [-3, -3, -3] and list(p.bounds_max) == [3, 6, 3]
    assert np.float32(p.h) == np.float32(0.1) and p.rest_density == 8000 and p.epsilon == 800
    assert p.iterations == 5 and p.scorr_n == 4 and np.float32(p.scorr_k) == np.float32(1e-4)
    assert np.float32(p.max_step) == np.float32(0.3) and np.float32(p.damping) == np.float32(0.996)
    assert mtlp.grid_dims(p) == (60, 90, 60)


def test_scene_params_workloads(mtlp):
    dam = mtlp.scene_params(4, 1_000_000)
    assert mtlp.grid_dims(dam) == (102, 102, 51)
    assert np.float32(dam.h) == np.float32(0.02) and dam.iterations == 4
    assert dam.rest_density == pytest.approx(6378 * 125, rel=1e-6)
    cube = mtlp.scene_params(3, 32768)
    assert mtlp.grid_dims(cube) == (39, 39, 39)
    assert cube.bounds_max[0] == pytest.approx(3.875) and cube.iterations == 3
    assert np.float32(cube.scorr_k) == np.float32(0.1) and cube.epsilon == 600


def test_generate_reference_cube_drop(mtlp):
    """Engine.cpp:530-548: side = ceil(cbrt(n)), y from 2 downwards, spacing h/2."""
    p = mtlp.default_params()
    x, v = mtlp.generate_scene(0, 1000, 0, p)
    assert (v == 0).all() and (x[:, 3] == 1).all()
    half = np.float32(0.5) * np.float32(9) * np.float32(0.05)
    assert x[0, 0] == -half and x[0, 1] == np.float32(2.0) and x[0, 2] == -half
    assert x[1, 0] == -half + np.float32(1) * np.float32(0.05)
    assert x[10, 2] == -half + np.float32(0.05)          # z advances after a row of 10
    assert x[100, 1] == np.float32(2.0) - np.float32(0.05)  # y after a 10 x 10 layer


def test_generate_reference_dam_break(mtlp):
    """Engine.cpp:463-491: 110 layers, square base, hugging the -x wall."""
    p = mtlp.default_params()
    n = 11000
    x, _ = mtlp.generate_scene(1, n, 0, p)
    side = 10  # ceil(sqrt(ceil(11000 / 110)))
    assert x[0, 0] == np.float32(-3.0) + np.float32(0.05) and x[0, 1] == np.float32(-2.95)
    assert x[side * side, 1] == np.float32(-2.95) + np.float32(0.05)
    assert x[:, 1].max() < 6.0


def test_generate_jittered_workloads(mtlp):
    n = 32768
    p = mtlp.scene_params(3, n)
    x, v = mtlp.generate_scene(3, n, 1, p)
    lattice = np.float32(0.775) + np.arange(32, dtype=np.float32) * np.float32(0.05)
    j = np.abs(x[:32, 0] - lattice)
    assert j.max() <= 0.01 * 0.1 * 1.0001 and j.max() > 0
    x2, _ = mtlp.generate_scene(3, n, 1, p)
    np.testing.assert_array_equal(x, x2)  # seeded
    x3, _ = mtlp.generate_scene(3, n, 2, p)
    assert not np.array_equal(x, x3)
    pd = mtlp.scene_params(4, 200_000)
    xd, _ = mtlp.generate_scene(4, 200_000, 1, pd)
    lo, hi = np.array(pd.bounds_min), np.array(pd.bounds_max)
    assert (xd[:, :3] >= lo - 0.01 * 0.02).all() and (xd[:, :3] <= hi).all()
    assert xd[:, 0].max() < 0.31 * hi[0]


def test_error_codes_without_gpu(mtlp):
    C = mtlp.capi
    L = C.lib()
    # no engine / bad arguments
    assert L.mtlp_step(None, None, ctypes.c_float(0.01)) == 1
    assert L.mtlp_destroy(None) == 1
    assert L.mtlp_get_buffer(None, 0, None, None) == 1
    assert L.mtlp_particle_count(None) == 0
    assert L.mtlp_exclusive_scan(None, None, None, None, None, 0) == 0   # n == 0: no-op
    assert L.mtlp_exclusive_scan(None, None, None, None, None, 10) == 1  # NULL buffers
    assert L.mtlp_exclusive_scan(None, None, 0x1004, 0x2000, 0x3000, 10) == 1  # misaligned
    assert L.mtlp_exclusive_scan(None, None, 0x1000, 0x2000, 0x3000, 0x80000000) == 2
    assert L.mtlp_scan_scratch_bytes(16384) == 8 + 16 and L.mtlp_scan_scratch_bytes(16385) == 16 + 16  # 16,384-element tiles
    # the allocation-free scan: limits checked before any device call
    assert L.mtlp_exclusive_scan_noalloc(None, None, None, 0) == 0              # n == 0: no-op
    assert L.mtlp_exclusive_scan_noalloc(None, 0x1000, 0x2000, (1 << 28) + 1) == 2  # above its maximum
    assert b"MTLP_NOALLOC_SCAN_MAX" in L.mtlp_last_error()
    assert L.mtlp_exclusive_scan_noalloc(None, None, None, 10) == 1             # NULL buffers
    assert L.mtlp_exclusive_scan_noalloc(None, 0x1004, 0x2000, 10) == 1         # misaligned
    # invalid parameters
    p = C.default_params()
    p.h = 0.0
    dims = (ctypes.c_uint32 * 3)()
    assert L.mtlp_grid_dims(ctypes.byref(p), dims) == 1
    p = C.default_params()
    p.cell_size = 1e-6
    assert L.mtlp_grid_dims(ctypes.byref(p), dims) == 2
    p = C.default_params()
    p.bounds_max[1] = p.bounds_min[1]
    assert L.mtlp_grid_dims(ctypes.byref(p), dims) == 1
    assert L.mtlp_scene_params(99, 10, ctypes.byref(p)) == 1
    assert b"scene" in L.mtlp_last_error()
    # creating an engine needs a device: a status code, not an abort
    h = ctypes.c_void_p()
    st = L.mtlp_create(0, None, ctypes.byref(h))
    if not _has_gpu():
        assert st in (1, 4)  # MTLP_E_INVALID or MTLP_E_HIP
        assert h.value is None
    elif st == 0:
        L.mtlp_destroy(h)


def _has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def test_package_fails_loudly_without_library(mtlp, monkeypatch, tmp_path):
    import importlib
    capi = importlib.import_module("mtlphys_amd._capi")
    monkeypatch.setattr(capi, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(capi, "_lib", None)
    with pytest.raises(RuntimeError, match="missing"):
        capi.lib()
