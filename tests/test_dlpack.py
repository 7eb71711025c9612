"""DLPack views: the ctypes-built capsule on host memory (CPU), and
Engine.tensor over device buffers (GPU)."""
import gc

import numpy as np
import pytest


def test_capsule_views_host_memory():
    import torch
    from mtlphys_amd import dlpack as D
    a = np.arange(12, dtype=np.float32).reshape(3, 4)
    before = D.live_views()
    t = torch.utils.dlpack.from_dlpack(D.capsule(a.ctypes.data, a.shape, np.float32, D.kDLCPU, owner=a))
    assert t.shape == (3, 4) and t.dtype == torch.float32
    a[1, 2] = 99.0
    assert t[1, 2].item() == 99.0  # same memory
    u = np.arange(7, dtype=np.uint32)
    t2 = torch.utils.dlpack.from_dlpack(D.capsule(u.ctypes.data, u.shape, np.uint32, D.kDLCPU, owner=u))
    assert t2.dtype == torch.uint32 and t2.tolist() == list(range(7))
    assert D.live_views() == before + 2
    del t, t2
    gc.collect()
    assert D.live_views() == before


@pytest.mark.gpu
def test_engine_tensor_aliases_buffers(mtlp):
    import torch
    p = mtlp.scene_params(3, 5000)
    x, v = mtlp.generate_scene(3, 5000, 1, p)
    with mtlp.Engine(0, p) as e:
        e.reset(5000, mtlp.capi.SCENE_EMPTY)
        e.upload_state(x, v)
        tx = e.tensor(mtlp.capi.BUF_POSITIONS)
        tn = e.tensor(mtlp.capi.BUF_NEIGHBOR_COUNTS)
        assert tx.is_cuda and tx.shape == (5000, 4) and tn.dtype == torch.uint32
        np.testing.assert_array_equal(tx.cpu().numpy(), x)
        e.step(1.0 / 120.0)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(tx.cpu().numpy(), e.positions_buffer().read())  # updated in place
        np.testing.assert_array_equal(tn.cpu().numpy(), e.neighbor_counts_buffer().read())
        tv = e.tensor(mtlp.capi.BUF_VELOCITIES)
        tv.zero_()
        torch.cuda.synchronize()
        assert not e.velocities_buffer().read().any()  # torch writes reach the engine
        del tx, tn, tv
