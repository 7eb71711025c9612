"""The allocation-free scan (mtlp_exclusive_scan_noalloc) against the
allocating one (mtlp_exclusive_scan, caller-owned scratch) through the C ABI:
bit-identical outputs on calls in a row, each smaller than the one before,
the first at the largest length the allocation-free scan accepts (its
working memory then holds exactly as many look-back words as tiles)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scan_alloc(mtlp, d_in, n):
    import torch
    d_out = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    scratch = torch.empty(mtlp.Engine.scan_scratch_bytes(n), dtype=torch.uint8, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    mtlp.capi.check(mtlp.capi.lib().mtlp_exclusive_scan(None, st, d_in.data_ptr(), d_out.data_ptr(),
                                                        scratch.data_ptr(), n), "mtlp_exclusive_scan")
    return d_out


def _scan_noalloc(mtlp, d_in, n):
    import torch
    d_out = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    mtlp.capi.check(mtlp.capi.lib().mtlp_exclusive_scan_noalloc(st, d_in.data_ptr(), d_out.data_ptr(), n),
                    "mtlp_exclusive_scan_noalloc")
    return d_out


def _ref_scan(a):
    out = np.zeros(len(a), np.uint64)
    out[1:] = np.cumsum(a[:-1].astype(np.uint64))
    return (out & 0xFFFFFFFF).astype(np.uint32)


def test_noalloc_scan_matches_allocating_scan(mtlp):
    import torch
    max_n = 1 << 28
    sizes = [max_n, 2_097_153, 530_605, 4097]
    print("noalloc scan sizes", sizes, flush=True)
    for k, n in enumerate(sizes):
        g = torch.Generator(device="cuda")
        g.manual_seed(100 + k)
        d_in = torch.randint(0, 8, (n,), dtype=torch.int32, device="cuda", generator=g)
        got = _scan_noalloc(mtlp, d_in, n)
        ref = _scan_alloc(mtlp, d_in, n)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), f"call {k}: n = {n}"
        if n <= 2_097_153:
            np.testing.assert_array_equal(got.cpu().numpy().view(np.uint32),
                                          _ref_scan(d_in.cpu().numpy().view(np.uint32)), err_msg=str(n))
        else:
            # exclusive scan of the maximum length: last = total - last input
            assert int(got[-1].item()) == int(d_in.sum().item()) - int(d_in[-1].item())
        del d_in, got, ref
        torch.cuda.empty_cache()
