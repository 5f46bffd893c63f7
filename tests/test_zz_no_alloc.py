"""hdfe_tall_dot_noalloc, the allocation-free variant of hdfe_tall_dot:
through the C ABI, its result equals the allocating entry point's bit for bit
on two calls in a row (the second smaller, so stale partials in the shared
scratch array would show), both agree with a plain FP64 product, and a call
whose partials exceed the scratch array is refused before any launch."""

import ctypes as C

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu


def _call(lib, name, K, a, b, A, B, w, out, stream=None):
    args = [K, a, b, A.data_ptr(), A.stride(0), B.data_ptr(), B.stride(0),
            w.data_ptr() if w is not None else None, out.data_ptr()]
    if name.endswith("noalloc"):
        args.append(stream)
    return getattr(lib, name)(*args)


@pytest.mark.skipif(not gpu_available(), reason="needs a GPU")
def test_noalloc_matches_the_allocating_entry_point():
    import torch
    from hdfe_stream_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(7)
    stream = torch.cuda.current_stream().cuda_stream
    # (rows, a, b, weighted): the second call is smaller than the first
    for K, a, b, weighted in [(3_000_000, 3, 5, True), (70_001, 2, 3, False), (1, 1, 1, False)]:
        print(f"tall_dot noalloc: K={K} a={a} b={b} weighted={weighted}", flush=True)
        A = torch.from_numpy(rng.standard_normal((K, a))).cuda()
        B = torch.from_numpy(rng.standard_normal((K, b))).cuda()
        w = torch.from_numpy(rng.uniform(0.5, 2.0, K)).cuda() if weighted else None
        ref = torch.empty((a, b), dtype=torch.float64, device="cuda")
        got = torch.full((a, b), np.nan, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        _lib.check(_call(lib, "hdfe_tall_dot", K, a, b, A, B, w, ref), "hdfe_tall_dot")
        _lib.check(_call(lib, "hdfe_tall_dot_noalloc", K, a, b, A, B, w, got, stream),
                   "hdfe_tall_dot_noalloc")
        torch.cuda.synchronize()
        r, g = ref.cpu().numpy(), got.cpu().numpy()
        assert np.array_equal(r, g), (K, a, b, np.abs(r - g).max())
        An, Bn = A.cpu().numpy(), B.cpu().numpy()
        dense = (An * (w.cpu().numpy()[:, None] if weighted else 1.0)).T @ Bn
        assert np.allclose(g, dense, rtol=1e-12, atol=1e-12 * max(1.0, np.abs(dense).max()))


@pytest.mark.skipif(not gpu_available(), reason="needs a GPU")
def test_noalloc_refuses_partials_beyond_its_scratch_array():
    import torch
    from hdfe_stream_amd import _lib
    lib = _lib.load()
    # 17 workgroups' partials of 256 x 256 pairs = 1,114,112 doubles > 1,048,576
    K, a, b = 17 * 4096, 256, 256
    A = torch.zeros((K, a), dtype=torch.float64, device="cuda")
    out = torch.zeros((a, b), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    st = _call(lib, "hdfe_tall_dot_noalloc", K, a, b, A, A, None, out)
    assert st == 1
    assert "scratch" in lib.hdfe_last_error().decode()
    # within the array: 16 workgroups' partials of 256 x 256 pairs
    K = 16 * 4096
    A = A[:K]
    ref = torch.zeros_like(out)
    _lib.check(_call(lib, "hdfe_tall_dot_noalloc", K, a, b, A, A, None, out), "noalloc")
    _lib.check(_call(lib, "hdfe_tall_dot", K, a, b, A, A, None, ref), "tall_dot")
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
