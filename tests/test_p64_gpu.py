"""64-bit row pointers on the device: S with 2^31 nonzeros or more (the
reference hands the solver int64 index arrays then, core/solve.py:169,
188-189).  The widened pointers of a small S run the same sums as its 32-bit
ones (bit for bit, split rows included) and agree with scipy's product and the
oracle's PCG; a banded S of 2,149,577,726 nonzeros against its closed form."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(O, d):
    p0 = O.pass0(d, "y", ["x1", "x2"], ["worker_id", "firm_id", "year"])
    c = O.pass1(p0)
    return O.build_S(c), O.jacobi(O.diag(c))


@pytest.fixture(scope="module")
def split_system(oracle_mod):
    """30,000 workers and 5,000 firms: the year rows hold more nonzeros than
    one workgroup takes (split rows), and P of eight columns fits an XCD's L2
    (the plain CSR form in both widths)."""
    from hdfe_stream_amd.simulate import akm_panel
    return _system(oracle_mod, akm_panel(30_000, 5_000, seed=5))


def test_widened_row_pointers_change_no_bit(split_system, oracle_mod):
    from hdfe_stream_amd.rocm import DeviceCSR
    S, dinv = split_system
    A32 = DeviceCSR.from_scipy(S, dinv)
    A64 = DeviceCSR.from_scipy(S, dinv)
    assert A32.row_pointer_bits() == 32
    assert A64.row_pointer_bits(64) == 64 and A64.row_pointer_bits() == 64
    assert A64.nlong() > 0
    rng = np.random.default_rng(11)
    for k in (1, 3, 8):
        X = rng.standard_normal((S.shape[0], k))
        print(f"p64 product k={k}", flush=True)
        Y64 = A64.matmat(X)
        np.testing.assert_array_equal(Y64, A32.matmat(X))
        np.testing.assert_array_equal(A64.matmat_f32(X), A32.matmat_f32(X))
        np.testing.assert_allclose(Y64, S @ X, rtol=1e-12, atol=1e-12 * np.abs(S.data).max())
    for k in (12, 64):
        X = rng.standard_normal((S.shape[0], k))
        print(f"p64 product k={k}", flush=True)
        np.testing.assert_allclose(A64.matmat(X), S @ X, rtol=1e-12,
                                   atol=1e-12 * np.abs(S.data).max())
    # the export gives back the same S
    S2, d2 = A64.export()
    np.testing.assert_array_equal(S2.indptr, S.indptr)
    np.testing.assert_array_equal(S2.indices, S.indices)
    np.testing.assert_array_equal(S2.data, S.data)
    b = S @ rng.standard_normal((S.shape[0], 3))
    for precision in ("fp64", "mixed"):
        print(f"p64 solve {precision}", flush=True)
        X32, it32, ok32, rel32 = A32.solve(b, tol=1e-11, precision=precision)
        X64, it64, ok64, rel64 = A64.solve(b, tol=1e-11, precision=precision)
        assert ok32 and ok64 and it32 == it64 and rel32 == rel64
        np.testing.assert_array_equal(X64, X32)
    Xc, _, okc, _ = oracle_mod.pcg(S, dinv, b, np.zeros(3, bool), None, 1e-11, 5000)
    X64, _, _, _ = A64.solve(b, tol=1e-11)
    assert okc
    assert np.max(np.abs(X64 - Xc)) < 1e-8 * np.max(np.abs(Xc))


def test_row_pointers_only_widen(split_system):
    from hdfe_stream_amd.rocm import DeviceCSR
    S, dinv = split_system
    A = DeviceCSR.from_scipy(S, dinv)
    A.row_pointer_bits(64)
    with pytest.raises(ValueError, match="only widen"):
        A.row_pointer_bits(32)


def test_more_than_2_31_nonzeros():
    """S = (4h + 2) I - B on L = 2^20 levels, B the ones of the band
    |i - j| <= h (h = 1,025) off the diagonal: 2,149,577,726 nonzeros.  With
    integer x every sum of S x is an integer below 2^24, so the device's
    FP64 and FP32 products must equal the closed form (4h + 3) x_i -
    sum_{|j - i| <= h} x_j exactly; a Jacobi-PCG solve of S x = S x_true
    returns x_true."""
    import torch
    from hdfe_stream_amd.rocm import DeviceCSR
    dev = torch.device("cuda", 0)
    L, h = 1 << 20, 1025
    diag = 4.0 * h + 2.0
    rows = np.arange(L)
    cnt = np.minimum(L - 1, rows + h) - np.maximum(0, rows - h) + 1
    indptr = np.zeros(L + 1, np.int64)
    np.cumsum(cnt, out=indptr[1:])
    nnz = int(indptr[-1])
    assert nnz == 2_149_577_726 and nnz >= 2**31
    print(f"p64 band: L={L} nnz={nnz}", flush=True)
    indices = torch.empty(nnz, dtype=torch.int32, device=dev)
    data = torch.empty(nnz, dtype=torch.float64, device=dev)
    off = torch.arange(-h, h + 1, device=dev)
    R = 16384
    for r0 in range(0, L, R):
        r = torch.arange(r0, r0 + R, device=dev)
        c = r[:, None] + off[None, :]
        m = (c >= 0) & (c < L)
        p0, p1 = int(indptr[r0]), int(indptr[r0 + R])
        indices[p0:p1] = c[m].to(torch.int32)
        data[p0:p1] = torch.where(c == r[:, None], diag, -1.0)[m].to(torch.float64)
        del c, m
    dinv = torch.full((L,), 1.0 / diag, dtype=torch.float64, device=dev)
    ip = torch.from_numpy(indptr).to(dev)
    torch.cuda.synchronize()
    A = DeviceCSR.from_device(L, ip, indices, data, dinv)
    del indices, data, ip
    torch.cuda.empty_cache()
    assert A.nnz == nnz and A.row_pointer_bits() == 64

    def closed_form(x):
        cs = np.zeros((L + 1, x.shape[1]))
        np.cumsum(x, axis=0, out=cs[1:])
        hi = np.minimum(L - 1, rows + h) + 1
        lo = np.maximum(0, rows - h)
        return (diag + 1.0) * x - (cs[hi] - cs[lo])

    base = ((rows * 7) % 11 - 5).astype(np.float64)
    for k in (1, 3, 12):
        X = np.stack([((rows * (3 + 2 * j)) % 11 - 5).astype(np.float64) for j in range(k)], 1)
        X[:, 0] = base
        want = closed_form(X)
        print(f"p64 band product k={k}", flush=True)
        np.testing.assert_array_equal(A.matmat(X), want)
        np.testing.assert_array_equal(A.matmat_f32(X), want.astype(np.float32))
    x_true = np.stack([base, -base], 1)
    b = closed_form(x_true)
    for precision in ("fp64", "mixed"):
        print(f"p64 band solve {precision}", flush=True)
        X, it, ok, rel = A.solve(b, tol=1e-10, precision=precision)
        assert ok and rel <= 1e-10 and it < 100
        assert np.max(np.abs(X - x_true)) < 1e-8 * np.max(np.abs(x_true))
    A.close()
    torch.cuda.empty_cache()
