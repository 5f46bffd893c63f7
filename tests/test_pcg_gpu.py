"""The device block PCG against the oracle's restatement of the reference's
_pcg (core/solve.py:66-140) on the same S, right-hand sides, zero columns and
starting guesses -- the pattern of the reference's
tests/test_gpu.py:56-81."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class System(tuple):
    """(S, dinv) with what normalising a solution needs."""


def _system(O, d):
    p0 = O.pass0(d, "y", ["x1", "x2"], ["worker_id", "firm_id", "year"])
    c = O.pass1(p0)
    nl = p0["n_levels"]
    comp, _ = O.rank_components(O.components(c, nl[0]), c["cnt"][:nl[0]])
    sysm = System((O.build_S(c), O.jacobi(O.diag(c))))
    sysm.norm = lambda X: O.normalize(X, comp, c["obs"], nl)
    return sysm


@pytest.fixture(scope="module")
def system(oracle_mod, small_panel):
    return _system(oracle_mod, small_panel)


@pytest.fixture(scope="module")
def long_system(oracle_mod):
    """Year rows with more nonzeros than one workgroup takes (split rows)."""
    from hdfe_stream_amd.simulate import akm_panel
    return _system(oracle_mod, akm_panel(60_000, 9_000, seed=3))


def _agree(sysm, Xd, Xc, precision, tol):
    """FP64: the same iterates, so raw solutions agree.  Mixed: solutions of
    the singular S agree up to its null space, which the reference's
    normalisation (core/solve.py:383-399) removes."""
    if precision == "mixed":
        Xd, Xc = sysm.norm(Xd), sysm.norm(Xc)
    return np.max(np.abs(Xd - Xc)) < tol * np.max(np.abs(Xc))


def _rhs(S, k, seed=3):
    rng = np.random.default_rng(seed)
    return S @ rng.standard_normal((S.shape[0], k))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_pcg_matches_the_oracle_pcg(oracle_mod, system, precision):
    from hdfe_stream_amd.rocm import explicit_solver
    S, dinv = system
    b = _rhs(S, 4)
    zero = np.array([False, True, False, False])
    b[:, 1] = 0.0
    x0 = np.zeros_like(b)
    x0[:, 2] = np.linalg.lstsq(S.toarray(), b[:, 2], rcond=None)[0]   # already solved
    solve, name = explicit_solver(S, dinv, 1e-11, 5000, precision=precision)
    assert name
    for start in (None, x0):
        X_cpu, it_cpu, ok_cpu, _ = oracle_mod.pcg(S, dinv, b, zero, start, 1e-11, 5000)
        X_dev, it_dev, ok_dev, rel = solve(b, zero, start)
        assert ok_cpu and ok_dev
        assert rel <= 1e-11
        if precision == "fp64":
            assert abs(it_cpu - it_dev) <= 1
        assert np.all(X_dev[:, 1] == 0.0)
        assert _agree(system, X_dev, X_cpu, precision, 1e-8)


@pytest.mark.parametrize("k", [1, 3, 8, 20, 256])
def test_device_spmm_is_the_operator(system, k):
    from hdfe_stream_amd.rocm import DeviceCSR
    S, dinv = system
    A = DeviceCSR.from_scipy(S, dinv)
    X = np.random.default_rng(k).standard_normal((S.shape[0], k))
    np.testing.assert_allclose(A.matmat(X), S @ X, rtol=1e-12, atol=1e-12 * np.abs(S.data).max())


@pytest.mark.parametrize("k", [1, 3, 40, 256])
def test_split_rows(long_system, k):
    """Split rows (the years, dense: each couples to most levels); for wide k
    they are taken by column in one pass over P."""
    from hdfe_stream_amd.rocm import DeviceCSR
    S, dinv = long_system
    A = DeviceCSR.from_scipy(S, dinv)
    assert A.nlong() > 0
    X = np.random.default_rng(k).standard_normal((S.shape[0], k))
    ref = S @ X
    np.testing.assert_allclose(A.matmat(X), ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())
    got32 = A.matmat_f32(X.astype(np.float32))
    assert np.max(np.abs(got32 - ref)) <= 1e-5 * np.abs(ref).max()
    np.testing.assert_array_equal(A.matmat(X), A.matmat(X))


@pytest.mark.parametrize("k", [3, 64])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_split_rows_solve(oracle_mod, long_system, precision, k):
    from hdfe_stream_amd.rocm import DeviceCSR
    S, dinv = long_system
    A = DeviceCSR.from_scipy(S, dinv)
    b = _rhs(S, k, seed=5)
    Xc, itc, okc, _ = oracle_mod.pcg(S, dinv, b, None, None, 1e-10, 5000)
    Xd, itd, okd, rel = A.solve(b, tol=1e-10, precision=precision)
    assert okc and okd and rel <= 1e-10
    # the device's solution meets the tolerance in the exact FP64 operator
    r = np.linalg.norm(b - S @ Xd, axis=0) / np.linalg.norm(b, axis=0)
    assert r.max() <= 1.5e-10
    if precision == "fp64":
        assert abs(itc - itd) <= 1
    # FP64 follows the oracle's iterates; mixed reaches another point of the
    # same residual ball, and on this system (years coupled to every level:
    # condition ~1e4 on range(S)) two such points can differ by ~1e-6 at 64
    # columns
    assert _agree(long_system, Xd, Xc, precision, 1e-7 if precision == "fp64" or k <= 8 else 1e-6)


@pytest.mark.parametrize("k", [16, 256])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_wide_blocks(oracle_mod, system, k, precision):
    from hdfe_stream_amd.rocm import DeviceCSR
    S, dinv = system
    A = DeviceCSR.from_scipy(S, dinv)
    b = _rhs(S, k, seed=k)
    Xc, itc, okc, _ = oracle_mod.pcg(S, dinv, b, None, None, 1e-10, 5000)
    Xd, itd, okd, rel = A.solve(b, tol=1e-10, precision=precision)
    assert okc and okd and rel <= 1e-10
    if precision == "fp64":
        assert abs(itc - itd) <= 1
    assert _agree(system, Xd, Xc, precision, 1e-7)


def test_all_zero_and_maxiter(system):
    from hdfe_stream_amd.rocm import DeviceCSR
    S, dinv = system
    A = DeviceCSR.from_scipy(S, dinv)
    b = _rhs(S, 2)
    X, it, ok, rel = A.solve(b, zero=[True, True])
    assert it == 0 and ok and np.all(X == 0)
    X, it, ok, rel = A.solve(b, tol=1e-14, maxiter=3)
    assert it == 3 and not ok and rel > 1e-14


def test_device_tensors_in_and_out(system):
    import torch
    from hdfe_stream_amd.rocm import DeviceCSR
    S, dinv = system
    A = DeviceCSR.from_scipy(S, dinv)
    b = _rhs(S, 3)
    Xh, ith, _, _ = A.solve(b, tol=1e-10)
    Xd, itd, ok, _ = A.solve(torch.from_numpy(b).cuda(), tol=1e-10)
    assert ok and itd == ith
    np.testing.assert_array_equal(Xd.cpu().numpy(), Xh)


@pytest.fixture(scope="module")
def big_system():
    """A symmetric, diagonally dominant S at the 5M-worker panel's size
    (400,008 levels): ~20 random couplings per level and 8 levels coupled to
    every other (rows of 400,000 nonzeros), so that P at k >= 1 outgrows an
    XCD's L2 and the small-k SpMM takes its column-blocked form."""
    import scipy.sparse as sp
    rng = np.random.default_rng(7)
    Ls, Ld = 400_000, 8
    L = Ls + Ld
    r = np.repeat(np.arange(Ls), 10)
    c = rng.integers(0, Ls, len(r))
    v = -rng.uniform(0.1, 1.0, len(r))
    rd = np.repeat(np.arange(Ls), Ld)
    cd = Ls + np.tile(np.arange(Ld), Ls)
    vd = -rng.uniform(0.01, 0.1, len(rd))
    # the long rows coupled among themselves too (as years are)
    ra, ca = np.triu_indices(Ld, 1)
    va = -rng.uniform(1.0, 5.0, len(ra))
    B = sp.csr_matrix((np.concatenate([v, vd, va]),
                       (np.concatenate([r, rd, Ls + ra]), np.concatenate([c, cd, Ls + ca]))),
                      shape=(L, L))
    S = (B + B.T).tocsr()
    S.setdiag(0.0)
    S.eliminate_zeros()
    S = (S + sp.diags(np.asarray(abs(S).sum(axis=1)).ravel() + 1.0)).tocsr()
    S.sort_indices()
    dinv = 1.0 / S.diagonal()
    return S, dinv


@pytest.mark.parametrize("k", [1, 3, 8])
def test_blocked_spmm_is_the_operator(big_system, k):
    from hdfe_stream_amd.rocm import DeviceCSR
    S, dinv = big_system
    A = DeviceCSR.from_scipy(S, dinv)
    assert A.nlong() == 8
    X = np.random.default_rng(k).standard_normal((S.shape[0], k))
    ref = S @ X
    np.testing.assert_allclose(A.matmat(X), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    got32 = A.matmat_f32(X.astype(np.float32))
    assert np.max(np.abs(got32 - ref)) <= 1e-5 * np.abs(ref).max()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_blocked_spmm_solve(oracle_mod, big_system, precision):
    from hdfe_stream_amd.rocm import DeviceCSR
    S, dinv = big_system
    A = DeviceCSR.from_scipy(S, dinv)
    b = _rhs(S, 3, seed=9)
    Xc, itc, okc, _ = oracle_mod.pcg(S, dinv, b, None, None, 1e-10, 5000)
    Xd, itd, okd, rel = A.solve(b, tol=1e-10, precision=precision)
    assert okc and okd and rel <= 1e-10
    if precision == "fp64":
        assert abs(itc - itd) <= 1
    assert np.max(np.abs(Xd - Xc)) <= 1e-8 * np.max(np.abs(Xc))


@pytest.mark.parametrize("k", [1, 3])
def test_autotuned_launch_shape_leaves_results_bit_identical(big_system, k):
    """The grouped form's row parts per XCD are chosen by timing; every choice
    partitions the same per-row sums, so the products are bit-identical."""
    from hdfe_stream_amd.rocm import DeviceCSR
    S, dinv = big_system
    A = DeviceCSR.from_scipy(S, dinv)
    A.small_form(k, "fp64", set=1)
    A.small_form(k, "mixed", set=1)
    X = np.random.default_rng(k).standard_normal((S.shape[0], k))
    first = A.matmat(X)
    tuned = A.blk_parts(k, "fp64")
    assert tuned in (64, 128, 256, 512, 1024)
    for parts in (64, 1024, 100):
        A.blk_parts(k, "fp64", set=parts)
        np.testing.assert_array_equal(A.matmat(X), first)
        A.blk_parts(k, "mixed", set=parts)
        a32 = A.matmat_f32(X.astype(np.float32))
        A.blk_parts(k, "mixed", set=128)
        np.testing.assert_array_equal(A.matmat_f32(X.astype(np.float32)), a32)


@pytest.mark.parametrize("k", [1, 3, 9])
def test_many_split_rows(k):
    """More split rows (20 dense levels, each coupled to every other level)
    than the finishing workgroup has waves: each split row's sum is then
    written by its own wave."""
    import scipy.sparse as sp
    from hdfe_stream_amd.rocm import DeviceCSR
    rng = np.random.default_rng(31)
    Ls, Ld = 10_000, 20
    r = np.repeat(np.arange(Ls), 6)
    c = rng.integers(0, Ls, len(r))
    rd = np.repeat(np.arange(Ls), Ld)
    cd = Ls + np.tile(np.arange(Ld), Ls)
    B = sp.csr_matrix((np.concatenate([-rng.uniform(0.1, 1, len(r)), -rng.uniform(0.01, 0.1, len(rd))]),
                       (np.concatenate([r, rd]), np.concatenate([c, cd]))), shape=(Ls + Ld, Ls + Ld))
    S = (B + B.T).tocsr()
    S.setdiag(0.0)
    S.eliminate_zeros()
    S = (S + sp.diags(np.asarray(abs(S).sum(axis=1)).ravel() + 1.0)).tocsr()
    S.sort_indices()
    A = DeviceCSR.from_scipy(S, 1.0 / S.diagonal())
    assert A.nlong() == Ld
    X = rng.standard_normal((S.shape[0], k))
    ref = S @ X
    np.testing.assert_allclose(A.matmat(X), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    b = ref
    Xd, _, ok, rel = A.solve(b, tol=1e-10)
    assert ok and rel <= 1e-10


@pytest.mark.parametrize("k", [1, 3, 8])
def test_tiled_and_grouped_forms(big_system, k):
    """The column-blocked product's two forms (groups of blocks with partial
    products and a combine; tiles of rows streamed block by block) are both
    S X, each the same bits on every call; the default is the tiled form,
    whose 16-bit block-local indices give the same bits as its 32-bit ones
    (form 3)."""
    from hdfe_stream_amd.rocm import DeviceCSR
    S, dinv = big_system
    A = DeviceCSR.from_scipy(S, dinv)
    X = np.random.default_rng(20 + k).standard_normal((S.shape[0], k))
    ref = S @ X
    out = {}
    for form in (1, 2, 3):
        A.small_form(k, "fp64", set=form)
        A.small_form(k, "mixed", set=form)
        out[form] = (A.matmat(X), A.matmat_f32(X.astype(np.float32)))
        np.testing.assert_allclose(out[form][0], ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
        assert np.max(np.abs(out[form][1] - ref)) <= 1e-5 * np.abs(ref).max()
        np.testing.assert_array_equal(A.matmat(X), out[form][0])
        np.testing.assert_array_equal(A.matmat_f32(X.astype(np.float32)), out[form][1])
    np.testing.assert_allclose(out[1][0], out[2][0], rtol=1e-13, atol=1e-13 * np.abs(ref).max())
    np.testing.assert_array_equal(out[3][0], out[2][0])
    np.testing.assert_array_equal(out[3][1], out[2][1])
    A.small_form(k, "fp64", set=0)
    A.small_form(k, "mixed", set=0)
    np.testing.assert_array_equal(A.matmat(X), out[2][0])
    np.testing.assert_array_equal(A.matmat_f32(X.astype(np.float32)), out[2][1])


@pytest.mark.parametrize("form", [1, 2])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_blocked_forms_solve(oracle_mod, big_system, form, precision):
    from hdfe_stream_amd.rocm import DeviceCSR
    S, dinv = big_system
    A = DeviceCSR.from_scipy(S, dinv)
    A.small_form(3, "fp64", set=form)
    A.small_form(3, "mixed", set=form)
    b = _rhs(S, 3, seed=19)
    Xc, itc, okc, _ = oracle_mod.pcg(S, dinv, b, None, None, 1e-10, 5000)
    Xd, itd, okd, rel = A.solve(b, tol=1e-10, precision=precision)
    assert okc and okd and rel <= 1e-10
    if precision == "fp64":
        assert abs(itc - itd) <= 1
    assert np.max(np.abs(Xd - Xc)) <= 1e-8 * np.max(np.abs(Xc))


@pytest.mark.parametrize("size", ["small", "big"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replays_match_the_eager_solve(system, big_system, size, precision):
    """Batches of CG iterations run eagerly on a key's first use, are captured
    as a graph on the second and replayed after, on a small system: every
    solve, with eager work and synchronisations between them, is
    bit-identical to the eager one.  The 400,008-level system runs its first
    two solves eagerly (a fit's matrix is solved once), then captures and
    replays its batches too, with the same bits on every solve."""
    import torch
    from hdfe_stream_amd.rocm import DeviceCSR
    S, dinv = (system[0], system[1]) if size == "small" else big_system
    A = DeviceCSR.from_scipy(S, dinv)
    b = _rhs(S, 3, seed=4)
    X = np.random.default_rng(2).standard_normal((S.shape[0], 3))
    outs, its = [], []
    for r in range(4 if size == "small" else 6):   # eager, capture, replay, ...
        Xs, it, ok, _ = A.solve(b, tol=1e-10, precision=precision)
        assert ok
        outs.append(Xs)
        its.append(it)
        A.matmat(X)                       # eager work between the replays
        torch.cuda.synchronize()
    for Xs, it in zip(outs[1:], its[1:]):
        np.testing.assert_array_equal(Xs, outs[0])
        assert it == its[0]
    assert A.graphs() > 0
    # the eager path with per-launch timing gives the same bits again
    A.timing(True)
    Xt, itt, _, _ = A.solve(b, tol=1e-10, precision=precision)
    A.timing(False)
    np.testing.assert_array_equal(Xt, outs[0])
    # another tolerance on the same graphs: the device-side parameters (the
    # mixed solve's FP32 inner tolerance may already reach it in one step)
    X8, it8, ok8, rel8 = A.solve(b, tol=1e-6, precision=precision)
    assert ok8 and rel8 <= 1e-6
    if precision == "fp64":
        assert it8 < its[0]


@pytest.mark.parametrize("k", [9, 64, 301, 1024])
def test_wide_spmm_forms_agree_with_scipy(k):
    """Both forms of the wide product (a wave per row; column slabs per XCD,
    whose pAp partials are bounded by the work buffer at k = 1024) against
    scipy's S @ X, with a split (long) row in the matrix."""
    import scipy.sparse as sp
    from hdfe_stream_amd.rocm import DeviceCSR
    rng = np.random.default_rng(k)
    L = 6000
    A = sp.random(L, L, density=0.005, random_state=k, format="csr")
    A = A + A.T
    A = A.tolil()
    A[7, :] = rng.standard_normal(L)          # one row longer than the split threshold
    A[:, 7] = A[7, :].T
    S = sp.csr_matrix(A) + sp.identity(L) * 50.0
    S.sort_indices()
    M = DeviceCSR.from_scipy(S, 1.0 / S.diagonal())
    X = rng.standard_normal((L, k))
    want = S @ X
    M.matmat(X)                                 # first use with k: the autotuner runs
    for mode in (0, 1):
        print(f"wide spmm k={k} form {mode}", flush=True)
        assert M.wide_mode(k, "fp64", set=mode) == mode
        got = M.matmat(X)
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max(), mode
        assert M.wide_mode(k, "mixed", set=mode) == mode
        got32 = M.matmat_f32(X.astype(np.float32))
        assert np.abs(got32 - want).max() <= 1e-5 * np.abs(want).max(), mode
