"""Device Poisson IRLS against the oracle's restatement of the reference's
IRLS, model by model in one batch; the split-BF16 MFMA product against the
same three products in FP64 (which also pins the matrix cores' lane layout)."""

import ctypes as C

import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

FE = ["worker_id", "firm_id", "year"]


def test_split_bf16_syrk_matches_fp64():
    import torch
    from hdfe_stream_amd import _lib
    rng = np.random.default_rng(0)
    Cp, kx = 96, 70
    n_m = [400, 1000, 37]
    pstart = np.concatenate(([0], np.cumsum([(n + 15) // 16 * 16 for n in n_m]))).astype(np.int64)
    npad = int(pstart[-1])
    S = np.zeros((npad, Cp), np.float32)
    for m, n in enumerate(n_m):
        S[pstart[m]:pstart[m] + n, :kx] = rng.standard_normal((n, kx)).astype(np.float32) * (1 + m)
    St = torch.from_numpy(S)
    hi = St.to(torch.bfloat16)
    lo = (St - hi.float()).to(torch.bfloat16)
    H, Lo = hi.double().numpy(), lo.double().numpy()
    A = torch.empty((len(n_m), Cp, Cp), dtype=torch.float64, device="cuda")
    hi_d, lo_d = hi.contiguous().cuda(), lo.contiguous().cuda()   # held: the launch reads them
    torch.cuda.synchronize()
    _lib.check(_lib.load().hdfe_glm_syrk(len(n_m), pstart.ctypes.data, Cp, hi_d.data_ptr(),
                                         lo_d.data_ptr(), 256, A.data_ptr()), "hdfe_glm_syrk")
    A = A.cpu().numpy()
    for m in range(len(n_m)):
        r = slice(int(pstart[m]), int(pstart[m + 1]))
        E = H[r].T @ H[r] + H[r].T @ Lo[r] + Lo[r].T @ H[r]
        assert np.abs(A[m] - E).max() <= 2e-6 * np.abs(E).max()
        np.testing.assert_array_equal(A[m], A[m].T)
        # the three products are within the split's rounding of S'S itself
        F = S[r].astype(np.float64).T @ S[r].astype(np.float64)
        assert np.abs(A[m] - F).max() <= 1e-4 * np.abs(F).max()


def _models(sizes, kx=4, seed=5):
    from hdfe_stream_amd.simulate import poisson_models
    return [poisson_models(1, lo=n, hi=n + 1, n_cov=kx, seed=seed + i)[0]
            for i, n in enumerate(sizes)]


def _oracle(O, d, tol=1e-10):
    kx = d["X"].shape[1]
    data = {"y": d["y"], "worker_id": d["worker_id"], "firm_id": d["firm_id"], "year": d["year"]}
    xs = []
    for j in range(kx):
        data[f"x{j}"] = d["X"][:, j].astype(np.float64)
        xs.append(f"x{j}")
    return O.fepois(data, "y", xs, FE, iwls_tol=tol, iwls_maxiter=50)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fepois_batch_matches_the_oracle(oracle_mod, precision):
    from hdfe_stream_amd.glm import fepois_batch
    O = oracle_mod
    ds = _models([3000, 9000, 5000])
    print(f"fepois batch of {len(ds)}, {precision}", flush=True)
    res = fepois_batch("y ~ X | worker_id + firm_id + year", ds, iwls_tol=1e-10,
                       iwls_maxiter=50, tol=1e-12, precision=precision)
    for d, r in zip(ds, res):
        ref = _oracle(O, d)
        assert r.converged and ref["converged"]
        assert r.n_obs == int(ref["keep"].sum())
        assert rel(r.beta, ref["beta"]) < 1e-7
        assert r.deviance == pytest.approx(ref["deviance"], rel=1e-9)
        eta = r.eta_source.cpu().numpy()
        assert np.all(np.isnan(eta[~ref["keep"]]))
        assert rel(eta[ref["keep"]], ref["eta"]) < 1e-7
        # the IRLS fixed point: the Poisson score of the covariates is zero
        y = ref["y"]
        mu = np.exp(eta[ref["keep"]])
        X = d["X"][ref["keep"]].astype(np.float64)
        assert np.all(np.abs(X.T @ (y - mu)) <= 1e-6 * (np.abs(X).T @ y))


def test_separated_levels_are_dropped(oracle_mod):
    from hdfe_stream_amd.glm import fepois_batch
    O = oracle_mod
    d = _models([6000], seed=11)[0]
    # a handful of workers with an outcome of zero on every row
    w = d["worker_id"]
    zero = np.isin(w, np.unique(w)[:25])
    d["y"] = np.where(zero, 0.0, d["y"])
    r = fepois_batch("y ~ X | worker_id + firm_id + year", [d], iwls_tol=1e-10,
                     iwls_maxiter=50, tol=1e-12, precision="fp64")[0]
    ref = _oracle(O, d)
    assert r.diagnostics["separation"]["levels"].get("worker_id", 0) >= 25
    assert r.n_obs == int(ref["keep"].sum())
    assert rel(r.beta, ref["beta"]) < 1e-7


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_every_model_meets_its_own_tolerance(oracle_mod, precision):
    """A small model whose right-hand sides are a thousandth of a large
    one's, batched with it: each model's own relative residual
    ||b_m - S_m x_m|| / ||b_m|| meets tol after every IRLS solve (not only
    the stacked one), and each model's slopes equal the oracle's IRLS on
    that model alone."""
    from hdfe_stream_amd.glm import fepois_batch
    O = oracle_mod
    ds = _models([40000, 800], seed=21)
    ds[1]["X"] = (ds[1]["X"] * 1e-3).astype(np.float32)
    print(f"fepois batch, uneven norms, {precision}", flush=True)
    tol = 1e-10
    res = fepois_batch("y ~ X | worker_id + firm_id + year", ds, iwls_tol=1e-10,
                       iwls_maxiter=50, tol=tol, precision=precision, check_residuals=True)
    for d, r in zip(ds, res):
        per_solve = r.diagnostics["model_rel_residual"]
        assert per_solve and max(per_solve) <= 2 * tol, per_solve
        ref = _oracle(O, d)
        assert r.converged and ref["converged"]
        assert rel(r.beta, ref["beta"]) < 1e-7
        assert r.deviance == pytest.approx(ref["deviance"], rel=1e-9)
