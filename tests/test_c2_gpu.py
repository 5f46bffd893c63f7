"""The 5,000,000-worker AKM fit (40M rows, 400,000 firms, 8 years, p_move
0.10, seed 2) on the device, checked through properties that hold whatever
the size: the normal equations (every fixed-effect level's residual sum and
X'e vanish to 1e-8 of the right-hand side they come from), the
normalisation, and mixed precision against FP64 to 1e-8.  The same
generator's 625,000-worker panel is checked against the oracle's FP64 fit."""

import numpy as np
import pytest

from conftest import gpu_available, rel

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

FML = "y ~ x1 + x2 | worker_id + firm_id + year"


@pytest.fixture(scope="module")
def c2():
    import torch
    from hdfe_stream_amd.estimator import feols_stream
    from hdfe_stream_amd.simulate import akm_panel
    print("C2: 5,000,000 workers x 8 years", flush=True)
    d = akm_panel(5_000_000, 400_000, seed=2)
    cols = {k: torch.from_numpy(d[k]).cuda() for k in ("worker_id", "firm_id", "year",
                                                        "x1", "x2", "y")}
    fits = {p: feols_stream(FML, cols, precision=p, tol=1e-10) for p in ("fp64", "mixed")}
    return d, fits


def _worker_demeaned(d, v):
    w = d["worker_id"]
    cnt = np.bincount(w)
    return v - (np.bincount(w, weights=v) / cnt)[w]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_c2_normal_equations(c2, precision):
    d, fits = c2
    r = fits[precision]
    assert r.solver_info["converged"]
    assert r.n_obs == 40_000_000 and r.n_components == 1
    e = r.resid()
    assert np.all(np.isfinite(e))
    yt = _worker_demeaned(d, d["y"])
    # each worker's residuals sum to zero (its effect is their mean)
    sw = np.bincount(d["worker_id"], weights=e)
    assert np.abs(sw).max() <= 1e-10 * np.abs(d["y"]).max() * 8
    # each firm's and each year's residual sum: D_d'e against D_d'(y - ybar_worker)
    for dim in ("firm_id", "year"):
        levels, _ = r.fixef(dim)
        code = np.searchsorted(np.asarray(levels), d[dim])
        res_sum = np.bincount(code, weights=e)
        rhs = np.bincount(code, weights=yt)
        assert np.linalg.norm(res_sum) <= 1e-8 * np.linalg.norm(rhs), dim
    # X'e against X'(y - ybar_worker)
    for x in ("x1", "x2"):
        assert abs(d[x] @ e) <= 1e-8 * np.linalg.norm(d[x]) * np.linalg.norm(yt), x


def test_c2_normalisation(c2):
    """One connected component: the firm effects have a zero row-weighted
    mean, the year effects too (the reference's _normalize)."""
    d, fits = c2
    for r in fits.values():
        for dim in ("firm_id", "year"):
            levels, vals = r.fixef(dim)
            n = np.bincount(np.searchsorted(np.asarray(levels), d[dim]), minlength=len(levels))
            vals = np.asarray(vals)
            assert abs(n @ vals) <= 1e-10 * (n @ np.abs(vals)), dim


def test_c2_mixed_matches_fp64(c2):
    d, fits = c2
    a, b = fits["fp64"], fits["mixed"]
    assert rel(b.beta, a.beta) <= 1e-8
    for dim in ("worker_id", "firm_id", "year"):
        va, vb = np.asarray(a.fixef(dim)[1]), np.asarray(b.fixef(dim)[1])
        assert np.abs(va - vb).max() <= 1e-8 * np.abs(va).max(), dim
    assert rel(b.resid(), a.resid()) <= 1e-8


def test_c2_sample_matches_the_oracle(oracle_mod):
    """The same generator at 625,000 workers (5M rows, 50,000 firms): the
    device fit against the oracle's FP64 fit."""
    import torch
    from hdfe_stream_amd.estimator import feols_stream
    from hdfe_stream_amd.simulate import akm_panel
    print("C2 sample: 625,000 workers vs the oracle", flush=True)
    d = akm_panel(625_000, 50_000, seed=2)
    ref = oracle_mod.fit(d, tol=1e-10)
    cols = {k: torch.from_numpy(d[k]).cuda() for k in ("worker_id", "firm_id", "year",
                                                        "x1", "x2", "y")}
    for precision in ("fp64", "mixed"):
        r = feols_stream(FML, cols, precision=precision, tol=1e-10)
        assert rel(r.beta, ref["beta"]) < 1e-8, precision
        G = r.Gamma.cpu().numpy()
        assert np.abs(G - ref["Gamma"]).max() <= 1e-8 * np.abs(ref["Gamma"]).max(), precision


def test_c2_stream_cg_matches_the_explicit_fit(c2):
    """The matrix-free fallback at the full size: the same coefficients and
    effects as the explicit S."""
    import torch
    from hdfe_stream_amd.estimator import feols_stream
    d, fits = c2
    print("C2 stream_cg", flush=True)
    cols = {k: torch.from_numpy(d[k]).cuda() for k in ("worker_id", "firm_id", "year",
                                                        "x1", "x2", "y")}
    r = feols_stream(FML, cols, precision="fp64", tol=1e-10, solver="stream_cg")
    a = fits["fp64"]
    assert r.solver_info["solver"] == "stream_cg" and r.solver_info["converged"]
    assert rel(r.beta, a.beta) <= 1e-8
    for dim in ("firm_id", "year"):
        va, vb = np.asarray(a.fixef(dim)[1]), np.asarray(r.fixef(dim)[1])
        assert np.abs(va - vb).max() <= 1e-8 * np.abs(va).max(), dim
