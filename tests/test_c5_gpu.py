"""The wide uneven Poisson batch: 32 models of 300 float32 covariates with
worker, firm and year effects, fitted as one `fepois_batch`.

* 32 models of 5,000 to 50,000 rows against the oracle's IRLS on each model
  alone (its answers kept in tests/golden/c5_poisson.npz: the oracle needs
  minutes for them), in FP64 and mixed precision, with every model's own
  relative residual within the tolerance after every solve;
* the configuration's own size (rows log-uniform in [50,000, 2,000,000],
  18.7M rows): at the fixed point every model's Poisson score X'(y - mu)
  vanishes."""

from pathlib import Path

import numpy as np
import pytest

from conftest import gpu_available, rel

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

GOLDEN = np.load(Path(__file__).parent / "golden" / "c5_poisson.npz")
FML = "y ~ X | worker_id + firm_id + year"


@pytest.mark.parametrize("precision,solver", [("fp64", "explicit"), ("mixed", "explicit"),
                                              ("fp64", "stream_cg"), ("mixed", "stream_cg")])
def test_c5_batch_matches_the_oracle_model_by_model(precision, solver):
    from hdfe_stream_amd.glm import fepois_batch
    from hdfe_stream_amd.simulate import poisson_models
    n_models, lo, hi, n_cov, seed = (int(v) for v in GOLDEN["spec"])
    ds = poisson_models(n_models, lo=lo, hi=hi, n_cov=n_cov, seed=seed)
    assert [len(d["y"]) for d in ds] == GOLDEN["rows"].tolist()
    print(f"C5 batch of {n_models} x {n_cov} covariates, {precision}, {solver}", flush=True)
    tol = 1e-10
    res = fepois_batch(FML, ds, iwls_tol=1e-10, iwls_maxiter=50, tol=tol, precision=precision,
                       check_residuals=True, solver=solver)
    for m, r in enumerate(res):
        assert r.converged, m
        assert r.n_obs == GOLDEN["kept"][m], m
        assert rel(r.beta, GOLDEN["beta"][m]) < 1e-7, (m, rel(r.beta, GOLDEN["beta"][m]))
        assert r.deviance == pytest.approx(GOLDEN["deviance"][m], rel=1e-9), m
        assert max(r.diagnostics["model_rel_residual"]) <= 2 * tol, m


def test_c5_full_size_score_vanishes():
    import torch
    from hdfe_stream_amd.glm import fepois_batch
    from hdfe_stream_amd.simulate import poisson_models_device
    print("C5 full size: 32 models, 50k-2M rows, 300 covariates", flush=True)
    ds = poisson_models_device(32, 50_000, 2_000_000, 300, seed=5)
    assert 18_000_000 < sum(len(d["y"]) for d in ds) < 20_000_000
    res = fepois_batch(FML, ds, iwls_tol=1e-8, precision="mixed")
    for m, (d, r) in enumerate(zip(ds, res)):
        assert r.converged, m
        eta = r.eta_source
        keep = ~torch.isnan(eta)
        mu = torch.exp(eta[keep])
        y = d["y"][keep].double()
        score = torch.zeros(d["X"].shape[1], dtype=torch.float64, device="cuda")
        scale = torch.zeros_like(score)
        X = d["X"][keep]
        for a in range(0, X.shape[0], 500_000):
            Xc = X[a:a + 500_000].double()
            score += Xc.T @ (y[a:a + 500_000] - mu[a:a + 500_000])
            scale += Xc.abs().T @ y[a:a + 500_000]
        worst = float((score.abs() / scale).max())
        assert worst <= 1e-6, (m, worst)
