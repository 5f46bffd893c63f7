"""The OLS batch (`feols_batch`): many uneven linear models with fixed
effects fitted at once on the device, each against the oracle's FP64 fit of
that model alone (the reference's feols_stream restated): slopes to 1e-9,
residual sums of squares, rows kept after singletons; weighted and not.  The
wide case (32 models x 300 covariates) against the oracle's answers kept in
tests/golden/c5_ols.npz."""

from pathlib import Path

import numpy as np
import pytest

from conftest import gpu_available, rel

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

FE = ["worker_id", "firm_id", "year"]


def _oracle(O, d, weights=None):
    data = {"y": d["y"], "worker_id": d["worker_id"], "firm_id": d["firm_id"],
            "year": d["year"]}
    xs = []
    for j in range(d["X"].shape[1]):
        data[f"x{j}"] = d["X"][:, j].astype(np.float64)
        xs.append(f"x{j}")
    if weights is not None:
        data["w"] = d[weights]
    return O.fit(data, "y", xs, FE, weights=None if weights is None else "w", tol=1e-12)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("weighted", [False, True])
def test_ols_batch_matches_the_oracle(oracle_mod, precision, weighted):
    from hdfe_stream_amd.glm import feols_batch
    from hdfe_stream_amd.simulate import ols_models
    ds = [ols_models(1, lo=n, hi=n + 1, n_cov=6, seed=40 + i)[0]
          for i, n in enumerate([3000, 12000, 700])]
    if weighted:
        for i, d in enumerate(ds):
            d["w"] = np.random.default_rng(i).uniform(0.5, 2.0, len(d["y"]))
    print(f"feols_batch precision={precision} weighted={weighted}", flush=True)
    res = feols_batch("y ~ X | worker_id + firm_id + year", ds, tol=1e-12,
                      precision=precision, weights="w" if weighted else None,
                      check_residuals=True)
    for d, r in zip(ds, res):
        ref = _oracle(oracle_mod, d, "w" if weighted else None)
        assert r.converged
        assert r.n_obs == len(ref["resid"])
        assert rel(r.beta, ref["beta"]) < 1e-9, rel(r.beta, ref["beta"])
        e = ref["resid"]
        w = ref["p0"]["w"] if weighted else np.ones(len(e))
        assert r.deviance == pytest.approx(float(w @ (e * e)), rel=1e-9)
        assert max(r.diagnostics["model_rel_residual"]) <= 2e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_wide_ols_batch_matches_the_oracle(precision):
    from hdfe_stream_amd.glm import feols_batch
    from hdfe_stream_amd.simulate import ols_models
    path = Path(__file__).parent / "golden" / "c5_ols.npz"
    if not path.exists():
        pytest.skip("tests/golden/c5_ols.npz not generated")
    G = np.load(path)
    n_models, lo, hi, n_cov, seed = (int(v) for v in G["spec"])
    ds = ols_models(n_models, lo=lo, hi=hi, n_cov=n_cov, seed=seed)
    print(f"wide feols_batch: {n_models} x {n_cov}, {precision}", flush=True)
    res = feols_batch("y ~ X | worker_id + firm_id + year", ds, tol=1e-12, precision=precision)
    for m, r in enumerate(res):
        assert r.converged, m
        assert r.n_obs == G["kept"][m], m
        assert rel(r.beta, G["beta"][m]) < 1e-9, (m, rel(r.beta, G["beta"][m]))
        assert r.deviance == pytest.approx(G["rss"][m], rel=1e-9), m
