"""`solve_normal_equations` / `predict` on the device against the dense
definition u = (Z'WZ)^+ Z'v on explicit dummies: the identified predictions
Z u and the covariate block to 1e-8, weighted and not; unknown levels raise
in the solve and give NaN in predict."""

import numpy as np
import pytest

from conftest import gpu_available, rel

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

FE = ["worker_id", "firm_id", "year"]


def _dense(d, w, V):
    cols = []
    for f in FE:
        u, inv = np.unique(d[f], return_inverse=True)
        D = np.zeros((len(inv), len(u)))
        D[np.arange(len(inv)), inv] = 1.0
        cols.append(D)
    Z = np.hstack(cols + [np.column_stack([d["x1"], d["x2"]])])
    A = Z.T @ (w[:, None] * Z)
    u = np.linalg.pinv(A, hermitian=True) @ (Z.T @ V)
    return Z @ u, u[-2:]


@pytest.mark.parametrize("weighted", [False, True])
def test_normal_equations_match_the_dense_solution(weighted):
    from hdfe_stream_amd.estimator import feols_stream
    from hdfe_stream_amd.simulate import akm_panel
    d = akm_panel(300, 30, seed=9)
    rng = np.random.default_rng(2)
    n = len(d["y"])
    d["v1"], d["v2"], d["v3"] = rng.standard_normal(n), rng.standard_normal(n), d["y"].copy()
    w = rng.uniform(0.5, 2.0, n) if weighted else np.ones(n)
    if weighted:
        d["w"] = w
    print(f"normal equations weighted={weighted}", flush=True)
    res = feols_stream("y ~ x1 + x2 | worker_id + firm_id + year", d, fixef_rm="none",
                       tol=1e-12, weights="w" if weighted else None)
    sol = res.solve_normal_equations(d, ["v1", "v2", "v3"], block=2, tol=1e-13)
    pred_dense, coef_dense = _dense(d, w, np.column_stack([d["v1"], d["v2"], d["v3"]]))
    pred = sol.predict(d).cpu().numpy()
    assert rel(pred, pred_dense) < 1e-8
    for j, nm in enumerate(["v1", "v2", "v3"]):
        assert rel(sol.coef[nm], coef_dense[:, j]) < 1e-8
    # v = W y gives the fit itself: its covariate block is beta
    if not weighted:
        np.testing.assert_allclose(sol.coef["v3"], res.beta, rtol=1e-8)
    lv, vals = sol.fixef("firm_id")
    assert vals.shape == (len(lv), 3)
    bad = {k: v.copy() for k, v in d.items()}
    bad["firm_id"][0] = 10_000_000
    with pytest.raises(ValueError, match="not in the estimation sample"):
        res.solve_normal_equations(bad, "v1")
    p = sol.predict(bad).cpu().numpy()
    assert np.isnan(p[0]).all() and np.isfinite(p[1:]).all()
