"""Pinning the oracle's Poisson IRLS: at its fixed point the Poisson score
equations hold -- X'(y - mu) = 0 and, for every level of every fixed effect,
sum (y - mu) = 0 -- which is what the maximum-likelihood estimate is; the
separation check drops exactly the all-zero levels."""

import numpy as np


def _model(n=3000, seed=2):
    from hdfe_stream_amd.simulate import poisson_models
    return poisson_models(1, lo=n, hi=n + 1, n_cov=4, seed=seed)[0]


def test_irls_fixed_point_solves_the_score_equations(oracle_mod):
    O = oracle_mod
    d = _model()
    data = {"y": d["y"], "worker_id": d["worker_id"], "firm_id": d["firm_id"],
            "year": d["year"]}
    xs = []
    for j in range(d["X"].shape[1]):
        data[f"x{j}"] = d["X"][:, j].astype(np.float64)
        xs.append(f"x{j}")
    r = O.fepois(data, "y", xs, ["worker_id", "firm_id", "year"], iwls_tol=1e-12,
                 iwls_maxiter=50)
    assert r["converged"]
    y = r["y"]
    mu = np.exp(r["eta"])
    sub = r["data"]
    X = np.column_stack([sub[x] for x in xs])
    scale = np.abs(X).T @ y
    assert np.all(np.abs(X.T @ (y - mu)) <= 1e-7 * scale)
    for f in ("worker_id", "firm_id", "year"):
        u, inv = np.unique(sub[f], return_inverse=True)
        s = np.bincount(inv, weights=y - mu)
        tot = np.bincount(inv, weights=y)
        assert np.all(np.abs(s) <= 1e-7 * np.maximum(tot, 1.0))
    # the deviance path falls
    assert all(b <= a + 1e-9 * abs(a) for a, b in zip(r["path"][1:], r["path"][2:]))


def test_separation_drops_all_zero_levels(oracle_mod):
    O = oracle_mod
    y = np.array([0, 0, 1, 2, 0, 3, 0, 0], float)
    a = np.array([0, 0, 1, 1, 2, 2, 3, 3])      # level 0 and 3 all zero
    b = np.array([0, 1, 0, 1, 0, 1, 0, 1])
    keep = O.poisson_separation(y, [a, b])
    assert keep.tolist() == [False, False, True, True, True, True, False, False]
    assert O.poisson_separation(np.ones(4), [a[:4], b[:4]]).all()
