"""Variance matrices on the device against the dense sandwich of the full
design (a dummy per level of every fixed effect, then the covariates): by
Frisch-Waugh-Lovell the coefficient block of (Z'WZ)^+ Z'W Omega W Z (Z'WZ)^+
equals B^-1 M B^-1 with the partialled covariates, so the meats are pinned by
an independent computation; K's fixed-effect part is the rank of the dummy
matrix."""

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

FE = ["worker_id", "firm_id", "year"]
FML = "y ~ x1 + x2 | worker_id + firm_id + year"


def _dummies(col):
    u, inv = np.unique(col, return_inverse=True)
    D = np.zeros((len(inv), len(u)))
    D[np.arange(len(inv)), inv] = 1.0
    return D


def _dense(d, w, clusters):
    D = np.hstack([_dummies(d[f]) for f in FE])
    X = np.column_stack([d["x1"], d["x2"]])
    Z = np.hstack([D, X])
    Ainv = np.linalg.pinv(Z.T @ (w[:, None] * Z), hermitian=True)
    coef = Ainv @ (Z.T @ (w * d["y"]))
    e = d["y"] - Z @ coef
    k_fe = np.linalg.matrix_rank(D)
    sc = Z * (w * e)[:, None]
    out = {"e": e, "k_fe": k_fe, "hc": (Ainv @ (sc.T @ sc) @ Ainv)[-2:, -2:],
           "bread": Ainv[-2:, -2:]}
    for name, code in clusters.items():
        S = _dummies(code).T @ sc
        out[name] = (Ainv @ (S.T @ S) @ Ainv)[-2:, -2:]
    return out


@pytest.mark.parametrize("weighted", [False, True])
def test_vcov_matches_the_dense_sandwich(weighted):
    from hdfe_stream_amd.estimator import feols_stream
    from hdfe_stream_amd.simulate import akm_panel
    d = akm_panel(400, 40, seed=12)
    n = len(d["y"])
    rng = np.random.default_rng(5)
    w = rng.uniform(0.5, 2.0, n) if weighted else np.ones(n)
    if weighted:
        d["w"] = w
    d["region"] = d["firm_id"] % 7
    print(f"vcov weighted={weighted}", flush=True)
    res = feols_stream(FML, d, fixef_rm="none", tol=1e-13, weights="w" if weighted else None)
    wf = np.array([str(a) + "_" + str(b) for a, b in zip(d["worker_id"], d["firm_id"])])
    dn = _dense(d, w, {"worker": d["worker_id"], "firm": d["firm_id"], "wf": wf,
                       "region": d["region"]})
    k, N = 2, n
    K = k + dn["k_fe"]
    # iid
    V, det = res.vcov("iid")
    assert det["K"] == K
    rss = float(w @ dn["e"] ** 2)
    np.testing.assert_allclose(V, rss / (N - K) * dn["bread"], rtol=1e-8)
    # hetero
    V, det = res.vcov("hetero")
    np.testing.assert_allclose(V, N / (N - K) * dn["hc"], rtol=1e-8)
    # CRV1 by worker: the worker effects are nested in the clusters
    G = len(np.unique(d["worker_id"]))
    V, det = res.vcov("CRV1:worker_id")
    Kc = k + dn["k_fe"] - G + 1
    assert det["nested"] == ["worker_id"] and det["Kc"] == Kc
    np.testing.assert_allclose(V, G / (G - 1) * (N - 1) / (N - Kc) * dn["worker"], rtol=1e-8)
    # CRV1 by a column of the data (regions of firms: the firm effects nest)
    Gr = len(np.unique(d["region"]))
    V, det = res.vcov({"CRV1": "region"}, data=d)
    Kr = k + dn["k_fe"] - len(np.unique(d["firm_id"])) + 1
    assert det["nested"] == ["firm_id"]
    np.testing.assert_allclose(V, Gr / (Gr - 1) * (N - 1) / (N - Kr) * dn["region"], rtol=1e-8)
    # two-way: worker + firm, inclusion-exclusion
    V, det = res.vcov("CRV1:worker_id+firm_id")
    F = len(np.unique(d["firm_id"]))
    Gm = min(G, F)
    nest = set(det["nested"])
    assert nest == {"worker_id", "firm_id"}
    assert det["Kc"] == k + dn["k_fe"] - G - F + 2
    meat = dn["worker"] + dn["firm"] - dn["wf"]
    np.testing.assert_allclose(V, Gm / (Gm - 1) * (N - 1) / (N - det["Kc"]) * meat, rtol=1e-8)
    se = res.se("hetero")
    assert np.all(se > 0)
