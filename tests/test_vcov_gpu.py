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


def _split_year_panel(n_workers=600, n_firms=60, seed=3):
    """Firms 0 .. F/2-1 hire only in years 0-3, the others only in 4-7: the
    (firm, year) pair then has two components while (worker, firm) has one,
    so fe_dof="exact" must take the largest pair's count."""
    rng = np.random.default_rng(seed)
    h = n_firms // 2
    wid = np.repeat(np.arange(n_workers), 8)
    year = np.tile(np.arange(8), n_workers)
    firm = np.where(year < 4, rng.integers(0, h, wid.size), rng.integers(h, n_firms, wid.size))
    x1, x2 = rng.standard_normal(wid.size), rng.standard_normal(wid.size)
    y = 0.3 * x1 - 0.2 * x2 + rng.standard_normal(wid.size)
    return {"worker_id": wid.astype(np.int64), "firm_id": firm.astype(np.int64),
            "year": year.astype(np.int64), "x1": x1, "x2": x2, "y": y}


def _components(a, b):
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    ua, ia = np.unique(a, return_inverse=True)
    ub, ib = np.unique(b, return_inverse=True)
    na, nb = len(ua), len(ub)
    g = coo_matrix((np.ones(len(ia)), (ia, na + ib)), shape=(na + nb, na + nb))
    return connected_components(g, directed=False)[0]


def test_pair_components_and_fe_dof():
    from hdfe_stream_amd.estimator import feols_stream
    d = _split_year_panel()
    print("pair components: split-year panel", flush=True)
    res = feols_stream(FML, d, fixef_rm="none", tol=1e-13)
    pc = res._vcov_engine.pair_components() if hasattr(res, "_vcov_engine") else None
    if pc is None:
        res.vcov("iid")
        pc = res._vcov_engine.pair_components()
    for (a, b), c in pc.items():
        assert c == _components(d[a], d[b]), (a, b)
    assert pc[("firm_id", "year")] == 2 and pc[("worker_id", "firm_id")] == 1
    D = np.hstack([_dummies(d[f]) for f in FE])
    total = D.shape[1]
    # exact: 2 components + one for the third dimension = the dummies' rank
    _, det = res.vcov("iid")
    assert det["K"] == 2 + total - 3 == 2 + np.linalg.matrix_rank(D)
    _, det = res.vcov("iid", fe_dof="pyfixest")
    assert det["K"] == 2 + total - 2
    assert feols_stream(FML, d, fixef_rm="none", fe_dof="pyfixest").vcov_details["K"] == \
        2 + total - 2


def _panel_pairs(cols, stream):
    """Every pair's components on the device for a panel of the columns
    `cols` (streamed: cols[stream]), with the columns in fit order."""
    import ctypes as C
    from hdfe_stream_amd import _lib
    from hdfe_stream_amd.estimator import Panel
    y = np.zeros(len(cols[0]))
    P = Panel(cols, [y], fixef_rm="none", stream_dim=stream)
    order = [cols[P.info["gdim"]]] + [cols[d] for d in P.dims]
    out = {}
    for a in range(len(order)):
        for b in range(a + 1, len(order)):
            v = C.c_int64()
            _lib.check(_lib.load().hdfe_panel_pair_components(P._h, a, b, C.byref(v)), "pairs")
            out[(a, b)] = int(v.value)
    return order, out


@pytest.mark.parametrize("case", ["few-level", "bitmap", "row-unions"])
def test_pair_components_every_path(case):
    """The three ways the device counts a pair's components: a side of at
    most 64 levels (adjacency of its levels), a modest bitmap of distinct
    edges, and unions per row when the bitmap would exceed 2^31 bits."""
    rng = np.random.default_rng({"few-level": 1, "bitmap": 2, "row-unions": 3}[case])
    nw = 25_000
    wid = np.repeat(np.arange(nw), 8)
    n = wid.size
    if case == "few-level":
        # years in two blocks never shared by a firm: (firm, year) has 2
        year = np.tile(np.arange(8), nw)
        firm = np.where(year < 4, rng.integers(0, 40, n), rng.integers(40, 80, n))
        region = rng.integers(0, 50, n)
        cols = [wid, firm, year, region]
    elif case == "bitmap":
        # levels above 64 on both sides, few enough for a bitmap; some
        # isolated blocks so that the counts exceed one
        firm = rng.integers(0, 300, n)
        spell = np.where(firm < 20, firm + 1000, rng.integers(0, 500, n))
        cols = [wid, firm, spell]
    else:
        a = rng.integers(0, 50_000, n)
        b = rng.integers(0, 50_000, n)
        cols = [wid, a, b]
    cols = [c.astype(np.int64) for c in cols]
    print(f"pair components: {case}", flush=True)
    order, got = _panel_pairs(cols, 0)
    for (a, b), c in got.items():
        assert c == _components(order[a], order[b]), (case, a, b, c)
