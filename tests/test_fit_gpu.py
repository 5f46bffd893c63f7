"""The device pipeline (passes 0 and 1, diag, rhs, S, components,
normalisation, assembly, beta, pass 2) against the oracle's restatement of the
reference on the same panels.  Integer results (rows kept, codes, cells, the
sparsity of S, components) must be identical; floating-point ones agree to
the stated tolerances."""

import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

FE = ["worker_id", "firm_id", "year"]


def _panel_with_singletons(seed=7):
    from hdfe_stream_amd.simulate import akm_panel
    d = akm_panel(3000, 300, seed=seed)
    rng = np.random.default_rng(seed)
    # drop rows at random so some workers and firms have a single row
    keep = rng.random(len(d["y"])) > 0.55
    return {k: v[keep] for k, v in d.items()}


def _check_fit(O, d, fe=FE, xs=("x1", "x2"), weights=None, precision="fp64", tol=1e-10,
               beta_tol=1e-8):
    from hdfe_stream_amd.estimator import StreamingHDFE
    ref = O.fit(d, "y", list(xs), list(fe), weights=weights, tol=tol)
    est = StreamingHDFE("y", list(xs), list(fe), tol=tol, weights=weights, precision=precision)
    res = est.fit(d)
    P = res.panel
    p0, c = ref["p0"], ref["cells"]
    # integer facts
    assert P.info["n"] == p0["n_obs"]
    assert P.info["C"] == c["n_cells"]
    assert P.info["M"] == len(c["n"])
    assert P.info["Gi"] == len(c["starts"]) - 1
    assert P.n_levels == p0["n_levels"]
    np.testing.assert_array_equal(P.array("igstart", np.int32), c["starts"])
    D = P.info["D"]
    ic = P.array("icodes", np.int32).reshape(D, -1).T
    np.testing.assert_array_equal(ic, c["codes"])
    np.testing.assert_allclose(P.array("in", np.float64), c["n"], rtol=1e-13)
    np.testing.assert_allclose(P.array("cnt", np.float64), c["cnt"], rtol=0, atol=0)
    np.testing.assert_allclose(P.array("obs", np.float64), c["obs"], rtol=1e-12)
    # the system
    np.testing.assert_allclose(P.array("diag", np.float64), ref["diag"], rtol=1e-11,
                               atol=1e-11 * np.abs(ref["diag"]).max())
    rhs = P.array("rhs", np.float64).reshape(ref["rhs"].shape)
    assert np.abs(rhs - ref["rhs"]).max() <= 1e-10 * np.abs(ref["rhs"]).max()
    S, dinv = P.matrix.export()
    Sr = ref["S"]
    np.testing.assert_array_equal(S.indptr, Sr.indptr)
    np.testing.assert_array_equal(S.indices, Sr.indices)
    assert np.abs(S.data - Sr.data).max() <= 1e-11 * np.abs(Sr.data).max()
    # components: identical labels
    np.testing.assert_array_equal(P.array("comp", np.int32), ref["comp"])
    assert res.n_components == ref["n_components"]
    # the solve and everything after it
    if precision == "fp64":
        its = res.solver_info["iterations"]
        its = its if isinstance(its, list) else [its]
        assert all(abs(a - b) <= 1 for a, b in zip(its, ref["iterations"]))
    assert res.solver_info["converged"]
    G = res.Gamma.cpu().numpy()
    assert np.abs(G - ref["Gamma"]).max() <= 1e-8 * np.abs(ref["Gamma"]).max()
    assert rel(res.beta, ref["beta"]) < beta_tol
    assert rel(res.A, ref["A"]) < 1e-9
    fe_order = res.fe_names
    gy = np.concatenate([res.fe_values[f] for f in fe_order[1:]])
    assert np.abs(gy - ref["gam_y"]).max() <= 1e-8 * max(np.abs(ref["gam_y"]).max(), 1)
    np.testing.assert_allclose(res.fe_values[fe_order[0]], ref["worker_effects"], rtol=0,
                               atol=1e-8)
    np.testing.assert_allclose(res.resid(), ref["resid_source_order"], rtol=0, atol=1e-8)
    return res, ref


def test_small_panel_fit(oracle_mod, small_panel):
    _check_fit(oracle_mod, small_panel)


def test_akm_20k_panel_fit(oracle_mod, c1_panel):
    res, ref = _check_fit(oracle_mod, c1_panel)
    # the largest connected set: component 0's firms, identical
    comp = res.panel.array("comp", np.int32)
    np.testing.assert_array_equal(np.flatnonzero(comp == 0), np.flatnonzero(ref["comp"] == 0))


def test_akm_20k_panel_mixed_precision(oracle_mod, c1_panel):
    _check_fit(oracle_mod, c1_panel, precision="mixed")


def test_singletons_are_dropped_like_the_reference(oracle_mod):
    d = _panel_with_singletons()
    res, ref = _check_fit(oracle_mod, d)
    assert res.panel.info["singletons"] == len(d["y"]) - ref["p0"]["n_obs"] > 0


def test_weighted_fit(oracle_mod):
    from hdfe_stream_amd.simulate import weighted_panel
    d = weighted_panel(3000, 300, seed=4)
    d["w"] = d["w"].astype(np.float64)
    _check_fit(oracle_mod, d, weights="w")


def test_one_other_dimension(oracle_mod, small_panel):
    _check_fit(oracle_mod, small_panel, fe=["worker_id", "firm_id"])


def test_three_other_dimensions(oracle_mod):
    from hdfe_stream_amd.simulate import akm_panel
    d = akm_panel(3000, 150, seed=9)
    d["region"] = d["firm_id"] % 37
    d["qtr"] = (d["year"] * 7 + d["worker_id"]) % 4
    _check_fit(oracle_mod, d, fe=["worker_id", "firm_id", "year", "region"])
    _check_fit(oracle_mod, d, fe=["year", "worker_id", "qtr", "firm_id"])


def test_many_sparse_rows_and_split_rows(oracle_mod):
    from hdfe_stream_amd.simulate import akm_panel
    d = akm_panel(60_000, 9_000, seed=3)
    res, ref = _check_fit(oracle_mod, d)
    assert res.panel.matrix.nlong() > 0


def test_rows_in_group_order_and_shuffled_build_the_same_panel():
    """Rows stored by worker take the within-group ranking; shuffled rows the
    radix sort.  Both must give the same sorted panel: the same cells, codes,
    system and fit (only duplicate cells' sums may round apart)."""
    from hdfe_stream_amd.estimator import StreamingHDFE
    from hdfe_stream_amd.simulate import akm_panel
    d = akm_panel(4000, 300, seed=12)
    d["year"] = d["year"] % 3            # repeat (worker, firm, year) cells
    perm = np.random.default_rng(1).permutation(len(d["y"]))
    shuffled = {k: v[perm] for k, v in d.items()}
    print("ordered vs shuffled rows", flush=True)
    fits = []
    for data in (d, shuffled):
        est = StreamingHDFE("y", ["x1", "x2"], FE, tol=1e-12, precision="fp64")
        fits.append(est.fit(data))
    a, b = fits
    for name in ("igstart", "icodes", "cstart"):
        np.testing.assert_array_equal(a.panel.array(name, np.int32), b.panel.array(name, np.int32))
    np.testing.assert_allclose(a.panel.array("diag", np.float64), b.panel.array("diag", np.float64),
                               rtol=1e-13)
    assert rel(a.beta, b.beta) < 1e-11
    for f in a.fe_names:
        np.testing.assert_allclose(a.fe_values[f], b.fe_values[f], rtol=0, atol=1e-9)


def test_a_wide_dense_dimension_on_short_groups(oracle_mod):
    """Forty levels of a dense dimension on groups of three rows: the dense
    levels' (level, column) slots outgrow the one-pass kernel's LDS, and the
    general path sums them."""
    from hdfe_stream_amd.simulate import akm_panel
    d = akm_panel(3000, 200, seed=21, n_years=3)
    d["month"] = np.random.default_rng(4).integers(0, 40, len(d["y"]))
    print("fit with a 40-level dense dimension", flush=True)
    _check_fit(oracle_mod, d, fe=["worker_id", "firm_id", "month"])


@pytest.mark.parametrize("weighted", [False, True])
def test_twelve_covariates(oracle_mod, weighted):
    """Thirteen variables (y and twelve regressors): the within-cell scatter
    and the between-cell assembly go by 8 x 8 tiles, the dense levels by the
    general path.  Repeated (worker, firm, year) cells give the within-cell
    part weight."""
    from hdfe_stream_amd.simulate import akm_panel
    d = akm_panel(3000, 250, seed=31)
    d["year"] = d["year"] % 3
    rng = np.random.default_rng(5)
    n = len(d["y"])
    xs = []
    for j in range(12):
        x = rng.standard_normal(n) + 0.3 * np.sin(d["firm_id"] * (j + 1.0)) + 0.1 * j * d["year"]
        d[f"z{j}"] = x
        d["y"] = d["y"] + 0.05 * (j - 5.5) * x
        xs.append(f"z{j}")
    if weighted:
        d["w"] = rng.uniform(0.5, 2.0, n)
    print(f"fit with twelve covariates, weighted={weighted}", flush=True)
    _check_fit(oracle_mod, d, xs=xs, weights="w" if weighted else None)


def test_level_values_contiguous_and_gapped(oracle_mod):
    """Each dimension's level values: the sorted distinct ids of the rows
    kept (worker ids filling a range, firm ids spaced by three, years with a
    gap)."""
    d = _panel_with_singletons(seed=13)
    d["firm_id"] = d["firm_id"] * 3 + 7
    d["year"] = np.where(d["year"] >= 4, d["year"] + 10, d["year"])
    res, ref = _check_fit(oracle_mod, d)
    keep = ref["p0"]["keep"]
    for f in FE:
        np.testing.assert_array_equal(res.fe_levels[f], np.unique(d[f][keep]))
        assert res.fe_levels[f].dtype == np.int64


def test_fit_after_the_memory_cache_is_released(oracle_mod, c1_panel):
    """The library's cached blocks go back to the driver (the backend's
    release()), and the next fit allocates afresh with the same result."""
    from hdfe_stream_amd import _lib
    from hdfe_stream_amd.estimator import feols_stream
    fml = "y ~ x1 + x2 | worker_id + firm_id + year"
    a = feols_stream(fml, c1_panel, tol=1e-10)
    del a
    freed = _lib.release_memory()
    assert freed >= 0
    b = feols_stream(fml, c1_panel, tol=1e-10)
    ref = oracle_mod.fit(c1_panel, tol=1e-10)
    assert rel(b.beta, ref["beta"]) < 1e-8
