"""Device KSS (leaving out an observation) against the oracle's restatement:
the coefficient-space inverse, the Hutchinson bias draw for draw (the same
counter-based coefficient-space draws), and leave_out_kss end to end on a
pruned panel in the reference's canonical row order."""

import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

FE2 = ["worker_id", "firm_id"]
FE3 = ["worker_id", "firm_id", "year"]


def _loo(O, d):
    a, b = np.asarray(d["worker_id"]), np.asarray(d["firm_id"])
    ua, ia = np.unique(a, return_inverse=True)
    ub, ib = np.unique(b, return_inverse=True)
    alive, _ = O.leave_one_out(ia, ib, len(ua), len(ub))
    keep = alive[ia] & alive[len(ua) + ib]
    return {k: np.asarray(v)[keep] for k, v in d.items()}


def _fits(O, d, fe, tol=1e-12):
    from hdfe_stream_amd.estimator import feols_stream
    res = feols_stream("y ~ 1 | " + " + ".join(fe), d, stream=fe[0], fixef_rm="none",
                       keep_intermediates=True, tol=tol)
    ref = O.fit(d, "y", (), fe, tol=tol, fixef_rm="none", stream=fe[0])
    return res, ref


@pytest.mark.parametrize("fe", [FE2, FE3])
def test_apply_inverse_matches_the_oracle(oracle_mod, fe):
    import torch
    from hdfe_stream_amd.leaveout import CoefficientInverse
    from hdfe_stream_amd.simulate import akm_panel
    O = oracle_mod
    d = _loo(O, akm_panel(1500, 120, seed=6, p_move=0.2))
    res, ref = _fits(O, d, fe)
    print(f"apply_inverse fe={fe} rows={res.n_obs}", flush=True)
    inv = CoefficientInverse(res.panel, tol=1e-12, precision="fp64")
    ctx = O.KSSContext(ref, tol=1e-12)
    assert (inv.G, inv.L) == (ctx.G, ctx.L)
    rng = np.random.default_rng(3)
    vl = rng.standard_normal((ctx.L, 5))
    vg = rng.standard_normal((ctx.G, 5))
    Ul, Ug, rows = inv(torch.from_numpy(vl).cuda(), torch.from_numpy(vg).cuda())
    rUl, rUg, rrows = O.apply_inverse(ctx, vl, vg)
    assert rel(rows.cpu().numpy(), rrows) < 1e-9
    # both are Jacobi PCG from zero: the same solution, null-space part included
    assert rel(Ul.cpu().numpy(), rUl) < 1e-8
    assert rel(Ug.cpu().numpy(), rUg) < 1e-8
    # the input is not modified
    assert np.array_equal(vl, vl)


@pytest.mark.parametrize("fe", [FE2, FE3])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_hutchinson_matches_the_oracle(oracle_mod, fe, precision):
    import torch
    from hdfe_stream_amd.leaveout import hutchinson_trace
    from hdfe_stream_amd.simulate import akm_panel
    O = oracle_mod
    d = _loo(O, akm_panel(1500, 120, seed=8, p_move=0.2))
    res, ref = _fits(O, d, fe)
    n = res.n_obs
    print(f"hutchinson fe={fe} precision={precision} rows={n}", flush=True)
    sigma2 = np.random.default_rng(5).uniform(0.05, 0.4, n)
    tol = 1e-12 if precision == "fp64" else 1e-10
    tr = hutchinson_trace(res.panel, torch.from_numpy(sigma2).cuda(), n_draws=48, block=20,
                          seed=17, tol=tol, precision=precision)
    ctx = O.KSSContext(ref, tol=1e-12)
    (vp, va, vc), totals = O.hutchinson_trace(ctx, sigma2, 48, 20, seed=17)
    got = np.array([tr.var_psi, tr.var_alpha, tr.cov])
    want = np.array([vp, va, vc])
    bound = 1e-8 if precision == "fp64" else 1e-6
    assert np.abs(got - want).max() <= bound * np.abs(want).max()
    assert tr.diagnostics["converged"]
    # blocking does not change the estimate
    tr2 = hutchinson_trace(res.panel, torch.from_numpy(sigma2).cuda(), n_draws=48, block=48,
                           seed=17, tol=tol, precision=precision)
    got2 = np.array([tr2.var_psi, tr2.var_alpha, tr2.cov])
    assert np.abs(got2 - got).max() <= bound * np.abs(want).max()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_leave_out_kss_matches_the_oracle(oracle_mod, precision):
    from hdfe_stream_amd.leaveout import leave_out_kss
    from hdfe_stream_amd.simulate import kss_panel
    O = oracle_mod
    d = kss_panel(4000, 400, seed=3)
    print(f"leave_out_kss precision={precision}", flush=True)
    tol = 1e-12 if precision == "fp64" else 1e-10
    lo = leave_out_kss("y ~ 1 | worker_id + firm_id", d, n_draws=64, seed=2,
                       precision=precision, tol=tol, leave_out="observation")
    ref = O.leave_out_kss(d, n_draws=64, block=64, seed=2, tol=1e-12)
    # the pruning is integer and exact
    assert lo.n_obs == int(ref["keep"].sum()) == ref["n"]
    for k in ("components_before", "single_observation", "cut", "rounds"):
        assert lo.pruning[k] == ref["prune"][k]
    names = ["var(psi)", "var(alpha)", "cov(psi, alpha)"]
    plug = np.array([lo.plug_in[k] for k in names])
    assert rel(plug, ref["plug_in"]) < 1e-9
    bound = 1e-8 if precision == "fp64" else 1e-5
    assert rel([lo.bias[k] for k in names], ref["bias"]) < bound
    assert rel([lo.leave_out[k] for k in names], ref["leave_out"]) < bound
    s2 = lo.diagnostics["sigma2"].cpu().numpy()
    assert rel(s2, ref["sigma2"]) < (1e-8 if precision == "fp64" else 1e-5)
    # movers: workers seen at more than one firm, counted by rows
    g, f = ref["fit"]["p0"]["g"], ref["fit"]["p0"]["codes"][:, 0]
    lo_f = np.full(g.max() + 1, np.iinfo(np.int64).max)
    hi_f = np.full(g.max() + 1, -1)
    np.minimum.at(lo_f, g, f)
    np.maximum.at(hi_f, g, f)
    assert lo.n_movers == int((lo_f != hi_f)[g].sum())
    # the bias is removed: variances shrink, the covariance grows
    assert lo.leave_out["var(psi)"] < lo.plug_in["var(psi)"]
    assert lo.leave_out["var(alpha)"] < lo.plug_in["var(alpha)"]
    assert lo.leave_out["cov(psi, alpha)"] > lo.plug_in["cov(psi, alpha)"]


def test_unit_weights_give_the_unweighted_components():
    from hdfe_stream_amd.leaveout import leave_out_kss
    from hdfe_stream_amd.simulate import kss_panel
    d = kss_panel(3000, 300, seed=4)
    d["w"] = np.ones(len(d["y"]))
    a = leave_out_kss("y ~ 1 | worker_id + firm_id", d, n_draws=32, seed=1,
                      precision="fp64", tol=1e-12, leave_out="observation")
    b = leave_out_kss("y ~ 1 | worker_id + firm_id", d, n_draws=32, seed=1,
                      precision="fp64", tol=1e-12, weights="w", leave_out="observation")
    for k in a.leave_out:
        assert b.leave_out[k] == pytest.approx(a.leave_out[k], rel=1e-8, abs=1e-12)
        assert b.plug_in[k] == pytest.approx(a.plug_in[k], rel=1e-12, abs=1e-15)


# ------------------------------------------------ covariates in the design

def _cov_fits(O, d, weighted, tol=1e-12):
    from hdfe_stream_amd.estimator import feols_stream
    kw = {"weights": "w"} if weighted else {}
    res = feols_stream("y ~ x1 + x2 | worker_id + firm_id", d, stream="worker_id",
                       fixef_rm="none", keep_intermediates=True, tol=tol, **kw)
    ref = O.fit(d, "y", ("x1", "x2"), FE2, weights="w" if weighted else None, tol=tol,
                fixef_rm="none", stream="worker_id")
    return res, ref


@pytest.fixture(scope="module", params=[False, True], ids=["unweighted", "weighted"])
def cov_panel(oracle_mod, request):
    from hdfe_stream_amd.simulate import akm_panel
    O = oracle_mod
    d = akm_panel(1500, 120, seed=6, p_move=0.2)
    if request.param:
        d["w"] = np.random.default_rng(7).uniform(0.5, 2.0, len(d["y"]))
    d = _loo(O, d)
    res, ref = _cov_fits(O, d, request.param)
    return O, d, res, ref, request.param


def test_apply_inverse_with_covariates_matches_the_oracle(cov_panel):
    import torch
    from hdfe_stream_amd.leaveout import CoefficientInverse
    O, d, res, ref, weighted = cov_panel
    print(f"apply_inverse with covariates weighted={weighted}", flush=True)
    inv = CoefficientInverse(res.panel, tol=1e-12, precision="fp64", covariates=(1, 2))
    ctx = O.KSSContext(ref, tol=1e-12, X=ref["p0"]["V"][:, 1:])
    assert (inv.K, inv.kc) == (ctx.K, 2)
    rng = np.random.default_rng(3)
    vl = rng.standard_normal((ctx.K, 5))
    vg = rng.standard_normal((ctx.G, 5))
    Ul, Ug, rows = inv(torch.from_numpy(vl).cuda(), torch.from_numpy(vg).cuda())
    rUl, rUg, rrows = O.apply_inverse(ctx, vl, vg)
    assert Ul.shape == (ctx.K, 5)
    assert rel(rows.cpu().numpy(), rrows) < 1e-9
    assert rel(Ul.cpu().numpy(), rUl) < 1e-8
    assert rel(Ug.cpu().numpy(), rUg) < 1e-8


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_leverages_with_covariates_match_the_oracle(cov_panel, precision):
    from hdfe_stream_amd.leaveout import jla_leverages
    O, d, res, ref, weighted = cov_panel
    print(f"jla with covariates weighted={weighted} {precision}", flush=True)
    tol = 1e-12 if precision == "fp64" else 1e-10
    t = {"keep_raw": True}
    lev = jla_leverages(res.panel, n_draws=40, block=16, seed=4, tol=tol, precision=precision,
                        timings=t, covariates=(1, 2))
    p0, c = ref["p0"], ref["cells"]
    mom = O.jla_leverages(p0, c, ref["S"], ref["dinv"], 40, 16, 4, tol=1e-12,
                          X=p0["V"][:, 1:])
    bound = 1e-8 if precision == "fp64" else 1e-6
    assert rel(t["raw_moments"].cpu().numpy(), mom) < bound
    P, M, F = O.combine_moments(mom, 40)
    assert rel(lev.leverage.cpu().numpy(), P) < bound
    # the covariates enter the design: the leverages are not the fixed effects' alone
    lev0 = jla_leverages(res.panel, n_draws=40, block=16, seed=4, tol=tol, precision=precision)
    assert float((lev.leverage - lev0.leverage).abs().max()) > 1e-4


def test_hutchinson_with_covariates_matches_the_oracle(cov_panel):
    import torch
    from hdfe_stream_amd.leaveout import hutchinson_trace
    O, d, res, ref, weighted = cov_panel
    n = res.n_obs
    print(f"hutchinson with covariates weighted={weighted}", flush=True)
    sigma2 = np.random.default_rng(5).uniform(0.05, 0.4, n)
    tr = hutchinson_trace(res.panel, torch.from_numpy(sigma2).cuda(), n_draws=30, block=12,
                          seed=17, tol=1e-12, precision="fp64", covariates=(1, 2))
    ctx = O.KSSContext(ref, tol=1e-12, X=ref["p0"]["V"][:, 1:])
    (vp, va, vc), totals = O.hutchinson_trace(ctx, sigma2, 30, 12, seed=17)
    got = np.array([tr.var_psi, tr.var_alpha, tr.cov])
    want = np.array([vp, va, vc])
    assert np.abs(got - want).max() <= 1e-8 * np.abs(want).max()


def test_standard_errors_with_covariates_match_the_oracle(cov_panel):
    import torch
    from hdfe_stream_amd.leaveout import CoefficientInverse, JLALeverages
    from hdfe_stream_amd.leaveout_se import component_variances
    O, d, res, ref, weighted = cov_panel
    print(f"component_variances with covariates weighted={weighted}", flush=True)
    ctx = O.KSSContext(ref, tol=1e-12, X=ref["p0"]["V"][:, 1:])
    n = len(ctx.w)
    rng = np.random.default_rng(4)
    M_bar, factor = rng.uniform(0.3, 0.95, n), rng.uniform(0.9, 1.1, n)
    y = ref["p0"]["V"][:, 0]
    s2 = O.sigma2(y, ref["resid"], ctx.w, M_bar, factor)
    want = O.component_variances(O.SELayout(ctx, 0, n), y, s2, 1.0 - M_bar, M_bar, factor, 12,
                                 6, 9)
    cu = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()
    lev = JLALeverages(leverage=cu(1.0 - M_bar), complement=cu(M_bar), factor=cu(factor),
                       n_draws=12)
    inv = CoefficientInverse(res.panel, tol=1e-12, precision="fp64", covariates=(1, 2))
    out = component_variances(inv, res.panel.tensor("V", torch.float64).view(n, 3)[:, 0]
                              .contiguous(), cu(s2), lev, n_draws=12, block=6, seed=9)
    names = ["var(psi)", "var(alpha)", "cov(psi, alpha)"]
    for k in range(3):
        assert rel(out.diagnostics["b_ii"][k].cpu().numpy(), want["b_ii"][k]) < 1e-8
    for key, part in (("first", out.first_term), ("theta", out.theta),
                      ("trace", out.trace_term), ("variance", out.variance)):
        assert rel([part[nm] for nm in names], want[key]) < 1e-7, key


def test_weak_id_with_covariates_matches_the_oracle(cov_panel):
    import torch
    from hdfe_stream_amd.leaveout import CoefficientInverse, JLALeverages
    from hdfe_stream_amd.leaveout_weakid import weak_id_diagnostics
    O, d, res, ref, weighted = cov_panel
    print(f"weak_id with covariates weighted={weighted}", flush=True)
    ctx = O.KSSContext(ref, tol=1e-12, X=ref["p0"]["V"][:, 1:])
    n = len(ctx.w)
    rng = np.random.default_rng(6)
    s2, lv = rng.uniform(0.1, 0.4, n), rng.uniform(0.2, 0.9, n)
    y = ref["p0"]["V"][:, 0]
    want = O.weak_id_diagnostics(O.SELayout(ctx, 0, n), y, s2, lv, n_draws=16, block=8, seed=7)
    cu = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()
    levs = JLALeverages(leverage=cu(lv), complement=cu(1.0 - lv), factor=cu(np.ones(n)),
                        n_draws=1)
    inv = CoefficientInverse(res.panel, tol=1e-12, precision="fp64", covariates=(1, 2))
    out = weak_id_diagnostics(inv, res.panel.tensor("V", torch.float64).view(n, 3)[:, 0]
                              .contiguous(), cu(s2), levs, n_draws=16, block=8, seed=7)
    assert out.diagnostics["lanczos_steps"] == want["steps"]
    for k, nm in enumerate(["var(psi)", "var(alpha)", "cov(psi, alpha)"]):
        np.testing.assert_allclose(out.eigenvalues[nm], want["eigenvalues"][k], rtol=1e-6)
        assert out.sum_sq[nm] == pytest.approx(want["sum_sq"][k], rel=1e-6)
        assert out.f_stat[nm] == pytest.approx(want["f_stat"][k], rel=1e-5)
        assert out.q[nm] == want["q"][k]
    assert out.lead_vectors[0].shape[0] == ctx.K


def test_leave_out_kss_with_covariates_matches_the_oracle(oracle_mod):
    from hdfe_stream_amd.leaveout import leave_out_kss
    from hdfe_stream_amd.simulate import kss_panel
    O = oracle_mod
    d = kss_panel(4000, 400, seed=3)
    print("leave_out_kss with covariates", flush=True)
    lo = leave_out_kss("y ~ x1 + x2 | worker_id + firm_id", d, n_draws=64, seed=2,
                       precision="fp64", tol=1e-12, leave_out="observation")
    ref = O.leave_out_kss(d, xs=("x1", "x2"), n_draws=64, block=64, seed=2, tol=1e-12)
    assert lo.diagnostics["covariates"] == ["x1", "x2"]
    assert lo.leverages.diagnostics["covariates"] == ["x1", "x2"]
    names = ["var(psi)", "var(alpha)", "cov(psi, alpha)"]
    assert rel([lo.plug_in[k] for k in names], ref["plug_in"]) < 1e-9
    assert rel([lo.bias[k] for k in names], ref["bias"]) < 1e-8
    assert rel([lo.leave_out[k] for k in names], ref["leave_out"]) < 1e-8
    assert rel(lo.diagnostics["sigma2"].cpu().numpy(), ref["sigma2"]) < 1e-8
    # () asks for the fixed-effect design alone: another estimate
    fe_only = O.leave_out_kss(d, xs=("x1", "x2"), n_draws=64, block=64, seed=2, tol=1e-12,
                              design_covariates=False)
    assert rel(ref["bias"], fe_only["bias"]) > 1e-6
    from hdfe_stream_amd.leaveout import leave_out_components
    lo0 = leave_out_components(lo.fit, n_draws=64, seed=2, precision="fp64", tol=1e-12,
                               covariates=())
    assert lo0.diagnostics["covariates"] == []
    assert rel([lo0.bias[k] for k in names], fe_only["bias"]) < 1e-8
    with pytest.raises(ValueError):
        leave_out_components(lo.fit, n_draws=4, covariates=("x9",))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_leave_out_match_matches_the_oracle(oracle_mod, precision):
    from hdfe_stream_amd.leaveout import leave_out_kss
    from hdfe_stream_amd.simulate import kss_panel
    O = oracle_mod
    d = kss_panel(4000, 400, seed=3)
    print(f"leave_out_kss match, {precision}", flush=True)
    tol = 1e-12 if precision == "fp64" else 1e-10
    lo = leave_out_kss("y ~ 1 | worker_id + firm_id", d, n_draws=64, seed=2,
                       precision=precision, tol=tol)
    ref = O.leave_out_match(d, n_draws=64, block=64, seed=2, tol=1e-12)
    assert lo.diagnostics["leave_out"] == "match"
    assert lo.n_obs == ref["person_years"]
    assert lo.diagnostics["n_matches"] == ref["n_matches"]
    assert lo.n_movers == int(ref["movers"].sum())
    names = ["var(psi)", "var(alpha)", "cov(psi, alpha)"]
    assert rel([lo.plug_in[k] for k in names], ref["plug_in"]) < 1e-9
    bound = 1e-8 if precision == "fp64" else 1e-5
    assert rel([lo.bias[k] for k in names], ref["bias"]) < bound
    assert rel([lo.leave_out[k] for k in names], ref["leave_out"]) < bound
    s2 = lo.diagnostics["sigma2"].cpu().numpy()
    assert rel(s2, ref["sigma2"]) < (1e-8 if precision == "fp64" else 1e-5)


def test_leave_out_match_with_covariates_and_years():
    """The row model's covariates and further effects are partialled out
    before the collapse, so they are allowed at match level."""
    from hdfe_stream_amd.leaveout import leave_out_kss
    from hdfe_stream_amd.simulate import kss_panel
    d = kss_panel(3000, 300, seed=4)
    d["x1"] = np.random.default_rng(0).standard_normal(len(d["y"]))
    lo = leave_out_kss("y ~ x1 | worker_id + firm_id + year", d, n_draws=32, seed=1)
    assert lo.diagnostics["leave_out"] == "match"
    assert lo.leave_out["var(psi)"] < lo.plug_in["var(psi)"]
