"""Device standard errors of the KSS components (leaveout_se.py) against the
oracle's restatement, draw for draw: with the oracle's own leverages and
sigma2 (every piece to solver tolerance), and end to end through
leave_out_kss(se=True)."""

import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

FE = ["worker_id", "firm_id"]


def _loo(O, d):
    a, b = np.asarray(d["worker_id"]), np.asarray(d["firm_id"])
    ua, ia = np.unique(a, return_inverse=True)
    ub, ib = np.unique(b, return_inverse=True)
    alive, _ = O.leave_one_out(ia, ib, len(ua), len(ub))
    keep = alive[ia] & alive[len(ua) + ib]
    return {k: np.asarray(v)[keep] for k, v in d.items()}


@pytest.mark.parametrize("weighted", [False, True])
def test_component_variances_match_the_oracle(oracle_mod, weighted):
    import torch
    from hdfe_stream_amd.estimator import feols_stream
    from hdfe_stream_amd.leaveout import CoefficientInverse, JLALeverages
    from hdfe_stream_amd.leaveout_se import component_variances
    from hdfe_stream_amd.simulate import kss_panel
    O = oracle_mod
    d = kss_panel(1200, 150, seed=5)
    if weighted:
        d["w"] = np.random.default_rng(2).uniform(0.5, 2.0, len(d["y"]))
    d = _loo(O, d)
    wkw = {"weights": "w"} if weighted else {}
    print(f"component_variances weighted={weighted}", flush=True)
    res = feols_stream("y ~ 1 | worker_id + firm_id", d, stream="worker_id", fixef_rm="none",
                       keep_intermediates=True, tol=1e-12, **wkw)
    f = O.fit(d, "y", (), FE, weights="w" if weighted else None, tol=1e-12, fixef_rm="none",
              stream="worker_id")
    ctx = O.KSSContext(f, tol=1e-12)
    n = len(ctx.w)
    assert res.n_obs == n
    rng = np.random.default_rng(4)
    M_bar = rng.uniform(0.3, 0.95, n)
    factor = rng.uniform(0.9, 1.1, n)
    s2 = O.sigma2(f["p0"]["V"][:, 0], f["resid"], ctx.w, M_bar, factor)
    seed, nd = 9, 20
    lay = O.SELayout(ctx, 0, n)
    ref = O.component_variances(lay, f["p0"]["V"][:, 0], s2, 1.0 - M_bar, M_bar, factor, nd, 7,
                                seed)
    inv = CoefficientInverse(res.panel, tol=1e-12, precision="fp64")
    cu = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()
    lev = JLALeverages(leverage=cu(1.0 - M_bar), complement=cu(M_bar), factor=cu(factor),
                       n_draws=nd)
    y = res.panel.tensor("V", torch.float64)
    out = component_variances(inv, y, cu(s2), lev, n_draws=nd, block=6, seed=seed)
    bii = out.diagnostics["b_ii"].cpu().numpy()
    for k in range(3):
        assert rel(bii[k], ref["b_ii"][k]) < 1e-8
    np.testing.assert_allclose(out.diagnostics["smoothed_sigma2"].cpu().numpy(), ref["s"],
                               rtol=1e-7, atol=1e-12 * np.abs(ref["s"]).max())
    names = ["var(psi)", "var(alpha)", "cov(psi, alpha)"]
    for key, part in (("first", out.first_term), ("theta", out.theta),
                      ("trace", out.trace_term), ("variance", out.variance)):
        got = np.array([part[nm] for nm in names])
        assert rel(got, ref[key]) < 1e-7, key
    for k, nm in enumerate(names):
        if ref["variance"][k] > 0:
            assert out.se[nm] == pytest.approx(np.sqrt(ref["variance"][k]), rel=1e-7)


def test_leave_out_kss_standard_errors(oracle_mod):
    from hdfe_stream_amd.leaveout import leave_out_kss
    from hdfe_stream_amd.simulate import kss_panel
    O = oracle_mod
    d = kss_panel(4000, 400, seed=3)
    print("leave_out_kss se=True", flush=True)
    lo = leave_out_kss("y ~ 1 | worker_id + firm_id", d, n_draws=64, seed=2, precision="fp64",
                       tol=1e-12, leave_out="observation", se=True, se_draws=32, block=16)
    ref = O.leave_out_kss(d, n_draws=64, block=64, seed=2, tol=1e-12)
    f = ref["fit"]
    ctx = O.KSSContext(f, tol=1e-12)
    n = ref["n"]
    lay = O.SELayout(ctx, 0, n)
    want = O.component_variances(lay, f["p0"]["V"][:, 0], ref["sigma2"], 1.0 - ref["M_bar"],
                                 ref["M_bar"], ref["factor"], 32, 16, 2 + O.SE_BII_OFFSET)
    names = ["var(psi)", "var(alpha)", "cov(psi, alpha)"]
    got = lo.se
    assert got is not None and got.n_draws == 32
    # the leverages and sigma2 agree to ~1e-9; the smoother's cells follow them
    for key, part in (("first", got.first_term), ("theta", got.theta),
                      ("variance", got.variance)):
        assert rel([part[nm] for nm in names], want[key]) < 1e-5, key
    # theta = y~'C y~ checks the point estimate: close to plug-in - bias
    for nm in names:
        assert got.theta[nm] == pytest.approx(lo.leave_out[nm], rel=0.1,
                                              abs=0.1 * abs(lo.plug_in[nm]))
        assert got.se[nm] > 0 and got.se[nm] <= got.se_conservative[nm]
    assert "standard errors" in lo.summary()


def test_without_the_trace_the_errors_are_conservative():
    from hdfe_stream_amd.leaveout import leave_out_kss
    from hdfe_stream_amd.simulate import kss_panel
    d = kss_panel(2000, 200, seed=6)
    lo = leave_out_kss("y ~ 1 | worker_id + firm_id", d, n_draws=16, seed=1,
                       leave_out="observation", se=True, se_trace=False)
    for nm, v in lo.se.se.items():
        assert lo.se.trace_term[nm] == 0.0
        assert v == pytest.approx(lo.se.se_conservative[nm], rel=1e-12)
    assert lo.se.n_draws == 0


@pytest.mark.parametrize("se_variance", ["person_year", "match"])
def test_leave_out_match_standard_errors(oracle_mod, se_variance):
    from hdfe_stream_amd.leaveout import leave_out_kss
    from hdfe_stream_amd.simulate import kss_panel
    O = oracle_mod
    d = kss_panel(4000, 400, seed=3)
    print(f"leave_out_kss match se_variance={se_variance}", flush=True)
    lo = leave_out_kss("y ~ 1 | worker_id + firm_id", d, n_draws=64, seed=2, precision="fp64",
                       tol=1e-12, se=True, se_draws=32, block=16, se_variance=se_variance)
    ref = O.leave_out_match(d, n_draws=64, block=64, seed=2, tol=1e-12, se=True, se_draws=32,
                            se_block=16, se_variance=se_variance)
    assert lo.diagnostics["leave_out"] == "match"
    s2 = lo.se.diagnostics["sigma2"].cpu().numpy()
    assert rel(s2, ref["se_sigma2"]) < 1e-6
    want = ref["se"]
    for key, part in (("first", lo.se.first_term), ("theta", lo.se.theta),
                      ("variance", lo.se.variance)):
        got = np.array([part["var(psi)"], part["cov(psi, alpha)"]])
        assert rel(got, want[key][[0, 2]]) < 1e-5, key
        assert np.isnan(part["var(alpha)"]) and np.isnan(want[key][1])
    assert np.isnan(lo.se.se["var(alpha)"])
    assert lo.se.se["var(psi)"] > 0


def test_weak_id_matches_the_oracle(oracle_mod):
    import torch
    from hdfe_stream_amd.estimator import feols_stream
    from hdfe_stream_amd.leaveout import CoefficientInverse, JLALeverages
    from hdfe_stream_amd.leaveout_weakid import weak_id_diagnostics
    from hdfe_stream_amd.simulate import kss_panel
    O = oracle_mod
    d = _loo(O, kss_panel(1200, 150, seed=5))
    print("weak_id_diagnostics", flush=True)
    res = feols_stream("y ~ 1 | worker_id + firm_id", d, stream="worker_id", fixef_rm="none",
                       keep_intermediates=True, tol=1e-12)
    f = O.fit(d, "y", (), FE, tol=1e-12, fixef_rm="none", stream="worker_id")
    ctx = O.KSSContext(f, tol=1e-12)
    n = len(ctx.w)
    rng = np.random.default_rng(6)
    s2, lev = rng.uniform(0.1, 0.4, n), rng.uniform(0.2, 0.9, n)
    y = f["p0"]["V"][:, 0]
    ref = O.weak_id_diagnostics(O.SELayout(ctx, 0, n), y, s2, lev, n_draws=32, block=8, seed=7)
    cu = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()
    levs = JLALeverages(leverage=cu(lev), complement=cu(1.0 - lev), factor=cu(np.ones(n)),
                        n_draws=1)
    inv = CoefficientInverse(res.panel, tol=1e-12, precision="fp64")
    out = weak_id_diagnostics(inv, res.panel.tensor("V", torch.float64), cu(s2), levs,
                              n_draws=32, block=8, seed=7)
    names = ["var(psi)", "var(alpha)", "cov(psi, alpha)"]
    assert out.diagnostics["lanczos_steps"] == ref["steps"]
    for k, nm in enumerate(names):
        np.testing.assert_allclose(out.eigenvalues[nm], ref["eigenvalues"][k], rtol=1e-6)
        assert out.sum_sq[nm] == pytest.approx(ref["sum_sq"][k], rel=1e-6)
        assert out.f_stat[nm] == pytest.approx(ref["f_stat"][k], rel=1e-5)
        assert out.q[nm] == ref["q"][k]
        np.testing.assert_allclose(out.ratios[nm], ref["ratios"][k], rtol=1e-5)
    assert out.diagnostics["n_draws"] == ref["n_draws"]
    assert "weak-identification" in out.summary()


# ------------------------------------------------ the q = 1 weak interval

def _device_to_oracle(res, f, Ul, Ug):
    """Coefficient blocks in the device's level/group order, moved to the
    oracle's through the rows (both hold the rows in the same order)."""
    import torch
    info = res.panel.info
    n = info["n"]
    g = res.panel.tensor("g", torch.int32).long().cpu().numpy()
    c0 = res.panel.tensor("codes", torch.int32).view(info["D"], n)[0].long().cpu().numpy()
    dl = int(res.panel.offs[0]) + c0
    ol = f["cells"]["offs"][0] + f["p0"]["codes"][:, 0]
    og = f["p0"]["g"]
    Ul, Ug = Ul.cpu().numpy(), Ug.cpu().numpy()
    L, G = int(ol.max()) + 1, int(og.max()) + 1
    out_l, out_g = np.zeros((L, Ul.shape[1])), np.zeros((G, Ug.shape[1]))
    out_l[ol] = Ul[dl]
    out_g[og] = Ug[g]
    return out_l, out_g


@pytest.mark.parametrize("centering", ["weighted", "reference"])
def test_weak_moments_match_the_oracle(oracle_mod, centering):
    import torch
    from hdfe_stream_amd.estimator import feols_stream
    from hdfe_stream_amd.leaveout import CoefficientInverse, JLALeverages
    from hdfe_stream_amd.leaveout_se import SELayout, eigen_rows, weak_moments
    from hdfe_stream_amd.simulate import kss_panel
    O = oracle_mod
    weighted = centering == "weighted"
    d = kss_panel(1200, 150, seed=5)
    if weighted:
        d["w"] = np.random.default_rng(2).uniform(0.5, 2.0, len(d["y"]))
    d = _loo(O, d)
    wkw = {"weights": "w"} if weighted else {}
    print(f"weak_moments centering={centering}", flush=True)
    res = feols_stream("y ~ 1 | worker_id + firm_id", d, stream="worker_id", fixef_rm="none",
                       keep_intermediates=True, tol=1e-12, **wkw)
    f = O.fit(d, "y", (), FE, weights="w" if weighted else None, tol=1e-12, fixef_rm="none",
              stream="worker_id")
    ctx = O.KSSContext(f, tol=1e-12)
    n = len(ctx.w)
    rng = np.random.default_rng(8)
    M_bar, factor = rng.uniform(0.3, 0.95, n), rng.uniform(0.9, 1.1, n)
    s2 = rng.uniform(0.05, 0.3, n)
    s = rng.uniform(-0.05, 0.4, (3, n))
    b_ii = rng.uniform(0.0, 0.02, (3, n))
    lams = np.array([0.3, 0.2, -0.1])
    stayers = rng.random(n) < 0.3 if centering == "reference" else None
    comps = (0, 2) if centering == "reference" else (0, 1, 2)
    cu = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()
    inv = CoefficientInverse(res.panel, tol=1e-12, precision="fp64")
    lay = SELayout(inv, 0, n)
    Ul_d = torch.randn((inv.L, 3), dtype=torch.float64, device="cuda",
                       generator=torch.Generator("cuda").manual_seed(1))
    Ug_d = torch.randn((inv.G, 3), dtype=torch.float64, device="cuda",
                       generator=torch.Generator("cuda").manual_seed(2))
    Ul, Ug = _device_to_oracle(res, f, Ul_d, Ug_d)
    olay = O.SELayout(ctx, 0, n)
    x1 = eigen_rows(lay, Ul_d, Ug_d).cpu().numpy()
    assert rel(x1, O.eigen_rows(olay, Ul, Ug)) < 1e-13
    y = f["p0"]["V"][:, 0]
    ref = O.weak_moments(olay, y, s2, s, 1.0 - M_bar, M_bar, factor, b_ii, lams, Ul, Ug, comps,
                         10, 4, 123, stayer_rows=stayers, centering=centering)
    lev = JLALeverages(leverage=cu(1.0 - M_bar), complement=cu(M_bar), factor=cu(factor),
                       n_draws=1)
    got = weak_moments(lay, res.panel.tensor("V", torch.float64), cu(s2), cu(s), lev,
                       cu(b_ii), lams, Ul_d, Ug_d, comps, n_draws=10, block=4, seed=123,
                       stayer_rows=None if stayers is None else torch.from_numpy(stayers).cuda(),
                       centering=centering)
    names = ["var(psi)", "var(alpha)", "cov(psi, alpha)"]
    assert set(got) == {names[k] for k in comps}
    for k in comps:
        p, q = got[names[k]], ref[k]
        for key in ("b1", "v_raw", "trace", "quadratic"):
            assert p[key] == pytest.approx(q[key], rel=1e-8, abs=1e-12), key
        np.testing.assert_allclose(p["sigma"], q["sigma"], rtol=1e-8, atol=1e-13)
        assert p["conservative"] == q["conservative"]


@pytest.fixture(scope="module")
def bottleneck():
    from hdfe_stream_amd.leaveout import leave_out_kss
    from hdfe_stream_amd.simulate import bottleneck_panel
    d = bottleneck_panel(n_bridge=6, seed=0)
    print("leave_out_kss on a bottleneck, se=True", flush=True)
    lo = leave_out_kss("y ~ 1 | worker_id + firm_id", d, n_draws=128, seed=3,
                       precision="fp64", tol=1e-12, leave_out="observation", se=True,
                       se_draws=64, block=16, diagnose_draws=64)
    return d, lo


def test_weak_interval_on_a_bottleneck(oracle_mod, bottleneck):
    import torch
    O = oracle_mod
    d, lo = bottleneck
    flagged = {k for k, q in lo.weak_id.q.items() if q >= 1}
    assert flagged, "the bottleneck panel should be weakly identified"
    assert set(lo.se.weak) == flagged
    # the moments again in the oracle, from the device's own inputs
    f = O.fit(d, "y", (), FE, tol=1e-12, fixef_rm="none", stream="worker_id")
    ctx = O.KSSContext(f, tol=1e-12)
    n = len(ctx.w)
    Ul, Ug = _device_to_oracle(lo.fit, f, *lo.weak_id.lead_vectors)
    lev = lo.leverages
    host = lambda t: t.detach().cpu().numpy()
    names = ["var(psi)", "var(alpha)", "cov(psi, alpha)"]
    comps = tuple(k for k, nm in enumerate(names) if nm in flagged)
    lams = np.array([lo.weak_id.eigenvalues[nm][0] for nm in names])
    ref = O.weak_moments(O.SELayout(ctx, 0, n), f["p0"]["V"][:, 0],
                         host(lo.diagnostics["sigma2"]),
                         host(lo.se.diagnostics["smoothed_sigma2"]), host(lev.leverage),
                         host(lev.complement), host(lev.factor), host(lo.se.diagnostics["b_ii"]),
                         lams, Ul, Ug, comps, 64, 16, 3 + O.SE_BII_OFFSET + O.WEAK_MOMENTS_OFFSET)
    for k in comps:
        p, q = lo.se.weak[names[k]], ref[k]
        for key in ("b1", "v_raw", "trace", "quadratic"):
            assert p[key] == pytest.approx(q[key], rel=1e-7, abs=1e-12), key
        np.testing.assert_allclose(p["sigma"], q["sigma"], rtol=1e-7, atol=1e-13)
        theta1 = lo.leave_out[names[k]] - p["lambda_1"] * (p["b1"] ** 2 - p["v_raw"])
        assert p["theta1"] == pytest.approx(theta1, rel=1e-12)
        olo, ohi, info = O.am_interval(p["lambda_1"], p["b1"], theta1, p["sigma"], 0.05)
        assert p["interval"] == pytest.approx((olo, ohi), rel=1e-9, abs=1e-12)
        assert p["kappa"] == pytest.approx(info["kappa"], rel=1e-9)
        assert p["interval"][0] < p["interval"][1] and p["confidence"] == 0.95
        # theta1 = y~'C2 y~ - exactly, for the leave-out sigma2 - so its
        # plug-in form and the quadratic form agree up to the bias estimate
        assert abs(p["quadratic"] - theta1) < 0.5 * abs(lo.plug_in[names[k]]) + 1e-3
    text = lo.summary()
    assert "q = 1 intervals" in text and "*" in text
    del torch


def test_weak_interval_switches_and_level(bottleneck):
    from hdfe_stream_amd.leaveout import leave_out_kss
    d, wide = bottleneck
    kw = dict(n_draws=128, seed=3, precision="fp64", tol=1e-12, leave_out="observation",
              se=True, se_draws=64, block=16, diagnose_draws=64)
    print("weak interval off / at 80%", flush=True)
    off = leave_out_kss("y ~ 1 | worker_id + firm_id", d, weak_interval=False, **kw)
    assert off.se.weak is None and "q = 1 intervals" not in off.summary()
    narrow = leave_out_kss("y ~ 1 | worker_id + firm_id", d, confidence=0.8, **kw)
    assert "80% interval" in narrow.summary()
    for nm, piece in narrow.se.weak.items():
        w_lo, w_hi = wide.se.weak[nm]["interval"]
        n_lo, n_hi = piece["interval"]
        assert n_hi - n_lo < w_hi - w_lo
        # same draws, same moments: only the level moved
        np.testing.assert_array_equal(piece["sigma"], wide.se.weak[nm]["sigma"])


def test_match_level_weak_interval_is_var_psi_alone():
    from hdfe_stream_amd.leaveout import leave_out_kss
    from hdfe_stream_amd.simulate import bottleneck_panel
    d = bottleneck_panel(n_bridge=6, seed=1)
    print("leave_out_kss match on a bottleneck", flush=True)
    lo = leave_out_kss("y ~ 1 | worker_id + firm_id", d, n_draws=64, seed=2, se=True,
                       se_draws=32, block=16, diagnose_draws=48, se_variance="match")
    if lo.weak_id.q["var(psi)"] >= 1:
        assert set(lo.se.weak) == {"var(psi)"}
        lo_, hi_ = lo.se.weak["var(psi)"]["interval"]
        assert lo_ < hi_
    else:
        assert not lo.se.weak


def test_tall_dot_matches_a_plain_product():
    import torch
    from hdfe_stream_amd.leaveout_se import tall_dot
    g = torch.Generator("cuda").manual_seed(3)
    for K, a, b in ((1, 1, 1), (1000, 3, 7), (200_003, 1, 40), (50_000, 20, 30), (7, 300, 2)):
        A = torch.randn((K, a), dtype=torch.float64, device="cuda", generator=g)
        B = torch.randn((K, b + 3), dtype=torch.float64, device="cuda", generator=g)[:, 2:2 + b]
        w = torch.rand(K, dtype=torch.float64, device="cuda", generator=g)
        print(f"tall_dot K={K} a={a} b={b}", flush=True)
        want = (A.cpu().T @ B.cpu()).numpy()
        got = tall_dot(A, B).cpu().numpy()
        np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-11 * np.abs(want).max())
        want_w = (A.cpu().T @ (w.cpu()[:, None] * B.cpu())).numpy()
        got_w = tall_dot(A, B, w).cpu().numpy()
        np.testing.assert_allclose(got_w, want_w, rtol=1e-11, atol=1e-11 * np.abs(want_w).max())
        # fixed-order sums: the same bits twice
        assert np.array_equal(tall_dot(A, B).cpu().numpy(), got)
    v = torch.randn(1000, dtype=torch.float64, device="cuda", generator=g)
    assert tall_dot(v, v).item() == pytest.approx(float((v * v).sum()), rel=1e-12)
