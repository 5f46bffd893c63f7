"""Pinning the oracle's leave-out (KSS) pieces against dense closed forms and
brute force: the coefficient-space inverse against the pseudo-inverse of the
full design, the Hutchinson moments against the dense quadratic form (draw by
draw, and in expectation against the exact trace), sigma2 against refitting
without each row, and the plug-in moments against numpy's covariance."""

import numpy as np
import pytest

FE2 = ["worker_id", "firm_id"]
FE3 = ["worker_id", "firm_id", "year"]


def _two_components(d):
    """The panel twice, ids of the second copy shifted so the copies share
    no worker or firm: a panel of two connected components."""
    out = {}
    for k, v in d.items():
        v = np.asarray(v)
        if k in ("worker_id", "firm_id"):
            out[k] = np.concatenate([v, v + v.max() + 1])
        else:
            out[k] = np.concatenate([v, v])
    return out


def _dense(O, f):
    """A = [D_g, D_o] (n, G + L) dense, the rows' weights and S_full."""
    Dg, Do = O.design(f["p0"], f["cells"])
    A = np.hstack([Dg.toarray(), Do.toarray()])
    return A


def _loo_panel(O, d):
    a, b = np.asarray(d["worker_id"]), np.asarray(d["firm_id"])
    ua, ia = np.unique(a, return_inverse=True)
    ub, ib = np.unique(b, return_inverse=True)
    alive, _ = O.leave_one_out(ia, ib, len(ua), len(ub))
    keep = alive[ia] & alive[len(ua) + ib]
    return {k: np.asarray(v)[keep] for k, v in d.items()}


@pytest.mark.parametrize("fe,copies", [(FE2, 1), (FE3, 1), (FE3, 2)])
def test_apply_inverse_is_the_pseudo_inverse(oracle_mod, fe, copies):
    """A u = A (A'A)^+ v and A'A u = the range part of v, for v given in
    coefficient space (inverse.py:311-374; tests/test_inverse.py:475-553 in
    the reference), with one or two connected components."""
    O = oracle_mod
    from hdfe_stream_amd.simulate import akm_panel
    d = akm_panel(200, 25, seed=8)
    if copies == 2:
        d = _two_components(d)
    f = O.fit(d, "y", (), fe, tol=1e-13, fixef_rm="none", stream="worker_id")
    assert f["n_components"] == copies
    ctx = O.KSSContext(f, tol=1e-13)
    A = _dense(O, f)
    G, L = ctx.G, ctx.L
    Sf = A.T @ A
    Sp = np.linalg.pinv(Sf)
    rng = np.random.default_rng(4)
    v = rng.standard_normal((G + L, 3))
    Ul, Ug, rows = O.apply_inverse(ctx, v[G:], v[:G])
    u = np.vstack([Ug, Ul])
    np.testing.assert_allclose(rows, A @ u, rtol=0, atol=1e-12)
    np.testing.assert_allclose(A @ u, A @ Sp @ v, rtol=0, atol=1e-8)
    np.testing.assert_allclose(Sf @ u, Sf @ Sp @ v, rtol=0, atol=1e-8)
    # the null space dropped has dimension components + extra dimensions
    assert ctx.q_levels.shape[1] == copies + len(fe) - 2
    assert np.linalg.matrix_rank(Sf) == G + L - ctx.q_levels.shape[1]


def _krylov_solution(A, w, G, u):
    """The solution that Jacobi-preconditioned CG started at zero reaches,
    from any solution u (G + L, k) of the full system.  Its iterates stay in
    D^-1 range(S_red) for the reduced S_red = D_o' W M_0 D_o and D = diag(S_red),
    so the levels part x satisfies N' D x = 0 for every null vector N of
    S_red: shift u along the full null space (whose levels parts span
    null(S_red)) until it does."""
    from scipy.linalg import null_space
    Nf = null_space(A * np.sqrt(w)[:, None])
    Dg, Do = A[:, :G], A[:, G:]
    M0 = np.diag(w) - (w[:, None] * Dg) @ np.linalg.pinv(Dg.T @ (w[:, None] * Dg)) @ (Dg.T * w)
    d = np.diag(Do.T @ M0 @ Do)
    Nl = Nf[G:]
    alpha = np.linalg.lstsq(Nl.T @ (d[:, None] * Nl), -(Nl.T @ (d[:, None] * u[G:])),
                            rcond=None)[0]
    return u + Nf @ alpha


def _dense_totals(ctx, A, sigma2, Z, psi_col):
    """The eight moments of Z and U2 = S^- A'W Sigma A S^- Z from dense
    matrices, summed over rows and Z's columns, U2 the solution CG reaches."""
    G, L = ctx.G, ctx.L
    w = ctx.w
    Sp = np.linalg.pinv(A.T @ (w[:, None] * A))
    M = Sp @ A.T @ ((w * sigma2)[:, None] * A) @ Sp
    Zc = np.vstack([Z[L:], Z[:L]])            # coefficient order (groups, levels)
    U2 = _krylov_solution(A, w, G, M @ Zc)
    lv = G + ctx.c["offs"][psi_col] + ctx.p0["codes"][:, psi_col]
    a1, a2, b1, b2 = Zc[lv], U2[lv], Zc[ctx.g], U2[ctx.g]
    wc = w[:, None]
    return np.array([w.sum() * Z.shape[1], (wc * a1 * a2).sum(), (wc * b1 * b2).sum(),
                     (wc * (a1 * b2 + b1 * a2)).sum(), (wc * a1).sum(), (wc * a2).sum(),
                     (wc * b1).sum(), (wc * b2).sum()])


@pytest.mark.parametrize("fe", [FE2, FE3])
def test_hutchinson_matches_the_dense_quadratic_form(oracle_mod, fe):
    """Draw by draw, the trace moments equal Z'Q S^+ A'W Sigma A S^+ Z from
    dense matrices, for the same coefficient-space Rademacher draws; their
    mean over many draws approaches the exact trace (leaveout.py:321-448)."""
    O = oracle_mod
    from hdfe_stream_amd.simulate import akm_panel
    d = _loo_panel(O, akm_panel(150, 12, seed=9, p_move=0.3))
    f = O.fit(d, "y", (), fe, tol=1e-13, fixef_rm="none", stream="worker_id")
    ctx = O.KSSContext(f, tol=1e-13)
    A = _dense(O, f)
    n = len(ctx.g)
    sigma2 = np.random.default_rng(2).uniform(0.1, 0.5, n)
    k = 40
    (vp, va, vc), totals = O.hutchinson_trace(ctx, sigma2, k, 16, seed=11)
    Z = O.rademacher(0, ctx.L + ctx.G, k, 11, 0)
    dense = _dense_totals(ctx, A, sigma2, Z, 0)
    np.testing.assert_allclose(totals, dense, rtol=1e-8, atol=1e-9)
    np.testing.assert_allclose([vp, va, vc], O.trace_components(dense, k, n),
                               rtol=1e-8, atol=1e-12)
    # blocking does not change the draws
    (vp2, va2, vc2), _ = O.hutchinson_trace(ctx, sigma2, k, 7, seed=11)
    np.testing.assert_allclose([vp2, va2, vc2], [vp, va, vc], rtol=1e-9, atol=1e-14)
    # expectation: with 4000 draws the three components are within 5
    # standard errors of the exact traces tr(Q M) (Q the centred quadratic
    # forms of psi, alpha and their symmetrised cross product)
    many = 4000
    Sp = np.linalg.pinv(A.T @ A)
    M = Sp @ A.T @ (sigma2[:, None] * A) @ Sp
    G, L = ctx.G, ctx.L
    lv = G + ctx.c["offs"][0] + ctx.p0["codes"][:, 0]
    Ep = np.zeros((n, G + L))
    Ep[np.arange(n), lv] = 1.0
    Ea = np.zeros((n, G + L))
    Ea[np.arange(n), ctx.g] = 1.0
    Cm = np.eye(n) / n - 1.0 / n ** 2
    Qs = [Ep.T @ Cm @ Ep, Ea.T @ Cm @ Ea, 0.5 * (Ep.T @ Cm @ Ea + Ea.T @ Cm @ Ep)]
    Zm = O.rademacher(0, L + G, many, 5, 0)
    Zc = np.vstack([Zm[L:], Zm[:L]])
    MZ = M @ Zc
    est = O.trace_components(_dense_totals(ctx, A, sigma2, Zm, 0), many, n)
    scale = n / O.moment_divisor(n, n)
    for Q, e in zip(Qs, est):
        per = np.einsum("ij,ij->j", Zc, Q @ MZ) * scale
        assert abs(e - np.trace(Q @ M) * scale) < 5 * per.std() / np.sqrt(many) + 1e-12


def test_sigma2_is_y_times_the_leave_one_out_residual(oracle_mod):
    """With exact leverages, (y_i - ybar) e_i / M_ii = (y_i - ybar) times the
    residual of row i from the fit without row i (leaveout.py:949-1016)."""
    O = oracle_mod
    from hdfe_stream_amd.simulate import akm_panel
    d = _loo_panel(O, akm_panel(120, 10, seed=12, p_move=0.35))
    f = O.fit(d, "y", (), FE2, tol=1e-13, fixef_rm="none", stream="worker_id")
    A = _dense(O, f)
    n = A.shape[0]
    P = np.einsum("ij,ji->i", A, np.linalg.pinv(A.T @ A) @ A.T)
    assert P.max() < 1 - 1e-9
    y = f["p0"]["V"][:, 0]
    s2 = O.sigma2(y, f["resid"], np.ones(n), 1.0 - P, np.ones(n))
    for i in np.random.default_rng(0).choice(n, 25, replace=False):
        m = np.ones(n, bool)
        m[i] = False
        theta = np.linalg.lstsq(A[m], y[m], rcond=None)[0]
        loo = y[i] - A[i] @ theta
        assert s2[i] == pytest.approx((y[i] - y.mean()) * loo, rel=1e-8, abs=1e-12)


def test_plug_in_is_the_sample_covariance(oracle_mod):
    """Unweighted, the plug-in moments are np.var / np.cov with ddof 1; with
    weights T (R - 1) / R divides (leaveout.py:807-832, leaveout_se.py:202-214)."""
    O = oracle_mod
    rng = np.random.default_rng(1)
    a, b = rng.standard_normal(500), rng.standard_normal(500)
    vp, va, cv = O.plug_in(a, b, np.ones(500), 500)
    assert vp == pytest.approx(np.var(a, ddof=1), rel=1e-13)
    assert va == pytest.approx(np.var(b, ddof=1), rel=1e-13)
    assert cv == pytest.approx(np.cov(a, b)[0, 1], rel=1e-12)
    w = rng.uniform(0.5, 2.0, 500)
    vp, _, _ = O.plug_in(a, b, w, 500)
    T = w.sum()
    mean = (w @ a) / T
    assert vp == pytest.approx((w @ (a - mean) ** 2) / (T * 499 / 500), rel=1e-12)


def test_leave_out_kss_approaches_the_exact_estimator(oracle_mod):
    """End to end on a small pruned panel: with many draws the bias term is
    close to the exact KSS bias tr(Q S^+ A' Sigma A S^+) built from the
    exact leverages (dense hat matrix), and leave-out = plug-in - bias
    (leaveout.py:634-805, 1069-1255)."""
    O = oracle_mod
    from hdfe_stream_amd.simulate import akm_panel
    d = akm_panel(200, 20, seed=21, p_move=0.3)
    r = O.leave_out_kss(d, n_draws=4000, block=500, seed=1)
    f = r["fit"]
    A = _dense(O, f)
    n = A.shape[0]
    Sp = np.linalg.pinv(A.T @ A)
    P = np.einsum("ij,ji->i", A, Sp @ A.T)
    s2 = O.sigma2(f["p0"]["V"][:, 0], f["resid"], np.ones(n), 1 - P, np.ones(n))
    M = Sp @ A.T @ (s2[:, None] * A) @ Sp
    ctx = O.KSSContext(f)
    G = ctx.G
    Ep = np.zeros((n, A.shape[1]))
    Ep[np.arange(n), G + ctx.c["offs"][0] + ctx.p0["codes"][:, 0]] = 1
    Ea = np.zeros((n, A.shape[1]))
    Ea[np.arange(n), ctx.g] = 1
    Cm = np.eye(n) / n - 1.0 / n ** 2
    exact = [np.trace(Q @ M) * n / (n - 1) for Q in
             (Ep.T @ Cm @ Ep, Ea.T @ Cm @ Ea, 0.5 * (Ep.T @ Cm @ Ea + Ea.T @ Cm @ Ep))]
    for b, e in zip(r["bias"], exact):
        assert abs(b - e) < 0.1 * abs(e)
    np.testing.assert_allclose(r["leave_out"], np.subtract(r["plug_in"], r["bias"]), rtol=1e-15)
    a = f["gam_y"][ctx.c["offs"][0] + ctx.p0["codes"][:, 0]]
    np.testing.assert_allclose(r["plug_in"][0], np.var(a, ddof=1), rtol=1e-12)


def test_match_level_plug_in_equals_the_row_level_one(oracle_mod):
    """Collapsing y~ = alpha + psi + e to match means and refitting with the
    spells' lengths as weights gives back the row fit's effects (the
    within-match sum of squares does not involve them), so the plug-in
    moments with the person-year divisor are the row level's
    (leaveout_match.py:63-88)."""
    O = oracle_mod
    from hdfe_stream_amd.simulate import kss_panel
    d = kss_panel(1500, 150, seed=5)
    r = O.leave_out_match(d, n_draws=8, block=8, seed=0)
    f = r["fit"]
    p0, c = f["p0"], f["cells"]
    a = f["gam_y"][c["offs"][0] + p0["codes"][:, 0]]
    b = f["worker_effects"][p0["g"]]
    row = O.plug_in(a, b, np.ones(len(a)), len(a))
    np.testing.assert_allclose(r["plug_in"], row, rtol=1e-7, atol=1e-12)


def test_stayer_sigma2_is_the_within_spell_leave_one_out_product(oracle_mod):
    """A stayer's single match has leverage 1, and sigma_for_stayers
    (leaveout_match.py:91-132) uses its person-years instead: the mean over
    the spell of (y~_it - ybar) times y~_it less the mean of the spell's
    other years."""
    O = oracle_mod
    from hdfe_stream_amd.simulate import kss_panel
    d = kss_panel(1500, 150, seed=5)
    r = O.leave_out_match(d, n_draws=8, block=8, seed=0)
    f = r["fit"]
    p0, c = f["p0"], f["cells"]
    g, pc = p0["g"], p0["codes"][:, 0]
    ytil = f["worker_effects"][g] + f["gam_y"][c["offs"][0] + pc] + f["resid"]
    ybar_all = ytil.mean()
    mg = r["match_fit"]["p0"]["g"]
    stay = np.flatnonzero(~r["movers"])[:20]
    for m in stay:
        rows = np.flatnonzero(g == mg[m])
        yy = ytil[rows]
        T = len(yy)
        loo = np.array([yy[i] - (yy.sum() - yy[i]) / (T - 1) for i in range(T)])
        want = np.mean((yy - ybar_all) * loo)
        assert r["sigma2"][m] == pytest.approx(want, rel=1e-9, abs=1e-14)


def test_match_level_bias_approaches_the_exact_trace(oracle_mod):
    """With many draws the match-level bias is close to the exact trace of
    the weighted collapsed design with the same sigma2, and the leverages to
    the exact square-root-weight ones."""
    O = oracle_mod
    from hdfe_stream_amd.simulate import kss_panel
    d = kss_panel(250, 25, seed=6)
    r = O.leave_out_match(d, n_draws=4000, block=500, seed=1)
    mf = r["match_fit"]
    ctx = O.KSSContext(mf)
    Dg, Do = O.design(mf["p0"], mf["cells"])
    A = np.hstack([Dg.toarray(), Do.toarray()])
    w = ctx.w
    nm = A.shape[0]
    G = ctx.G
    s2 = r["sigma2"]
    Sp = np.linalg.pinv(A.T @ (w[:, None] * A))
    lev = np.einsum("ij,ji->i", np.sqrt(w)[:, None] * A, Sp @ (A.T * np.sqrt(w)))
    assert abs((1 - r["M_bar"]).mean() - lev.mean()) < 0.01
    M = Sp @ A.T @ ((w * s2)[:, None] * A) @ Sp
    Ep = np.zeros((nm, A.shape[1]))
    Ep[np.arange(nm), G + ctx.c["offs"][0] + ctx.p0["codes"][:, 0]] = 1
    Ea = np.zeros((nm, A.shape[1]))
    Ea[np.arange(nm), ctx.g] = 1
    W = w.sum()
    Cm = np.diag(w) / W - np.outer(w, w) / W ** 2
    R = r["person_years"]
    scale = W / O.moment_divisor(W, R)
    exact = [np.trace(Q @ M) * scale for Q in
             (Ep.T @ Cm @ Ep, Ea.T @ Cm @ Ea, 0.5 * (Ep.T @ Cm @ Ea + Ea.T @ Cm @ Ep))]
    for b, e in zip(r["bias"], exact):
        assert abs(b - e) < 0.1 * abs(e)
