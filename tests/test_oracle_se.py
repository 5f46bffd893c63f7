"""The oracle's KSS standard errors (leaveout_se.py) pinned by dense algebra
on a small weighted panel: the kernel C of each component built from
explicit matrices -- X~ = W^(1/2) A, S^- = pinv(X~'X~), B = X~ S^- Q S^- X~',
M = I - X~ S^- X~' -- and every streamed piece checked against it with the
same draws, plus the estimator's identity y~'C y~ = plug-in - bias."""

import numpy as np
import pytest

from conftest import rel

FE = ["worker_id", "firm_id"]


def _loo(O, d):
    a, b = np.asarray(d["worker_id"]), np.asarray(d["firm_id"])
    ua, ia = np.unique(a, return_inverse=True)
    ub, ib = np.unique(b, return_inverse=True)
    alive, _ = O.leave_one_out(ia, ib, len(ua), len(ub))
    keep = alive[ia] & alive[len(ua) + ib]
    return {k: np.asarray(v)[keep] for k, v in d.items()}


@pytest.fixture(scope="module", params=["fixed effects", "covariates"])
def dense(oracle_mod, request):
    """The design [D_g, D_o] or, with covariates, [D_g, D_o, X] (x1, x2 in
    the model and in the leave-out design: the border)."""
    from hdfe_stream_amd.simulate import kss_panel
    O = oracle_mod
    d = kss_panel(220, 30, seed=12)
    d["w"] = np.random.default_rng(3).uniform(0.5, 2.0, len(d["y"]))
    d = _loo(O, d)
    xs = ("x1", "x2") if request.param == "covariates" else ()
    f = O.fit(d, "y", xs, FE, weights="w", tol=1e-13, fixef_rm="none", stream="worker_id")
    X = f["p0"]["V"][:, 1:] if xs else None
    ctx = O.KSSContext(f, tol=1e-13, X=X)
    n = len(ctx.w)
    lay = O.SELayout(ctx, 0, n)
    # coefficients: groups, levels, covariates
    A = np.hstack([ctx.Dg.toarray(), ctx.Do.toarray(), ctx.X])
    Xt = lay.root[:, None] * A
    Sm = np.linalg.pinv(Xt.T @ Xt)
    P = Xt @ Sm @ Xt.T
    M = np.eye(n) - P
    G, L = ctx.G, ctx.K
    psi = np.zeros((n, G + L))
    psi[np.arange(n), G + lay.psi_level] = 1.0
    alp = np.zeros((n, G + L))
    alp[np.arange(n), ctx.g] = 1.0
    T = lay.total

    def centred(Z):
        return Z - (ctx.w @ Z)[None, :] / T

    Lp = lay.root[:, None] * centred(psi) / np.sqrt(lay.norm)
    La = lay.root[:, None] * centred(alp) / np.sqrt(lay.norm)
    Q = [Lp.T @ Lp, La.T @ La, 0.5 * (Lp.T @ La + La.T @ Lp)]
    B = [Xt @ Sm @ q @ Sm @ Xt.T for q in Q]
    bvec = np.stack([np.diag(Bk) / np.diag(M) for Bk in B])
    Cm = [B[k] - 0.5 * (np.diag(bvec[k]) @ M + M @ np.diag(bvec[k])) for k in range(3)]
    return dict(O=O, f=f, ctx=ctx, lay=lay, Xt=Xt, Sm=Sm, M=M, Lp=Lp, La=La, B=B, b=bvec,
                C=Cm, d=d, n=n, A=A)


def test_the_kernel_has_a_zero_diagonal(dense):
    for Ck in dense["C"]:
        assert np.abs(np.diag(Ck)).max() < 1e-10 * np.abs(Ck).max()


@pytest.mark.parametrize("per_component", [False, True])
def test_apply_c_is_the_dense_kernel(dense, per_component):
    O, lay, n = dense["O"], dense["lay"], dense["n"]
    rng = np.random.default_rng(5)
    v = rng.standard_normal((n, 3, 4) if per_component else (n, 4))
    cv = O.apply_c(lay, v, dense["b"], per_component=per_component)
    for k in range(3):
        want = dense["C"][k] @ (v[:, k, :] if per_component else v)
        assert rel(cv[:, k, :], want) < 1e-8


def test_rank_two_term(dense):
    O, lay, n = dense["O"], dense["lay"], dense["n"]
    u = np.random.default_rng(6).standard_normal((3, n))
    v = np.random.default_rng(7).standard_normal((n, 2))
    cv = O.apply_c(lay, v, dense["b"], rank_two=(u, 0.3))
    one = np.ones(n)
    for k in range(3):
        K = dense["C"][k] + 0.3 * (np.outer(one, u[k]) + np.outer(u[k], one))
        assert rel(cv[:, k, :], K @ v) < 1e-8


def test_m_times(dense):
    O, lay = dense["O"], dense["lay"]
    b = np.random.default_rng(8).standard_normal((3, dense["n"]))
    assert rel(O.m_times(lay, b), (dense["M"] @ b.T).T) < 1e-8


def test_bii_rows_with_the_same_draws(dense):
    O, lay, n = dense["O"], dense["lay"], dense["n"]
    nd, seed = 24, 11
    got = O.bii_rows(lay, nd, 10, seed)
    q = O.rademacher(0, n, nd, seed, 0)
    E = q @ q.T / nd
    R1 = dense["Xt"] @ dense["Sm"] @ dense["Lp"].T
    R2 = dense["Xt"] @ dense["Sm"] @ dense["La"].T
    want = [((R1 @ E) * R1).sum(1), ((R2 @ E) * R2).sum(1), ((R1 @ E) * R2).sum(1)]
    for k in range(3):
        assert rel(got[k], want[k]) < 1e-8
    # the expectation over draws is the diagonal of B (E[qq'] = I)
    exact = [R1 @ R1.T, R2 @ R2.T, 0.5 * (R1 @ R2.T + R2 @ R1.T)]
    for k in range(3):
        np.testing.assert_allclose(np.diag(exact[k]), np.diag(dense["B"][k]), rtol=1e-8,
                                   atol=1e-14)


def test_y_c_y_is_plug_in_minus_bias(dense):
    """With exact leverages and B_ii, theta = y~'C y~ equals the plug-in less
    sum_i B_ii sigma2_i, sigma2_i the leave-out variance (leaveout_se.py:10-14)."""
    O, lay, f, ctx, n = dense["O"], dense["lay"], dense["f"], dense["ctx"], dense["n"]
    M = dense["M"]
    y = f["p0"]["V"][:, 0]
    comp = np.diag(M)
    s2 = O.sigma2(y, f["resid"], ctx.w, comp, np.ones(n))
    res = O.component_variances(lay, y, s2, np.diag(np.eye(n) - M), comp, np.ones(n), 8, 8, 0,
                                trace=False, b_ii=np.stack([np.diag(B) for B in dense["B"]]))
    c = f["cells"]
    a = f["gam_y"][c["offs"][0] + f["p0"]["codes"][:, 0]]
    bb = f["worker_effects"][f["p0"]["g"]]
    plug = np.array(O.plug_in(a, bb, ctx.w, n))
    bias = np.array([np.diag(B) @ s2 for B in dense["B"]])
    assert rel(res["theta"], plug - bias) < 1e-8
    # and the quadratic forms themselves: y~'B y~ is the plug-in
    yt = lay.root * (y - ctx.w @ y / ctx.w.sum())
    assert rel([yt @ B @ yt for B in dense["B"]], plug) < 1e-8


def test_trace_term_with_the_same_draws(dense):
    O, lay, n = dense["O"], dense["lay"], dense["n"]
    rng = np.random.default_rng(9)
    s2 = rng.uniform(0.1, 0.5, n)
    f = dense["f"]
    nd, seed = 6, 21
    res = O.component_variances(lay, f["p0"]["V"][:, 0], s2, rng.uniform(0.1, 0.9, n),
                                np.diag(dense["M"]), np.ones(n), nd, 4, seed,
                                b_ii=np.stack([np.diag(B) for B in dense["B"]]), smooth_bins=0)
    # smooth_bins=0: no smoother, s = sigma2 itself
    z = O.rademacher(0, n, nd, seed + O.SE_TRACE_OFFSET, 0)
    Om = np.sqrt(np.maximum(s2, 0.0))[:, None] * z
    want = np.array([(s2[:, None] * (dense["C"][k] @ Om) ** 2).sum() / nd for k in range(3)])
    assert rel(res["trace"], 2.0 * want) < 1e-8
    # the first term
    yt = lay.root * (f["p0"]["V"][:, 0] - dense["ctx"].w @ f["p0"]["V"][:, 0] / lay.total)
    first = np.array([4.0 * s2 @ (dense["C"][k] @ yt) ** 2 for k in range(3)])
    assert rel(res["first"], first) < 1e-8
    np.testing.assert_allclose(res["variance"], res["first"] - res["trace"], rtol=1e-12)


def test_smoother(oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(1)
    n = 5000
    lev, bii = rng.uniform(0, 1, n), rng.exponential(1.0, n)
    assert O.auto_bins(n) == round(n ** (1 / 6))
    # a constant is reproduced exactly, everywhere
    t = O.fit_smoother(np.full(n, 0.7), lev, bii)
    np.testing.assert_allclose(O.smooth_eval(t, lev, bii), 0.7, rtol=1e-14)
    # cell means: each row's cell is found by its ranks, and the means are
    # those of the rows in the cell
    s2 = rng.standard_normal(n)
    kp, kb, means, nb = O.fit_smoother(s2, lev, bii, n_bins=4)
    take = np.arange(n)
    np.testing.assert_array_equal(kp[:-1] <= kp[1:], True)
    assert kp[0] == lev.min() and kb[0] == bii.min()
    cp = np.minimum((O.smoother_rank(kp, lev) * 4).astype(int), 3)
    cb = np.minimum((O.smoother_rank(kb, bii) * 4).astype(int), 3)
    for i in range(4):
        for j in range(4):
            sel = (cp == i) & (cb == j)
            assert means[i, j] == pytest.approx(s2[sel].mean(), rel=1e-12)
    # quartile cells of independent uniform ranks hold about n / 16 rows each
    counts = np.bincount(cp * 4 + cb, minlength=16)
    assert counts.min() > n / 16 * 0.7
    del take
    # the evaluation interpolates between cell centres: at a centre (half way
    # between two knots, as the rank is linear there) it is the cell's mean
    mid_p = 0.5 * (kp[1] + kp[2])
    mid_b = 0.5 * (kb[2] + kb[3])
    assert O.smooth_eval((kp, kb, means, 4), np.array([mid_p]), np.array([mid_b]))[0] == \
        pytest.approx(means[1, 2], rel=1e-9)


# ------------------------------------------------ weak-identification diagnostic

def _dense_A(dense, k):
    """S^- Q_k in the coefficient order (groups, levels)."""
    lay = dense["lay"]
    Sfull = dense["Xt"].T @ dense["Xt"]
    Lp, La = dense["Lp"], dense["La"]
    Q = [Lp.T @ Lp, La.T @ La, 0.5 * (Lp.T @ La + La.T @ Lp)][k]
    return np.linalg.pinv(Sfull) @ Q, Sfull, Q, lay


def test_lanczos_finds_the_leading_eigenvalues(dense):
    O = dense["O"]
    out = O.weak_id_diagnostics(dense["lay"], dense["f"]["p0"]["V"][:, 0],
                                np.random.default_rng(1).uniform(0.1, 0.4, dense["n"]),
                                np.random.default_rng(2).uniform(0.2, 0.9, dense["n"]),
                                n_draws=24, seed=5, tol=1e-10, max_iter=400)
    G = dense["ctx"].G
    for k in range(3):
        A, Sfull, Q, lay = _dense_A(dense, k)
        ev = np.linalg.eigvals(A).real
        ev = ev[np.argsort(-np.abs(ev))][:3]
        got = out["eigenvalues"][k]
        np.testing.assert_allclose(got[:len(ev)], ev[:len(got)], rtol=1e-7,
                                   atol=1e-9 * abs(ev[0]))
        # the lead vector: an eigenvector of S^- Q, S-normalized, up to a part
        # in S's null space (the streamed block's recovery leaves one; Q, X~u
        # and u'Su do not see it)
        u = np.concatenate([out["lead"][1][:, k], out["lead"][0][:, k]])
        assert u @ Sfull @ u == pytest.approx(1.0, rel=1e-6)
        wS, VS = np.linalg.eigh(Sfull)
        null = VS[:, wS < 1e-9 * wS.max()]
        u_range = u - null @ (null.T @ u)
        assert rel(A @ u, got[0] * u_range) < 1e-6
        # sum lambda^2: exact part plus a Monte Carlo remainder, within its error
        exact = np.trace(A @ A)
        assert abs(out["sum_sq"][k] - exact) <= 5 * out["sum_sq_se"][k] + 1e-9 * exact
        r = np.array(got) ** 2 / out["sum_sq"][k]
        np.testing.assert_allclose(out["ratios"][k], r, rtol=1e-12)
    del G


def test_first_stage_f(dense):
    O, lay, ctx, n = dense["O"], dense["lay"], dense["ctx"], dense["n"]
    rng = np.random.default_rng(3)
    s2, lev = rng.uniform(0.1, 0.4, n), rng.uniform(0.2, 0.9, n)
    y = dense["f"]["p0"]["V"][:, 0]
    out = O.weak_id_diagnostics(lay, y, s2, lev, n_draws=8, seed=1)
    tab = O.fit_smoother_1d(s2, lev)
    sm = O.smooth_eval_1d(tab, lev)
    yt = lay.root * (y - ctx.w @ y / ctx.w.sum())
    A = dense["A"]
    for k in range(3):
        u = np.concatenate([out["lead"][1][:, k], out["lead"][0][:, k]])
        x1 = lay.root * (A @ u)
        assert out["f_stat"][k] == pytest.approx((x1 @ yt) ** 2 / ((x1 ** 2) @ sm), rel=1e-10)


def test_ritz_filter_drops_a_lone_spurious_value(oracle_mod):
    O = oracle_mod
    # a tridiagonal whose first row decouples: the trailing block's values
    # are eigenvalues of T with its first row and column deleted, lone ones,
    # unseen by the starting vector -- spurious, dropped; 2.0 is genuine
    alpha = np.array([2.0, 5.0, 1.0, 3.0])
    beta = np.array([1e-14, 0.5, 0.2, 1e-3])
    theta, bounds, S = O.ritz_top(alpha, beta, 4)
    np.testing.assert_allclose(theta, [2.0], rtol=1e-12)
    # coupled, every value is genuine
    theta, _b, _S = O.ritz_top(alpha, np.array([0.7, 0.5, 0.2, 1e-3]), 4)
    assert len(theta) == 4
    # merged copies: two equal eigenvalues count once
    theta, _b, _S = O.ritz_top(np.array([1.0, 1.0]), np.array([0.0, 1e-3]), 4)
    assert len(theta) == 1 and theta[0] == pytest.approx(1.0)


# ------------------------------------------------ the q = 1 weak interval

def _lead(dense, k):
    """The leading eigenvector of S^- Q_k, S-normalised, split into the
    level and group blocks, and its eigenvalue."""
    A, Sfull, _Q, _lay = _dense_A(dense, k)
    ev, V = np.linalg.eig(A)
    j = int(np.argmax(np.abs(ev.real)))
    u = V[:, j].real
    u = u / np.sqrt(u @ Sfull @ u)
    G = dense["ctx"].G
    return float(ev[j].real), u[G:], u[:G]


def _c2(dense, k, lam, x1):
    B2 = dense["B"][k] - lam * np.outer(x1, x1)
    b2 = np.diag(B2) / np.diag(dense["M"])
    M = dense["M"]
    return B2 - 0.5 * (np.diag(b2) @ M + M @ np.diag(b2)), b2


def test_eigen_rows_are_x_tilde_u(dense):
    O, lay = dense["O"], dense["lay"]
    lam, ul, ug = _lead(dense, 0)
    x1 = O.eigen_rows(lay, ul[:, None], ug[:, None])[:, 0]
    assert rel(x1, dense["Xt"] @ np.concatenate([ug, ul])) < 1e-12
    # u1'S u1 = 1 makes x1 a unit vector, and the Rayleigh quotient is lambda
    assert x1 @ x1 == pytest.approx(1.0, rel=1e-9)
    assert x1 @ dense["B"][0] @ x1 == pytest.approx(lam, rel=1e-8)


@pytest.mark.parametrize("per_component", [False, True])
def test_deflated_kernel_is_dense_c2(dense, per_component):
    O, lay, n = dense["O"], dense["lay"], dense["n"]
    leads = [_lead(dense, k) for k in range(3)]
    lams = np.array([t[0] for t in leads])
    x1 = np.stack([O.eigen_rows(lay, t[1][:, None], t[2][:, None])[:, 0] for t in leads], 1)
    C2, b2 = zip(*[_c2(dense, k, lams[k], x1[:, k]) for k in range(3)])
    rng = np.random.default_rng(11)
    v = rng.standard_normal((n, 3, 4) if per_component else (n, 4))
    u = rng.standard_normal((3, n))
    cv = O.apply_c(lay, v, np.stack(b2), per_component=per_component, rank_two=(u, 0.2),
                   deflate=(lams, x1))
    one = np.ones(n)
    for k in range(3):
        K = C2[k] + 0.2 * (np.outer(one, u[k]) + np.outer(u[k], one))
        want = K @ (v[:, k, :] if per_component else v)
        assert rel(cv[:, k, :], want) < 1e-8
        assert np.abs(np.diag(C2[k])).max() < 1e-12 * np.abs(C2[k]).max()


def test_weak_moments_dense(dense):
    O, lay, ctx, n = dense["O"], dense["lay"], dense["ctx"], dense["n"]
    leads = [_lead(dense, k) for k in range(3)]
    lams = np.array([t[0] for t in leads])
    Ul = np.stack([t[1] for t in leads], 1)
    Ug = np.stack([t[2] for t in leads], 1)
    x1 = dense["Xt"] @ np.vstack([Ug, Ul])
    y = dense["f"]["p0"]["V"][:, 0]
    yt = lay.root * (y - ctx.w @ y / ctx.w.sum())
    M = dense["M"]
    # the exact leave-out sigma2 of the observation: y~_i (M y~)_i / M_ii
    s2 = yt * (M @ yt) / np.diag(M)
    rng = np.random.default_rng(5)
    s = rng.uniform(-0.05, 0.4, (3, n))            # some negative: the floor matters
    b_ii = np.stack([np.diag(B) for B in dense["B"]])
    nd, seed = 12, 77
    out = O.weak_moments(lay, y, s2, s, 1.0 - np.diag(M), np.diag(M), np.ones(n), b_ii, lams,
                         Ul, Ug, (0, 1, 2), nd, 5, seed)
    sf = np.maximum(s, 0.0)
    Z = O.rademacher(0, n, nd, seed, 0)
    names = range(3)
    for k in names:
        C2, _ = _c2(dense, k, lams[k], x1[:, k])
        c2y = C2 @ yt
        p = out[k]
        assert p["b1"] == pytest.approx(x1[:, k] @ yt, rel=1e-10)
        assert p["v_raw"] == pytest.approx((x1[:, k] ** 2) @ s2, rel=1e-10)
        assert p["sigma"][0, 0] == pytest.approx((x1[:, k] ** 2) @ sf[k], rel=1e-10)
        assert p["sigma"][0, 1] == pytest.approx(2 * (x1[:, k] * sf[k]) @ c2y, rel=1e-7)
        Om = Z * np.sqrt(sf[k])[:, None]
        tr = (sf[k][:, None] * (C2 @ Om) ** 2).sum() / nd
        assert p["trace"] == pytest.approx(tr, rel=1e-7)
        s22 = 4 * sf[k] @ c2y ** 2 - 2 * tr
        if not p["conservative"]:
            assert p["sigma"][1, 1] == pytest.approx(s22, rel=1e-7)
        # theta1 = y~'C2 y~ = y~'C y~ - lambda (b1^2 - V[b1] raw) when sigma2
        # is the leave-out one: the deflation moves lambda b1^2 out of the
        # quadratic form and lambda sum x1^2 sigma2 out of its bias
        assert p["quadratic"] == pytest.approx(yt @ c2y, rel=1e-8)
        theta = yt @ dense["C"][k] @ yt
        assert p["quadratic"] == pytest.approx(
            theta - lams[k] * (p["b1"] ** 2 - p["v_raw"]), rel=1e-7, abs=1e-10 * abs(theta))


def test_am_cdf_is_pinned(oracle_mod):
    from scipy import integrate, special
    O = oracle_mod
    # the limits: |N(0,1)| at kappa = 0, the chi(2) radius as kappa -> inf
    for z in (0.7, 1.96, 2.6):
        assert O.am_cdf(z, 0.0) == pytest.approx(special.erf(z / np.sqrt(2)), abs=1e-15)
        assert O.am_cdf(z, 1e9) == pytest.approx(1 - np.exp(-z * z / 2), abs=1e-8)
        # first order in small kappa: rho ~ b + kappa a^2 / 2, so the law
        # moves by kappa E[a^2] / 2 times b's density 2 phi(z)
        k = 1e-5
        first = special.erf(z / np.sqrt(2)) - k * np.exp(-z * z / 2) / np.sqrt(2 * np.pi)
        assert O.am_cdf(z, k) == pytest.approx(first, abs=1e-9)
    # the region itself, integrated in two dimensions, at moderate curvature
    for z, k in ((1.3, 0.4), (2.1, 2.5)):
        c = 1.0 / k
        hn = lambda x: np.sqrt(2 / np.pi) * np.exp(-0.5 * x * x)
        top = lambda b: np.sqrt(max((z + c) ** 2 - (b + c) ** 2, 0.0))
        want, _ = integrate.dblquad(lambda a, b: hn(a) * hn(b), 0.0, z, lambda b: 0.0, top,
                                    epsabs=1e-12, epsrel=1e-10)
        assert O.am_cdf(z, k) == pytest.approx(want, abs=1e-9)
    # and by simulation
    rng = np.random.default_rng(0)
    a, b = np.abs(rng.standard_normal((2, 2_000_000)))
    for z, k in ((1.5, 0.5), (2.2, 3.0)):
        c = 1.0 / k
        p = np.mean(np.sqrt(a * a + (b + c) ** 2) - c <= z)
        assert abs(O.am_cdf(z, k) - p) < 5 * np.sqrt(p * (1 - p) / a.size)


def test_am_critical_value_limits(oracle_mod):
    O = oracle_mod
    assert O.am_critical_value(0.0) == pytest.approx(1.959963984540054, abs=1e-12)
    assert O.am_critical_value(1e9) == pytest.approx(np.sqrt(-2 * np.log(0.05)), abs=1e-7)
    z = O.am_critical_value(1.5, 0.1)
    assert O.am_cdf(z, 1.5) == pytest.approx(0.9, abs=1e-12)
    prev = 0.0
    for k in (0.01, 0.1, 1.0, 10.0):
        cur = O.am_critical_value(k)
        assert cur > prev
        prev = cur


def test_am_interval_is_the_projected_ellipse(oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(2)
    p = np.linspace(0.0, 2 * np.pi, 400_001)
    for _ in range(6):
        lam, b1, th = rng.normal() * rng.choice([0.05, 1.0, 20.0]), rng.normal(), rng.normal()
        R = rng.normal(size=(2, 2))
        S = R @ R.T + 0.1 * np.eye(2)
        lo, hi, info = O.am_interval(lam, b1, th, S)
        z = info["z"]
        Lc = np.linalg.cholesky(S)
        b = b1 + z * Lc[0, 0] * np.cos(p)
        t = th + z * (Lc[1, 0] * np.cos(p) + Lc[1, 1] * np.sin(p))
        g = lam * b * b + t
        assert lo == pytest.approx(g.min(), abs=1e-9 * (1 + abs(g).max()))
        assert hi == pytest.approx(g.max(), abs=1e-9 * (1 + abs(g).max()))
        assert lo <= lam * b1 ** 2 + th <= hi
    # a linear map (lambda = 0): theta1 -/+ z sqrt(V[theta1]) at the |N| critical value
    S = np.array([[0.3, 0.1], [0.1, 0.5]])
    lo, hi, info = O.am_interval(0.0, 0.4, 1.0, S)
    assert info["kappa"] == 0.0 and info["z"] == pytest.approx(1.959963984540054)
    assert (lo, hi) == pytest.approx((1.0 - info["z"] * np.sqrt(0.5),
                                      1.0 + info["z"] * np.sqrt(0.5)), abs=1e-12)
