"""The oracle's covariate border (inverse.py:182-231): with covariates in the
leave-out design the coefficient space is [X, fixed effects] and the solves
go through the Schur complement on the covariate corner.  Pinned against the
dense pseudo-inverse of the full design A = [D_g, D_o, X]: the
coefficient-space inverse, the JLA projections draw by draw, and the
Hutchinson moments draw by draw."""

import numpy as np
import pytest
from scipy.linalg import null_space

FE = ["worker_id", "firm_id"]


def _loo(O, d):
    a, b = np.asarray(d["worker_id"]), np.asarray(d["firm_id"])
    ua, ia = np.unique(a, return_inverse=True)
    ub, ib = np.unique(b, return_inverse=True)
    alive, _ = O.leave_one_out(ia, ib, len(ua), len(ub))
    keep = alive[ia] & alive[len(ua) + ib]
    return {k: np.asarray(v)[keep] for k, v in d.items()}


@pytest.fixture(scope="module", params=[False, True], ids=["unweighted", "weighted"])
def design(oracle_mod, request):
    from hdfe_stream_amd.simulate import akm_panel
    O = oracle_mod
    d = akm_panel(160, 14, seed=9, p_move=0.3)
    if request.param:
        d["w"] = np.random.default_rng(5).uniform(0.5, 2.0, len(d["y"]))
    d = _loo(O, d)
    f = O.fit(d, "y", ("x1", "x2"), FE, weights="w" if request.param else None, tol=1e-13,
              fixef_rm="none", stream="worker_id")
    X = f["p0"]["V"][:, 1:]
    ctx = O.KSSContext(f, tol=1e-13, X=X)
    A = np.hstack([ctx.Dg.toarray(), ctx.Do.toarray(), X])
    return O, f, ctx, A


def _krylov_solution(ctx, A, u):
    """The representative of a solution u (G + K, k) that the reduced CG
    started at zero reaches: its level part x satisfies N'D x = 0 for the
    null vectors N of the full design, D = diag of the reduced S."""
    G, L, w = ctx.G, ctx.L, ctx.w
    Nf = null_space(A * np.sqrt(w)[:, None])
    assert np.abs(Nf[G + L:]).max() < 1e-10          # the covariates are identified
    d = ctx.S.diagonal()
    Nl = Nf[G:G + L]
    alpha = np.linalg.lstsq(Nl.T @ (d[:, None] * Nl), -(Nl.T @ (d[:, None] * u[G:G + L])),
                            rcond=None)[0]
    return u + Nf @ alpha


def test_apply_inverse_with_covariates_is_the_pseudo_inverse(design):
    O, f, ctx, A = design
    G, K = ctx.G, ctx.K
    assert K == ctx.L + 2
    w = ctx.w
    Sf = A.T @ (w[:, None] * A)
    rng = np.random.default_rng(4)
    t = rng.standard_normal((G + K, 3))
    v = Sf @ t                                        # in the range
    U, Ug, rows = O.apply_inverse(ctx, v[G:], v[:G])
    assert U.shape == (K, 3)
    u = np.vstack([Ug, U])
    np.testing.assert_allclose(rows, A @ u, rtol=0, atol=1e-12)
    np.testing.assert_allclose(A @ u, A @ t, rtol=0, atol=1e-8)
    np.testing.assert_allclose(U[ctx.L:], t[G + ctx.L:], rtol=1e-8, atol=1e-10)
    # a missing covariate block reads as zero
    U2, _, rows2 = O.apply_inverse(ctx, np.vstack([v[G:G + ctx.L], np.zeros((2, 3))]), v[:G])
    U3, _, rows3 = O.apply_inverse(ctx, v[G:G + ctx.L], v[:G])
    np.testing.assert_array_equal(rows2, rows3)


def test_jla_with_covariates_projects_with_the_full_hat_matrix(design):
    """Draw by draw, the five moments equal those of H r for the hat matrix
    H = W^1/2 A (A'WA)^+ A' W^1/2 of the full design (leaveout.py:130-198)."""
    O, f, ctx, A = design
    p0, c = f["p0"], f["cells"]
    X = ctx.X
    nd, seed = 12, 31
    mom = O.jla_leverages(p0, c, f["S"], f["dinv"], nd, 5, seed, tol=1e-13, X=X)
    w = ctx.w
    r = np.sqrt(w)
    H = (r[:, None] * A) @ np.linalg.pinv(A.T @ (w[:, None] * A)) @ (A.T * r)
    R = O.rademacher(0, len(w), nd, seed, 0)
    P = H @ R
    Pm, Mm = P * P, (R - P) ** 2
    want = np.stack([Pm.sum(1), Mm.sum(1), (Pm * Pm).sum(1), (Mm * Mm).sum(1),
                     (Pm * Mm).sum(1)])
    np.testing.assert_allclose(mom, want, rtol=1e-8, atol=1e-10)
    # the covariates change the operator: without them it is another one
    mom0 = O.jla_leverages(p0, c, f["S"], f["dinv"], nd, 5, seed, tol=1e-13)
    assert np.abs(mom0[0] - mom[0]).max() > 1e-3


def test_hutchinson_with_covariates_draw_by_draw(design):
    """The eight trace moments equal those of U2 = S^+ A'W Sigma A S^+ Z
    from dense matrices, Z drawn at positions levels, groups, covariates."""
    O, f, ctx, A = design
    G, L, K = ctx.G, ctx.L, ctx.K
    n = len(ctx.g)
    w = ctx.w
    sigma2 = np.random.default_rng(2).uniform(0.1, 0.5, n)
    k = 9
    _, totals = O.hutchinson_trace(ctx, sigma2, k, 4, seed=13)
    Z = O.rademacher(0, L + G + 2, k, 13, 0)
    Zc = np.vstack([Z[L:L + G], Z[:L], Z[L + G:]])      # groups, levels, covariates
    Sp = np.linalg.pinv(A.T @ (w[:, None] * A))
    M = Sp @ A.T @ ((w * sigma2)[:, None] * A) @ Sp
    U2 = _krylov_solution(ctx, A, M @ Zc)
    lv = G + ctx.c["offs"][0] + ctx.p0["codes"][:, 0]
    a1, a2, b1, b2 = Zc[lv], U2[lv], Zc[ctx.g], U2[ctx.g]
    wc = w[:, None]
    dense = np.array([w.sum() * k, (wc * a1 * a2).sum(), (wc * b1 * b2).sum(),
                      (wc * (a1 * b2 + b1 * a2)).sum(), (wc * a1).sum(), (wc * a2).sum(),
                      (wc * b1).sum(), (wc * b2).sum()])
    np.testing.assert_allclose(totals, dense, rtol=1e-8, atol=1e-9)
    del K
