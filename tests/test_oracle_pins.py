"""The oracle's whole fit pinned to the reference's own published answers:
the reference's benchmark panels (`simulate_akm`, restated draw for draw in
numpy) fitted as `log_earn ~ age_squared + age_cubed | worker_id + firm_id +
year` with singletons dropped must give the coefficients the reference
reports for them (benchmarks/results/akm.csv, the hdfe_stream (explicit)
rows; committed in tests/golden/akm_coefficients.json)."""

import json
from pathlib import Path

import numpy as np
import pytest

GOLDEN = json.loads((Path(__file__).parent / "golden" / "akm_coefficients.json").read_text())
FITS = {f["n_workers"]: f for f in GOLDEN["fits"]}
FML_X = ("age_squared", "age_cubed")
FE3 = ("worker_id", "firm_id", "year")


@pytest.mark.parametrize("n_workers", [25_000, 100_000])
def test_simulate_akm_reproduces_the_reference_panel_sizes(n_workers):
    from hdfe_stream_amd.simulate import simulate_akm
    d = simulate_akm(n_workers)
    assert len(d["log_earn"]) == FITS[n_workers]["n_rows"]
    assert d["worker_id"].min() == 1_000_000 and np.all((d["worker_id"] - 1_000_000) % 7 == 0)


@pytest.mark.parametrize("n_workers", [25_000, 100_000])
def test_oracle_fit_matches_the_reference_benchmark_coefficients(oracle_mod, n_workers):
    from hdfe_stream_amd.simulate import simulate_akm
    d = simulate_akm(n_workers)
    f = oracle_mod.fit(d, "log_earn", FML_X, FE3, tol=1e-10)
    want = np.array([FITS[n_workers]["coef_age_squared"], FITS[n_workers]["coef_age_cubed"]])
    err = np.abs(f["beta"] - want) / np.abs(want)
    assert np.all(err < 1e-9), err


def test_a_plausible_generator_slip_breaks_the_pin(oracle_mod):
    """The pin is sharp: the same panel with the noise drawn from the main
    stream instead of seed + 1 moves the coefficients far beyond 1e-9."""
    from hdfe_stream_amd.simulate import simulate_akm
    d = simulate_akm(25_000)
    rng = np.random.default_rng(0)
    n = len(d["log_earn"])
    d["log_earn"] = d["log_earn"] + rng.normal(0, 0.3, n) - \
        np.random.default_rng(1).normal(0, 0.3, n)
    f = oracle_mod.fit(d, "log_earn", FML_X, FE3, tol=1e-10)
    want = np.array([FITS[25_000]["coef_age_squared"], FITS[25_000]["coef_age_cubed"]])
    assert np.max(np.abs(f["beta"] - want) / np.abs(want)) > 1e-6


XHDFE = json.loads((Path(__file__).parent / "golden" / "xhdfe_match.json").read_text())


def anchor_panel(demeaned: bool, oracle):
    """The reference's anchor panel (tests/test_leaveout.py:2236-2244):
    simulate_akm(200, 20, seed=11) pruned to its leave-one-out connected
    set; the outcome raw or demeaned over the pruned rows."""
    from hdfe_stream_amd.simulate import simulate_akm
    d = simulate_akm(200, 20, seed=11)
    ua, ia = np.unique(d["worker_id"], return_inverse=True)
    ub, ib = np.unique(d["firm_id"], return_inverse=True)
    alive, info = oracle.leave_one_out(ia, ib, len(ua), len(ub))
    keep = alive[ia] & alive[len(ua) + ib]
    y = d["log_earn"][keep]
    return {"worker_id": d["worker_id"][keep], "firm_id": d["firm_id"][keep],
            "y": y - y.mean() if demeaned else y}, info


@pytest.mark.parametrize("outcome", ["raw", "demeaned"])
def test_oracle_match_leave_out_reproduces_xhdfe(oracle_mod, outcome):
    """The oracle's match-level leave-out with exact leverages gives xhdfe's
    KSS var(psi) on the anchor panel to 1e-10: pruning, the collapse to
    matches, the stayers' sigma, the reference centring and the person-year
    divisor are all pinned by it."""
    data, info = anchor_panel(outcome == "demeaned", oracle_mod)
    assert info["components_before"] == 2
    out = oracle_mod.leave_out_match(data, exact=True, tol=1e-13)
    assert out["leave_out"][0] == pytest.approx(XHDFE["kss_var_psi"][outcome], abs=1e-10)


def test_weighted_centring_would_miss_xhdfe(oracle_mod):
    """The pin separates the two centrings: centring sigma2 at the weighted
    mean of the match outcomes (not the reference's plain mean of sqrt(T)
    ybar) moves the raw-outcome estimate far beyond 1e-10."""
    data, _ = anchor_panel(False, oracle_mod)
    out = oracle_mod.leave_out_match(data, exact=True, tol=1e-13)
    mf, movers = out["match_fit"], out["movers"]
    T = mf["p0"]["w"]
    yv = mf["p0"]["V"][:, 0]
    s2 = out["sigma2"].copy()
    ybw = float(T @ yv / T.sum())
    s2[movers] = (T * (yv - ybw) * mf["resid"] / out["M_bar"])[movers]
    hat = oracle_mod._exact_match_design(mf["p0"]["g"], mf["p0"]["codes"][:, 0], T,
                                         out["person_years"])
    moved = out["plug_in"][0] - float((T * hat["B"][0]) @ s2)
    assert abs(moved - XHDFE["kss_var_psi"]["raw"]) > 1e-6


def _two_way_design(n_workers, n_firms, n_years, p_move, seed):
    from hdfe_stream_amd.simulate import akm_panel
    d = akm_panel(n_workers, n_firms, n_years=n_years, p_move=p_move, seed=seed)
    n = len(d["y"])
    wi = np.unique(d["worker_id"], return_inverse=True)[1]
    fi = np.unique(d["firm_id"], return_inverse=True)[1]
    A = np.zeros((n, wi.max() + 1 + fi.max() + 1))
    A[np.arange(n), wi] = 1.0
    A[np.arange(n), wi.max() + 1 + fi] = 1.0
    return A


def test_nonlinearity_factor_removes_the_bias_of_one_over_m(oracle_mod):
    """The JLA non-linearity factor F (leaveout.py:240-253): with p = 32
    Rademacher draws, F_i / M_bar_i is an unbiased estimate of 1 / M_ii to
    second order while 1 / M_bar_i is biased upwards.  On a 300-row two-way
    design with exact M_ii, over 2,000 independent draw sets, the mean of
    F / M_bar matches 1 / M_ii within its standard error (per row and
    pooled over rows), the naive 1 / M_bar misses by dozens of standard
    errors, and so does F with the sign of its variance term V or of its bias
    term B flipped."""
    O = oracle_mod
    A = _two_way_design(100, 25, 3, 0.5, 2)
    n = A.shape[0]
    assert n <= 300
    P = A @ np.linalg.pinv(A.T @ A) @ A.T
    M_exact = 1.0 - np.diag(P)
    assert 0.3 < M_exact.min() and M_exact.max() < 0.7
    p, S = 32, 2000
    rng = np.random.default_rng(0)
    est = {k: np.empty((S, n)) for k in ("factor", "naive", "flip_V", "flip_B")}
    for s in range(S):
        r = rng.integers(0, 2, (n, p)) * 2.0 - 1.0
        mom = np.zeros((5, n))
        O.jla_moments(np.ascontiguousarray(r), np.ascontiguousarray(P @ r), mom)
        P_bar, M_bar, F = O.combine_moments(mom, p)
        # the mutants: V and B as the 2021 note defines them, one sign flipped
        _, _, mPP, mMM, mPM = (mom[i] / p for i in range(5))
        V = (M_bar ** 2 * mPP + P_bar ** 2 * mMM - 2 * P_bar * M_bar * mPM) / p
        B = (M_bar * mPP - P_bar * mMM + (M_bar - P_bar) * mPM) / p
        assert np.allclose(F, 1.0 - V / M_bar ** 2 + B / M_bar, rtol=1e-12, atol=1e-14)
        est["factor"][s] = F / M_bar
        est["naive"][s] = 1.0 / M_bar
        est["flip_V"][s] = (1.0 + V / M_bar ** 2 + B / M_bar) / M_bar
        est["flip_B"][s] = (1.0 - V / M_bar ** 2 - B / M_bar) / M_bar

    def z_scores(e):
        per_row = (e.mean(0) - 1.0 / M_exact) / (e.std(0, ddof=1) / np.sqrt(S))
        pooled = (e * M_exact[None, :]).mean(1) - 1.0     # one number per draw set
        return per_row, pooled.mean() / (pooled.std(ddof=1) / np.sqrt(S))

    per_row, pooled = z_scores(est["factor"])
    assert np.max(np.abs(per_row)) < 4.5 and abs(pooled) < 3.0, (per_row, pooled)
    for mutant in ("naive", "flip_V", "flip_B"):
        _, z = z_scores(est[mutant])
        assert z > 8.0, (mutant, z)
