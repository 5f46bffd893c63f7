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
