"""The q = 1 interval's host arithmetic (am_interval.py) against the oracle:
the statistic's law over the whole range of curvature, its quantile, and
the projection of the ellipse, plus the error behaviour."""

import numpy as np
import pytest


@pytest.mark.parametrize("kappa", [0.0, 1e-12, 1e-6, 1e-3, 0.05, 0.5, 1.0, 4.0, 50.0, 1e6, 1e12])
def test_am_cdf_matches_the_oracle(oracle_mod, kappa):
    from hdfe_stream_amd.am_interval import am_cdf
    for z in (0.05, 0.8, 1.6, 1.96, 2.3, 2.45, 3.5):
        assert am_cdf(z, kappa) == pytest.approx(oracle_mod.am_cdf(z, kappa), abs=2e-13)
    assert am_cdf(0.0, kappa) == 0.0 and am_cdf(-1.0, kappa) == 0.0


def test_am_critical_value_matches_the_oracle(oracle_mod):
    from hdfe_stream_amd.am_interval import am_critical_value
    for alpha in (0.01, 0.05, 0.1):
        for kappa in (0.0, 1e-9, 0.02, 0.3, 2.0, 30.0, 1e9):
            assert am_critical_value(kappa, alpha) == pytest.approx(
                oracle_mod.am_critical_value(kappa, alpha), abs=1e-11)
    with pytest.raises(ValueError):
        am_critical_value(1.0, 1.5)


def test_am_interval_matches_the_oracle(oracle_mod):
    from hdfe_stream_amd.am_interval import am_interval
    rng = np.random.default_rng(9)
    for _ in range(40):
        lam = rng.normal() * rng.choice([1e-3, 0.1, 1.0, 30.0])
        b1, th = rng.normal(scale=2.0), rng.normal()
        R = rng.normal(size=(2, 2))
        S = R @ R.T + rng.choice([1e-3, 0.1]) * np.eye(2)
        alpha = rng.choice([0.05, 0.1])
        lo, hi, info = am_interval(lam, b1, th, S, alpha)
        olo, ohi, oinfo = oracle_mod.am_interval(lam, b1, th, S, alpha)
        scale = 1.0 + abs(olo) + abs(ohi)
        assert lo == pytest.approx(olo, abs=1e-10 * scale)
        assert hi == pytest.approx(ohi, abs=1e-10 * scale)
        for key in ("kappa", "z", "rho"):
            assert info[key] == pytest.approx(oinfo[key], rel=1e-10, abs=1e-14)


def test_am_interval_needs_a_positive_definite_covariance():
    from hdfe_stream_amd.am_interval import am_interval
    with pytest.raises(np.linalg.LinAlgError):
        am_interval(0.5, 1.0, 0.2, np.array([[1.0, 2.0], [2.0, 1.0]]))
