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
