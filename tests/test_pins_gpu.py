"""Device fits pinned to the reference's own published answers: the
reference's benchmark panels (`simulate_akm` restated in numpy) fitted on the
GPU, in FP64 and in mixed precision, give the coefficients the reference
reports for them (benchmarks/results/akm.csv, hdfe_stream (explicit) rows),
from 25,000 workers up to the 5,000,000-worker, 42.5M-row panel."""

import json
from pathlib import Path

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

GOLDEN = json.loads((Path(__file__).parent / "golden" / "akm_coefficients.json").read_text())
FITS = {f["n_workers"]: f for f in GOLDEN["fits"]}
FML = "log_earn ~ age_squared + age_cubed | worker_id + firm_id + year"


def _check(n_workers, precisions=("fp64", "mixed")):
    import torch
    from hdfe_stream_amd.estimator import feols_stream
    from hdfe_stream_amd.simulate import simulate_akm
    d = simulate_akm(n_workers)
    assert len(d["log_earn"]) == FITS[n_workers]["n_rows"]
    cols = {k: torch.from_numpy(v).cuda() for k, v in d.items()}
    want = np.array([FITS[n_workers]["coef_age_squared"], FITS[n_workers]["coef_age_cubed"]])
    for precision in precisions:
        print(f"akm.csv pin: {n_workers} workers, {precision}", flush=True)
        res = feols_stream(FML, cols, precision=precision, tol=1e-10)
        assert res.solver_info["converged"]
        err = np.abs(np.asarray(res.beta) - want) / np.abs(want)
        print(f"  beta {res.beta}, rel err {err}", flush=True)
        assert np.all(err < 1e-9), (precision, err)


@pytest.mark.parametrize("n_workers", [25_000, 100_000, 400_000, 1_000_000])
def test_device_fit_matches_the_reference_benchmark_coefficients(n_workers):
    _check(n_workers)


def test_device_fit_at_five_million_workers_matches_the_reference():
    """The reference's 5M-worker benchmark panel (42,498,893 rows)."""
    _check(5_000_000)


XHDFE = json.loads((Path(__file__).parent / "golden" / "xhdfe_match.json").read_text())


@pytest.mark.parametrize("outcome", ["raw", "demeaned"])
def test_device_match_leave_out_on_the_xhdfe_anchor_panel(oracle_mod, outcome):
    """The device's match-level leave-out on the reference's anchor panel:
    draw for draw equal to the oracle's (4,096 JLA and Hutchinson draws), and
    within the draws' Monte Carlo error of xhdfe's exact KSS var(psi); the
    oracle's exact-leverage version reproduces xhdfe to 1e-10 (CPU test)."""
    from hdfe_stream_amd.leaveout import leave_out_kss
    from test_oracle_pins import anchor_panel
    data, _ = anchor_panel(outcome == "demeaned", oracle_mod)
    print(f"xhdfe anchor panel, {outcome}", flush=True)
    lo = leave_out_kss("y ~ 1 | worker_id + firm_id", data, n_draws=4096, seed=1,
                       precision="fp64", tol=1e-12)
    ref = oracle_mod.leave_out_match(data, n_draws=4096, block=256, seed=1, tol=1e-12)
    names = ["var(psi)", "var(alpha)", "cov(psi, alpha)"]
    got = np.array([lo.leave_out[k] for k in names])
    assert np.max(np.abs(got - np.array(ref["leave_out"]))) < 1e-8 * np.max(np.abs(got))
    exact = oracle_mod.leave_out_match(data, exact=True, tol=1e-13)
    assert exact["leave_out"][0] == pytest.approx(XHDFE["kss_var_psi"][outcome], abs=1e-10)
    assert abs(got[0] - XHDFE["kss_var_psi"][outcome]) < 0.004
