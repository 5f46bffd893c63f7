"""Device JLA leverages against the oracle's restatement on the same panel,
seed and row order: the draws are the same counter-based vectors, so the
moments agree draw for draw, to the solver tolerance."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _panel(d, fe):
    from hdfe_stream_amd.estimator import Panel
    return Panel([d[f] for f in fe], [d["y"]], sort_vars=True)


@pytest.mark.parametrize("fe", [["worker_id", "firm_id", "year"], ["worker_id", "firm_id"]])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_jla_moments_match_the_oracle(oracle_mod, fe, precision):
    from hdfe_stream_amd.leaveout import jla_leverages
    from hdfe_stream_amd.simulate import akm_panel
    O = oracle_mod
    d = akm_panel(1500, 120, seed=6, p_move=0.2)
    P = _panel(d, fe)
    P.system(with_rhs=False)
    tm = {"keep_raw": True}
    # mixed precision stops on the true FP64 residual, whose rounding floor on
    # these systems is near 1e-12; FP64 CG stops on its recursive residual
    tol = 1e-12 if precision == "fp64" else 1e-10
    lev = jla_leverages(P, n_draws=96, block=40, seed=11, tol=tol, precision=precision,
                        timings=tm)
    p0 = O.pass0(d, "y", [], fe)
    c = O.pass1(p0)
    S = O.build_S(c)
    dinv = O.jacobi(O.diag(c))
    ref = O.jla_leverages(p0, c, S, dinv, n_draws=96, block=40, seed=11, tol=1e-12)
    raw = tm["raw_moments"].cpu().numpy()
    assert np.abs(raw - ref).max() <= 1e-7 * np.abs(ref).max()
    Pb, Mb, F = O.combine_moments(ref, 96)
    np.testing.assert_allclose(lev.leverage.cpu().numpy(), Pb, rtol=0, atol=1e-8)
    np.testing.assert_allclose(lev.factor.cpu().numpy(), F, rtol=0, atol=1e-7)
    assert lev.diagnostics["converged"]
    assert abs(lev.diagnostics["mean_leverage"] - Pb.mean()) < 1e-9


def test_blocking_does_not_change_the_moments():
    from hdfe_stream_amd.leaveout import jla_leverages
    from hdfe_stream_amd.simulate import akm_panel
    d = akm_panel(1500, 120, seed=7)
    P = _panel(d, ["worker_id", "firm_id", "year"])
    a = {"keep_raw": True}
    b = {"keep_raw": True}
    jla_leverages(P, n_draws=70, block=70, seed=3, tol=1e-12, precision="fp64", timings=a)
    jla_leverages(P, n_draws=70, block=23, seed=3, tol=1e-12, precision="fp64", timings=b)
    ra, rb = a["raw_moments"].cpu().numpy(), b["raw_moments"].cpu().numpy()
    assert np.abs(ra - rb).max() <= 1e-9 * np.abs(ra).max()


def test_device_rademacher_matches_the_golden_table():
    import torch
    from pathlib import Path
    import ctypes as C
    from hdfe_stream_amd import _lib
    g = np.load(Path(__file__).parent / "golden" / "rademacher.npz")
    n, k = g["table"].shape
    out = torch.empty((n, k), dtype=torch.float64, device="cuda")
    _lib.check(_lib.load().hdfe_rademacher(int(g["ordinal"]), n, k, int(g["seed"]),
                                           int(g["first_draw"]), out.data_ptr()), "rademacher")
    np.testing.assert_array_equal(out.cpu().numpy(), g["table"])
