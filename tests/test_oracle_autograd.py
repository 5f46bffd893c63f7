"""Pinning the dense differentiable restatement of a weighted fixed-effect
fit (the checker of the device adjoint): its values against the oracle's
sparse fit and plug-in moments, its autograd gradients against central
finite differences of itself."""

import numpy as np
import pytest


def _panel(nw=60, nf=8, seed=2):
    from hdfe_stream_amd.simulate import weighted_panel
    return weighted_panel(nw, nf, seed=seed, p_move=0.3)


def _dense_inputs(O, d, fe):
    import torch
    y = torch.tensor(d["y"], dtype=torch.float64)
    w = torch.tensor(d["w"], dtype=torch.float64)
    X = torch.tensor(np.column_stack([d["x1"], d["x2"]]), dtype=torch.float64)
    codes = [np.unique(d[f], return_inverse=True)[1] for f in fe]
    return y, w, X, codes


def test_dense_components_match_the_sparse_fit(oracle_mod):
    O = oracle_mod
    d = _panel()
    fe = ["worker_id", "firm_id", "year"]
    y, w, X, codes = _dense_inputs(O, d, fe)
    out = O.dense_fe_components(y, w, X, codes).numpy()
    f = O.fit(d, "y", ("x1", "x2"), fe, weights="w", tol=1e-13, fixef_rm="none",
              stream="worker_id")
    np.testing.assert_allclose(out[:2], f["beta"], rtol=1e-9)
    p0, c = f["p0"], f["cells"]
    a = f["gam_y"][c["offs"][0] + p0["codes"][:, 0]]
    b = f["worker_effects"][p0["g"]]
    comps = O.plug_in(a, b, p0["w"], len(a))
    np.testing.assert_allclose(out[2:], comps, rtol=1e-8, atol=1e-12)


def test_dense_gradients_match_finite_differences(oracle_mod):
    import torch
    O = oracle_mod
    d = _panel(40, 6, seed=3)
    fe = ["worker_id", "firm_id", "year"]
    y, w, X, codes = _dense_inputs(O, d, fe)
    y.requires_grad_(True)
    w.requires_grad_(True)
    g = torch.tensor([0.7, -1.3, 2.0, 0.5, -1.1], dtype=torch.float64)
    loss = (O.dense_fe_components(y, w, X, codes) * g).sum()
    gy, gw = torch.autograd.grad(loss, (y, w))
    rng = np.random.default_rng(0)
    for i in rng.choice(len(y), 12, replace=False):
        for t, grad, h in ((y, gy, 1e-5), (w, gw, 1e-5)):
            base = t.detach().clone()
            up, dn = base.clone(), base.clone()
            up[i] += h
            dn[i] -= h
            args = [y.detach(), w.detach()]
            k = 0 if t is y else 1
            args[k] = up
            fu = (O.dense_fe_components(args[0], args[1], X, codes) * g).sum()
            args[k] = dn
            fd = (O.dense_fe_components(args[0], args[1], X, codes) * g).sum()
            num = float((fu - fd) / (2 * h))
            assert num == pytest.approx(float(grad[i]), rel=1e-6, abs=1e-9)
