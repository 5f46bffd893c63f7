"""The device adjoint of a weighted fixed-effect fit (fe_components) against
the dense FP64 autograd restatement on a 2,000-worker / 200-firm panel
(vector-Jacobian products in y and w for all outputs), and against central
FP64 finite differences of the device fit itself on a larger panel."""

import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

FML = "y ~ x1 + x2 | worker_id + firm_id + year"
FE = ["worker_id", "firm_id", "year"]


def _inputs(d):
    import torch
    y = torch.tensor(d["y"], dtype=torch.float64, device="cuda", requires_grad=True)
    w = torch.tensor(d["w"], dtype=torch.float32, device="cuda", requires_grad=True)
    return y, w


def test_adjoint_matches_the_dense_gradient(oracle_mod):
    import torch
    from hdfe_stream_amd.autograd import fe_components
    from hdfe_stream_amd.simulate import weighted_panel
    O = oracle_mod
    d = weighted_panel(2000, 200, seed=4)
    print("adjoint vs dense, 2000 workers / 200 firms", flush=True)
    y, w = _inputs(d)
    out = fe_components(FML, d, y=y, weights=w, tol=1e-13)
    g = torch.tensor([0.7, -1.3, 2.0, 0.5, -1.1], dtype=torch.float64, device="cuda")
    (out * g).sum().backward()
    yc = torch.tensor(d["y"], dtype=torch.float64, requires_grad=True)
    wc = torch.tensor(d["w"].astype(np.float64), requires_grad=True)
    X = torch.tensor(np.column_stack([d["x1"], d["x2"]]), dtype=torch.float64)
    codes = [np.unique(d[f], return_inverse=True)[1] for f in FE]
    ref = O.dense_fe_components(yc, wc, X, codes)
    (ref * g.cpu()).sum().backward()
    assert rel(out.detach().cpu().numpy(), ref.detach().numpy()) < 1e-9
    assert rel(y.grad.cpu().numpy(), yc.grad.numpy()) < 1e-7
    assert rel(w.grad.double().cpu().numpy(), wc.grad.numpy()) < 1e-5   # w.grad is float32
    assert w.grad.dtype == torch.float32 and y.grad.dtype == torch.float64


@pytest.mark.parametrize("which", [0, 2, 4])
def test_adjoint_matches_finite_differences(which):
    import torch
    from hdfe_stream_amd.autograd import fe_components
    from hdfe_stream_amd.simulate import weighted_panel
    d = weighted_panel(50_000, 5_000, seed=5)
    print(f"adjoint vs finite differences, output {which}", flush=True)
    y = torch.tensor(d["y"], dtype=torch.float64, device="cuda", requires_grad=True)
    w = torch.tensor(d["w"].astype(np.float64), device="cuda", requires_grad=True)
    out = fe_components(FML, d, y=y, weights=w, tol=1e-13)
    out[which].backward()
    rng = np.random.default_rng(which)
    for i in rng.choice(len(d["y"]), 6, replace=False):
        for t, grad, h in ((y, y.grad, 0.05), (w, w.grad, 0.05)):
            vals = []
            for sgn in (1.0, -1.0):
                tt = t.detach().clone()
                tt[i] += sgn * h
                args = (tt, w.detach()) if t is y else (y.detach(), tt)
                vals.append(float(fe_components(FML, d, y=args[0], weights=args[1],
                                                tol=1e-13)[which]))
            num = (vals[0] - vals[1]) / (2 * h)
            an = float(grad[i])
            assert abs(num - an) <= 1e-4 * abs(an) + 1e-11, (i, t is y, num, an)


def test_unweighted_gradient_in_y():
    import torch
    from hdfe_stream_amd.autograd import fe_components
    from hdfe_stream_amd.simulate import akm_panel
    d = akm_panel(3000, 300, seed=9)
    y = torch.tensor(d["y"], dtype=torch.float64, device="cuda", requires_grad=True)
    out = fe_components("y ~ x1 | worker_id + firm_id", d, y=y, tol=1e-13)
    out[1].backward()     # var(psi)
    # the derivative of a variance of fitted effects in y sums to zero over
    # rows (adding a constant to y moves only the effects' level)
    assert abs(float(y.grad.sum())) < 1e-10 * float(y.grad.abs().sum())
