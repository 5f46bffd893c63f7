"""The matrix-free S (the reference's stream_cg fallback, on the device):
its product equals the explicit S's, fits through it equal the explicit
fits, and the fallback is taken -- and reported in the solver info -- when
the explicit S would not fit the memory headroom (stubbed here through
max_memory_gb, as the reference's tests stub the free memory,
tests/test_gpu.py:139-143)."""

import numpy as np
import pytest

from conftest import gpu_available, rel

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

FML = "y ~ x1 + x2 | worker_id + firm_id + year"


@pytest.mark.parametrize("weighted", [False, True])
def test_stream_product_equals_the_explicit_product(c1_panel, weighted):
    import torch
    from hdfe_stream_amd.estimator import Panel
    d = c1_panel
    cols = [torch.from_numpy(d[k]).cuda() for k in ("worker_id", "firm_id", "year")]
    w = None
    if weighted:
        w = torch.from_numpy(np.random.default_rng(3).uniform(0.5, 2.0, len(d["y"]))).cuda()
    P = Panel(cols, [torch.from_numpy(d["y"]).cuda()], weights=w)
    P.system(with_rhs=True)
    explicit = P.matrix
    L = P.info["L"]
    rng = np.random.default_rng(1)
    print(f"stream product weighted={weighted}", flush=True)
    for k in (1, 3, 12, 64, 301):
        # k > 8 takes the wide form (group means; the years, a dense dimension,
        # summed per range of groups)
        X = torch.from_numpy(rng.standard_normal((L, k))).cuda()
        want = explicit.matmat(X).cpu().numpy()
        stream = P.stream_matrix()
        got = stream.matmat(X).cpu().numpy()
        assert np.abs(got - want).max() <= 1e-11 * np.abs(want).max(), k
        got32 = stream.matmat_f32(X.cpu().numpy())
        assert np.abs(got32 - want).max() <= 1e-5 * np.abs(want).max(), k
    P.close()


def test_rebound_operator_equals_a_new_one(c1_panel):
    """A reweighting keeps the cells (IRLS): hdfe_stream_rebind keeps the
    operator's plan and takes the new cell weights, its products equal a new
    operator's bit for bit; an operator over other cells is not rebound."""
    import ctypes as C
    import torch
    from hdfe_stream_amd import _lib
    from hdfe_stream_amd.estimator import Panel
    d = c1_panel
    cols = [torch.from_numpy(d[k]).cuda() for k in ("worker_id", "firm_id", "year")]
    P = Panel(cols, [torch.from_numpy(d["y"]).cuda()],
              weights=torch.from_numpy(np.random.default_rng(2).uniform(0.5, 2.0, len(d["y"]))).cuda())
    P.system(with_rhs=False, budget=0)
    A = P.stream_matrix()
    n, L = P.info["n"], P.info["L"]
    rng = np.random.default_rng(4)
    V = torch.from_numpy(rng.standard_normal((n, P.info["m"]))).cuda().contiguous()
    w = torch.from_numpy(rng.uniform(0.2, 3.0, n)).cuda()
    torch.cuda.synchronize()
    _lib.check(_lib.load().hdfe_panel_reweight(P._h, V.data_ptr(), w.data_ptr()),
               "hdfe_panel_reweight")
    P.system(with_rhs=False, budget=0)
    print("stream rebind", flush=True)
    B = P.stream_matrix(reuse=True)
    assert B is A                       # the same operator, rebound
    fresh = P.stream_matrix()            # a new one over the same cells
    assert fresh is not A
    for k in (3, 301):
        X = torch.from_numpy(rng.standard_normal((L, k))).cuda()
        np.testing.assert_array_equal(A.matmat(X).cpu().numpy(), fresh.matmat(X).cpu().numpy())
    # other cells: the rebind leaves the operator as it was
    Q = Panel([c[: len(c) // 2] for c in cols], [torch.from_numpy(d["y"][: len(d["y"]) // 2]).cuda()])
    Q.system(with_rhs=False, budget=0)
    ok = C.c_int(7)
    _lib.check(_lib.load().hdfe_stream_rebind(
        A._h, Q.info["L"], Q.info["Gi"], Q.info["M"], Q.info["D"], Q.ptr("igstart")[0],
        Q.ptr("icodes")[0], Q.info["M"], Q.ptr("in")[0], P._mf_dinv.data_ptr(), C.byref(ok)),
        "hdfe_stream_rebind")
    assert ok.value == 0
    Q.close()
    P.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_stream_cg_fit_equals_the_explicit_fit(c1_panel, precision):
    from hdfe_stream_amd.estimator import feols_stream
    print(f"stream_cg fit {precision}", flush=True)
    a = feols_stream(FML, c1_panel, precision=precision, tol=1e-11)
    b = feols_stream(FML, c1_panel, precision=precision, tol=1e-11, solver="stream_cg")
    assert b.solver_info["solver"] == "stream_cg" and "fallback" not in b.solver_info
    assert b.solver_info["converged"]
    assert rel(b.beta, a.beta) < 1e-9
    assert np.abs(b.Gamma.cpu().numpy() - a.Gamma.cpu().numpy()).max() <= \
        1e-8 * np.abs(a.Gamma.cpu().numpy()).max()


def test_auto_falls_back_when_s_exceeds_the_headroom(c1_panel):
    from hdfe_stream_amd.estimator import feols_stream
    a = feols_stream(FML, c1_panel, tol=1e-11)
    assert a.solver_info["solver"] == "explicit" and "fallback" not in a.solver_info
    b = feols_stream(FML, c1_panel, tol=1e-11, max_memory_gb=1e-4)
    assert b.solver_info["solver"] == "stream_cg"
    assert "headroom" in b.solver_info["fallback"]
    assert rel(b.beta, a.beta) < 1e-9
    with pytest.raises(MemoryError, match="stream_cg"):
        feols_stream(FML, c1_panel, tol=1e-11, solver="explicit", max_memory_gb=1e-4)
