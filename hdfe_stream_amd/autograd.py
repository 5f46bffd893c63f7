"""Differentiable fixed-effect regressions on the device.

`fe_components(fml, data, y=..., weights=...)` fits `y ~ x | alpha + psi
[+ more]` on the device and returns one tensor

    [beta_1, ..., beta_kx, var(psi), var(alpha), cov(psi, alpha)]

(the coefficients and the weighted plug-in variance components of the
fitted effects, leaveout.py:807-832) as a torch.autograd.Function of the
outcome y and the row weights w.  Its backward is hand-written: one adjoint
solve u = (Z'WZ)^- g with the fit's own operator, then

    dL/dy_i = w_i z_i'u,      dL/dw_i = e_i z_i'u + (the components' own w_i term)

for the gradient g = dL/dtheta of the loss in the fit's coefficients theta.
The reference has no autograd; the same quantities are what its adjoint hook
`solve_normal_equations` / `predict` (normal_equations.py:65-231) returns:
"w_i z_i'u_j is the derivative of the weighted sum of predictions that
right-hand side j stands for with respect to that row's outcome".

The covariate block uses the Schur complement on X: u_X = A_xx^-1 (g_X -
X'W F T^- g_F), z_i'u = (F T^- g_F)_i + x~_i'u_X with x~ = M_F x (the
covariates partialled on the fixed effects) and A_xx = X'W M_F X, the fit's
own V'M_D V block (core/inference.py:109-137).  T^- g_F is the
coefficient-space inverse of the leave-out path (`apply_inverse`,
inverse.py:311-374).  The variance components are invariant to the effects'
normalisation, so their gradient lies in the range of the fixed-effect Gram
matrix and the null-space projection drops nothing.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib

__all__ = ["fe_components", "FEComponents"]


@dataclass
class _Spec:
    yname: str
    xs: list
    fe: list
    psi: str
    data: dict
    tol: float
    maxiter: int
    precision: str
    fixef_rm: str


def _components(res, pi):
    """(six weighted sums, the components and their means) of a fit."""
    import torch
    lib = _lib.load()
    panel = res.panel
    info = panel.info
    n, G = info["n"], info["G"]
    gstart, _ = panel.ptr("gstart")
    codes, _ = panel.ptr("codes")
    w = panel.ptr("w")[0] if info["weighted"] else None
    s6 = np.empty(6)
    torch.cuda.synchronize()
    _lib.check(lib.hdfe_kss_plugin(G, gstart, codes + 4 * pi * n, int(panel.offs[pi]),
                                   res.gam_y.data_ptr(), res.geff.data_ptr(), w, s6.ctypes.data),
               "hdfe_kss_plugin")
    total, sa, sb, saa, sbb, sab = s6
    ma, mb = sa / total, sb / total
    Dp = total * (n - 1.0) / n if n > 1 else total
    scale = total / Dp
    comps = (scale * (saa / total - ma * ma), scale * (sbb / total - mb * mb),
             scale * (sab / total - ma * mb))
    return comps, {"psibar": ma, "alphabar": mb, "W": total, "Dp": Dp}


class FEComponents:
    """The torch.autograd.Function (built lazily, torch is imported late)."""
    _fn = None

    @classmethod
    def fn(cls):
        if cls._fn is None:
            cls._fn = _make_function()
        return cls._fn


def _make_function():
    import torch

    from .estimator import StreamingHDFE
    from .leaveout import CoefficientInverse

    class _F(torch.autograd.Function):
        @staticmethod
        def forward(ctx, y, w, spec):
            cols = dict(spec.data)
            cols[spec.yname] = y.detach()
            wname = None
            if w is not None:
                wname = "__weights__"
                cols[wname] = w.detach().to(torch.float64)
            est = StreamingHDFE(spec.yname, spec.xs, spec.fe, weights=wname, stream=spec.fe[0],
                                tol=spec.tol, maxiter=spec.maxiter, precision=spec.precision,
                                fixef_rm=spec.fixef_rm, keep_intermediates=False)
            res = est.fit(cols)
            others = res.fe_names[1:]
            pi = others.index(spec.psi)
            comps, means = _components(res, pi)
            ctx.res, ctx.pi, ctx.comps, ctx.means, ctx.spec = res, pi, comps, means, spec
            ctx.has_w = w is not None
            ctx.ydtype, ctx.wdtype = y.dtype, (w.dtype if w is not None else None)
            out = torch.tensor(list(res.beta) + list(comps), dtype=torch.float64, device="cuda")
            return out

        @staticmethod
        def backward(ctx, gout):
            res, pi = ctx.res, ctx.pi
            panel = res.panel
            info = panel.info
            lib = _lib.load()
            n, G, L, m = info["n"], info["G"], info["L"], info["m"]
            kx = m - 1
            g = gout.detach().to(torch.float64).cpu().numpy()
            g_beta, (a_psi, a_alpha, a_cov) = g[:kx], g[kx:kx + 3]
            w = panel.ptr("w")[0] if info["weighted"] else None
            gstart, _ = panel.ptr("gstart")
            codes, _ = panel.ptr("codes")
            V, _ = panel.ptr("V")
            direct = rows = None
            xwr = np.zeros(kx)
            if a_psi or a_alpha or a_cov:
                Vpsi, Valpha, Cov = ctx.comps
                mm = ctx.means
                coef = np.array([a_psi, a_alpha, a_cov, mm["psibar"], mm["alphabar"], mm["Dp"],
                                 mm["W"], Vpsi, Valpha, Cov])
                r_psi = torch.empty((n, 1), dtype=torch.float64, device="cuda")
                r_alpha = torch.empty((n, 1), dtype=torch.float64, device="cuda")
                direct = torch.empty(n, dtype=torch.float64, device="cuda")
                torch.cuda.synchronize()
                _lib.check(lib.hdfe_adj_component_rows(
                    G, gstart, codes + 4 * pi * n, int(panel.offs[pi]), res.gam_y.data_ptr(),
                    res.geff.data_ptr(), coef.ctypes.data, r_psi.data_ptr(), r_alpha.data_ptr(),
                    direct.data_ptr()), "hdfe_adj_component_rows")
                inv = CoefficientInverse(panel, tol=ctx.spec.tol, maxiter=ctx.spec.maxiter,
                                         precision=ctx.spec.precision)
                lv = inv.reduce(1, R=r_psi)
                o, nl = int(panel.offs[pi]), panel.n_levels[pi]
                g_levels = torch.zeros((L, 1), dtype=torch.float64, device="cuda")
                g_levels[o:o + nl] = lv[o:o + nl]
                g_groups = inv.group_sums(r_alpha)
                _, _, rows = inv(g_levels, g_groups)
                rows = rows.reshape(n).contiguous()
                if kx:
                    xc = np.arange(1, m, dtype=np.int32)
                    torch.cuda.synchronize()
                    _lib.check(lib.hdfe_adj_xwr(n, m, V, xc.ctypes.data, kx, w, rows.data_ptr(),
                                                xwr.ctypes.data), "hdfe_adj_xwr")
            Xt = None
            uX = np.zeros(max(kx, 1))
            if kx:
                A = res.A[1:, 1:]
                uX = np.linalg.solve(A, g_beta - xwr)
                Xt = torch.empty((n, kx), dtype=torch.float64, device="cuda")
                xc = np.arange(1, m, dtype=np.int32)
                torch.cuda.synchronize()
                _lib.check(lib.hdfe_panel_partial(panel._h, C.c_void_p(res.Gamma.data_ptr()),
                                                  xc.ctypes.data, kx, C.c_void_p(Xt.data_ptr())),
                           "hdfe_panel_partial")
            gy = torch.empty(n, dtype=torch.float64, device="cuda")
            gw = torch.empty(n, dtype=torch.float64, device="cuda")
            uXc = np.ascontiguousarray(uX, np.float64)
            e = res.resid_sorted.contiguous()
            torch.cuda.synchronize()
            _lib.check(lib.hdfe_adj_gradients(
                n, kx, None if rows is None else rows.data_ptr(),
                None if Xt is None else Xt.data_ptr(), uXc.ctypes.data, w, e.data_ptr(),
                None if direct is None else direct.data_ptr(), gy.data_ptr(), gw.data_ptr()),
                "hdfe_adj_gradients")
            # sorted rows -> the caller's rows; rows dropped by the fit get 0
            perm = panel.tensor("perm", torch.int32).long()
            n_in = info["n_in"]
            grad_y = torch.zeros(n_in, dtype=torch.float64, device="cuda")
            grad_y[perm] = gy
            grad_w = None
            if ctx.has_w:
                grad_w = torch.zeros(n_in, dtype=torch.float64, device="cuda")
                grad_w[perm] = gw
                grad_w = grad_w.to(ctx.wdtype)
            return grad_y.to(ctx.ydtype), grad_w, None

    return _F


def fe_components(fml: str, data, y=None, weights=None, psi: str | None = None,
                  tol: float = 1e-12, maxiter: int = 5000, precision: str = "fp64",
                  fixef_rm: str = "singleton"):
    """[beta..., var(psi), var(alpha), cov(psi, alpha)] of the weighted fit of
    `fml`, differentiable in `y` and `weights` (device tensors in data's row
    order; `y` defaults to data's outcome column).  alpha is the formula's
    first fixed effect (streamed), psi the second unless named."""
    import torch

    from .estimator import _column, _to_device, parse_formula
    yname, xs, fe = parse_formula(fml)
    if len(fe) < 2:
        raise ValueError("fe_components needs two fixed effects (alpha and psi)")
    psi = psi or fe[1]
    if psi == fe[0] or psi not in fe:
        raise ValueError(f"psi must be one of {fe[1:]}; got {psi!r}")
    if y is None:
        y = _to_device(_column(data, yname), torch.float64)
    cols = {c: _column(data, c) for c in dict.fromkeys(list(xs) + list(fe))}
    spec = _Spec(yname, list(xs), list(fe), psi, cols, tol, maxiter, precision, fixef_rm)
    return FEComponents.fn().apply(y, weights, spec)
