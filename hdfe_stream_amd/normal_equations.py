"""A fit's normal equations solved for right-hand sides built from rows, on
the device (the reference's `HDFEResult.solve_normal_equations`,
normal_equations.py:115-231, and `NormalEquationsSolution.predict`, :65-108).

For row vectors v_j (columns of `data`) the solutions are

    u_j = (Z'WZ)^- Z'v_j,

Z the model's full design (a dummy per level of every fixed effect, then the
covariates), W the fit's weights: a coefficient for every level and every
covariate.  The level effects are identified only up to the model's
normalisation; `predict` returns what is identified, z'u_j for a row with
design z.  For an estimation row i, w_i z_i'u_j is the derivative with
respect to y_i of the weighted sum of predictions v_j stands for -- the
adjoint hook the reference exposes.

Z'v is formed on the device (sums of v per level and per group, X'v), the
solve is the fit's own inverse (`CoefficientInverse`: the streamed dimension
absorbed, the covariates by their Schur border, S's null space dropped first
as inverse.py:376-460 does), in blocks of `block` right-hand sides.
"""

from __future__ import annotations

import numpy as np

from .estimator import _column


class NormalEquationsSolution:
    """The solutions u_j of `solve_normal_equations` (normal_equations.py:
    21-131): `names` the right-hand sides, `coef` {"coefficient": the fit's
    coefficient names, name: the covariate block of u_name}."""

    def __init__(self, result, names, U, Ug, coefnames):
        self._result = result
        self.names = list(names)
        self._U = U          # (K, m): the non-streamed levels, then the covariates
        self._Ug = Ug        # (G, m): the streamed dimension's groups
        L = result.panel.info["L"]
        C = U[L:].cpu().numpy()
        self.coef = {"coefficient": list(coefnames)}
        for j, n in enumerate(self.names):
            self.coef[n] = C[:, j].copy()

    def fixef(self, name: str):
        """(level values, solutions (levels, m)) of one fixed effect, in the
        order of the fit's `fixef(name)`."""
        res = self._result
        if name not in res.fe_names:
            raise KeyError(f"{name!r} is not a fixed effect of this model; it has "
                           + ", ".join(map(repr, res.fe_names)))
        levels = res.fe_levels[name]
        if name == res.fe_names[0]:
            return levels, self._Ug.cpu().numpy()
        t = res.fe_names.index(name) - 1
        o, nl = int(res.panel.offs[t]), int(res.panel.n_levels[t])
        return levels, self._U[o:o + nl].cpu().numpy()

    def predict(self, newdata):
        """z'u_j for every row of `newdata` (its fixed effects and
        covariates), (rows, m) float64 on the device; NaN for a row with a
        level not in the estimation sample."""
        import torch
        res = self._result
        codes, ok = _row_codes(res, newdata)
        out = self._Ug[codes[0].clamp(min=0)].clone()
        for t in range(1, len(codes)):
            o = int(res.panel.offs[t - 1])
            out += self._U[o + codes[t].clamp(min=0)]
        kc = len(res.coefnames)
        if kc:
            X = _covariates(res, newdata)
            out += X @ self._U[res.panel.info["L"]:]
        out[~ok] = float("nan")
        return out


def _row_codes(res, data):
    """Each row's code in every fixed effect of the fit (the fit's level
    order; -1 where the level is unknown), and the rows with every level
    known."""
    import torch
    codes = []
    ok = None
    for name in res.fe_names:
        lv = torch.as_tensor(np.asarray(res.fe_levels[name])).to("cuda", torch.int64)
        col = _column(data, name)
        ids = (col if torch.is_tensor(col) else torch.from_numpy(np.asarray(col))).to(
            "cuda", torch.int64)
        order = torch.argsort(lv)
        srt = lv[order]
        pos = torch.searchsorted(srt, ids).clamp(max=max(len(srt) - 1, 0))
        hit = srt[pos] == ids if len(srt) else torch.zeros_like(ids, dtype=torch.bool)
        code = torch.where(hit, order[pos], torch.full_like(pos, -1))
        codes.append(code)
        ok = hit if ok is None else ok & hit
    return codes, ok


def _covariates(res, data):
    import torch
    cols = []
    for name in res.coefnames:
        c = _column(data, name)
        cols.append((c if torch.is_tensor(c) else torch.from_numpy(np.asarray(c))).to(
            "cuda", torch.float64))
    return torch.stack(cols, 1)


def solve_normal_equations(result, data, columns, block: int = 16, tol: float = 1e-10,
                           maxiter: int = 5000, precision: str = "fp64"):
    """u_j = (Z'WZ)^- Z'v_j for the columns v_j of `data` (normal_equations.py:
    115-231): `data` holds the model's fixed effects and covariates and the
    columns; every row's levels must be in the estimation sample."""
    import torch
    from .leaveout import CoefficientInverse
    if isinstance(columns, str):
        columns = [columns]
    columns = list(columns)
    if not columns or len(set(columns)) != len(columns):
        raise ValueError("columns must name one or more distinct columns")
    if not isinstance(block, int) or isinstance(block, bool) or block < 1:
        raise ValueError("block must be a positive integer")
    if result.panel is None or not result.fe_names:
        raise ValueError("the fit has no device panel or no fixed effects")
    panel = result.panel
    if panel.matrix is None:
        raise ValueError("the fit's solver is matrix-free; the normal equations need its "
                         "explicit S")
    codes, ok = _row_codes(result, data)
    if not bool(ok.all()):
        raise ValueError(f"{int((~ok).sum()):,} rows of data have a fixed-effect level that "
                         "is not in the estimation sample")
    V = torch.stack([(c if torch.is_tensor(c) else torch.from_numpy(np.asarray(c))).to(
        "cuda", torch.float64) for c in (_column(data, n) for n in columns)], 1)
    info = panel.info
    L, G = info["L"], info["G"]
    kc = len(result.coefnames)
    X = _covariates(result, data) if kc else None
    inv = CoefficientInverse(panel, tol=tol, maxiter=maxiter, precision=precision,
                             covariates=tuple(range(1, kc + 1)))
    U = torch.empty((L + kc, len(columns)), dtype=torch.float64, device="cuda")
    Ug = torch.empty((G, len(columns)), dtype=torch.float64, device="cuda")
    for j0 in range(0, len(columns), block):
        j1 = min(len(columns), j0 + block)
        v = V[:, j0:j1]
        # Z'v: the sums of v per level of every non-streamed dimension, per
        # group of the streamed one, and X'v
        lv = torch.zeros((L + kc, j1 - j0), dtype=torch.float64, device="cuda")
        for t in range(1, len(codes)):
            lv.index_add_(0, int(panel.offs[t - 1]) + codes[t], v)
        if kc:
            lv[L:] = X.T @ v
        gv = torch.zeros((G, j1 - j0), dtype=torch.float64, device="cuda")
        gv.index_add_(0, codes[0], v)
        u, ug, _ = inv(lv, gv, want_rows=False)
        U[:u.shape[0], j0:j1] = u
        if u.shape[0] < L + kc:
            U[u.shape[0]:, j0:j1] = 0.0
        Ug[:, j0:j1] = ug
    if not inv.converged:
        raise RuntimeError(f"the normal equations' solves did not converge "
                           f"(max relative residual {inv.max_rel:.3g})")
    return NormalEquationsSolution(result, columns, U, Ug, result.coefnames)
