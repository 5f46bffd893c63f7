"""Leave-out (KSS) estimation on the device: JLA leverages by random
projection and the Hutchinson bias term, each solved as blocks of many
right-hand sides, and `leave_out_kss` leaving out an observation.

Mirrors hdfe_stream's `jla_leverages` (leaveout.py:130-198) and
`_combine_moments` (:209-282) over the operator `project_rows`
(inverse.py:240-309): per block of draws, a row pass reduces the Rademacher
vectors onto the levels, one block solve with k right-hand sides gives the
reduced solution, and a second row pass projects back to the rows and adds
the five running moments.  The draws are keyed on the sorted row ordinal, as
in the reference, so a panel whose rows are in the reference's canonical
order (sorted by group, codes and variables) gets the same draws.

The bias term mirrors `hutchinson_trace` / `_trace_block` (leaveout.py:
321-448) over `apply_inverse` (inverse.py:311-460): coefficient-space
Rademacher draws, two applications of the inverse per block, eight
cross-moments.  `leave_out_kss` mirrors leaveout.py:1069-1255 with
leave_out="observation" and `leave_out_components` (:634-805) with the
default stayer rule ("own") and weighted centring.

Leaving out a match (the reference's default) follows leaveout_match.py:
the row model's worker and firm effects plus residual, collapsed to spells,
refitted with the spells' lengths as weights, and KSS on that fit with the
within-match stayer rule and the reference centring.

Not mirrored yet: covariates in the leave-out design (the Schur border,
inverse.py:182-231) -- at observation level a model with covariates is
refused (at match level the row model may have them: they are partialled out
before the collapse).  The components' standard errors are in
`leaveout_se`, the weak-identification diagnostic in `leaveout_weakid`.
"""

from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .leaveout_se import tall_dot

RHS_ZERO_TOL = 1e-11      # core/solve.py:44
TRACE_SEED_OFFSET = 1_000_003     # leaveout.py:730


@dataclass
class JLALeverages:
    """Leverages, their complements and the non-linearity factor, per row in
    the panel's sorted row order (device tensors), as the reference's
    JLALeverages (leaveout.py)."""
    leverage: object
    complement: object
    factor: object
    n_draws: int
    diagnostics: dict = field(default_factory=dict)

    @property
    def n_rows(self):
        return int(self.leverage.shape[0])


class RowPasses:
    """The row-pass state of a Panel (level-to-rows CSRs)."""

    def __init__(self, panel):
        self.panel = panel
        info = panel.info
        D = info["D"]
        codes, _ = panel.ptr("codes")
        offs = np.ascontiguousarray(panel.offs, np.int64)
        nlev = np.ascontiguousarray(panel.n_levels, np.int64)
        h = C.c_void_p()
        _lib.check(_lib.load().hdfe_rows_create(info["n"], info["G"], D, codes, offs.ctypes.data,
                                                nlev.ctypes.data, C.byref(h)), "hdfe_rows_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().hdfe_rows_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _rows(panel):
    rp = getattr(panel, "_row_passes", None)
    if rp is None:
        rp = panel._row_passes = RowPasses(panel)
    return rp


class DesignBorder:
    """The covariate border of the leave-out design (inverse.py:182-231).

    With covariates X beside the fixed effects, partialling out the streamed
    groups leaves the system [[S, T_dx], [T_dx', T_xx]] over (levels,
    covariates), with T_dx = D_o'W M_0 X and T_xx = X'W M_0 X (M_0 the
    weighted within-group projector).  The covariate corner is handled by a
    Schur complement, computed once per design and kept on the panel:
    Z = S^- T_dx (k solves; a column is zero when its norm is at most
    1e-11 sqrt(sum w x^2)) and Schur^+ = pinv(T_xx - T_dx'Z).  A reduced
    right-hand side (r, r_x) then solves as C = Schur^+ (r_x - Z'r),
    U = S^- r - Z C.  Device tensors: X (n, k) the covariates, Xd (n, k) less
    their weighted group means, Xg (G, k) their weighted group sums."""

    def __init__(self, panel, cols, tol, maxiter, precision):
        import torch
        info = panel.info
        n = info["n"]
        self.cols = tuple(int(c) for c in cols)
        self.kc = len(self.cols)
        if self.kc > 64:
            raise ValueError("at most 64 covariates in the leave-out design")
        inv0 = CoefficientInverse(panel, tol, maxiter, precision)
        V = panel.tensor("V", torch.float64).view(n, info["m"])
        self.X = V[:, list(self.cols)].contiguous()
        g = panel.tensor("g", torch.int32).long()
        self.Xg = inv0.group_sums(self.X)
        Wg = torch.where(inv0.Wg > 0, inv0.Wg, torch.ones_like(inv0.Wg))
        self.Xd = (self.X - (self.Xg / Wg[:, None])[g]).contiguous()
        w = panel.tensor("w", torch.float64) if info["weighted"] else None
        T_dx = inv0.reduce(self.kc, R=self.Xd, levels_only=True)
        T_xx = tall_dot(self.Xd, self.Xd, w).cpu().numpy()
        wx2 = (self.X * self.X) if w is None else w[:, None] * self.X * self.X
        scale = np.sqrt(wx2.sum(0).cpu().numpy())
        torch.cuda.synchronize()
        zero = _col_norms(T_dx, info["L"], self.kc) <= RHS_ZERO_TOL * scale
        Z, it, conv, rel = panel.matrix.solve(T_dx, zero, None, tol=tol, maxiter=maxiter,
                                              precision=precision)
        torch.cuda.synchronize()
        self.Z = Z
        schur = T_xx - tall_dot(T_dx, Z).cpu().numpy()
        self.schur_inv = torch.from_numpy(np.linalg.pinv(schur, hermitian=True)).cuda()
        self.info = {"iterations": it, "converged": bool(conv), "max_rel_residual": rel}

    def solve(self, rhs, rhs_x, U):
        """(U - Z C, C) for the reduced solution U = S^- rhs."""
        C = self.schur_inv @ (rhs_x - tall_dot(self.Z, rhs))
        return (U - self.Z @ C).contiguous(), C.contiguous()


def design_border(panel, cols, tol=1e-10, maxiter=5000, precision="mixed"):
    """The panel's `DesignBorder` for covariate columns `cols` (V's columns,
    1-based: column 0 is the outcome), built on first use; None for none."""
    cols = tuple(int(c) for c in cols)
    if not cols:
        return None
    cache = getattr(panel, "_borders", None)
    if cache is None:
        cache = panel._borders = {}
    if cols not in cache:
        cache[cols] = DesignBorder(panel, cols, tol, maxiter, precision)
    return cache[cols]


def design_columns(result, covariates=None):
    """V's columns of the covariates in the leave-out design: the fit's own
    by default (leaveout.py:111-127), none for (), or those named."""
    names = list(result.coefnames)
    if covariates is None:
        return tuple(range(1, len(names) + 1))
    cols = []
    for c in covariates:
        if c not in names:
            raise ValueError(f"{c!r} is not a covariate of the fit; it has {names}")
        cols.append(names.index(c) + 1)
    return tuple(cols)


def jla_leverages(panel, n_draws: int = 250, block: int | None = None, seed: int = 0,
                  tol: float = 1e-10, maxiter: int = 5000, precision: str = "mixed",
                  sqrt_weights: bool | None = None, timings: dict | None = None,
                  covariates=()) -> JLALeverages:
    """Every row's leverage by random projection (leaveout.py:130-198).

    `panel` is a fitted estimator's Panel (its rows, S and solver stay on the
    device).  `block` draws share one block solve (default: all of them, up
    to 1024).  With weights the leverages are those of the square-root-weight
    hat matrix, as in the reference.  `covariates` are V's columns of the
    covariates that join the design (the border, inverse.py:263-281)."""
    import torch
    block = min(int(n_draws), 1024) if block is None else int(block)
    if block < 1 or n_draws < 1:
        raise ValueError("n_draws and block must be at least 1")
    if panel.matrix is None:
        panel.system(with_rhs=False)
    bd = design_border(panel, covariates, tol, maxiter, precision)
    lib = _lib.load()
    info = panel.info
    n, L = info["n"], info["L"]
    weighted = bool(info["weighted"])
    sqrtw = int(weighted if sqrt_weights is None else sqrt_weights)
    rp = _rows(panel)
    g, _ = panel.ptr("g")
    gstart, _ = panel.ptr("gstart")
    codes, _ = panel.ptr("codes")
    offs_d, _ = panel.ptr("offs")
    w, _ = panel.ptr("w") if weighted else (None, 0)
    # the zero-column scale: sqrt(sum_i w_i r_ij^2) (inverse.py:179), which is
    # sqrt(n) in the square-root metric and sqrt(sum w) otherwise
    if sqrtw or not weighted:
        scale = np.sqrt(float(n))
    else:
        scale = float(np.sqrt(panel.array("w", np.float64).sum()))
    moments = torch.zeros((5, n), dtype=torch.float64, device="cuda")
    drawn, blocks, its, conv, rel = 0, 0, [], True, 0.0
    t_solve = t_rows = 0.0
    while drawn < n_draws:
        size = min(block, n_draws - drawn)
        rhs = torch.empty((L, size), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _lib.check(lib.hdfe_rows_center(rp._h, g, gstart, w, sqrtw, seed, drawn, size,
                                        rhs.data_ptr()), "hdfe_rows_center")
        bn = np.empty(size)
        _lib.check(lib.hdfe_col_norms(rhs.data_ptr(), L, size, bn.ctypes.data), "hdfe_col_norms")
        zero = bn <= RHS_ZERO_TOL * scale
        rhs_x = Cb = None
        if bd is not None:
            rhs_x = torch.empty((bd.kc, size), dtype=torch.float64, device="cuda")
            _lib.check(lib.hdfe_rows_cov_center(rp._h, bd.kc, bd.Xd.data_ptr(), w, sqrtw, seed,
                                                drawn, size, rhs_x.data_ptr()),
                       "hdfe_rows_cov_center")
        t1 = time.perf_counter()
        U, it, c, r = panel.matrix.solve(rhs, zero, None, tol=tol, maxiter=maxiter,
                                         precision=precision)
        if bd is not None:
            U, Cb = bd.solve(rhs, rhs_x, U)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        _lib.check(lib.hdfe_rows_project_moments(
            rp._h, gstart, codes, w, sqrtw, seed, drawn, size, offs_d, U.data_ptr(),
            0 if bd is None else bd.kc, None if bd is None else bd.Xd.data_ptr(),
            None if Cb is None else Cb.data_ptr(), moments.data_ptr()),
            "hdfe_rows_project_moments")
        t3 = time.perf_counter()
        t_rows += (t1 - t0) + (t3 - t2)
        t_solve += t2 - t1
        its.append(it)
        conv &= c
        rel = max(rel, r)
        drawn += size
        blocks += 1
    raw = moments.clone() if timings is not None and timings.get("keep_raw") else None
    stats = np.empty(4)
    _lib.check(lib.hdfe_jla_combine(n, n_draws, moments.data_ptr(), stats.ctypes.data),
               "hdfe_jla_combine")
    if timings is not None:
        timings.update(solve_seconds=t_solve, row_pass_seconds=t_rows,
                       rhs_solves_per_second=n_draws / t_solve if t_solve > 0 else float("nan"))
        if raw is not None:
            timings["raw_moments"] = raw
    return JLALeverages(
        leverage=moments[0], complement=moments[1], factor=moments[2], n_draws=n_draws,
        diagnostics={"blocks": blocks, "min_leverage": float(stats[0]),
                     "max_leverage": float(stats[1]), "mean_leverage": float(stats[2]),
                     "leave_one_out_connected": bool(stats[1] < 1.0),
                     "raw_sum_deviation": float(stats[3]), "solver": "explicit",
                     "converged": bool(conv), "iterations": its, "max_rel_residual": rel,
                     "precision": precision, "covariates": list(covariates)})


# ------------------------------------------------------------ bias term

def _col_norms(X, rows, k):
    out = np.empty(k)
    if rows:
        _lib.check(_lib.load().hdfe_col_norms(X.data_ptr(), rows, k, out.ctypes.data),
                   "hdfe_col_norms")
    else:
        out[:] = 0.0
    return out


class CoefficientInverse:
    """u = S^- v for coefficient-space blocks of a fitted Panel
    (`apply_inverse`, inverse.py:311-374): a level block and a streamed block
    (G, k), device tensors.  With covariates in the design (`covariates`,
    V's columns) the level block is (K, k), K = L + k_cov: the fixed-effect
    levels, then the covariates; without, K = L."""

    def __init__(self, panel, tol=1e-10, maxiter=5000, precision="mixed", covariates=()):
        import torch
        if panel.matrix is None:
            panel.system(with_rhs=False)
        self.panel = panel
        self.tol, self.maxiter, self.precision = tol, maxiter, precision
        info = panel.info
        self.n, self.G, self.L, self.D = info["n"], info["G"], info["L"], info["D"]
        self.ncomp = max(int(info["ncomp"]), 1)
        self.g, _ = panel.ptr("g")
        self.gstart, _ = panel.ptr("gstart")
        self.codes, _ = panel.ptr("codes")
        self.offs_d, _ = panel.ptr("offs")
        self.lcomp, nc = panel.ptr("comp")
        if not nc:
            raise RuntimeError("the panel's components are not computed")
        self.w = panel.ptr("w")[0] if info["weighted"] else None
        self.offs = np.ascontiguousarray(panel.offs, np.int64)
        self.nlev = np.ascontiguousarray(panel.n_levels, np.int64)
        self.Wg = torch.empty(self.G, dtype=torch.float64, device="cuda")
        _lib.check(_lib.load().hdfe_kss_group_weight(self.G, self.gstart, self.w,
                                                     self.Wg.data_ptr()),
                   "hdfe_kss_group_weight")
        self.rows_state = _rows(panel)
        self.iterations, self.converged, self.max_rel = [], True, 0.0
        self.border = design_border(panel, covariates, tol, maxiter, precision)
        self.kc = 0 if self.border is None else self.border.kc
        self.K = self.L + self.kc
        if self.kc:
            self.g_long = panel.tensor("g", torch.int32).long()
            self.Wg_safe = torch.where(self.Wg > 0, self.Wg, torch.ones_like(self.Wg))

    def drop_null(self, levels, streamed):
        """In place: the component along S's null space removed
        (inverse.py:376-460)."""
        k = levels.shape[1]
        _lib.check(_lib.load().hdfe_kss_drop_null(
            self.G, self.L, self.D, self.ncomp, k, self.gstart, self.codes, self.lcomp,
            self.offs.ctypes.data, self.nlev.ctypes.data, streamed.data_ptr(),
            levels.data_ptr()), "hdfe_kss_drop_null")

    def reduce(self, k, R=None, V=None, x=None, levels_only=False):
        """A'W diag(x) R onto the level block for rows R (n, k), or A'W rho
        for the spread streamed block V (G, k), rho_i = V[g(i)] / W_g: D_o'
        then, with covariates and not `levels_only`, X' (K rows)."""
        import torch
        kc = 0 if levels_only else self.kc
        out = torch.empty((self.L + kc, k), dtype=torch.float64, device="cuda")
        _lib.check(_lib.load().hdfe_kss_reduce(
            self.rows_state._h, k, self.g, self.w, None if x is None else x.data_ptr(),
            None if R is None else R.data_ptr(), None if V is None else V.data_ptr(),
            self.Wg.data_ptr(), out.data_ptr()), "hdfe_kss_reduce")
        if kc:
            bd = self.border
            if R is not None:
                wx = torch.ones(self.n, dtype=torch.float64, device="cuda") if self.w is None \
                    else self.panel.tensor("w", torch.float64)
                if x is not None:
                    wx = wx * x
                out[self.L:] = tall_dot(bd.X, R, wx)
            else:
                out[self.L:] = tall_dot(bd.Xg, V, 1.0 / self.Wg_safe)
        return out

    def group_sums(self, R, x=None):
        import torch
        k = R.shape[1]
        out = torch.empty((self.G, k), dtype=torch.float64, device="cuda")
        _lib.check(_lib.load().hdfe_kss_group_sums(
            self.G, k, self.gstart, self.w, None if x is None else x.data_ptr(), R.data_ptr(),
            out.data_ptr()), "hdfe_kss_group_sums")
        return out

    def __call__(self, levels, streamed, want_rows=True):
        """(U (K, k), U_streamed (G, k), rows (A u)_i (n, k) or None).
        `levels` has L or K rows (a missing covariate block is zero)."""
        import torch
        levels = levels.to(torch.float64)
        cov = levels[self.L:] if (self.kc and levels.shape[0] > self.L) else None
        levels = levels[:self.L].contiguous().clone()
        streamed = streamed.to(torch.float64).contiguous().clone()
        k = levels.shape[1]
        torch.cuda.synchronize()
        self.drop_null(levels, streamed)
        scale = np.sqrt(_col_norms(levels, self.L, k) ** 2 + _col_norms(streamed, self.G, k) ** 2)
        red = self.reduce(k, V=streamed)
        rhs = (levels - red[:self.L]).contiguous()
        torch.cuda.synchronize()
        zero = _col_norms(rhs, self.L, k) <= RHS_ZERO_TOL * scale
        U, it, conv, rel = self.panel.matrix.solve(rhs, zero, None, tol=self.tol,
                                                   maxiter=self.maxiter,
                                                   precision=self.precision)
        Cb = None
        if self.kc:
            rhs_x = -red[self.L:] if cov is None else cov - red[self.L:]
            U, Cb = self.border.solve(rhs, rhs_x, U)
        torch.cuda.synchronize()
        self.iterations.append(it)
        self.converged &= conv
        self.max_rel = max(self.max_rel, rel)
        Ug = torch.empty((self.G, k), dtype=torch.float64, device="cuda")
        rows = (torch.empty((self.n, k), dtype=torch.float64, device="cuda") if want_rows
                else None)
        _lib.check(_lib.load().hdfe_kss_project(
            self.G, self.D, k, self.gstart, self.codes, self.n, self.offs_d, self.w,
            self.Wg.data_ptr(), U.data_ptr(), streamed.data_ptr(),
            None if rows is None else rows.data_ptr(), Ug.data_ptr()), "hdfe_kss_project")
        if Cb is None:
            return U, Ug, rows
        # the covariates' share: h gains x_i'C, so the streamed block loses
        # its weighted group mean, and the rows gain (x_i - x_bar_g)'C
        bd = self.border
        shift = (bd.Xg @ Cb) / self.Wg_safe[:, None]
        Ug -= shift
        if rows is not None:
            rows += bd.Xd @ Cb
        return torch.cat([U, Cb]), Ug, rows


def _moment_divisor(total, rows):
    """leaveout_se.py:202-214: T (R - 1) / R for R person-year rows."""
    return total * (rows - 1.0) / rows if rows > 1 else total


@dataclass
class TraceEstimate:
    """The three bias terms (leaveout.py:284-310)."""
    var_psi: float
    var_alpha: float
    cov: float
    n_draws: int
    diagnostics: dict = field(default_factory=dict)

    def as_dict(self):
        return {"var(psi)": self.var_psi, "var(alpha)": self.var_alpha,
                "cov(psi, alpha)": self.cov}


def hutchinson_trace(panel, sigma2, n_draws: int = 250, block: int | None = None,
                     seed: int = 0, psi: int = 0, tol: float = 1e-10, maxiter: int = 5000,
                     precision: str = "mixed", timings: dict | None = None,
                     person_years: int | None = None, covariates=()) -> TraceEstimate:
    """tr(Q S^- A'W Omega A S^-) for var(psi), var(alpha) and their
    covariance by Hutchinson's estimator (leaveout.py:321-448).

    `sigma2` is the per-row variance estimate (device, the panel's sorted row
    order); `psi` the index of the non-streamed dimension whose variance is
    wanted (0: the first).  Per block of draws: Z (levels, then groups) from
    the counter-based generator, u = S^- Z with its rows, v = A'W Sigma A u,
    u2 = S^- v, and the eight moments of (Z, u2).  With `covariates` (V's
    columns) in the design the draws take positions levels, groups, then
    covariates (leaveout.py:410-419)."""
    import torch
    block = min(int(n_draws), 256) if block is None else int(block)
    if block < 1 or n_draws < 1:
        raise ValueError("n_draws and block must be at least 1")
    inv = CoefficientInverse(panel, tol, maxiter, precision, covariates=covariates)
    if not 0 <= psi < inv.D:
        raise ValueError(f"psi must index a non-streamed dimension (0..{inv.D - 1})")
    lib = _lib.load()
    L, G = inv.L, inv.G
    sig = sigma2.to(torch.float64).contiguous()
    psi_codes = inv.codes + 4 * psi * inv.n
    totals = np.zeros(8)
    drawn, blocks = 0, 0
    t0 = time.perf_counter()
    while drawn < n_draws:
        size = min(block, n_draws - drawn)
        Z = torch.empty((L + G + inv.kc, size), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        _lib.check(lib.hdfe_rademacher(0, L + G + inv.kc, size, seed, drawn, Z.data_ptr()),
                   "hdfe_rademacher")
        Zl, Zg = (Z[:L], Z[L:L + G]) if not inv.kc else \
            (torch.cat([Z[:L], Z[L + G:]]), Z[L:L + G].contiguous())
        _, _, rows = inv(Zl, Zg)
        vl = inv.reduce(size, R=rows, x=sig)
        vg = inv.group_sums(rows, x=sig)
        del rows
        Ul, Ug, _ = inv(vl, vg, want_rows=False)
        m8 = np.empty(8)
        torch.cuda.synchronize()
        _lib.check(lib.hdfe_kss_trace_moments(G, size, inv.gstart, psi_codes,
                                              int(inv.offs[psi]), inv.w, Zl.data_ptr(),
                                              Zg.data_ptr(), Ul.data_ptr(), Ug.data_ptr(),
                                              m8.ctypes.data), "hdfe_kss_trace_moments")
        totals += m8
        drawn += size
        blocks += 1
    seconds = time.perf_counter() - t0
    weight, a1a2, b1b2, cross, a1, a2, b1, b2 = totals / n_draws
    var_psi = a1a2 / weight - (a1 / weight) * (a2 / weight)
    var_alpha = b1b2 / weight - (b1 / weight) * (b2 / weight)
    cov = 0.5 * (cross / weight - (a1 / weight) * (b2 / weight) - (b1 / weight) * (a2 / weight))
    scale = weight / _moment_divisor(weight, inv.n if person_years is None else person_years)
    if timings is not None:
        timings.update(trace_seconds=seconds, trace_solves=2 * n_draws,
                       trace_iterations=list(inv.iterations))
    return TraceEstimate(
        var_psi=float(scale * var_psi), var_alpha=float(scale * var_alpha),
        cov=float(scale * cov), n_draws=n_draws,
        diagnostics={"blocks": blocks, "n_obs": float(weight), "totals": totals,
                     "converged": bool(inv.converged), "max_rel_residual": inv.max_rel,
                     "iterations": list(inv.iterations), "precision": precision})


# ------------------------------------------------------- the components

@dataclass
class LeaveOutComponents:
    """KSS leave-out variance components with the plug-in ones and the bias
    (leaveout.py:452-520): dicts keyed "var(psi)", "var(alpha)",
    "cov(psi, alpha)"."""
    plug_in: dict
    bias: dict
    leave_out: dict
    n_obs: int
    n_movers: int
    leverages: JLALeverages
    trace: TraceEstimate
    se: object = None
    weak_id: object = None
    diagnostics: dict = field(default_factory=dict)
    pruning: dict | None = None
    fit: object = None
    match_fit: object = None

    def as_dict(self):
        return dict(self.leave_out)

    def summary(self) -> str:
        """The components; with standard errors, the normal intervals at the
        run's confidence (flagged * where the diagnostic finds q >= 1) and
        the q = 1 intervals (leaveout.py:519-627)."""
        from scipy.special import ndtri
        level = self.diagnostics.get("confidence", 0.95)
        zn = float(ndtri(0.5 + level / 2.0))
        lines = [f"KSS leave-out components ({self.n_obs:,} observations, "
                 f"{self.n_movers:,} movers)",
                 f"  {'':18s}{'plug-in':>12s}{'bias':>12s}{'leave-out':>12s}"
                 + (f"{'se':>11s}{f'{level:.0%} interval':>28s}" if self.se is not None
                    else "")]
        for k in self.leave_out:
            row = (f"  {k:18s}{self.plug_in[k]:12.6f}{self.bias[k]:12.6f}"
                   f"{self.leave_out[k]:12.6f}")
            if self.se is not None:
                se = self.se.se[k]
                row += f"{se:11.6f}"
                if np.isfinite(se):
                    weak = self.weak_id is not None and self.weak_id.q[k] >= 1
                    lo, hi = self.leave_out[k] - zn * se, self.leave_out[k] + zn * se
                    row += f"{f'[{lo:.6f}, {hi:.6f}]':>28s}" + ("*" if weak else "")
            lines.append(row)
        weak = (self.se.weak or {}) if self.se is not None else {}
        shown = [k for k in self.leave_out if "interval" in weak.get(k, {})]
        if self.se is not None and self.weak_id is not None and any(
                self.weak_id.q[k] >= 1 for k in self.leave_out):
            lines.append("  * q >= 1: weakly identified, the normal interval understates "
                         "the uncertainty")
        if shown:
            lines += ["", f"q = 1 intervals, valid under weak identification "
                          f"({weak[shown[0]]['confidence']:.0%})",
                      f"  {'':18s}{'leave-out':>12s}{'interval':>28s}{'kappa':>9s}"
                      f"{'z':>8s}{'vs normal':>11s}"]
            for k in shown:
                piece = weak[k]
                lo, hi = piece["interval"]
                normal = 2.0 * zn * self.se.se[k]
                ratio = (hi - lo) / normal if np.isfinite(normal) and normal > 0 else np.nan
                lines.append(f"  {k:18s}{self.leave_out[k]:12.6f}"
                             f"{f'[{lo:.6f}, {hi:.6f}]':>28s}{piece['kappa']:9.3f}"
                             f"{piece['z']:8.4f}{ratio:10.2f}x")
            if any(piece.get("conservative") for piece in weak.values()):
                lines.append("  V[theta1] fell back to its first term (conservative) for "
                             "at least one component")
            if any(self.weak_id.q[k] >= 2 for k in shown):
                lines.append("  q >= 2 somewhere: the q = 1 interval ignores the second "
                             "direction and may undercover")
        if self.se is not None:
            lines += ["", self.se.summary()]
        if self.weak_id is not None:
            lines += ["", self.weak_id.summary()]
        return "\n".join(lines)


def _movers(panel, psi):
    """Rows whose group is seen at more than one level of dimension `psi`
    (leaveout.py:928-947), as a device bool tensor."""
    import torch
    info = panel.info
    n, G = info["n"], info["G"]
    g = panel.tensor("g", torch.int32).long()
    codes = panel.tensor("codes", torch.int32).view(info["D"], n)[psi].long()
    lo = torch.full((G,), np.iinfo(np.int64).max, dtype=torch.int64, device="cuda")
    hi = torch.full((G,), -1, dtype=torch.int64, device="cuda")
    lo.scatter_reduce_(0, g, codes, "amin")
    hi.scatter_reduce_(0, g, codes, "amax")
    return (lo != hi)[g]


def leave_out_components(result, n_draws: int = 250, block: int | None = None, seed: int = 0,
                         psi: str | None = None, tol: float = 1e-10, maxiter: int = 5000,
                         precision: str = "mixed", timings: dict | None = None,
                         stayers: str = "own", stayer_sigma2=None, centering: str = "weighted",
                         person_years: int | None = None, se: bool = False, se_draws=None,
                         se_trace: bool = True, smooth_bins=None,
                         se_variance: str = "person_year", match_within=None,
                         diagnose=None, diagnose_draws: int = 64, weak_interval: bool = True,
                         confidence: float = 0.95, covariates=None) -> LeaveOutComponents:
    """KSS components for a fit (leaveout.py:634-805): JLA leverages, the
    leverage check, the plug-in moments, sigma2, the Hutchinson bias with
    seed + 1_000_003, and plug-in minus bias.

    stayers "own" (every row its own sigma2) or "within_match" with
    `stayer_sigma2` (device, the fit's sorted rows; used on stayers' rows,
    leaveout.py:1012-1016); centering "weighted" or "reference"
    (leaveout.py:996-1003); `person_years` the R of the divisor T (R - 1) / R
    when the rows are matches (leaveout_se.py:216-220).  With "within_match"
    only movers' rows are held to the leverage check (leaveout.py:712-724).
    With `se`, the standard errors of `leaveout_se.component_variances`
    (seed + 5_500_011, `se_draws` draws, default n_draws; leaveout.py:
    759-775).  Leaving out a match (stayers "within_match") the variance of
    each match row is `_match_se_sigma2`'s (leaveout.py:848-894): sigma2 with
    weighted centring on movers' matches; "person_year" adds it over T to
    `match_within` (device, each match's within-match variance, which a
    stayer's match gets alone), "match" takes it with the stayers' values; b
    is zero on stayers' matches and var(alpha) gets no standard error.
    `diagnose` (default: `se`) runs the weak-identification diagnostic with
    `diagnose_draws` trace draws and seed + 9_900_019 (leaveout.py:749-758).
    With `se` and `weak_interval`, each component it flags (q >= 1) also
    gets KSS's q = 1 interval at `confidence` in se.weak[name]["interval"]:
    the Andrews-Mikusheva projection of the ellipse of (b1, theta1), theta1
    = leave-out - lambda_1 (b1^2 - V[b1] raw) (leaveout.py:777-793).
    `covariates` join the fixed effects in the leave-out design: the fit's
    own by default, () for the fixed effects alone, or a list of names
    (leaveout.py:111-127); the solves then go through the border's Schur
    complement (`DesignBorder`)."""
    if not 0.0 < confidence < 1.0:
        raise ValueError("confidence must lie in (0, 1)")
    if stayers not in ("own", "within_match"):
        raise ValueError("stayers must be 'own' or 'within_match' on the device")
    if (stayers == "within_match") != (stayer_sigma2 is not None):
        raise ValueError("stayers='within_match' needs stayer_sigma2")
    if centering not in ("weighted", "reference"):
        raise ValueError("centering must be 'weighted' or 'reference'")
    import torch
    panel = result.panel
    info = panel.info
    cols = design_columns(result, covariates)
    cov_names = [result.coefnames[c - 1] for c in cols]
    alpha = result.fe_names[0]
    others = result.fe_names[1:]
    psi = psi or others[0]
    if psi not in others:
        raise ValueError(f"psi must be a non-streamed dimension, one of {others}; got {psi!r}")
    pi = others.index(psi)
    n, G = info["n"], info["G"]
    R = n if person_years is None else int(person_years)
    t = {} if timings is None else timings
    t0 = time.perf_counter()
    lev = jla_leverages(panel, n_draws=n_draws, block=block, seed=seed, tol=tol,
                        maxiter=maxiter, precision=precision, timings=t, covariates=cols)
    lev.diagnostics["covariates"] = cov_names
    t1 = time.perf_counter()
    movers = _movers(panel, pi)
    match_level = stayers == "within_match"
    if match_level:
        worst = float(lev.leverage[movers].max()) if bool(movers.any()) else 0.0
    else:
        worst = lev.diagnostics["max_leverage"]
    if worst >= 1.0 - 1e-9:
        unit = "match" if match_level else "observation"
        raise ValueError(
            f"maximum leverage is {worst:.6f}: the panel is not leave-one-out connected "
            f"(leaving out one {unit} at a time), so the leave-out residual is "
            f"undefined for at least one {unit}. Refit with "
            "connected='leave_one_out', or use leave_out_kss, which does")
    lib = _lib.load()
    gstart, _ = panel.ptr("gstart")
    codes, _ = panel.ptr("codes")
    w = panel.ptr("w")[0] if info["weighted"] else None
    gam = result.gam_y.contiguous()
    geff = result.geff.contiguous()
    s6 = np.empty(6)
    torch.cuda.synchronize()
    _lib.check(lib.hdfe_kss_plugin(G, gstart, codes + 4 * pi * n, int(panel.offs[pi]),
                                   gam.data_ptr(), geff.data_ptr(), w, s6.ctypes.data),
               "hdfe_kss_plugin")
    total, sa, sb, saa, sbb, sab = s6
    ma, mb = sa / total, sb / total
    scale = total / _moment_divisor(total, R)
    plug_in = {"var(psi)": scale * (saa / total - ma * ma),
               "var(alpha)": scale * (sbb / total - mb * mb),
               "cov(psi, alpha)": scale * (sab / total - ma * mb)}
    V, _ = panel.ptr("V")
    sigma2 = torch.empty(n, dtype=torch.float64, device="cuda")
    resid = result.resid_sorted.contiguous()
    ybar = C.c_double()
    _lib.check(lib.hdfe_kss_sigma2(n, info["m"], V, w, 1 if centering == "reference" else 0,
                                   resid.data_ptr(), lev.complement.data_ptr(),
                                   lev.factor.data_ptr(), sigma2.data_ptr(), C.byref(ybar)),
               "hdfe_kss_sigma2")
    if match_level:
        sigma2 = torch.where(movers, sigma2, stayer_sigma2.to(torch.float64))
    t2 = time.perf_counter()
    trace = hutchinson_trace(panel, sigma2, n_draws=n_draws, block=block,
                             seed=seed + TRACE_SEED_OFFSET, psi=pi, tol=tol, maxiter=maxiter,
                             precision=precision, timings=t, person_years=R, covariates=cols)
    t3 = time.perf_counter()
    bias = trace.as_dict()
    leave_out = {k: plug_in[k] - bias[k] for k in plug_in}
    if info["weighted"]:
        wt = panel.tensor("w", torch.float64)
        sigma2_mean = float((sigma2 / wt).mean())
    else:
        sigma2_mean = float(sigma2.mean())
    weak_id = None
    if bool(se) if diagnose is None else diagnose:
        from .leaveout_weakid import SEED_OFFSET, weak_id_diagnostics
        y_d = panel.tensor("V", torch.float64).view(n, info["m"])[:, 0].contiguous()
        weak_id = weak_id_diagnostics(CoefficientInverse(panel, tol, maxiter, precision,
                                                         covariates=cols), y_d,
                                      sigma2, lev, psi=pi, rows=R, n_draws=int(diagnose_draws),
                                      block=8 if block is None else int(block),
                                      seed=seed + SEED_OFFSET)
        weak_id.diagnostics.update(psi=psi, alpha=alpha, covariates=cov_names)
    errors = None
    if se:
        from .leaveout_se import BII_SEED_OFFSET, component_variances
        if se_variance not in ("person_year", "match"):
            raise ValueError("se_variance must be 'person_year' or 'match'")
        inv = CoefficientInverse(panel, tol, maxiter, precision, covariates=cols)
        y = panel.tensor("V", torch.float64).view(n, info["m"])[:, 0].contiguous()
        sigma2_se, options = sigma2, {}
        if match_level:
            blk = torch.empty(n, dtype=torch.float64, device="cuda")
            yb = C.c_double()
            torch.cuda.synchronize()
            _lib.check(lib.hdfe_kss_sigma2(n, info["m"], V, w, 0, resid.data_ptr(),
                                           lev.complement.data_ptr(), lev.factor.data_ptr(),
                                           blk.data_ptr(), C.byref(yb)), "hdfe_kss_sigma2")
            if se_variance == "match":
                sigma2_se = torch.where(movers, blk, stayer_sigma2.to(torch.float64))
            else:
                if match_within is None:
                    raise ValueError("se_variance='person_year' needs the within-match "
                                     "variances")
                wt = panel.tensor("w", torch.float64)
                mw = match_within.to(torch.float64)
                sigma2_se = torch.where(movers, mw + blk / wt, mw)
            options = {"stayer_rows": ~movers, "centering": centering, "reported": (0, 2),
                       "sigma2_point": sigma2}
        errors = component_variances(
            inv, y, sigma2_se.contiguous(), lev, psi=pi, rows=R,
            n_draws=n_draws if se_draws is None else int(se_draws), block=block,
            seed=seed + BII_SEED_OFFSET, trace=se_trace, smooth_bins=smooth_bins,
            weak_id=weak_id if weak_interval else None, **options)
        errors.diagnostics.update(psi=psi, alpha=alpha, covariates=cov_names,
                                  se_variance=se_variance if match_level else None,
                                  sigma2=sigma2_se)
        if errors.weak:
            from .am_interval import am_interval
            for name, piece in errors.weak.items():
                theta1 = leave_out[name] - piece["lambda_1"] * (piece["b1"] ** 2
                                                                - piece["v_raw"])
                lo, hi, extra = am_interval(piece["lambda_1"], piece["b1"], theta1,
                                            piece["sigma"], 1.0 - confidence)
                piece.update(theta1=float(theta1), interval=(lo, hi), confidence=confidence,
                             **extra)
    t4 = time.perf_counter()
    t.update(leverage_seconds=t1 - t0, sigma2_seconds=t2 - t1, bias_seconds=t3 - t2,
             se_seconds=t4 - t3)
    return LeaveOutComponents(
        plug_in=plug_in, bias=bias, leave_out=leave_out, n_obs=n,
        n_movers=int(movers.sum()), leverages=lev, trace=trace, se=errors, weak_id=weak_id,
        diagnostics={"psi": psi, "alpha": alpha, "stayers": stayers, "max_leverage": worst,
                     "confidence": confidence,
                     "leave_out": "match" if match_level else "observation",
                     "centering": centering, "sigma2_mean": sigma2_mean,
                     "ybar": float(ybar.value), "covariates": cov_names,
                     "n_levels": {psi: panel.n_levels[pi], alpha: G},
                     "sigma2": sigma2})


def _leave_out_match(fit, psi, n_draws, block, seed, tol, maxiter, precision, timings,
                     se=False, se_draws=None, se_trace=True, smooth_bins=None,
                     se_variance="person_year", diagnose=None, diagnose_draws=64,
                     weak_interval=True, confidence=0.95):
    """KSS leaving out a whole (worker, firm) match (leaveout_match.py:
    156-214): the row fit's y~ = the worker effect + the firm effect + the
    residual collapsed to one row per match (its weighted mean, weight the
    spell's total weight; leaveout_match.py:63-88), the weighted fit of
    y ~ 1 | alpha + psi on the matches, sigma_for_stayers on the stayers'
    person-years (leaveout_match.py:91-132), then `leave_out_components` with
    the within-match stayer rule, the reference centring and the person-year
    divisor."""
    import torch

    from .estimator import StreamingHDFE
    panel = fit.panel
    info = panel.info
    n, D = info["n"], info["D"]
    others = fit.fe_names[1:]
    pi = others.index(psi)
    if pi != 0:
        raise NotImplementedError("leaving out a match needs psi to be the second fixed effect "
                                  "of the formula")
    alpha = fit.fe_names[0]
    g = panel.tensor("g", torch.int32).long()
    c0 = panel.tensor("codes", torch.int32).view(D, n)[0].long()
    w = panel.tensor("w", torch.float64) if info["weighted"] else \
        torch.ones(n, dtype=torch.float64, device="cuda")
    off0 = int(panel.offs[0])
    ytil = fit.geff[g] + fit.gam_y[off0 + c0] + fit.resid_sorted
    head = torch.ones(n, dtype=torch.bool, device="cuda")
    head[1:] = (g[1:] != g[:-1]) | (c0[1:] != c0[:-1])
    mid = torch.cumsum(head.long(), 0) - 1
    nm = int(mid[-1]) + 1
    T = torch.zeros(nm, dtype=torch.float64, device="cuda").index_add_(0, mid, w)
    ybar = torch.zeros(nm, dtype=torch.float64, device="cuda").index_add_(0, mid, w * ytil) / T
    first = torch.nonzero(head, as_tuple=True)[0]
    glev = panel.tensor("glevels", torch.int64)
    lev0 = panel.tensor("levels0", torch.int64)
    mdata = {alpha: glev[g[first]], psi: lev0[c0[first]], "__ybar__": ybar, "__T__": T}
    mfit = StreamingHDFE("__ybar__", [], [alpha, psi], weights="__T__", stream=alpha,
                         fixef_rm="none", keep_intermediates=True, tol=tol,
                         maxiter=maxiter).fit(mdata)
    mp = mfit.panel
    # the match fit factorises the same workers and firms: the same codes
    if not (torch.equal(mp.tensor("glevels", torch.int64), glev)
            and torch.equal(mp.tensor("levels0", torch.int64), lev0)):
        raise RuntimeError("internal: the match fit's levels differ from the row fit's")
    # sigma_for_stayers on the person-year rows
    ybar_all = float((w * ytil).sum() / w.sum())
    mn = mp.info["n"]
    mg = mp.tensor("g", torch.int32).long()
    fitted = mfit.geff[g] + mfit.gam_y[int(mp.offs[0]) + c0]
    W_worker = torch.zeros(info["G"], dtype=torch.float64, device="cuda").index_add_(0, g, w)[g]
    r = (ytil - ybar_all) * (ytil - fitted) / (1.0 - w / W_worker)
    firms = torch.zeros(info["G"], dtype=torch.long, device="cuda").index_add_(
        0, g[first], torch.ones_like(first))
    stay_row = firms[g] == 1
    num = torch.zeros(nm, dtype=torch.float64, device="cuda").index_add_(
        0, mid[stay_row], (w * w * r)[stay_row])
    den = torch.zeros(nm, dtype=torch.float64, device="cuda").index_add_(
        0, mid[stay_row], w[stay_row])
    st_tab = torch.where(den > 0, num / torch.where(den > 0, den, torch.ones_like(den)),
                         torch.full_like(num, float("nan")))
    # the table's rows in the match fit's sorted order
    perm = mp.tensor("perm", torch.int32).long()[:mn]
    st_sorted = st_tab[perm].contiguous()
    within = None
    if se and se_variance == "person_year":
        # each match's within-match variance of y~, (1/T) sum (y~ - its mean)^2
        # over the spell's rows, unweighted (leaveout_match.py:135-153)
        cnt = torch.zeros(nm, dtype=torch.float64, device="cuda").index_add_(
            0, mid, torch.ones_like(ytil))
        mu = torch.zeros(nm, dtype=torch.float64, device="cuda").index_add_(0, mid, ytil) / cnt
        within = (torch.zeros(nm, dtype=torch.float64, device="cuda").index_add_(
            0, mid, (ytil - mu[mid]) ** 2) / cnt)[perm].contiguous()
    comps = leave_out_components(mfit, n_draws=n_draws, block=block, seed=seed, psi=psi,
                                 tol=tol, maxiter=maxiter, precision=precision, timings=timings,
                                 stayers="within_match", stayer_sigma2=st_sorted,
                                 centering="reference", person_years=n, se=se,
                                 se_draws=se_draws, se_trace=se_trace, smooth_bins=smooth_bins,
                                 se_variance=se_variance, match_within=within,
                                 diagnose=diagnose, diagnose_draws=diagnose_draws,
                                 weak_interval=weak_interval, confidence=confidence)
    comps.n_obs = n
    comps.diagnostics["n_matches"] = mn
    comps.diagnostics["person_years"] = n
    comps.match_fit = mfit
    return comps


def leave_out_kss(fml: str, data, *, psi: str | None = None, n_draws: int = 250,
                  seed: int = 0, block: int | None = None, prune: bool = True,
                  leave_out: str = "match", stayers: str | None = None,
                  centering: str = "reference",
                  precision: str = "mixed", tol: float = 1e-10, maxiter: int = 5000,
                  timings: dict | None = None, se: bool = False, se_draws=None,
                  se_trace: bool = True, smooth_bins=None, se_variance: str = "person_year",
                  diagnose=None, diagnose_draws: int = 64, weak_interval: bool = True,
                  confidence: float = 0.95, **options) -> LeaveOutComponents:
    """Kline-Saggio-Solvsten leave-out variance components, end to end, on
    the device (leaveout.py:1069-1255): prune to the largest leave-one-out
    connected set of (alpha, psi), fit with alpha streamed and no singleton
    drop, then leave out a match (the default, `_leave_out_match`) or an
    observation (`leave_out_components`).

    `alpha` is the formula's first fixed effect and `psi` (default: the
    second) the dimension whose variance is wanted.  At observation level the
    stayer rule is "own" and the centring weighted; at match level the
    within-match rule and the reference centring (the reference's defaults
    there).  `se`, `se_draws`, `se_trace` and `smooth_bins` ask for the
    components' standard errors (leaveout_se.py); leaving out a match,
    `se_variance` is how each match's error variance is estimated
    ("person_year", the reference implementation's, unweighted only, or
    "match") and var(alpha) gets none (leaveout.py:1136-1170).  `diagnose`
    (default: `se`) and `diagnose_draws` run the weak-identification
    diagnostic (leaveout_weakid.py); with `se` and `weak_interval` every
    component it flags gets the q = 1 interval at `confidence`, the level of
    the normal intervals too (leaveout.py:1144-1158)."""
    import torch

    from .connected import leave_one_out
    from .estimator import StreamingHDFE, _column, parse_formula

    if leave_out not in ("match", "observation"):
        raise ValueError("leave_out must be 'match' or 'observation'")
    if centering not in ("reference", "weighted"):
        raise ValueError("centering must be 'reference' or 'weighted'")
    if leave_out == "match" and stayers is not None:
        raise ValueError("stayers applies when leaving out an observation; leaving out a "
                         "match uses LeaveOutTwoWay's within-match rule")
    if leave_out == "match" and centering != "reference":
        raise NotImplementedError("leaving out a match uses the reference centring on the "
                                  "device")
    if stayers not in (None, "own"):
        raise NotImplementedError("only stayers='own' is implemented on the device")
    if se_variance not in ("match", "person_year"):
        raise ValueError("se_variance must be 'match' or 'person_year'")
    if (leave_out == "match" and se and se_variance == "person_year"
            and options.get("weights") is not None):
        raise ValueError("se_variance='person_year' is LeaveOutTwoWay's, which has no user "
                         "weights; use se_variance='match' with a weighted fit")
    if "|" not in fml:
        raise ValueError(
            f"{fml!r} has no fixed effects; leave-out estimation needs at least two "
            "dimensions, as in 'y ~ x | worker_id + firm_id'")
    y, xs, fe = parse_formula(fml)
    if len(fe) < 2:
        raise ValueError(f"{fml!r} has one fixed effect; leave-out estimation needs at least "
                         "two, one of which is streamed")
    alpha = fe[0]
    psi = psi or fe[1]
    if psi == alpha or psi not in fe:
        raise ValueError(f"psi must be one of the fixed effects other than {alpha!r} "
                         f"(which plays the part of alpha); got {psi!r} from {fe}")
    options.setdefault("stream", alpha)
    options.setdefault("fixef_rm", "none")
    options.setdefault("tol", tol)
    options.setdefault("maxiter", maxiter)
    pruning = None
    t = {} if timings is None else timings
    t0 = time.perf_counter()
    if prune:
        keep, pinfo = leave_one_out(_column(data, alpha), _column(data, psi))
        idx = torch.nonzero(keep, as_tuple=True)[0]
        cols = [y] + list(xs) + list(fe) + ([options["weights"]] if options.get("weights")
                                            else [])
        sub = {}
        for c in dict.fromkeys(cols):
            v = _column(data, c)
            if not isinstance(v, torch.Tensor):
                v = torch.from_numpy(np.ascontiguousarray(v)).cuda()
            sub[c] = v.cuda().index_select(0, idx)
        data = sub
        pruning = {"fe": [alpha, psi], **pinfo}
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    fit = StreamingHDFE(y, xs, fe, keep_intermediates=True, **options).fit(data)
    t2 = time.perf_counter()
    if leave_out == "match":
        comps = _leave_out_match(fit, psi, n_draws, block, seed, tol, maxiter, precision, t,
                                 se=se, se_draws=se_draws, se_trace=se_trace,
                                 smooth_bins=smooth_bins, se_variance=se_variance,
                                 diagnose=diagnose, diagnose_draws=diagnose_draws,
                                 weak_interval=weak_interval, confidence=confidence)
    else:
        comps = leave_out_components(fit, n_draws=n_draws, block=block, seed=seed, psi=psi,
                                     tol=tol, maxiter=maxiter, precision=precision, timings=t,
                                     se=se, se_draws=se_draws, se_trace=se_trace,
                                     smooth_bins=smooth_bins, diagnose=diagnose,
                                     diagnose_draws=diagnose_draws,
                                     weak_interval=weak_interval, confidence=confidence)
    t.update(prune_seconds=t1 - t0, fit_seconds=t2 - t1,
             total_seconds=time.perf_counter() - t0)
    comps.pruning = pruning
    comps.fit = fit
    return comps
