"""Leave-out (KSS) estimation on the device: JLA leverages by random
projection, solved as blocks of many right-hand sides.

Mirrors hdfe_stream's `jla_leverages` (leaveout.py:130-198) and
`_combine_moments` (:209-282) over the operator `project_rows`
(inverse.py:240-309): per block of draws, a row pass reduces the Rademacher
vectors onto the levels, one block solve with k right-hand sides gives the
reduced solution, and a second row pass projects back to the rows and adds
the five running moments.  The draws are keyed on the sorted row ordinal, as
in the reference, so a panel whose rows are in the reference's canonical
order (sorted by group, codes and variables) gets the same draws.

Not mirrored yet: covariates in the design (the Schur border, inverse.py:
182-231); the leverages here are those of the fixed-effect design.
"""

from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib

RHS_ZERO_TOL = 1e-11      # core/solve.py:44


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


def jla_leverages(panel, n_draws: int = 250, block: int | None = None, seed: int = 0,
                  tol: float = 1e-10, maxiter: int = 5000, precision: str = "mixed",
                  sqrt_weights: bool | None = None, timings: dict | None = None) -> JLALeverages:
    """Every row's leverage by random projection (leaveout.py:130-198).

    `panel` is a fitted estimator's Panel (its rows, S and solver stay on the
    device).  `block` draws share one block solve (default: all of them, up
    to 1024).  With weights the leverages are those of the square-root-weight
    hat matrix, as in the reference."""
    import torch
    block = min(int(n_draws), 1024) if block is None else int(block)
    if block < 1 or n_draws < 1:
        raise ValueError("n_draws and block must be at least 1")
    if panel.matrix is None:
        panel.system(with_rhs=False)
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
        t1 = time.perf_counter()
        U, it, c, r = panel.matrix.solve(rhs, zero, None, tol=tol, maxiter=maxiter,
                                         precision=precision)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        _lib.check(lib.hdfe_rows_project_moments(rp._h, gstart, codes, w, sqrtw, seed, drawn, size,
                                                 offs_d, U.data_ptr(), moments.data_ptr()),
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
                     "precision": precision, "covariates": []})
