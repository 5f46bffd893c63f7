"""Poisson regression with high-dimensional fixed effects on the device:
`fepois_stream` (api.py:144-240) and `fepois_batch`, many independent
models in one go.

The algorithm is the reference's IRLS (glm.py:743-838, core/irls.py):
Poisson's start eta = log((y + ybar) / 2), then per step one fused row pass
(mu, W = mu, working response z = eta + y / mu - 1, deviance), one weighted
fit of [z, X] on the fixed effects (diag, S and the reduced right-hand sides
from the new weights, a block CG with the previous solution as the start),
the slopes, the fixed effects and the group effects; a step that raises the
deviance is halved back toward the current coefficients (up to 30 times);
convergence when |dev - dev_old| / (0.1 + |dev|) < iwls_tol, or when a later
step no longer lowers the deviance.  The separation check drops the levels
whose outcome is zero on every row, one pass over the dimensions
(glm.py:319-370), before the singleton drop.

Different from the reference, on purpose, for the device:
- every model of a batch is one stacked panel whose levels are kept apart
  per model: one S (block diagonal), one block CG for all models' [z, X]
  columns, one launch per row pass;
- the slopes' Hessian X~'W X~ is a split-BF16 product on the matrix cores
  and the slope update the Newton step beta += H^-1 X~'W (z~ - X~ beta) with
  the FP64 score: the converged slopes solve the FP64 normal equations (the
  steps' path differs from the reference's exact A_xx beta = A_xz by the
  Hessian's rounding);
- each model stops at its own convergence (its coefficients are frozen while
  the others go on).

Not mirrored: logit/probit, offsets, prior weights, standard errors, the
bias correction, Newton's method for models with one fixed effect.
"""

from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib

MAX_HALVINGS = 30          # glm.py:96
DEV_ROUNDING = 1e-12       # glm.py:100


@dataclass
class GLMResult:
    """One model's Poisson fit."""
    coefnames: list
    beta: np.ndarray
    deviance: float
    loglik: float
    iterations: int
    converged: bool
    n_obs: int
    fe_names: list
    fe_levels: dict
    fe_values: dict
    deviance_path: list = field(default_factory=list)
    diagnostics: dict = field(default_factory=dict)
    eta_source: object = None       # linear predictor per input row (NaN if dropped)

    def coef(self):
        return dict(zip(self.coefnames, self.beta.tolist()))

    def fixef(self, dim):
        return self.fe_levels[dim], self.fe_values[dim]


def _design(fml, data):
    """(y name, covariate blocks [(names, 2-D array)], fe names) of one model:
    a covariate term may be a 2-D column (n, k), named term[0..k-1]."""
    from .estimator import _column, parse_formula
    yname, xs, fe = parse_formula(fml)
    blocks = []
    for x in xs:
        col = _column(data, x)
        arr = col
        if getattr(arr, "ndim", 1) == 2:
            k = arr.shape[1]
            blocks.append(([f"{x}[{j}]" for j in range(k)], arr))
        else:
            blocks.append(([x], arr))
    return yname, blocks, fe


def _null_groups(panel, lmodel, n_models, ncomp):
    """Each level's group of S's null space: S = D_o'W M_0 D_o vanishes on
    every vector that is constant over the first solved dimension's levels
    of one connected component (each worker's rows lie in one component),
    and on every vector constant over one model's levels of a further
    dimension (every row of the model moves alike).  These indicators are
    disjoint, so the null space is spanned by orthogonal group indicators:
    component c of the first dimension is group c, dimension t >= 1 of model
    m is group ncomp + m (D - 1) + t - 1.  Device (L,) int64."""
    import torch
    D = len(panel.n_levels)
    gid = torch.empty(panel.info["L"], dtype=torch.int64, device="cuda")
    o0, n0 = int(panel.offs[0]), int(panel.n_levels[0])
    gid[o0:o0 + n0] = panel.tensor("comp", torch.int32)[:n0].long()
    for t in range(1, D):
        o, nl = int(panel.offs[t]), int(panel.n_levels[t])
        gid[o:o + nl] = ncomp + lmodel[o:o + nl] * (D - 1) + (t - 1)
    return gid


def _group_sums(ids, X, ng):
    """(ng, k) sums of the rows of X (L, k) by group id (L,), each group's
    rows in their order: the rows sorted by group (a stable sort), then one
    reduction per contiguous group (an index_add_ over few groups serialises
    on atomics: 144 ms at 234,046 x 301 into 32 groups)."""
    import torch
    if ng > 1024:
        out = torch.zeros((ng, X.shape[1]), dtype=X.dtype, device=X.device)
        out.index_add_(0, ids, X)
        return out
    order = torch.argsort(ids, stable=True)
    counts = torch.bincount(ids, minlength=ng).tolist()
    Xs = X.index_select(0, order)
    out = torch.zeros((ng, X.shape[1]), dtype=X.dtype, device=X.device)
    at = 0
    for gi, c in enumerate(counts):
        if c:
            out[gi] = Xs[at:at + c].sum(0)
        at += c
    return out


def _to_range(rhs, null_id):
    """rhs less its part in S's null space (the group means of `null_id`,
    whose indicators span it): the rounding of D_o'W M_0 v leaves the sums
    over a null group nonzero by a few ulps of their terms, and no x can
    reduce a model's own relative residual below that part."""
    import torch
    ng = int(null_id.max()) + 1
    sums = _group_sums(null_id, rhs, ng)
    cnt = torch.bincount(null_id, minlength=ng).to(torch.float64).clamp_(min=1)
    return rhs - (sums / cnt[:, None])[null_id]


def _per_model_solve(matrix, rhs, x0, lmodel, n_models, tol, maxiter, precision, info):
    """Solve the block-diagonal stacked system so that every model meets `tol`
    on its own relative residual ||r_m|| / ||b_m||, column by column.

    The models' systems are independent, so each model's block of every
    column is scaled to unit norm (its solution scales alike and is scaled
    back).  A column then has ||b'|| = sqrt(k) for its k models with a
    nonzero block, and stopping the stacked CG at ||r'|| <= (tol / sqrt(k))
    ||b'|| = tol bounds every model's ||r'_m|| = ||r_m|| / ||b_m|| by tol: no
    model is judged against the stacked norm of the others.  A model's zero
    block stays zero.  Returns (X, iterations, converged, max_rel) with
    max_rel the stacked scaled residual."""
    import torch
    sq = _group_sums(lmodel, rhs * rhs, n_models)
    norm = sq.sqrt()
    inv = torch.where(norm > 0, 1.0 / torch.where(norm > 0, norm, torch.ones_like(norm)),
                      torch.zeros_like(norm))
    live = max(int((norm > 0).sum(0).max()), 1)
    scale = inv[lmodel]
    x0s = None if x0 is None else x0 * scale
    b = rhs * scale
    X, its, ok, rel = matrix.solve(b, None, x0s, tol=tol / np.sqrt(live),
                                   maxiter=maxiter, precision=precision, info=info)
    if not ok and precision == "mixed":
        # the refinement stopped at FP32's reach above the tightened
        # tolerance: FP64 CG finishes from its solution
        X, its2, ok, rel = matrix.solve(b, None, X, tol=tol / np.sqrt(live),
                                        maxiter=max(maxiter - its, 1), precision="fp64")
        its += its2
        info["fp64_finish_iterations"] = its2
    return X * norm[lmodel], its, ok, rel


def _model_residuals(matrix, rhs, X, lmodel, n_models):
    """The largest ||b_m - S_m x_m|| / ||b_m|| over columns, per model (host)."""
    import torch
    R = rhs - matrix.matmat(X)
    num = _group_sums(lmodel, R * R, n_models)
    den = _group_sums(lmodel, rhs * rhs, n_models)
    rel = torch.where(den > 0, (num / torch.where(den > 0, den, torch.ones_like(den))).sqrt(),
                      num.sqrt())
    return rel.max(1).values.cpu().numpy()


def fepois_batch(fml: str, datasets: list, iwls_tol: float = 1e-8, iwls_maxiter: int = 25,
                 separation_check: bool = True, tol: float = 1e-10, maxiter: int = 5000,
                 precision: str = "auto", rows_per_syrk: int = 32768,
                 timings: dict | None = None, check_residuals: bool = False,
                 weights: str | None = None, solver: str = "auto") -> list:
    """Fit `fml` by Poisson IRLS on each dataset, all on the device at once.
    Returns one GLMResult per dataset.  Every model's weighted fit of [z, X]
    on its fixed effects meets `tol` on its own relative residual (the
    reference solves each model alone, core/solve.py:66-140); with
    `check_residuals` each model's true residuals b - S x are measured after
    every solve and reported (diagnostics["model_rel_residual"]).  `solver`:
    "explicit" builds S at every reweighting; "stream_cg" solves through the
    matrix-free S over the cells (the reference's stream_cg); "auto" takes
    the matrix-free S for more than eight columns of [z, X].  `precision`:
    "fp64", "mixed" (FP32 CG with FP64 refinement) or "auto" (FP64 through
    the matrix-free S, mixed through the explicit one)."""
    return _glm_batch(fml, datasets, "poisson", iwls_tol, iwls_maxiter, separation_check, tol,
                      maxiter, precision, rows_per_syrk, timings, check_residuals, weights,
                      solver)


def feols_batch(fml: str, datasets: list, weights: str | None = None, tol: float = 1e-10,
                maxiter: int = 5000, precision: str = "auto", rows_per_syrk: int = 32768,
                max_refinements: int = 6, timings: dict | None = None,
                check_residuals: bool = False, solver: str = "auto") -> list:
    """OLS with fixed effects on many datasets at once (feols_stream,
    api.py:22-107, of each; the wide-design path of the reference's
    assembly="rows": core/inference.py:109-137, `_nb_residualize_rows` and
    V'V).  The same machinery as `fepois_batch` with the identity link: one
    weighted fit of [y, X] on the fixed effects (every model to `tol` on its
    own residual), then the slopes by Newton steps with the split-BF16
    Hessian X~'W X~ on the matrix cores and the FP64 score X~'W (y~ - X~
    beta), repeated until they stop moving (|d beta| <= 1e-13 |beta|), so
    that they solve the FP64 normal equations.  `deviance` is the weighted
    residual sum of squares; `weights` names a column of analytic weights;
    `solver` as fepois_batch's."""
    return _glm_batch(fml, datasets, "gaussian", 1e-13, max_refinements, False, tol, maxiter,
                      precision, rows_per_syrk, timings, check_residuals, weights, solver)


def _glm_batch(fml, datasets, family, iwls_tol, iwls_maxiter, separation_check, tol, maxiter,
               precision, rows_per_syrk, timings, check_residuals, weights, solver="auto"):
    import torch

    from .estimator import Panel, _column
    from .leaveout import RowPasses
    if not 0 < iwls_tol < 1:
        raise ValueError("iwls_tol must be in (0, 1)")
    if iwls_maxiter < 1:
        raise ValueError("iwls_maxiter must be at least 1")
    if solver not in ("auto", "explicit", "stream_cg"):
        raise ValueError("solver must be 'auto', 'explicit' or 'stream_cg'")
    if precision not in ("auto", "fp64", "mixed"):
        raise ValueError("precision must be 'auto', 'fp64' or 'mixed'")
    lib = _lib.load()
    t_start = time.perf_counter()
    dev = "cuda"
    M = len(datasets)
    if M < 1:
        raise ValueError("no data")
    per = []
    for d in datasets:
        yname, blocks, fe = _design(fml, d)
        if len(fe) < 2:
            raise ValueError("device batch fits need at least two fixed effects")
        y = torch.as_tensor(np.asarray(_column(d, yname)) if not torch.is_tensor(
            _column(d, yname)) else _column(d, yname)).to(dev, torch.float64)
        if family == "poisson":
            if float(y.min()) < 0:
                raise ValueError("Poisson regression needs a nonnegative dependent variable; "
                                 f"its minimum is {float(y.min())}")
            if float(y.max()) <= 0:
                raise ValueError("the dependent variable is zero on every row")
        wt = None
        if weights is not None:
            col = _column(d, weights)
            wt = torch.as_tensor(np.asarray(col) if not torch.is_tensor(col) else col).to(
                dev, torch.float64)
            if float(wt.min()) < 0:
                raise ValueError("weights must be nonnegative")
        xs = [torch.as_tensor(b).to(dev, torch.float32).reshape(len(y), -1)
              for _, b in blocks]
        X = torch.cat(xs, dim=1).contiguous() if xs else torch.zeros((len(y), 0), device=dev)
        ids = [torch.as_tensor(np.asarray(_column(d, f)) if not torch.is_tensor(_column(d, f))
                               else _column(d, f)).to(dev, torch.int64) for f in fe]
        names = [n for nm, _ in blocks for n in nm]
        per.append({"y": y, "X": X, "ids": ids, "names": names, "fe": fe, "n_in": len(y),
                    "w": wt})
    kx = per[0]["X"].shape[1]
    if any(p["X"].shape[1] != kx for p in per):
        raise ValueError("every model of a batch needs the same covariates")
    if kx < 1 or kx > 511:
        raise ValueError("1 to 511 covariates")
    if solver == "auto":
        # the [z, X] block of more than eight columns through the matrix-free
        # S (group means, no S built at each reweighting: C5's S has 133M
        # nonzeros from 18.7M cells); a narrow one through the explicit S
        solver = "stream_cg" if kx + 1 > 8 else "explicit"
    if precision == "auto":
        # through the wide matrix-free S an FP32 product costs 0.6 of an FP64
        # one, and the refinements' restarts and FP64 residuals cost more than
        # that saves (C5: 2.32 s FP64 against 2.76 s mixed); the explicit S's
        # narrow block keeps the mixed solve
        precision = "fp64" if solver == "stream_cg" else "mixed"
    fe_names = per[0]["fe"]
    nfe = len(fe_names)
    # separation: levels whose outcome is zero on every row, one pass over
    # the dimensions on the rows still kept (glm.py:346-367)
    for p in per:
        keep = torch.ones(len(p["y"]), dtype=torch.bool, device=dev)
        dropped = {}
        if separation_check and bool((p["y"] == 0).any()):
            for d, f in enumerate(fe_names):
                u, inv = torch.unique(p["ids"][d], return_inverse=True)
                # a level is separated when every kept row has y = 0 (y >= 0):
                # marks of the levels with a kept row and with a positive one
                # (a max-scatter or a weighted count over a few year levels
                # serialised on atomics)
                seen = torch.zeros(len(u), dtype=torch.bool, device=dev)
                seen[inv[keep]] = True
                pos = torch.zeros(len(u), dtype=torch.bool, device=dev)
                pos[inv[keep & (p["y"] > 0)]] = True
                bad = seen & ~pos
                if bool(bad.any()):
                    dropped[f] = int(bad.sum())
                    keep &= ~bad[inv]
        p["keep"] = keep
        p["separation"] = {"checked": bool(separation_check), "levels": dropped,
                           "observations": int((~keep).sum())}
    # the stacked panel: every dimension's codes offset per model
    stacked, src_model, src_row = [[] for _ in range(nfe)], [], []
    offs_model = np.zeros((M, nfe), np.int64)
    level_ids = [[None] * nfe for _ in range(M)]
    tot = np.zeros(nfe, np.int64)
    for m, p in enumerate(per):
        rows = torch.nonzero(p["keep"], as_tuple=True)[0]
        for d in range(nfe):
            u, inv = torch.unique(p["ids"][d][rows], return_inverse=True)
            level_ids[m][d] = u
            offs_model[m, d] = tot[d]
            stacked[d].append(inv + int(tot[d]))
            tot[d] += len(u)
        src_model.append(torch.full((len(rows),), m, dtype=torch.int32, device=dev))
        src_row.append(rows)
    fe_cols = [torch.cat(s) for s in stacked]
    smodel = torch.cat(src_model)
    y_all = torch.cat([p["y"][r] for p, r in zip(per, src_row)])
    f_all = None if weights is None else torch.cat([p["w"][r] for p, r in zip(per, src_row)])
    X_all = torch.cat([p["X"][r] for p, r in zip(per, src_row)])
    n_stack = len(y_all)
    ones = torch.ones(n_stack, dtype=torch.float64, device=dev)
    panel = Panel(fe_cols, [y_all], weights=ones, fixef_rm="singleton", stream_dim=0)
    info = panel.info
    n, G, L, D = info["n"], info["G"], info["L"], info["D"]
    perm = panel.tensor("perm", torch.int32).long()[:n]
    rmodel = smodel[perm].contiguous()
    y_s = y_all[perm].contiguous()
    f_s = None if f_all is None else f_all[perm].contiguous()
    X_s = X_all[perm].contiguous()
    del X_all
    rm_host = rmodel.cpu().numpy()
    rs = np.searchsorted(rm_host, np.arange(M + 1), side="left").astype(np.int64)
    if np.any(np.diff(rm_host) < 0):
        raise RuntimeError("internal: the stacked panel's rows are not grouped by model")
    gstart_t = panel.tensor("gstart", torch.int32)
    gs_host = gstart_t.cpu().numpy().astype(np.int64)
    gmodel = rmodel[gstart_t[:-1].long()].contiguous()
    gb_model = np.searchsorted(gs_host[:-1], rs, side="left").astype(np.int64)  # first group per model
    codes_t = panel.tensor("codes", torch.int32).view(D, n)
    lmodel = torch.zeros(L, dtype=torch.int32, device=dev)
    for t in range(D):
        lmodel[int(panel.offs[t]) + codes_t[t].long()] = rmodel
    # pass 1, run again at every reweighting, reallocates the group starts:
    # the device pointers are read afresh after each (ptr())
    ptr = {}
    lmodel_l = lmodel.long()
    model_rel = []

    def refresh_ptrs():
        for k in ("g", "gstart", "codes", "offs"):
            ptr[k] = panel.ptr(k)[0]
    refresh_ptrs()
    n_m = np.diff(rs)

    def wmean(v, m):
        if not n_m[m]:
            return 1.0
        if f_s is None:
            return float(v[rs[m]:rs[m + 1]].mean())
        fw = f_s[rs[m]:rs[m + 1]]
        return float((fw * v[rs[m]:rs[m + 1]]).sum() / fw.sum().clamp(min=1e-300))
    ybar = torch.tensor([wmean(y_s, m) for m in range(M)], dtype=torch.float64, device=dev)
    fam = {"poisson": 0, "gaussian": 1}[family]
    # launch lists
    # row chunks of 2,048 (one workgroup each: enough workgroups to keep
    # every CU's waves busy; 65,536-row chunks left ~300 workgroups)
    ch_model, ch_b, ch_e = [], [], []
    for m in range(M):
        for b in range(int(rs[m]), int(rs[m + 1]), 2048):
            ch_model.append(m)
            ch_b.append(b)
            ch_e.append(min(int(rs[m + 1]), b + 2048))
    ch_model = np.array(ch_model, np.int32)
    ch_b = np.array(ch_b, np.int64)
    ch_e = np.array(ch_e, np.int64)
    wg_model, wg_gb, wg_ge = [], [], []
    for m in range(M):
        for gb in range(int(gb_model[m]), int(gb_model[m + 1]), 512):
            wg_model.append(m)
            wg_gb.append(gb)
            wg_ge.append(min(int(gb_model[m + 1]), gb + 512))
    wg_model = np.array(wg_model, np.int32)
    wg_gb = np.array(wg_gb, np.int64)
    wg_ge = np.array(wg_ge, np.int64)
    Cp = (kx + 31) // 32 * 32
    pstart = np.zeros(M + 1, np.int64)
    pstart[1:] = np.cumsum((n_m + 15) // 16 * 16)
    npad = int(pstart[-1])
    hi = torch.zeros((npad, Cp), dtype=torch.bfloat16, device=dev)
    lo = torch.zeros((npad, Cp), dtype=torch.bfloat16, device=dev)
    Cc = kx + 1
    vbar = torch.empty((G, Cc), dtype=torch.float64, device=dev)
    Abuf = torch.empty((M, Cp, Cp), dtype=torch.float64, device=dev)
    rp = RowPasses(panel)
    Vz = torch.empty((n, 1), dtype=torch.float64, device=dev)

    def row_pass(beta, fe, a, init, family_code=None):
        w = torch.empty(n, dtype=torch.float64, device=dev)
        z = torch.empty(n, dtype=torch.float64, device=dev)
        part = np.zeros((len(ch_model), 5))
        torch.cuda.synchronize()
        _lib.check(lib.hdfe_glm_rows(
            len(ch_model), ch_model.ctypes.data, ch_b.ctypes.data, ch_e.ctypes.data, int(init),
            fam if family_code is None else family_code,
            ybar.data_ptr(), y_s.data_ptr(), X_s.data_ptr(), kx, ptr["g"], ptr["codes"], n, D,
            ptr["offs"], None if init else fe.data_ptr(), None if init else a.data_ptr(),
            None if init else beta.data_ptr(), None if f_s is None else f_s.data_ptr(),
            w.data_ptr(), z.data_ptr(), part.ctypes.data), "hdfe_glm_rows")
        sums = np.zeros((M, 5))
        np.add.at(sums, ch_model, part)
        return w, z, sums

    tm = {"row_pass": 0.0, "system": 0.0, "rhs": 0.0, "solve": 0.0, "partial": 0.0,
          "syrk": 0.0, "update": 0.0}

    def clock(key, t0):
        torch.cuda.synchronize()
        tm[key] += time.perf_counter() - t0

    t0 = time.perf_counter()
    w, z, sums = row_pass(None, None, None, True)
    clock("row_pass", t0)
    slg = sums[:, 4].copy()
    dev_cur = sums[:, 0].copy()
    beta = torch.zeros((M, kx), dtype=torch.float64, device=dev)
    fe = torch.zeros(L, dtype=torch.float64, device=dev)
    a = torch.zeros(G, dtype=torch.float64, device=dev)
    init = np.ones(M, bool)
    active = np.ones(M, bool)
    converged = np.zeros(M, bool)
    iters = np.zeros(M, np.int64)
    halvings = np.zeros(M, np.int64)
    paths = [[float(dev_cur[m])] for m in range(M)]
    Gamma = None
    ncomp = None
    solver_its = []
    for it in range(1, iwls_maxiter + 1):
        # the weighted fit of [z, X] on the fixed effects
        t0 = time.perf_counter()
        Vz[:, 0] = z
        torch.cuda.synchronize()
        _lib.check(lib.hdfe_panel_reweight(panel._h, Vz.data_ptr(), w.data_ptr()),
                   "hdfe_panel_reweight")
        refresh_ptrs()
        if solver == "stream_cg":
            panel.system(with_rhs=False, budget=0)
            panel.stream_matrix(reuse=True)
        else:
            panel.system(with_rhs=False)
        if ncomp is None:
            ncomp = panel.components()
            null_id = _null_groups(panel, lmodel_l, M, ncomp)
        clock("system", t0)
        t0 = time.perf_counter()
        rhs = torch.empty((L, Cc), dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        _lib.check(lib.hdfe_glm_rhs(rp._h, G, ptr["gstart"], ptr["g"], z.data_ptr(), X_s.data_ptr(),
                                    kx, w.data_ptr(), vbar.data_ptr(), rhs.data_ptr()),
                   "hdfe_glm_rhs")
        clock("rhs", t0)
        t0 = time.perf_counter()
        sinfo = {}
        rhs = _to_range(rhs, null_id)
        Gamma, its, ok, rel = _per_model_solve(panel.matrix, rhs, Gamma, lmodel_l, M, tol,
                                               maxiter, precision, sinfo)
        solver_its.append((its, sinfo.get("refinement_steps", 0), bool(ok), float(rel)))
        if check_residuals:
            model_rel.append(_model_residuals(panel.matrix, rhs, Gamma, lmodel_l, M))
        clock("solve", t0)
        # the slopes: Newton step with the split-BF16 Hessian and FP64 score
        t0 = time.perf_counter()
        gpart = np.zeros((len(wg_model), kx))
        torch.cuda.synchronize()
        _lib.check(lib.hdfe_glm_partial(
            len(wg_model), wg_model.ctypes.data, wg_gb.ctypes.data, wg_ge.ctypes.data,
            ptr["gstart"], rs.ctypes.data, pstart.ctypes.data, M, Cp, z.data_ptr(),
            X_s.data_ptr(), kx, w.data_ptr(), ptr["codes"], n, D, ptr["offs"], Gamma.data_ptr(),
            beta.data_ptr(), hi.data_ptr(), lo.data_ptr(), gpart.ctypes.data),
            "hdfe_glm_partial")
        clock("partial", t0)
        t0 = time.perf_counter()
        _lib.check(lib.hdfe_glm_syrk(M, pstart.ctypes.data, Cp, hi.data_ptr(), lo.data_ptr(),
                                     int(rows_per_syrk), Abuf.data_ptr()), "hdfe_glm_syrk")
        clock("syrk", t0)
        t0 = time.perf_counter()
        score = np.zeros((M, kx))
        np.add.at(score, wg_model, gpart)
        H = Abuf[:, :kx, :kx].cpu().numpy()
        bcur = beta.cpu().numpy()
        bnew = bcur.copy()
        for m in range(M):
            if active[m]:
                bnew[m] = bcur[m] + np.linalg.solve(H[m], score[m])
        beta_new = torch.from_numpy(bnew).to(dev)
        fe_new = torch.empty(L, dtype=torch.float64, device=dev)
        a_new = torch.empty(G, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        _lib.check(lib.hdfe_glm_coefficients(
            L, G, kx, Gamma.data_ptr(), lmodel.data_ptr(), gmodel.data_ptr(), ptr["gstart"],
            z.data_ptr(), X_s.data_ptr(), w.data_ptr(), ptr["codes"], n, D, ptr["offs"],
            beta_new.data_ptr(), fe_new.data_ptr(), a_new.data_ptr()), "hdfe_glm_coefficients")
        clock("update", t0)

        def toward(tv):
            """Per model: current + t (new - current); t = 0 freezes a model."""
            tt = torch.from_numpy(np.ascontiguousarray(tv, np.float64)).to(dev)
            out_b = torch.empty_like(beta)
            out_f = torch.empty_like(fe)
            out_a = torch.empty_like(a)
            torch.cuda.synchronize()
            _lib.check(lib.hdfe_glm_toward(M, None, kx, tt.data_ptr(), beta.data_ptr(),
                                           beta_new.data_ptr(), out_b.data_ptr()),
                       "hdfe_glm_toward")
            _lib.check(lib.hdfe_glm_toward(L, lmodel.data_ptr(), 1, tt.data_ptr(),
                                           fe.data_ptr(), fe_new.data_ptr(), out_f.data_ptr()),
                       "hdfe_glm_toward")
            _lib.check(lib.hdfe_glm_toward(G, gmodel.data_ptr(), 1, tt.data_ptr(),
                                           a.data_ptr(), a_new.data_ptr(), out_a.data_ptr()),
                       "hdfe_glm_toward")
            return out_b, out_f, out_a

        t = active.astype(np.float64)
        t0 = time.perf_counter()
        cb, cf, ca = toward(t)
        wc, zc, sums = row_pass(cb, cf, ca, False)
        clock("row_pass", t0)
        dev_new = sums[:, 0].copy()
        tol_acc = max(iwls_tol, DEV_ROUNDING)

        def accept(m):
            return np.isfinite(dev_new[m]) and (
                dev_new[m] < dev_cur[m]
                or abs(dev_new[m] - dev_cur[m]) / (0.1 + abs(dev_cur[m])) < tol_acc)

        # least squares: the Newton step contracts whatever the deviance's
        # rounding says, so it is always taken
        bad = [] if fam == 1 else \
            [m for m in range(M) if active[m] and not init[m] and not accept(m)]
        for m in range(M):
            if active[m] and init[m] and not np.isfinite(dev_new[m]):
                raise FloatingPointError(f"model {m}: the first IRLS step gives a "
                                         "non-finite deviance")
        nh = 0
        while bad:
            if nh >= MAX_HALVINGS:
                raise RuntimeError(f"models {bad}: IRLS step-halving failed")
            nh += 1
            for m in bad:
                t[m] /= 2
                halvings[m] += 1
            t0 = time.perf_counter()
            cb, cf, ca = toward(t)
            wc, zc, sums = row_pass(cb, cf, ca, False)
            clock("row_pass", t0)
            for m in bad:
                dev_new[m] = sums[m, 0]
            bad = [m for m in bad if not (np.isfinite(dev_new[m]) and dev_new[m] < dev_cur[m])]
        for m in range(M):
            if not active[m]:
                continue
            if fam == 1:
                # the slopes' relative change: the Newton steps refine them
                # to the FP64 normal equations
                step = float((cb[m] - beta[m]).abs().max())
                crit = step / max(float(cb[m].abs().max()), 1e-300)
                flat = False
            else:
                crit = abs(dev_new[m] - dev_cur[m]) / (0.1 + abs(dev_cur[m]))
                flat = (not init[m]) and dev_new[m] >= dev_cur[m]
            iters[m] = it
            paths[m].append(float(dev_new[m]))
            dev_cur[m] = dev_new[m]
            init[m] = False
            if crit < iwls_tol or flat:
                converged[m] = True
                active[m] = False
        beta, fe, a, w, z = cb, cf, ca, wc, zc
        if not active.any():
            break
    # the reported effects: per model, each non-streamed dimension after the
    # first at obs-weighted mean zero, the first per component (the
    # normalisation of core/solve.py:383-399), the group effects absorbing
    # the shifts; the linear predictor is unchanged
    eta_s = _eta(y_s, fe, a, beta, row_pass, fam)
    fe_h = fe.cpu().numpy()
    a_h = a.cpu().numpy()
    lmod_h = lmodel.cpu().numpy()
    obs = panel.array("obs", np.float64)
    gmod_h = gmodel.cpu().numpy()
    for t_ in range(1, D):
        o, nl = int(panel.offs[t_]), panel.n_levels[t_]
        seg = slice(o, o + nl)
        for m in range(M):
            sel = np.flatnonzero(lmod_h[seg] == m)
            if not len(sel):
                continue
            wts = obs[seg][sel]
            c = float(wts @ fe_h[seg][sel] / max(wts.sum(), 1e-300))
            fe_h[o + sel] -= c
            a_h[gmod_h == m] += c
    o0, nl0 = int(panel.offs[0]), panel.n_levels[0]
    comp = panel.array("comp", np.int32)
    wts = obs[o0:o0 + nl0]
    ncmp = int(comp.max()) + 1 if len(comp) else 0
    cw = np.bincount(comp, weights=wts, minlength=ncmp)
    cm = np.bincount(comp, weights=wts * fe_h[o0:o0 + nl0], minlength=ncmp) / np.where(cw > 0, cw, 1)
    fe_h[o0:o0 + nl0] -= cm[comp]
    # each component's shift goes to its groups: a group's component is its
    # first-dimension level's
    g0codes = codes_t[0][gstart_t[:-1].long()].cpu().numpy()
    a_h += cm[comp[g0codes]]
    eta_src_all = torch.full((n_stack,), float("nan"), dtype=torch.float64, device=dev)
    eta_src_all[perm] = eta_s
    perm_h = perm.cpu().numpy()
    out = []
    stack_off = np.concatenate(([0], np.cumsum([len(r) for r in src_row])))
    glevels = panel.array("glevels", np.int64)
    lv_arrays = [panel.array(f"levels{t_}", np.int64) for t_ in range(D)]
    seconds = time.perf_counter() - t_start
    for m, p in enumerate(per):
        e_src = torch.full((p["n_in"],), float("nan"), dtype=torch.float64, device=dev)
        e_src[src_row[m]] = eta_src_all[int(stack_off[m]):int(stack_off[m + 1])]
        levels, values = {}, {}
        gsel = np.flatnonzero(gmod_h == m)
        ids0 = level_ids[m][0].cpu().numpy()
        levels[fe_names[0]] = ids0[glevels[gsel] - offs_model[m, 0]]
        values[fe_names[0]] = a_h[gsel]
        for t_ in range(D):
            d = panel.dims[t_]
            o, nl = int(panel.offs[t_]), panel.n_levels[t_]
            sel = np.flatnonzero(lmod_h[o:o + nl] == m)
            idsd = level_ids[m][d].cpu().numpy()
            levels[fe_names[d]] = idsd[lv_arrays[t_][sel] - offs_model[m, d]]
            values[fe_names[d]] = fe_h[o + sel]
        nm = int(n_m[m])
        out.append(GLMResult(
            coefnames=p["names"], beta=beta[m].cpu().numpy(), deviance=float(dev_cur[m]),
            loglik=float(sums[m, 1] - slg[m]) if fam == 0 else float("nan"),
            iterations=int(iters[m]),
            converged=bool(converged[m]), n_obs=nm, fe_names=[fe_names[0]] +
            [fe_names[d] for d in panel.dims], fe_levels=levels, fe_values=values,
            deviance_path=paths[m], eta_source=e_src,
            diagnostics={"separation": p["separation"], "halvings": int(halvings[m]),
                         "solver": solver, "precision": precision,
                         "solver_iterations": solver_its, "batch": M,
                         "model_rel_residual": [float(r[m]) for r in model_rel],
                         "seconds_batch": round(seconds, 4)}))
    if timings is not None:
        timings.update({k: round(v, 4) for k, v in tm.items()})
        timings["total"] = round(seconds, 4)
        timings["steps"] = int(iters.max())
        timings["solver_iterations"] = solver_its
        timings["rows"] = int(n)
        timings["levels"] = int(L)
        timings["groups"] = int(G)
    rp.close()
    panel.close()
    return out


def _eta(y_s, fe, a, beta, row_pass, fam=0):
    """The linear predictor per sorted row at (beta, fe, a) (the row pass's
    family 2 writes it to w)."""
    w, _, _ = row_pass(beta, fe, a, False, 2)
    return w


def fepois_stream(fml: str, data, iwls_tol: float = 1e-8, iwls_maxiter: int = 25,
                  separation_check: bool = True, device: str = "rocm", **options) -> GLMResult:
    """Poisson regression with fixed effects on the device (api.py:144-240,
    with device='rocm')."""
    if device != "rocm":
        raise ValueError("device must be 'rocm' (this package is the MI355X backend)")
    return fepois_batch(fml, [data], iwls_tol=iwls_tol, iwls_maxiter=iwls_maxiter,
                        separation_check=separation_check, **options)[0]
