"""CPU oracle of the hdfe-stream solve path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module, and only as the checker (or the timed CPU baseline).  The
product (the `hdfe_stream_amd` package) never imports it.

It restates, in numpy and in the C of `hdfe_oracle.c` (loaded with ctypes),
the reference's FP64 algorithm:

  pass 0   singleton restriction      core/passes.py:859-904
           streamed-dimension choice  core/estimator.py:371-385
           level codes in sorted order, rows sorted by (group, codes)
                                      core/passes.py:311-394
  pass 1   cells, global-mean centring, identifying groups, obs/cnt, the
           within-cell scatter        core/passes.py:397-589
  step 2   diag, rhs, explicit S, block Jacobi-PCG, zero-column test,
           normalisation              core/solve.py:44-49, 66-194, 294-399
  step 3   V' M_D V from cells, beta  core/inference.py:109-137, 833-876
  pass 2   worker effects, residuals  core/kernels_base.py:543-619
  graph    components and ranking     core/kernels_base.py:155-174,
                                      core/passes.py:933-979
           leave-one-out pruning      core/connected.py:503-562
           standalone components: (k0, k1)-core peel, union-find labels in
           root order, ranking        core/connected.py:378-493,
                                      core/kernels_graph.py:187-262
  KSS      Rademacher draws, centre/project rows, JLA moments and their
           combination               core/kernels_groups.py:52-77,
                                      kernels_inverse.py:34-108,
                                      inverse.py:162-309, leaveout.py:130-282
           apply_inverse, null space, Hutchinson bias, sigma2, plug-in,
           leave_out_kss leaving out an observation
                                      inverse.py:311-460, kernels_inverse.py:
                                      111-231, leaveout.py:321-448, 634-1016,
                                      1069-1255, leaveout_se.py:202-214
  GLM      Poisson IRLS: separation, start, weighted fits, step-halving,
           convergence                glm.py:319-370, 743-838, core/irls.py:
                                      100-144, core/families.py:69-100
  autograd a dense FP64 torch (CPU) restatement of the weighted fit and its
           plug-in components, differentiated by torch's autograd: the check
           of the device adjoint (the reference has no autograd)

Each piece is pinned in tests/test_oracle.py against what the mathematics
fixes (dense closed forms, scipy routines, brute force) and the golden
vectors under tests/golden/.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np
import scipy.sparse as sp

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
RHS_ZERO_TOL = 1e-11     # core/solve.py:44

_lib = None


def build(force: bool = False) -> Path:
    """Compile hdfe_oracle.c into liboracle.so with gcc and OpenMP."""
    src = HERE / "hdfe_oracle.c"
    if force or not LIB.exists() or LIB.stat().st_mtime < src.stat().st_mtime:
        subprocess.check_call(["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", "-o", str(LIB),
                               str(src), "-lm"])
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        _lib.orc_s_count.restype = C.c_int64
        _lib.orc_bfs_components.restype = C.c_int64
        _lib.orc_mix64.restype = C.c_uint64
        _lib.orc_mix64.argtypes = [C.c_uint64]
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def set_threads(n: int) -> None:
    lib().orc_set_threads(C.c_int(int(n)))


def num_threads() -> int:
    return int(lib().orc_num_threads())


# ------------------------------------------------------------------ pass 0

def singleton_restriction(fe_cols: list[np.ndarray]) -> np.ndarray:
    """Boolean mask of the rows kept after dropping singletons: dimensions
    checked in turn, each one's singletons dropped before the next is
    counted, until every dimension has been checked since the last drop
    (core/passes.py:859-904)."""
    n = len(fe_cols[0])
    keep = np.ones(n, bool)
    dims = len(fe_cols)
    found, quiet, i = False, 0, 0
    while quiet < dims - int(found):
        col = fe_cols[i % dims]
        vals = col[keep]
        u, inv, cnt = np.unique(vals, return_inverse=True, return_counts=True)
        bad = cnt[inv] == 1
        if bad.any():
            idx = np.flatnonzero(keep)
            keep[idx[bad]] = False
            found = True
            quiet = 0
        else:
            quiet += 1
        i += 1
    return keep


def pass0(data: dict, y: str, xs: list[str], fe: list[str], weights: str | None = None,
          fixef_rm: str = "singleton", sort_vars: bool = True, stream: str | None = None) -> dict:
    """Rows after the restrictions, with level codes in sorted level order,
    sorted by (streamed group, other codes[, variables])."""
    cols = [np.asarray(data[f]) for f in fe]
    keep = singleton_restriction(cols) if fixef_rm == "singleton" else np.ones(len(cols[0]), bool)
    cols = [c[keep] for c in cols]
    names = [y] + [x for x in xs if x != y]
    V = np.column_stack([np.asarray(data[v], np.float64)[keep] for v in names])
    w = np.asarray(data[weights], np.float64)[keep] if weights else None
    card = [len(np.unique(c)) for c in cols]
    gi = (fe.index(stream) if stream is not None
          else max(range(len(fe)), key=lambda d: (card[d], -d)))     # estimator.py:371-385
    others = [d for d in range(len(fe)) if d != gi]
    levels, codes = [], []
    for d in others:
        u, inv = np.unique(cols[d], return_inverse=True)
        levels.append(u)
        codes.append(inv.astype(np.int64))
    gvals, ginv = np.unique(cols[gi], return_inverse=True)
    keys = [ginv] + codes
    if sort_vars:
        keys = keys + [V[:, j] for j in range(V.shape[1])]
    order = np.lexsort(tuple(reversed(keys)))
    return {
        "names": names, "fe": [fe[gi]] + [fe[d] for d in others], "n_obs": int(keep.sum()),
        "keep": keep, "order": order,
        "g": ginv[order].astype(np.int64), "g_levels": gvals,
        "codes": (np.column_stack([c[order] for c in codes]) if codes
                  else np.zeros((len(order), 0), np.int64)),
        "levels": levels, "n_levels": [len(u) for u in levels],
        "V": V[order], "w": None if w is None else w[order],
    }


# ------------------------------------------------------------------ pass 1

def pass1(p0: dict) -> dict:
    """Cells, identifying cells and the pass-1 accumulators
    (core/passes.py:397-589)."""
    g, codes, V, w = p0["g"], p0["codes"], p0["V"], p0["w"]
    n_rows, D = codes.shape
    m = V.shape[1]
    means = V.mean(axis=0)                                  # unweighted, as pass 0's scan
    raw_ss = ((V - means) ** 2).sum(axis=0)
    Z = V - means
    same = g[1:] == g[:-1]
    for d in range(D):
        same &= codes[1:, d] == codes[:-1, d]
    cs = np.flatnonzero(np.concatenate(([True], ~same)))
    bounds = np.append(cs, n_rows)
    run = np.diff(bounds).astype(np.float64)
    wz = Z if w is None else Z * w[:, None]
    sums = np.add.reduceat(wz, cs, axis=0) if n_rows else np.zeros((0, m))
    n = np.add.reduceat(w, cs) if w is not None else run
    cg = g[cs]
    ccodes = codes[cs]
    offs = np.concatenate(([0], np.cumsum(p0["n_levels"])[:-1])).astype(np.int64)
    L = int(sum(p0["n_levels"]))
    obs = np.zeros(L)
    cnt = np.zeros(L)
    for d in range(D):
        obs += np.bincount(offs[d] + ccodes[:, d], weights=n, minlength=L)
        cnt += np.bincount(offs[d] + ccodes[:, d], weights=run, minlength=L)
    # within-cell scatter (assembly="cells", kernels_base.py:374-422)
    mean_c = sums / n[:, None]
    dev = Z - np.repeat(mean_c, run.astype(np.int64), axis=0)
    W = (dev * (w[:, None] if w is not None else 1.0)).T @ dev
    # identifying groups: more than one cell
    gstart = np.flatnonzero(np.concatenate(([True], cg[1:] != cg[:-1])))
    ncell = np.diff(np.append(gstart, len(cg)))
    keep_g = ncell > 1
    keep_c = np.repeat(keep_g, ncell)
    kept = ncell[keep_g]
    starts = np.concatenate(([0], np.cumsum(kept))).astype(np.int64)
    return {
        "means": means, "raw_ss": raw_ss, "obs": obs, "cnt": cnt, "W_within": W,
        "offs": offs, "L": L, "D": D, "m": m, "n_cells": len(cs),
        "starts": starts, "codes": np.ascontiguousarray(ccodes[keep_c], dtype=np.uint32),
        "n": np.ascontiguousarray(n[keep_c]), "sums": np.ascontiguousarray(sums[keep_c]),
        "ident_groups": np.flatnonzero(keep_g),
    }


# ------------------------------------------------------------------ step 2

def diag(c: dict) -> np.ndarray:
    out = np.zeros(c["L"])
    lib().orc_diag(C.c_int64(len(c["starts"]) - 1), _p(c["starts"]), C.c_int32(c["D"]),
                   _p(c["codes"]), _p(c["n"]), _p(c["offs"]), _p(out))
    return out


def rhs(c: dict, sums: np.ndarray | None = None) -> np.ndarray:
    s = np.ascontiguousarray(c["sums"] if sums is None else sums, dtype=np.float64)
    m = s.shape[1]
    out = np.zeros((c["L"], m))
    lib().orc_rhs(C.c_int64(len(c["starts"]) - 1), _p(c["starts"]), C.c_int32(c["D"]),
                  _p(c["codes"]), _p(c["n"]), C.c_int32(m), _p(s), _p(c["offs"]), _p(out))
    return out


def build_S(c: dict) -> sp.csr_matrix:
    """The explicit S = D_o' M_0 D_o (core/solve.py:148-194)."""
    L = c["L"]
    G = len(c["starts"]) - 1
    rnnz = np.zeros(L, np.int64)
    nnz = lib().orc_s_count(C.c_int64(G), _p(c["starts"]), C.c_int32(c["D"]), _p(c["codes"]),
                            _p(c["offs"]), C.c_int64(L), _p(rnnz))
    indptr = np.zeros(L + 1, np.int64)
    np.cumsum(rnnz, out=indptr[1:])
    indices = np.empty(max(nnz, 1), np.int32)
    data = np.empty(max(nnz, 1))
    lib().orc_s_fill(C.c_int64(G), _p(c["starts"]), C.c_int32(c["D"]), _p(c["codes"]),
                     _p(c["n"]), _p(c["offs"]), C.c_int64(L), _p(indptr), _p(indices),
                     _p(data))
    S = sp.csr_matrix((data[:nnz], indices[:nnz], indptr), shape=(L, L))
    S.has_sorted_indices = True
    return S


def jacobi(d: np.ndarray) -> np.ndarray:
    """core/solve.py:142-146."""
    pos = d > 1e-12
    return np.where(pos, 1.0 / np.where(pos, d, 1.0), 0.0)


def pcg(S: sp.csr_matrix, dinv: np.ndarray, b: np.ndarray, zero=None, x0=None,
        tol: float = 1e-10, maxiter: int = 5000):
    """(X, iterations, converged, max_rel) -- core/solve.py:66-140."""
    b = np.ascontiguousarray(np.asarray(b, np.float64).reshape(S.shape[0], -1))
    L, k = b.shape
    X = np.empty_like(b)
    indptr = np.ascontiguousarray(S.indptr, np.int64)
    indices = np.ascontiguousarray(S.indices, np.int32)
    data = np.ascontiguousarray(S.data, np.float64)
    z = None if zero is None else np.ascontiguousarray(np.asarray(zero, np.uint8))
    xs = None if x0 is None else np.ascontiguousarray(np.asarray(x0, np.float64))
    it = C.c_int64(0)
    conv = C.c_int(0)
    rel = C.c_double(0.0)
    lib().orc_pcg(C.c_int64(L), _p(indptr), _p(indices), _p(data),
                  _p(np.ascontiguousarray(dinv, np.float64)), C.c_int64(k), _p(b),
                  None if z is None else _p(z), None if xs is None else _p(xs),
                  C.c_double(tol), C.c_int64(maxiter), _p(X), C.byref(it), C.byref(conv),
                  C.byref(rel))
    return X, int(it.value), bool(conv.value), float(rel.value)


def csr_matmat(S: sp.csr_matrix, X: np.ndarray) -> np.ndarray:
    X = np.ascontiguousarray(X, np.float64).reshape(S.shape[0], -1)
    out = np.empty_like(X)
    lib().orc_csr_matmat(C.c_int64(S.shape[0]), _p(np.ascontiguousarray(S.indptr, np.int64)),
                         _p(np.ascontiguousarray(S.indices, np.int32)),
                         _p(np.ascontiguousarray(S.data, np.float64)), C.c_int64(X.shape[1]),
                         _p(X), _p(out))
    return out


def stream_matvec(c: dict, P: np.ndarray) -> np.ndarray:
    P = np.ascontiguousarray(P, np.float64).reshape(c["L"], -1)
    out = np.empty_like(P)
    lib().orc_stream_matvec(C.c_int64(len(c["starts"]) - 1), _p(c["starts"]), C.c_int32(c["D"]),
                            _p(c["codes"]), _p(c["n"]), _p(c["offs"]), C.c_int64(c["L"]),
                            C.c_int64(P.shape[1]), _p(P), _p(out))
    return out


def components(c: dict, n_levels: int, col: int = 0) -> np.ndarray:
    parent = np.empty(n_levels, np.int64)
    lib().orc_components(C.c_int64(len(c["starts"]) - 1), _p(c["starts"]), C.c_int32(c["D"]),
                         _p(c["codes"]), C.c_int32(col), C.c_int64(n_levels), _p(parent))
    return parent


def rank_components(root: np.ndarray, cnt_f1: np.ndarray):
    """Labels ranked by rows, ties to the smallest fe[1] code
    (core/passes.py:933-950)."""
    nl = len(root)
    is_root = root == np.arange(nl)
    dense = (np.cumsum(is_root) - 1)[root]
    rows = np.bincount(dense, weights=cnt_f1, minlength=int(is_root.sum()))
    order = np.argsort(-rows, kind="stable")
    rank = np.empty(len(order), np.int64)
    rank[order] = np.arange(len(order))
    return rank[dense], len(order)


def normalize(G: np.ndarray, comp: np.ndarray, obs: np.ndarray, n_levels: list[int]) -> np.ndarray:
    """fe[1] effects: obs-weighted mean zero per component; the others: a
    global obs-weighted mean of zero (core/solve.py:383-399)."""
    G = np.array(G, dtype=np.float64, copy=True)
    o = 0
    nl = n_levels[0]
    w = obs[o:o + nl]
    cw = np.bincount(comp, weights=w, minlength=nl)
    cw[cw == 0] = 1.0
    blk = G[o:o + nl]
    mean = np.column_stack([np.bincount(comp, weights=w * blk[:, j], minlength=nl)
                            for j in range(G.shape[1])]) / cw[:, None]
    G[o:o + nl] -= mean[comp]
    o += nl
    for nl in n_levels[1:]:
        w = obs[o:o + nl]
        G[o:o + nl] -= (w @ G[o:o + nl]) / max(w.sum(), 1.0)
        o += nl
    return G


def assemble(c: dict, Gamma: np.ndarray) -> np.ndarray:
    m = c["m"]
    acc = np.zeros((m, m))
    lib().orc_assemble(C.c_int64(len(c["starts"]) - 1), _p(c["starts"]), C.c_int32(c["D"]),
                       _p(c["codes"]), _p(c["n"]), C.c_int32(m), _p(c["sums"]), _p(c["offs"]),
                       _p(np.ascontiguousarray(Gamma, np.float64)), _p(acc))
    return c["W_within"] + acc


def pass2(p0: dict, c: dict, gam_y: np.ndarray, beta: np.ndarray, xcols: list[int]):
    """Worker effects (per group) and residuals (per sorted row)."""
    g, codes = p0["g"], p0["codes"]
    n = len(g)
    starts = np.concatenate(([0], np.flatnonzero(g[1:] != g[:-1]) + 1, [n])).astype(np.int64)
    w = p0["w"] if p0["w"] is not None else np.ones(n)
    y = np.ascontiguousarray(p0["V"][:, 0])
    X = np.ascontiguousarray(p0["V"][:, xcols]) if xcols else np.zeros((n, 0))
    G = len(starts) - 1
    geff = np.empty(G)
    e = np.empty(n)
    lib().orc_pass2(C.c_int64(G), _p(starts), C.c_int32(codes.shape[1]),
                    _p(np.ascontiguousarray(codes, np.int64)), _p(c["offs"]),
                    _p(np.ascontiguousarray(w, np.float64)), _p(y), C.c_int32(X.shape[1]),
                    _p(np.ascontiguousarray(X, np.float64)),
                    _p(np.ascontiguousarray(beta, np.float64)),
                    _p(np.ascontiguousarray(gam_y, np.float64)), _p(geff), _p(e))
    return geff, e


def fit(data: dict, y: str = "y", xs=("x1", "x2"), fe=("worker_id", "firm_id", "year"),
        weights: str | None = None, tol: float = 1e-10, maxiter: int = 5000,
        rhs_block: int = 8, fixef_rm: str = "singleton", stream: str | None = None) -> dict:
    """The whole FP64 fit of `y ~ xs | fe`, as the reference runs it."""
    p0 = pass0(data, y, list(xs), list(fe), weights, fixef_rm=fixef_rm, stream=stream)
    c = pass1(p0)
    L = c["L"]
    m = c["m"]
    d = diag(c)
    b = rhs(c)
    S = build_S(c)
    dinv = jacobi(d)
    nl = p0["n_levels"]
    comp, n_comp = rank_components(components(c, nl[0]), c["cnt"][:nl[0]])
    Gamma = np.zeros((L, m))
    its = []
    for j0 in range(0, m, rhs_block):
        j1 = min(m, j0 + rhs_block)
        bb = b[:, j0:j1]
        scale = np.sqrt(np.maximum(c["raw_ss"][j0:j1], 0.0))
        zero = np.linalg.norm(bb, axis=0) <= RHS_ZERO_TOL * scale
        X, it, conv, rel = pcg(S, dinv, bb, zero, None, tol, maxiter)
        Gamma[:, j0:j1] = normalize(X, comp, c["obs"], nl)
        its.append(it)
    A = assemble(c, Gamma)
    xj = list(range(1, m))
    beta = np.linalg.solve(A[np.ix_(xj, xj)], A[xj, 0]) if xj else np.zeros(0)
    gam_y = Gamma[:, 0] - Gamma[:, xj] @ beta
    geff, e = pass2(p0, c, gam_y, beta, xj)
    resid_src = np.full(len(p0["keep"]), np.nan)
    resid_src[np.flatnonzero(p0["keep"])[p0["order"]]] = e
    return {"resid_source_order": resid_src, "beta": beta, "Gamma": Gamma, "gam_y": gam_y,
            "A": A, "S": S, "diag": d, "dinv": dinv, "rhs": b, "comp": comp,
            "n_components": n_comp, "iterations": its, "p0": p0, "cells": c,
            "worker_effects": geff, "resid": e}


# ------------------------------------------------------------------ KSS

def mix64(x: int) -> int:
    return int(lib().orc_mix64(C.c_uint64(x)))


def rademacher(ordinal: int, n_rows: int, draws: int, seed: int, first_draw: int) -> np.ndarray:
    out = np.empty((n_rows, draws))
    lib().orc_rademacher(C.c_int64(ordinal), C.c_int64(n_rows), C.c_int64(draws),
                         C.c_uint64(seed), C.c_int64(first_draw), _p(out))
    return out


def center_rows(starts, codes, offs, w, X, R, L):
    G = len(starts) - 1
    R = np.ascontiguousarray(R, np.float64)
    m = R.shape[1]
    kc = X.shape[1]
    acc = np.zeros((L, m))
    acc_x = np.zeros((kc, m))
    lib().orc_center_rows(C.c_int64(G), _p(np.ascontiguousarray(starts, np.int64)),
                          C.c_int32(codes.shape[1]), _p(np.ascontiguousarray(codes, np.int64)),
                          _p(np.ascontiguousarray(offs, np.int64)),
                          _p(np.ascontiguousarray(w, np.float64)), C.c_int32(kc),
                          _p(np.ascontiguousarray(X, np.float64)), C.c_int32(m), _p(R), _p(acc),
                          _p(acc_x))
    return acc, acc_x


def project_rows(starts, codes, offs, w, X, Cm, R, U):
    G = len(starts) - 1
    R = np.ascontiguousarray(R, np.float64)
    m = R.shape[1]
    kc = X.shape[1]
    out = np.empty_like(R)
    lib().orc_project_rows(C.c_int64(G), _p(np.ascontiguousarray(starts, np.int64)),
                           C.c_int32(codes.shape[1]), _p(np.ascontiguousarray(codes, np.int64)),
                           _p(np.ascontiguousarray(offs, np.int64)),
                           _p(np.ascontiguousarray(w, np.float64)), C.c_int32(kc),
                           _p(np.ascontiguousarray(X, np.float64)),
                           _p(np.ascontiguousarray(Cm, np.float64)), C.c_int32(m), _p(R),
                           _p(np.ascontiguousarray(U, np.float64)), _p(out))
    return out


def jla_moments(r: np.ndarray, proj: np.ndarray, moments: np.ndarray) -> None:
    n, k = r.shape
    lib().orc_jla_moments(C.c_int64(n), C.c_int32(k), _p(np.ascontiguousarray(r)),
                          _p(np.ascontiguousarray(proj)), _p(moments))


def combine_moments(moments: np.ndarray, n_draws: int):
    """(P_bar, M_bar, factor) -- leaveout.py:209-282 (the per-row algebra)."""
    P, M, mPP, mMM, mPM = (np.asarray(moments[i], np.float64) / n_draws for i in range(5))
    total = P + M
    safe = total > 0
    P_bar = np.divide(P, total, out=np.zeros_like(P), where=safe)
    M_bar = 1.0 - P_bar
    V = (M_bar ** 2 * mPP + P_bar ** 2 * mMM - 2 * P_bar * M_bar * mPM) / n_draws
    B = (M_bar * mPP - P_bar * mMM + (M_bar - P_bar) * mPM) / n_draws
    usable = M_bar > 0
    factor = np.ones_like(M_bar)
    np.divide(V, M_bar ** 2, out=factor, where=usable)
    factor = np.where(usable, 1.0 - factor + np.divide(B, M_bar, out=np.zeros_like(B),
                                                       where=usable), 1.0)
    return P_bar, M_bar, factor


def border(p0: dict, c: dict, S, dinv, X: np.ndarray, tol: float = 1e-10,
           maxiter: int = 5000):
    """The covariate border of the reduced system (inverse.py:182-231):
    T_dx = D_o'W M_0 X and T_xx = X'W M_0 X from one centring pass of X
    itself, Z = S^- T_dx (k solves; a column is zero when its norm is at most
    1e-11 sqrt(sum w x^2)), and pinv of the Schur complement T_xx - T_dx'Z.
    Returns (Z, Schur^+)."""
    g, codes = p0["g"], p0["codes"]
    n = len(g)
    starts = np.concatenate(([0], np.flatnonzero(g[1:] != g[:-1]) + 1, [n])).astype(np.int64)
    w = np.ones(n) if p0["w"] is None else np.asarray(p0["w"], np.float64)
    X = np.ascontiguousarray(X, np.float64)
    T_dx, T_xx = center_rows(starts, codes, c["offs"], w, X, X, c["L"])
    scale = np.sqrt(w @ (X * X))
    zero = np.linalg.norm(T_dx, axis=0) <= RHS_ZERO_TOL * scale
    Z, _, _, _ = pcg(S, dinv, T_dx, zero, None, tol, maxiter)
    return Z, np.linalg.pinv(T_xx - T_dx.T @ Z, hermitian=True)


def jla_leverages(p0: dict, c: dict, S, dinv, n_draws: int, block: int, seed: int = 0,
                  tol: float = 1e-10, maxiter: int = 5000, X=None):
    """Leverages by the reference's JLA: per block, centre, solve, project,
    accumulate the five moments (leaveout.py:130-198, inverse.py:240-309).
    With weights the operator is the square-root-weight one,
    W^1/2 A S^- A' W^1/2: the draws enter divided by sqrt(w) and the
    projection leaves multiplied by it (inverse.py:299-305, leaveout.py:179).
    With covariates X (n, k) the design is [X, fixed effects]: the border's
    Schur complement on the covariate corner (inverse.py:263-281)."""
    g, codes = p0["g"], p0["codes"]
    n = len(g)
    starts = np.concatenate(([0], np.flatnonzero(g[1:] != g[:-1]) + 1, [n])).astype(np.int64)
    weighted = p0["w"] is not None
    w = np.ones(n) if not weighted else np.asarray(p0["w"], np.float64)
    root = np.sqrt(w)
    X0 = np.zeros((n, 0)) if X is None else np.ascontiguousarray(X, np.float64)
    bz = border(p0, c, S, dinv, X0, tol, maxiter) if X0.shape[1] else None
    moments = np.zeros((5, n))
    drawn = 0
    while drawn < n_draws:
        size = min(block, n_draws - drawn)
        R = rademacher(0, n, size, seed, drawn)
        Rin = R / root[:, None] if weighted else R
        acc, acc_x = center_rows(starts, codes, c["offs"], w, X0, Rin, c["L"])
        scale = np.sqrt(w @ (Rin * Rin))
        zero = np.linalg.norm(acc, axis=0) <= RHS_ZERO_TOL * scale
        U, _, _, _ = pcg(S, dinv, acc, zero, None, tol, maxiter)
        Cm = np.zeros((0, size))
        if bz is not None:
            Cm = bz[1] @ (acc_x - bz[0].T @ acc)
            U = U - bz[0] @ Cm
        proj = project_rows(starts, codes, c["offs"], w, X0, Cm, Rin, U)
        if weighted:
            proj = proj * root[:, None]
        jla_moments(R, proj, moments)
        drawn += size
    return moments


# ------------------------------------------------------- KSS bias term

def design(p0: dict, c: dict):
    """The sparse incidence matrices of the sorted rows: D_g (n, G) of the
    streamed groups and D_o (n, L) of the other dimensions' levels."""
    g, codes = p0["g"], p0["codes"]
    n, D = codes.shape
    G = int(g.max()) + 1 if n else 0
    rows = np.arange(n)
    Dg = sp.csr_matrix((np.ones(n), (rows, g)), shape=(n, G))
    Do = sp.csr_matrix((np.ones(n * D), (np.repeat(rows, D),
                                        (codes + c["offs"][None, :]).reshape(-1))),
                       shape=(n, c["L"]))
    return Dg, Do


class KSSContext:
    """What the bias term reads from a fit: incidences, weights, group
    weights, the reduced S with its Jacobi inverse, and the null-space basis
    of the full coefficient space (inverse.py:376-433)."""

    def __init__(self, fitres: dict, tol: float = 1e-10, maxiter: int = 5000, X=None):
        p0, c = fitres["p0"], fitres["cells"]
        self.p0, self.c = p0, c
        self.Dg, self.Do = design(p0, c)
        n = self.Dg.shape[0]
        self.w = np.ones(n) if p0["w"] is None else np.asarray(p0["w"], np.float64)
        self.Wg = self.Dg.T @ self.w
        self.S, self.dinv = fitres["S"], fitres["dinv"]
        self.tol, self.maxiter = tol, maxiter
        self.G, self.L = self.Dg.shape[1], c["L"]
        self.g = p0["g"]
        # basis: per component (+1 on its groups, -1 on its first-dimension
        # levels); per further dimension (+1 first dimension, -1 that one)
        comp, ncomp = fitres["comp"], fitres["n_components"]
        nl, offs = p0["n_levels"], c["offs"]
        gcomp = comp[p0["codes"][np.r_[0, np.flatnonzero(np.diff(self.g)) + 1], 0]]
        vecs = []
        for k in range(ncomp):
            lv = np.zeros(self.L)
            lv[offs[0]:offs[0] + nl[0]] = -(comp == k).astype(float)
            vecs.append(np.concatenate([(gcomp == k).astype(float), lv]))
        for t in range(1, len(nl)):
            lv = np.zeros(self.L)
            lv[offs[0]:offs[0] + nl[0]] = 1.0
            lv[offs[t]:offs[t] + nl[t]] = -1.0
            vecs.append(np.concatenate([np.zeros(self.G), lv]))
        q, _ = np.linalg.qr(np.column_stack(vecs))
        self.q_groups, self.q_levels = q[:self.G], q[self.G:]
        # covariates of the design: the level block of coefficient space is
        # [levels (L); covariates (k)], K = L + k rows
        self.X = np.zeros((n, 0)) if X is None else np.ascontiguousarray(X, np.float64)
        self.kc = self.X.shape[1]
        self.K = self.L + self.kc
        self.border = (border(p0, c, self.S, self.dinv, self.X, tol, maxiter) if self.kc
                       else None)


def drop_null(ctx: KSSContext, levels: np.ndarray, streamed: np.ndarray):
    """inverse.py:446-460: the component along the null space removed."""
    overlap = ctx.q_groups.T @ streamed + ctx.q_levels.T @ levels
    return levels - ctx.q_levels @ overlap, streamed - ctx.q_groups @ overlap


def apply_inverse(ctx: KSSContext, levels: np.ndarray, streamed: np.ndarray):
    """u = S^- v for a coefficient-space block (inverse.py:311-374): drop the
    null component, reduce the spread streamed block onto the levels and the
    covariates (kernels_inverse.py:111-152), solve with the border's Schur
    complement, then recover the streamed block and the rows (A u)_i
    (kernels_inverse.py:156-188).  `levels` has L rows, or K = L + k with
    the covariate block below (missing: zero).  Returns (U (K, m), U_groups,
    rows)."""
    levels = np.asarray(levels, np.float64)
    lev, cov = levels[:ctx.L], levels[ctx.L:]
    lev, streamed = drop_null(ctx, lev, np.asarray(streamed, np.float64))
    scale = np.sqrt((lev * lev).sum(axis=0) + (streamed * streamed).sum(axis=0))
    Wg = np.where(ctx.Wg > 0, ctx.Wg, 1.0)
    rho = (streamed / Wg[:, None])[ctx.g]
    rhs = lev - ctx.Do.T @ (ctx.w[:, None] * rho)
    zero = np.linalg.norm(rhs, axis=0) <= RHS_ZERO_TOL * scale
    U, _, _, _ = pcg(ctx.S, ctx.dinv, rhs, zero, None, ctx.tol, ctx.maxiter)
    Cm = np.zeros((ctx.kc, levels.shape[1]))
    if ctx.kc:
        rhs_x = -(ctx.X.T @ (ctx.w[:, None] * rho))
        if cov.shape[0]:
            rhs_x = rhs_x + cov
        Z, sinv = ctx.border
        Cm = sinv @ (rhs_x - Z.T @ rhs)
        U = U - Z @ Cm
    h = ctx.Do @ U + ctx.X @ Cm
    u0 = (ctx.Dg.T @ (ctx.w[:, None] * (rho - h))) / Wg[:, None]
    return np.vstack([U, Cm]), u0, h + u0[ctx.g]


def trace_moments(ctx: KSSContext, psi_col: int, Zl, Zg, Ul, Ug) -> np.ndarray:
    """The eight cross-moments summed over rows and draws
    (kernels_inverse.py:191-231)."""
    lv = ctx.c["offs"][psi_col] + ctx.p0["codes"][:, psi_col]
    a1, a2, b1, b2 = Zl[lv], Ul[lv], Zg[ctx.g], Ug[ctx.g]
    w = ctx.w[:, None]
    return np.array([w.sum() * Zl.shape[1], (w * a1 * a2).sum(), (w * b1 * b2).sum(),
                     (w * (a1 * b2 + b1 * a2)).sum(), (w * a1).sum(), (w * a2).sum(),
                     (w * b1).sum(), (w * b2).sum()])


def moment_divisor(total: float, rows: int) -> float:
    """leaveout_se.py:202-214: total (R - 1) / R for R rows."""
    return total * (rows - 1.0) / rows if rows > 1 else total


def trace_components(totals: np.ndarray, n_draws: int, rows: int):
    """(var_psi, var_alpha, cov) from the summed moments (leaveout.py:378-390)."""
    weight, a1a2, b1b2, cross, a1, a2, b1, b2 = totals / n_draws
    var_psi = a1a2 / weight - (a1 / weight) * (a2 / weight)
    var_alpha = b1b2 / weight - (b1 / weight) * (b2 / weight)
    cov = 0.5 * (cross / weight - (a1 / weight) * (b2 / weight) - (b1 / weight) * (a2 / weight))
    scale = weight / moment_divisor(weight, rows)
    return scale * var_psi, scale * var_alpha, scale * cov


def hutchinson_trace(ctx: KSSContext, sigma2: np.ndarray, n_draws: int, block: int,
                     seed: int, psi_col: int = 0, n_rows: int | None = None):
    """tr(Q S^- A'W Omega A S^-) for the three components by Hutchinson's
    estimator with coefficient-space Rademacher draws keyed on (position,
    draw), levels first, then groups, then covariates (leaveout.py:321-448)."""
    L, G = ctx.L, ctx.G
    totals = np.zeros(8)
    drawn = 0
    while drawn < n_draws:
        size = min(block, n_draws - drawn)
        # positions: levels, groups, then covariates (leaveout.py:410-419)
        Z = rademacher(0, L + G + ctx.kc, size, seed, drawn)
        Zl, Zg = np.vstack([Z[:L], Z[L + G:]]), Z[L:L + G]
        _, _, rows = apply_inverse(ctx, Zl, Zg)
        weighted = rows * sigma2[:, None]
        vl = np.vstack([ctx.Do.T @ (ctx.w[:, None] * weighted),
                        ctx.X.T @ (ctx.w[:, None] * weighted)])
        vg = ctx.Dg.T @ (ctx.w[:, None] * weighted)
        Ul, Ug, _ = apply_inverse(ctx, vl, vg)
        totals += trace_moments(ctx, psi_col, Zl, Zg, Ul, Ug)
        drawn += size
    return trace_components(totals, n_draws, len(ctx.g) if n_rows is None else n_rows), totals


def sigma2(y: np.ndarray, e: np.ndarray, w: np.ndarray, complement: np.ndarray,
           factor: np.ndarray) -> np.ndarray:
    """w_i (y_i - ybar_w) e_i / M_ii * factor_i (leaveout.py:1000-1003)."""
    ybar = float(w @ y / w.sum())
    return w * (y - ybar) * e / complement * factor


def plug_in(a: np.ndarray, b: np.ndarray, w: np.ndarray, rows: int):
    """Weighted var(a), var(b), cov(a, b) over rows, divided by n - 1
    (leaveout.py:807-832)."""
    t = w.sum()
    ma, mb = (w @ a) / t, (w @ b) / t
    scale = t / moment_divisor(t, rows)
    return (scale * ((w @ (a * a)) / t - ma * ma), scale * ((w @ (b * b)) / t - mb * mb),
            scale * ((w @ (a * b)) / t - ma * mb))


def leave_out_kss(data: dict, y: str = "y", xs=(), alpha: str = "worker_id",
                  psi: str = "firm_id", others=(), n_draws: int = 250, block: int = 256,
                  seed: int = 0, tol: float = 1e-10, design_covariates: bool = True) -> dict:
    """KSS leaving out an observation, end to end (leaveout.py:1069-1255,
    634-805): leave-one-out pruning of the (alpha, psi) graph, the fit with
    alpha streamed and no singleton drop, JLA leverages, sigma2 ("own"
    stayers, weighted centring), the Hutchinson bias with seed + 1_000_003,
    and the components plug-in minus bias.  Unweighted.  The model's
    covariates `xs` join the leave-out design unless `design_covariates`
    is False (leaveout.py:111-127)."""
    a_ids, b_ids = np.asarray(data[alpha]), np.asarray(data[psi])
    ua, ia = np.unique(a_ids, return_inverse=True)
    ub, ib = np.unique(b_ids, return_inverse=True)
    alive, prune = leave_one_out(ia, ib, len(ua), len(ub))
    keep = alive[ia] & alive[len(ua) + ib]
    sub = {k: np.asarray(v)[keep] for k, v in data.items()
           if hasattr(v, "__len__") and len(v) == len(keep)}
    fe = [alpha, psi] + list(others)
    f = fit(sub, y, xs, fe, tol=tol, fixef_rm="none", stream=alpha)
    p0, c = f["p0"], f["cells"]
    n = len(p0["g"])
    X = p0["V"][:, 1:] if (len(xs) and design_covariates) else None
    moments = jla_leverages(p0, c, f["S"], f["dinv"], n_draws, block, seed, tol, X=X)
    _, M_bar, factor = combine_moments(moments, n_draws)
    ctx = KSSContext(f, tol, X=X)
    w = ctx.w
    s2 = sigma2(p0["V"][:, 0], f["resid"], w, M_bar, factor)
    psi_col = 0
    a = f["gam_y"][c["offs"][psi_col] + p0["codes"][:, psi_col]]
    b = f["worker_effects"][p0["g"]]
    plug = plug_in(a, b, w, n)
    bias, totals = hutchinson_trace(ctx, s2, n_draws, block, seed + 1_000_003, psi_col)
    return {"plug_in": plug, "bias": bias,
            "leave_out": tuple(p - q for p, q in zip(plug, bias)), "keep": keep,
            "prune": prune, "fit": f, "sigma2": s2, "M_bar": M_bar, "factor": factor,
            "moments": moments, "totals": totals, "n": n}


def leave_out_match(data: dict, y: str = "y", xs=(), alpha: str = "worker_id",
                    psi: str = "firm_id", others=(), n_draws: int = 250, block: int = 256,
                    seed: int = 0, tol: float = 1e-10, se: bool = False, se_draws=None,
                    se_block: int = 16, se_variance: str = "person_year",
                    exact: bool = False) -> dict:
    """KSS leaving out a whole match (leaveout_match.py:156-214, 63-132;
    leaveout.py:634-805 with stayers "within_match" and the reference
    centring): prune as leave_out_kss, fit the row model, collapse y~ = the
    worker effect + the firm effect + the residual to one row per (worker,
    firm) match (its weighted mean, weight = the spell's total weight), fit
    y ~ 1 | alpha + psi on the matches with those weights, JLA leverages in
    the square-root-weight metric, sigma2 = (sqrt(T) ybar - its mean over
    matches) sqrt(T) e / M_ii * factor for movers' matches and
    sigma_for_stayers for stayers', the Hutchinson bias, the plug-in moments
    with the person-year divisor.  Unweighted rows.  With `se`, the standard
    errors of var(psi) and the covariance (leaveout.py:759-775, 848-894):
    sigma2 recomputed with weighted centring on movers' matches, per match
    the within-match variance plus that over T ("person_year") or that alone
    with the stayers' values ("match"), b zero on stayers' matches, and the
    reference-centred kernel's rank-two term.  With `exact` (small panels)
    the leverages and the bias come from the dense pseudo-inverse of the
    weighted match design instead of JLA and Hutchinson draws (the exact
    counterpart of prototypes/kss_exact.py:398-486)."""
    a_ids, b_ids = np.asarray(data[alpha]), np.asarray(data[psi])
    ua, ia = np.unique(a_ids, return_inverse=True)
    ub, ib = np.unique(b_ids, return_inverse=True)
    alive, prune = leave_one_out(ia, ib, len(ua), len(ub))
    keep = alive[ia] & alive[len(ua) + ib]
    sub = {k: np.asarray(v)[keep] for k, v in data.items()
           if hasattr(v, "__len__") and len(v) == len(keep)}
    f = fit(sub, y, xs, [alpha, psi] + list(others), tol=tol, fixef_rm="none", stream=alpha)
    p0, c = f["p0"], f["cells"]
    g, pc = p0["g"], p0["codes"][:, 0]
    n = len(g)
    w_row = np.ones(n)
    ytil = f["worker_effects"][g] + f["gam_y"][c["offs"][0] + pc] + f["resid"]
    head = np.concatenate(([True], (g[1:] != g[:-1]) | (pc[1:] != pc[:-1])))
    mid = np.cumsum(head) - 1
    nm = int(mid[-1]) + 1
    T = np.bincount(mid, weights=w_row, minlength=nm)
    ybar = np.bincount(mid, weights=w_row * ytil, minlength=nm) / T
    first = np.flatnonzero(head)
    R = n
    mdata = {alpha: p0["g_levels"][g[first]], psi: p0["levels"][0][pc[first]], "y": ybar,
             "w": T}
    mf = fit(mdata, "y", (), [alpha, psi], weights="w", tol=tol, fixef_rm="none", stream=alpha)
    mp0, mc = mf["p0"], mf["cells"]
    mg, mpc = mp0["g"], mp0["codes"][:, 0]
    # the match fit keeps the table's (worker, firm) order
    assert np.array_equal(mg, g[first]) and np.array_equal(mpc, pc[first])
    if exact:
        hat = _exact_match_design(mg, mpc, mp0["w"], R)
        moments = None
        M_bar, factor = 1.0 - hat["P"], np.ones(nm)
    else:
        moments = jla_leverages(mp0, mc, mf["S"], mf["dinv"], n_draws, block, seed, tol)
        _, M_bar, factor = combine_moments(moments, n_draws)
    P_bar = 1.0 - M_bar
    nf = np.bincount(g, minlength=int(g.max()) + 1)          # rows per worker
    firms = np.bincount(mg, minlength=int(g.max()) + 1)      # matches per worker
    movers = firms[mg] > 1
    worst = float(P_bar[movers].max()) if movers.any() else 0.0
    a = mf["gam_y"][mc["offs"][0] + mpc]
    b = mf["worker_effects"][mg]
    Tm = mp0["w"]
    plug = plug_in(a, b, Tm, R)
    root = np.sqrt(Tm)
    yt = root * mp0["V"][:, 0]
    s2 = np.full(nm, np.nan)
    s2[movers] = ((yt[movers] - yt.mean()) * root[movers] * mf["resid"][movers]
                  / M_bar[movers] * factor[movers])
    # sigma_for_stayers (leaveout_match.py:91-132) on the person-year rows
    ybar_all = float(w_row @ ytil / w_row.sum())
    fitted = mf["worker_effects"][g] + mf["gam_y"][mc["offs"][0] + pc]
    Wk = np.bincount(g, weights=w_row)[g]
    r = (ytil - ybar_all) * (ytil - fitted) / (1.0 - w_row / Wk)
    st_rows = firms[g] == 1
    num = np.bincount(mid[st_rows], weights=(w_row ** 2 * r)[st_rows], minlength=nm)
    den = np.bincount(mid[st_rows], weights=w_row[st_rows], minlength=nm)
    s2[~movers] = num[~movers] / den[~movers]
    ctx = KSSContext(mf, tol)
    if exact:
        bias = tuple(float((Tm * b_ii) @ s2) for b_ii in hat["B"])
        totals = None
    else:
        bias, totals = hutchinson_trace(ctx, s2, n_draws, block, seed + 1_000_003, 0,
                                        n_rows=R)
    del nf
    out = {"plug_in": plug, "bias": bias,
           "leave_out": tuple(p - q for p, q in zip(plug, bias)), "keep": keep,
           "prune": prune, "fit": f, "match_fit": mf, "sigma2": s2, "M_bar": M_bar,
           "factor": factor, "moments": moments, "person_years": R, "n_matches": nm,
           "movers": movers, "max_leverage": worst}
    if se:
        yv = mp0["V"][:, 0]
        ybw = float(Tm @ yv / Tm.sum())
        blk = Tm * (yv - ybw) * mf["resid"] / M_bar * factor
        if se_variance == "match":
            sig = np.where(movers, blk, s2)
        else:
            within = (np.bincount(mid, weights=(ytil - ybar[mid]) ** 2, minlength=nm)
                      / np.bincount(mid, minlength=nm))
            sig = np.where(movers, within + blk / Tm, within)
        lay = SELayout(ctx, 0, R)
        out["se"] = component_variances(lay, yv, sig, P_bar, M_bar, factor,
                                        n_draws if se_draws is None else se_draws, se_block,
                                        seed + SE_BII_OFFSET, stayer_rows=~movers,
                                        centering="reference", reported=(0, 2))
        out["se_sigma2"] = sig
    return out


def _exact_match_design(worker, firm, T, person_years):
    """Exact leverages and bias kernels of the weighted match regression
    y_m = alpha_worker + psi_firm (weights T, the spells' lengths), densely:
    P_m = T_m a_m' S^+ a_m with S = A'TA, and B_m = a_m' S^+ Q S^+ a_m for
    the three person-year moments' Q (var(psi), var(alpha), cov), each the
    T-weighted, T-centred second moment divided by person_years - 1
    (prototypes/kss_exact.py:175-221, 398-486).  Both are invariant to the
    choice of generalised inverse, since A's null direction (one on the
    workers, minus one on the firms) is also Q's."""
    nw, nfm = int(worker.max()) + 1, int(firm.max()) + 1
    nm = len(worker)
    A = np.zeros((nm, nw + nfm))
    A[np.arange(nm), worker] = 1.0
    A[np.arange(nm), nw + firm] = 1.0
    S = A.T @ (T[:, None] * A)
    Sp = np.linalg.pinv(S, hermitian=True)
    AS = A @ Sp
    P = T * np.einsum("ij,ij->i", AS, A)
    total = T.sum()
    norm = total * (person_years - 1) / person_years
    Wc = A[:, :nw] - (T @ A[:, :nw]) / total
    Jc = A[:, nw:] - (T @ A[:, nw:]) / total
    K = nw + nfm
    Qpsi = np.zeros((K, K))
    Qpsi[nw:, nw:] = Jc.T @ (T[:, None] * Jc) / norm
    Qalpha = np.zeros((K, K))
    Qalpha[:nw, :nw] = Wc.T @ (T[:, None] * Wc) / norm
    Qcov = np.zeros((K, K))
    Qcov[nw:, :nw] = Jc.T @ (T[:, None] * Wc) / (2 * norm)
    Qcov[:nw, nw:] = Qcov[nw:, :nw].T
    B = [np.einsum("ij,jk,ik->i", AS, Q, AS) for Q in (Qpsi, Qalpha, Qcov)]
    return {"P": P, "B": B}


# ------------------------------------------------- differentiable fits

def dense_fe_components(y, w, X, codes, psi: int = 1):
    """[beta..., var(psi), var(alpha), cov(psi, alpha)] of the weighted least
    squares fit of y on [X, dummies of every fixed effect] as dense FP64
    torch operations on the CPU, so that torch's autograd differentiates it
    in y and w.  `codes` are 0-based level codes per dimension, alpha the
    first; one dummy of every further dimension is dropped (identification),
    which shifts the effects by constants and leaves the components
    unchanged.  The components are the weighted plug-in moments over rows
    with the divisor W (R - 1) / R (leaveout.py:807-832,
    leaveout_se.py:202-214)."""
    import torch
    n = y.shape[0]
    kx = X.shape[1] if X is not None else 0
    cols = [X] if kx else []
    for d, c in enumerate(codes):
        c = torch.as_tensor(np.asarray(c), dtype=torch.long)
        D = torch.zeros((n, int(c.max()) + 1), dtype=torch.float64)
        D[torch.arange(n), c] = 1.0
        cols.append(D[:, 1:] if d > 0 else D)
    Z = torch.cat(cols, dim=1)
    A = Z.T @ (w[:, None] * Z)
    theta = torch.linalg.solve(A, Z.T @ (w * y))
    bounds = np.cumsum([0] + [b.shape[1] for b in cols])
    fitted = [cols[j] @ theta[bounds[j]:bounds[j + 1]] for j in range(len(cols))]
    off = 1 if kx else 0
    alpha = fitted[off]
    ps = fitted[off + psi]
    W = w.sum()
    Dp = W * (n - 1.0) / n
    ma, mp = (w @ alpha) / W, (w @ ps) / W
    vpsi = (w @ ((ps - mp) ** 2)) / Dp
    valpha = (w @ ((alpha - ma) ** 2)) / Dp
    cov = (w @ ((ps - mp) * (alpha - ma))) / Dp
    return torch.cat([theta[:kx], torch.stack([vpsi, valpha, cov])])


# ---------------------------------------------------------- Poisson IRLS

def poisson_separation(y: np.ndarray, fe_cols: list) -> np.ndarray:
    """Rows kept by the separation check: one pass over the dimensions, each
    dropping the levels whose outcome is zero on every kept row
    (glm.py:346-367, Poisson)."""
    keep = np.ones(len(y), bool)
    if not np.any(y == 0):
        return keep
    for col in fe_cols:
        u, inv = np.unique(col, return_inverse=True)
        ymax = np.full(len(u), -np.inf)
        np.maximum.at(ymax, inv[keep], y[keep])
        bad = ymax == 0
        keep &= ~bad[inv]
    return keep


def fepois(data: dict, y: str, xs, fe, iwls_tol: float = 1e-8, iwls_maxiter: int = 25,
           tol: float = 1e-12) -> dict:
    """Poisson IRLS with fixed effects as the reference runs it (glm.py:743-838,
    core/irls.py:100-144, families.py:69-100): separation check, singleton
    drop, the start eta = log((y + ybar) / 2), then per step the weighted fit
    of z = eta + y / mu - 1 on [X, fixed effects] with W = mu (the sparse FP64
    fit above, its fitted values z - resid the new linear predictor), halving
    of steps that raise the deviance, and |dev - dev_old| / (0.1 + |dev|) <
    iwls_tol.  Returns beta, the linear predictor and deviance per kept row
    (source order), the kept-row mask and the step count."""
    yv = np.asarray(data[y], np.float64)
    cols = [np.asarray(data[f]) for f in fe]
    keep = poisson_separation(yv, cols)
    keep[keep] &= singleton_restriction([c[keep] for c in cols])
    sub = {k: np.asarray(v)[keep] for k, v in data.items()
           if hasattr(v, "__len__") and len(v) == len(keep)}
    yk = sub[y].astype(np.float64)

    def deviance(eta):
        mu = np.exp(eta)
        xlogy = np.where(yk != 0, yk * np.log(np.where(yk != 0, yk, 1.0)), 0.0)
        return float(np.sum(2.0 * (xlogy - yk * eta - yk + mu)))

    eta = np.log((yk + yk.mean()) / 2.0)
    dev = deviance(eta)
    beta = np.zeros(len(xs))
    init = True
    path = [dev]
    it = 0
    converged = False
    for it in range(1, iwls_maxiter + 1):
        mu = np.exp(eta)
        w = mu
        z = eta + yk / mu - 1.0
        d = dict(sub, __z__=z, __w__=w)
        f = fit(d, "__z__", list(xs), list(fe), weights="__w__", tol=tol, fixef_rm="none",
                stream=fe[0])
        eta_new = z - f["resid_source_order"]
        beta_new = f["beta"]
        dev_new = deviance(eta_new)
        if not init and not (np.isfinite(dev_new) and (
                dev_new < dev or abs(dev_new - dev) / (0.1 + abs(dev)) < max(iwls_tol, 1e-12))):
            t = 1.0
            target_eta, target_beta = eta_new, beta_new
            for _ in range(30):
                t /= 2
                eta_new = eta + t * (target_eta - eta)
                beta_new = beta + t * (target_beta - beta)
                dev_new = deviance(eta_new)
                if np.isfinite(dev_new) and dev_new < dev:
                    break
            else:
                raise RuntimeError("IRLS step-halving failed")
        crit = abs(dev_new - dev) / (0.1 + abs(dev))
        flat = (not init) and dev_new >= dev
        eta, beta, dev, init = eta_new, beta_new, dev_new, False
        path.append(dev)
        if crit < iwls_tol or flat:
            converged = True
            break
    return {"beta": beta, "eta": eta, "deviance": dev, "keep": keep, "iterations": it,
            "converged": converged, "path": path, "y": yk, "data": sub}


# ------------------------------------------------------------------ graph

def articulation_points(indptr, indices, alive):
    nv = len(indptr) - 1
    cut = np.zeros(nv, np.uint8)
    lib().orc_articulation_points(C.c_int64(nv), _p(np.ascontiguousarray(indptr, np.int64)),
                                  _p(np.ascontiguousarray(indices, np.int64)), _p(cut),
                                  _p(np.ascontiguousarray(alive, np.uint8)))
    return cut.astype(bool)


def bfs_components(indptr, indices, alive):
    nv = len(indptr) - 1
    label = np.empty(nv, np.int64)
    nc = lib().orc_bfs_components(C.c_int64(nv), _p(np.ascontiguousarray(indptr, np.int64)),
                                  _p(np.ascontiguousarray(indices, np.int64)),
                                  _p(np.ascontiguousarray(alive, np.uint8)), _p(label))
    return label, int(nc)


def pair_graph(a: np.ndarray, b: np.ndarray, n0: int, n1: int, alive: np.ndarray):
    """CSR adjacency in both directions of the distinct pairs (a, b), the
    second dimension's vertices from n0 (core/kernels_graph.py:313-341)."""
    keys = np.unique(a.astype(np.int64) * n1 + b.astype(np.int64))
    ka, kb = keys // n1, n0 + keys % n1
    ok = alive[ka] & alive[kb]
    ka, kb = ka[ok], kb[ok]
    nv = n0 + n1
    deg = np.bincount(ka, minlength=nv) + np.bincount(kb, minlength=nv)
    indptr = np.concatenate(([0], np.cumsum(deg))).astype(np.int64)
    indices = np.empty(indptr[-1], np.int64)
    pos = indptr[:-1].copy()
    for x, y in zip(ka, kb):        # in key order, as _nb_pairs_fill
        indices[pos[x]] = y
        pos[x] += 1
        indices[pos[y]] = x
        pos[y] += 1
    return keys, indptr, indices


def leave_one_out(a: np.ndarray, b: np.ndarray, n0: int, n1: int, alive=None, rank_by=None):
    """Leave-one-out connected set of the (worker a, firm b) rows: the
    alive mask over the n0 + n1 vertices and the counts
    (core/connected.py:503-562).  `alive` on entry (default all) is the set
    left by min_degree; rank_by None ranks components by rows, 0 or 1 by the
    levels of that dimension (:525-536)."""
    nv = n0 + n1
    alive = np.ones(nv, bool) if alive is None else np.asarray(alive, bool).copy()
    keys, indptr, indices = pair_graph(a, b, n0, n1, alive)
    kk = a.astype(np.int64) * n1 + b.astype(np.int64)
    _, counts = np.unique(kk, return_counts=True)
    ka, kb = keys // n1, n0 + keys % n1

    def keep_largest():
        label, nc = bfs_components(indptr, indices, alive)
        if nc == 0:
            return 0
        if rank_by is None:
            ok = (label[ka] >= 0) & (label[kb] >= 0)
            size = np.bincount(label[ka][ok], weights=counts[ok], minlength=nc)
        else:
            lab = label[:n0] if rank_by == 0 else label[n0:]
            size = np.bincount(lab[lab >= 0], minlength=nc)
        alive[:] &= label == int(np.argmax(size))
        return nc

    out = {"components_before": keep_largest()}
    ok = alive[ka] & alive[kb]
    rows0 = np.bincount(ka[ok], weights=counts[ok], minlength=n0)
    single = (rows0 == 1) & alive[:n0]
    out["single_observation"] = int(single.sum())
    if out["single_observation"]:
        alive[:n0] &= ~single
        keep_largest()
    cut_total, rounds = 0, 0
    while True:
        cut = articulation_points(indptr, indices, alive.astype(np.uint8))[:n0] & alive[:n0]
        n_cut = int(cut.sum())
        if not n_cut:
            break
        cut_total += n_cut
        rounds += 1
        alive[:n0] &= ~cut
        keep_largest()
    out.update(cut=cut_total, rounds=rounds)
    return alive, out


def peel(keys: np.ndarray, n0: int, n1: int, k0: int, k1: int, alive: np.ndarray) -> int:
    """The (k0, k1)-core of the distinct pairs keys = a * n1 + b, in place on
    alive (core/kernels_graph.py:187-221): every level with fewer alive
    partners than its dimension's k goes, all at once per round; returns the
    rounds that removed something."""
    ka, kb = keys // n1, n0 + keys % n1
    thr = np.where(np.arange(n0 + n1) < n0, k0, k1)
    rounds = 0
    while True:
        ok = alive[ka] & alive[kb]
        deg = (np.bincount(ka[ok], minlength=n0 + n1) + np.bincount(kb[ok], minlength=n0 + n1))
        drop = alive & (deg < thr)
        if not drop.any():
            return rounds
        alive &= ~drop
        rounds += 1


def union_labels(keys: np.ndarray, n0: int, n1: int, alive: np.ndarray):
    """Union-find over the alive pairs, each root the smaller vertex
    (core/kernels_graph.py:224-240), then labels 0, 1, ... in the order of
    the roots, -1 for dead vertices (:243-262).  Pure Python: small graphs."""
    nv = n0 + n1
    parent = np.arange(nv, dtype=np.int64)

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for key in keys.tolist():
        a, b = key // n1, n0 + key % n1
        if not (alive[a] and alive[b]):
            continue
        ra, rb = find(a), find(b)
        if ra != rb:
            if ra < rb:
                parent[rb] = ra
            else:
                parent[ra] = rb
    label = np.full(nv, -1, np.int64)
    nc = 0
    for x in range(nv):
        r = find(x)
        if not alive[x]:
            continue
        if r == x:
            label[x] = nc
            nc += 1
        else:
            label[x] = label[r]
    return label, nc


def connected_components(a: np.ndarray, b: np.ndarray, n0: int, n1: int, min_degree=(1, 1),
                         leave_one_out_set: bool = False, rank_by=None):
    """Components of the graph of the level codes (a, b) of the rows
    (core/connected.py:378-493): the (k0, k1)-core first, then the
    leave-one-out set or union-find, then the ranking by size (rows, or the
    levels of dimension rank_by), stable so that ties go to the component
    with the smallest first-dimension level.  Returns (comp per vertex,
    sizes (nc, 3): rows and levels of each dimension, diagnostics)."""
    kk = a.astype(np.int64) * n1 + b.astype(np.int64)
    keys, counts = np.unique(kk, return_counts=True)
    nv = n0 + n1
    alive = np.ones(nv, bool)
    diag = {"rows": int(len(a)), "distinct_pairs": int(len(keys))}
    ka, kb = keys // n1, n0 + keys % n1

    def rows_alive(al):
        return int(counts[al[ka] & al[kb]].sum())

    if tuple(min_degree) != (1, 1):
        diag["pruning_rounds"] = peel(keys, n0, n1, min_degree[0], min_degree[1], alive)
        diag["pruned_levels"] = (int(n0 - alive[:n0].sum()), int(n1 - alive[n0:].sum()))
        diag["pruned_rows"] = len(a) - rows_alive(alive)
    if leave_one_out_set:
        before = alive.copy()
        alive, lo = leave_one_out(a, b, n0, n1, alive=alive, rank_by=rank_by)
        lo["levels"] = (int((before[:n0] & ~alive[:n0]).sum()),
                        int((before[n0:] & ~alive[n0:]).sum()))
        lo["rows"] = rows_alive(before) - rows_alive(alive)
        diag["leave_one_out"] = lo
        label = np.where(alive, 0, -1).astype(np.int64)
        nc = int(alive.any())
    else:
        label, nc = union_labels(keys, n0, n1, alive)
    ok = (label[ka] >= 0) & (label[kb] >= 0)
    rows = np.bincount(label[ka][ok], weights=counts[ok], minlength=nc).astype(np.int64)
    c0 = np.bincount(label[:n0][label[:n0] >= 0], minlength=nc)
    c1 = np.bincount(label[n0:][label[n0:] >= 0], minlength=nc)
    key = rows if rank_by is None else (c0 if rank_by == 0 else c1)
    order = np.argsort(-key, kind="stable")
    rank = np.empty(nc, np.int64)
    rank[order] = np.arange(nc)
    comp = np.where(label >= 0, rank[np.maximum(label, 0)], -1)
    sizes = np.stack([rows[order], c0[order], c1[order]], axis=1) if nc else np.zeros((0, 3), np.int64)
    return comp, sizes, diag


# ------------------------------------------------------ KSS standard errors
#
# leaveout_se.py:1-49, 128-950: in the square-root-weight metric X~ = W^(1/2) A,
# the kernel of component k is C_k = B_k - (diag(b_k) M + M diag(b_k)) / 2 with
# B_k = X~ S^- Q_k S^- X~', M = I - X~ S^- X~', b_k = B_k,ii / M_ii (times the
# leverage pass's factor), and V-hat = 4 sum_i s_i (C y~)_i^2 - 2 tr(C Om C Om).

SE_COMPONENTS = ("var(psi)", "var(alpha)", "cov(psi, alpha)")
SE_QUANTILE_SAMPLE = 2_000_000     # leaveout_se.py:62
SE_BII_OFFSET = 5_500_011          # leaveout.py:772
SE_TRACE_OFFSET = 7_700_017        # leaveout_se.py:752


class SELayout:
    """Weight totals per level and group, the grand total, the divisor and
    the psi block (leaveout_se.py:135-180)."""

    def __init__(self, ctx: KSSContext, psi_col: int, rows: int):
        self.ctx = ctx
        w = ctx.w
        self.level_w = ctx.Do.T @ w
        self.group_w = ctx.Dg.T @ w
        self.total = float(self.group_w.sum())
        self.norm = moment_divisor(self.total, rows)
        o = int(ctx.c["offs"][psi_col])
        self.psi = slice(o, o + int(ctx.p0["n_levels"][psi_col]))
        self.psi_level = o + ctx.p0["codes"][:, psi_col]
        self.root = np.sqrt(w)
        self.inv_root = np.divide(1.0, self.root, out=np.zeros_like(self.root),
                                  where=self.root > 0)


def se_reduce(lay: SELayout, R: np.ndarray):
    """X~' R in coefficient space: (levels and covariates (K), groups) of
    sum_i sqrt(w_i) R_i (the reference feeds R / sqrt(w) to reductions that
    apply w)."""
    ctx = lay.ctx
    x = lay.root[:, None] * R
    return np.vstack([ctx.Do.T @ x, ctx.X.T @ x]), ctx.Dg.T @ x


def bii_rows(lay: SELayout, n_draws: int, block: int, seed: int) -> np.ndarray:
    """B_ii of the three components by random projection, (3, n)
    (leaveout_se.py:228-327): per draw q, the two solves of L_psi'q and
    L_alpha'q (weight-centred selectors over sqrt(norm)), read back at the
    rows in the metric, squared and crossed."""
    ctx = lay.ctx
    n, L, G = len(ctx.w), ctx.L, ctx.G
    out = np.zeros((3, n))
    drawn = 0
    sq = np.sqrt(lay.norm)
    while drawn < n_draws:
        size = min(block, n_draws - drawn)
        q = rademacher(0, n, size, seed, drawn)
        lv, gs = se_reduce(lay, q)
        grand = lv[lay.psi].sum(axis=0) / lay.total
        rl = np.zeros((L, 2 * size))
        rl[lay.psi, :size] = (lv[lay.psi] - np.outer(lay.level_w[lay.psi], grand)) / sq
        rg = np.zeros((G, 2 * size))
        rg[:, size:] = (gs - np.outer(lay.group_w, grand)) / sq
        _, _, rows = apply_inverse(ctx, rl, rg)
        r1 = lay.root[:, None] * rows[:, :size]
        r2 = lay.root[:, None] * rows[:, size:]
        out[0] += (r1 * r1).sum(axis=1)
        out[1] += (r2 * r2).sum(axis=1)
        out[2] += (r1 * r2).sum(axis=1)
        drawn += size
    return out / n_draws


def apply_q(lay: SELayout, k: int, Ul: np.ndarray, Ug: np.ndarray):
    """Q_k u in coefficient space (leaveout_se.py:339-385): the variances by
    weight-centred scaling, the covariance by crossing the two dimensions
    over the rows."""
    ctx = lay.ctx
    s = Ul.shape[1]
    ol, og = np.zeros((ctx.K, s)), np.zeros((ctx.G, s))
    lw = lay.level_w[lay.psi]
    psi_mean = (lw @ Ul[lay.psi]) / lay.total
    alpha_mean = (lay.group_w @ Ug) / lay.total
    if k == 0:
        ol[lay.psi] = lw[:, None] * (Ul[lay.psi] - psi_mean) / lay.norm
    elif k == 1:
        og[:] = lay.group_w[:, None] * (Ug - alpha_mean) / lay.norm
    else:
        w = ctx.w[:, None]
        cross_l = ctx.Do.T @ (w * Ug[ctx.g])
        cross_g = ctx.Dg.T @ (w * Ul[lay.psi_level])
        ol[lay.psi] = (cross_l[lay.psi] - lw[:, None] * alpha_mean) / (2 * lay.norm)
        og[:] = (cross_g - lay.group_w[:, None] * psi_mean) / (2 * lay.norm)
    return ol, og


def apply_c(lay: SELayout, v: np.ndarray, b: np.ndarray, per_component: bool = False,
            rank_two=None, deflate=None) -> np.ndarray:
    """C_k v for the three components, (n, 3, s) (leaveout_se.py:426-596):
    v is (n, s) shared or (n, 3, s) one per component; b (3, n).  Solve 1:
    S^- X~'[v, b_k v]; Q_k on the input block; solve 2; then
    C v = B v - (b .* M v + M (b .* v)) / 2, plus rank_two = (u (3, n), c):
    c (1 u_k' + u_k 1') v.  deflate = (lambdas (3,), x1 (n, 3)) applies C2,
    built from B2 = B - lambda_1 x1 x1' (leaveout_se.py:497-499, 579-580)."""
    ctx = lay.ctx
    n_in = 3 if per_component else 1
    s = v.shape[-1]
    part = (lambda k: v[:, k, :]) if per_component else (lambda k: v)
    stack = np.concatenate([part(k) for k in range(n_in)]
                           + [part(k) * b[k][:, None] for k in range(3)], axis=1)
    lv, gs = se_reduce(lay, stack)
    U0l, U0g, base = apply_inverse(ctx, lv, gs)
    ql, qg = np.zeros((ctx.K, 3 * s)), np.zeros((ctx.G, 3 * s))
    inc = lambda k: slice((k if per_component else 0) * s, ((k if per_component else 0) + 1) * s)
    bc = lambda k: slice((n_in + k) * s, (n_in + k + 1) * s)
    for k in range(3):
        ol, og = apply_q(lay, k, U0l[:, inc(k)], U0g[:, inc(k)])
        ql[:, k * s:(k + 1) * s] = ol
        qg[:, k * s:(k + 1) * s] = og
    _, _, thq = apply_inverse(ctx, ql, qg)
    root = lay.root[:, None]
    cv = np.empty((len(ctx.w), 3, s))
    for k in range(3):
        vk = part(k)
        Mv = vk - root * base[:, inc(k)]
        Bv = root * thq[:, k * s:(k + 1) * s]
        if deflate is not None:
            Bv = Bv - deflate[0][k] * np.outer(deflate[1][:, k], deflate[1][:, k] @ vk)
        Mbv = vk * b[k][:, None] - root * base[:, bc(k)]
        cv[:, k, :] = Bv - 0.5 * (b[k][:, None] * Mv + Mbv)
        if rank_two is not None:
            u, c = rank_two
            cv[:, k, :] += c * ((u[k] @ vk)[None, :] + np.outer(u[k], vk.sum(axis=0)))
    return cv


def m_times(lay: SELayout, b: np.ndarray) -> np.ndarray:
    """u_k = M b_k for the three row vectors (leaveout_se.py:387-424)."""
    lv, gs = se_reduce(lay, b.T)
    _, _, base = apply_inverse(lay.ctx, lv, gs)
    return b - (lay.root[:, None] * base).T


def auto_bins(n: int, dims: int = 2) -> int:
    """Bins per axis, n^(1/6) in two dimensions (leaveout_se.py:955-964)."""
    power = 1.0 / 6.0 if dims == 2 else 1.0 / 3.0
    return int(np.clip(round(n ** power), 2, 500))


def smoother_knots(values: np.ndarray, n_bins: int, take: np.ndarray) -> np.ndarray:
    """Quantile knots made strictly increasing (leaveout_se.py:967-973)."""
    q = np.quantile(np.asarray(values[take], np.float64), np.linspace(0.0, 1.0, n_bins + 1))
    q = np.maximum.accumulate(q)
    span = max(float(q[-1] - q[0]), 1e-300)
    return q + np.arange(len(q)) * span * 1e-12


def smoother_rank(knots: np.ndarray, values: np.ndarray) -> np.ndarray:
    return np.interp(np.asarray(values, np.float64), knots, np.linspace(0.0, 1.0, len(knots)))


def smoother_sample(n: int) -> np.ndarray:
    """The rows the knots are taken from (leaveout_se.py:1021-1023)."""
    if n <= SE_QUANTILE_SAMPLE:
        return np.arange(n)
    return np.sort(np.random.default_rng(0).choice(n, SE_QUANTILE_SAMPLE, replace=False))


def fit_smoother(sigma2, leverage, b_ii, n_bins=None):
    """Cell means of sigma2 on the quantile grid of (P_ii, B_ii)
    (leaveout_se.py:982-1035): (knots_p, knots_b, means, n_bins)."""
    n = len(sigma2)
    nb = auto_bins(n, 2) if n_bins is None else int(n_bins)
    if nb < 1:
        return None
    take = smoother_sample(n)
    kp, kb = smoother_knots(leverage, nb, take), smoother_knots(b_ii, nb, take)
    cp = np.minimum((smoother_rank(kp, leverage) * nb).astype(np.int64), nb - 1)
    cb = np.minimum((smoother_rank(kb, b_ii) * nb).astype(np.int64), nb - 1)
    flat = cp * nb + cb
    tot = np.bincount(flat, weights=np.asarray(sigma2, np.float64), minlength=nb * nb)
    cnt = np.bincount(flat, minlength=nb * nb).astype(np.float64)
    overall = float(tot.sum() / max(cnt.sum(), 1.0))
    means = np.where(cnt > 0, tot / np.maximum(cnt, 1.0), overall).reshape(nb, nb)
    return kp, kb, means, nb


def smooth_eval(table, leverage, b_ii) -> np.ndarray:
    """Bilinear between cell centres (leaveout_se.py:1038-1054)."""
    kp, kb, means, nb = table

    def at(knots, values):
        f = smoother_rank(knots, values) * nb - 0.5
        j0 = np.clip(np.floor(f).astype(np.int64), 0, nb - 2)
        return j0, np.clip(f - j0, 0.0, 1.0)

    i0, a = at(kp, leverage)
    j0, c = at(kb, b_ii)
    return ((1.0 - a) * (1.0 - c) * means[i0, j0] + a * (1.0 - c) * means[i0 + 1, j0]
            + (1.0 - a) * c * means[i0, j0 + 1] + a * c * means[i0 + 1, j0 + 1])


def component_variances(lay: SELayout, y: np.ndarray, sigma2: np.ndarray, leverage,
                        complement, factor, n_draws: int, block: int, seed: int,
                        trace: bool = True, smooth_bins=None, b_ii=None, stayer_rows=None,
                        centering: str = "weighted", reported=(0, 1, 2)) -> dict:
    """V-hat and the standard errors of the three components
    (leaveout_se.py:599-784); `seed` is the one component_variances receives
    (the caller's seed + 5_500_011), the trace's draws use seed + 7_700_017."""
    ctx = lay.ctx
    n = len(ctx.w)
    if b_ii is None:
        b_ii = bii_rows(lay, n_draws, block, seed)
    b = b_ii * factor[None, :] / complement[None, :]
    if stayer_rows is not None:
        b[:, stayer_rows] = 0.0
    rank_two = (m_times(lay, b), 0.5 / n) if centering == "reference" else None
    tables = [fit_smoother(sigma2, leverage, b_ii[k], smooth_bins) for k in range(3)]
    s = np.stack([sigma2 if t is None else smooth_eval(t, leverage, b_ii[k])
                  for k, t in enumerate(tables)])
    w = ctx.w
    ybar = float(w @ y / w.sum())
    yt = lay.root * (y - (ybar if centering == "weighted" else 0.0))
    cy = apply_c(lay, yt[:, None], b, rank_two=rank_two)[:, :, 0]
    first = np.array([4.0 * s[k] @ cy[:, k] ** 2 for k in range(3)])
    theta = np.array([yt @ cy[:, k] for k in range(3)])
    total = np.zeros(3)
    if trace:
        drawn = 0
        while drawn < n_draws:
            size = min(block, n_draws - drawn)
            z = rademacher(0, n, size, seed + SE_TRACE_OFFSET, drawn)
            v = np.stack([z * np.sqrt(np.maximum(s[k], 0.0))[:, None] for k in range(3)], axis=1)
            cv = apply_c(lay, v, b, per_component=True, rank_two=rank_two)
            for k in range(3):
                total[k] += s[k] @ (cv[:, k, :] ** 2).sum(axis=1)
            drawn += size
        total /= n_draws
    variance = first - 2.0 * total
    for k in range(3):
        if k not in reported:
            variance[k] = first[k] = total[k] = theta[k] = np.nan
    return {"variance": variance, "first": first, "trace": 2.0 * total, "theta": theta,
            "se": np.where(variance > 0, np.sqrt(np.abs(variance)), np.nan),
            "se_conservative": np.where(first > 0, np.sqrt(np.abs(first)), np.nan),
            "b_ii": b_ii, "b": b, "s": s}


# -------------------------------------------- weak-identification diagnostic
#
# leaveout_weakid.py: Lanczos on Q S^-, self-adjoint in <x, y> = x' S^- y, the
# three components in lockstep (one solve of three columns per step), from
# x0 = Q r; the leading Ritz values with Cullum-Willoughby filtering; the Ritz
# vectors by retracing; sum lambda^2 = tr((S^- Q)^2) by deflated Hutchinson;
# KSS's q rule and the first-stage F of the leading direction.

WEAK_ID_THRESHOLD = 0.1            # leaveout_weakid.py:67
WEAKID_SEED_OFFSET = 9_900_019     # leaveout.py:758
WEAKID_SUM_OFFSET = 3_300_007      # leaveout_weakid.py:651
SCRATCH_BYTES = 32 << 20           # core/estimator.py:77 (scratch_mb)


def q_block(lay: SELayout, comps, Ul: np.ndarray, Ug: np.ndarray):
    """Q applied column by column, column j with component comps[j]
    (leaveout_weakid.py:164-179)."""
    comps = np.asarray(comps)
    ol, og = np.zeros((lay.ctx.K, Ul.shape[1])), np.zeros_like(Ug)
    for k in range(3):
        cols = np.flatnonzero(comps == k)
        if len(cols):
            a, b = apply_q(lay, k, Ul[:, cols], Ug[:, cols])
            ol[:, cols], og[:, cols] = a, b
    return ol, og


def _cdot(a, b):
    return sum((x * y).sum(axis=0) for x, y in zip(a, b))


def krylov_plan(lay: SELayout, max_iter: int):
    """Steps per component and whether its basis is kept for full
    reorthogonalization (leaveout_weakid.py:186-215)."""
    n_psi = lay.psi.stop - lay.psi.start
    G, K = lay.ctx.G, lay.ctx.K
    dims = [max(n_psi - 1, 0), max(K + 1, 1), max(2 * (n_psi - 1), 0)]
    support = [n_psi, G, n_psi + G]
    caps = [min(max_iter, d) for d in dims]
    reorth = [cap * 2 * rows * 8 <= SCRATCH_BYTES for cap, rows in zip(caps, support)]
    return caps, reorth


def lanczos(lay: SELayout, n_iter: int, seed: int, ritz=None, check=None, every: int = 20,
            caps=None, reorth=None):
    """The three-component recurrence (leaveout_weakid.py:217-383); blocks are
    (levels (L, 3), groups (G, 3)).  Returns (alphas, betas, lengths, steps),
    or with `ritz` the Ritz combinations (X, Y) of 3k columns."""
    ctx = lay.ctx
    L, G = ctx.L, ctx.G
    psi = lay.psi
    caps = [n_iter] * 3 if caps is None else list(caps)
    reorth = [False] * 3 if reorth is None else list(reorth)
    comps = [0, 1, 2]

    def support(c, blocks):
        parts = []
        if c in (0, 2):
            parts.append(blocks[0][psi, c])
        if c in (1, 2):
            parts.append(blocks[1][:, c])
        return np.concatenate(parts) if len(parts) > 1 else parts[0].copy()

    def project(blocks):
        for c in (0, 2):
            blocks[0][psi, c] -= blocks[0][psi, c].mean()
        for c in (1, 2):
            blocks[1][:, c] -= blocks[1][:, c].mean()

    def scatter(c, flat, blocks):
        at = 0
        if c in (0, 2):
            n = psi.stop - psi.start
            blocks[0][psi, c] -= flat[at:at + n]
            at += n
        if c in (1, 2):
            blocks[1][:, c] -= flat[at:]

    # x0 = Q r, r Rademacher keyed on (position, column): levels then groups
    rl = rademacher(0, L, 3, seed, 0)
    rg = rademacher(L, G, 3, seed, 0)
    ql, qg = q_block(lay, comps, rl, rg)
    x = [ql, qg]
    project(x)
    y = list(apply_inverse(ctx, x[0], x[1])[:2])
    norm = np.sqrt(np.maximum(_cdot(x, y), 0.0))
    live = (norm > 0) & (np.array(caps) > 0)
    for blk in (*x, *y):
        blk[:, live] /= norm[live]
        blk[:, ~live] = 0.0
    x_prev = [np.zeros((ctx.K, 3)), np.zeros((G, 3))]
    beta_prev = np.zeros(3)
    basis = [([], []) for _ in comps]
    alphas, betas = [[] for _ in comps], [[] for _ in comps]
    lengths = np.zeros(3, dtype=int)
    acc = None
    if ritz is not None:
        k = ritz[0].shape[1]
        acc = ([np.zeros((ctx.K, 3 * k)), np.zeros((G, 3 * k))],
               [np.zeros((ctx.K, 3 * k)), np.zeros((G, 3 * k))])
    beta_max = np.zeros(3)
    steps = 0
    for j in range(n_iter):
        if not live.any():
            break
        if (check is not None and j > 0 and j % every == 0
                and check([np.array(a) for a in alphas], [np.array(b) for b in betas],
                          lengths, j)):
            break
        steps = j + 1
        if acc is not None:
            for c in np.flatnonzero(live):
                if j < len(ritz[c]):
                    cols = slice(c * k, (c + 1) * k)
                    for ax, ay, xb, yb in zip(acc[0], acc[1], x, y):
                        ax[:, cols] += np.outer(xb[:, c], ritz[c][j])
                        ay[:, cols] += np.outer(yb[:, c], ritz[c][j])
        for c in np.flatnonzero(live):
            if reorth[c]:
                basis[c][0].append(support(c, x))
                basis[c][1].append(support(c, y))
        ql, qg = q_block(lay, comps, y[0], y[1])
        w = [ql, qg]
        for wb, xp in zip(w, x_prev):
            wb -= beta_prev * xp
        alpha = _cdot(w, y)
        for wb, xb in zip(w, x):
            wb -= alpha * xb
        fix = _cdot(w, y)
        for wb, xb in zip(w, x):
            wb -= fix * xb
        alpha = alpha + fix
        project(w)
        for c in np.flatnonzero(live):
            if not reorth[c] or not basis[c][0]:
                continue
            for _ in range(2):
                wc = support(c, w)
                Xs, Ys = np.column_stack(basis[c][0]), np.column_stack(basis[c][1])
                scatter(c, Xs @ (Ys.T @ wc), w)
        project(w)
        z = list(apply_inverse(ctx, w[0], w[1])[:2])
        beta = np.sqrt(np.maximum(_cdot(w, z), 0.0))
        beta_max = np.maximum(beta_max, beta)
        for c in np.flatnonzero(live):
            alphas[c].append(float(alpha[c]))
            betas[c].append(float(beta[c]))
            lengths[c] = j + 1
        live = live & (beta > 1e-8 * np.maximum(beta_max, 1e-300)) & (lengths < np.array(caps))
        x_prev = x
        x = [wb.copy() for wb in w]
        y = z
        for blk in (*x, *y):
            blk[:, live] /= beta[live]
            blk[:, ~live] = 0.0
        beta_prev = np.where(live, beta, 0.0)
    if acc is not None:
        return acc
    return [np.array(a) for a in alphas], [np.array(b) for b in betas], lengths, steps


def ritz_top(alpha: np.ndarray, beta: np.ndarray, top: int):
    """The leading genuine Ritz values, their error bounds and eigenvectors:
    ghosts merged, lone spurious values dropped by Cullum-Willoughby
    (leaveout_weakid.py:385-433)."""
    from scipy.linalg import eigh_tridiagonal, eigvalsh_tridiagonal
    m = len(alpha)
    if m == 0:
        return np.array([]), np.array([]), np.zeros((0, 0))
    off = beta[:m - 1]
    theta, S = eigh_tridiagonal(alpha, off)
    bounds = np.abs(beta[m - 1] * S[-1, :])
    norm = max(float(np.abs(theta).max()), 1e-300)
    merge_tol = 1e-8 * norm
    order = np.argsort(theta)
    keep = []
    for idx in order:
        if keep and abs(theta[idx] - theta[keep[-1]]) <= merge_tol:
            if bounds[idx] < bounds[keep[-1]]:
                keep[-1] = idx
            continue
        keep.append(idx)
    mult = {}
    for idx in order:
        head = min(keep, key=lambda q: abs(theta[q] - theta[idx]))
        mult[head] = mult.get(head, 0) + 1
    if m > 1:
        hat = eigvalsh_tridiagonal(alpha[1:], off[1:]) if m > 2 else alpha[1:]
        spur = 1e-11 * norm
        keep = [q for q in keep if mult[q] > 1 or np.min(np.abs(hat - theta[q])) > spur]
    keep = sorted(keep, key=lambda q: -abs(theta[q]))[:top]
    return theta[keep], bounds[keep], S[:, keep]


def s_orthonormalize(X, Y, k: int):
    """Modified Gram-Schmidt of each component's k Ritz vectors in
    <x, y> = x'S^- y (leaveout_weakid.py:436-465)."""
    for c in range(3):
        base = c * k
        for a in range(k):
            ca = base + a
            for p in range(a):
                cp = base + p
                coef = sum(xb[:, ca] @ yb[:, cp] for xb, yb in zip(X, Y))
                for xb, yb in zip(X, Y):
                    xb[:, ca] -= coef * xb[:, cp]
                    yb[:, ca] -= coef * yb[:, cp]
            n2 = sum(xb[:, ca] @ yb[:, ca] for xb, yb in zip(X, Y))
            lead = sum(xb[:, base] @ yb[:, base] for xb, yb in zip(X, Y))
            if n2 <= 1e-20 * max(lead, 1e-300) or n2 <= 0:
                for blk in (*X, *Y):
                    blk[:, ca] = 0.0
                continue
            for blk in (*X, *Y):
                blk[:, ca] *= 1.0 / np.sqrt(n2)
    return X, Y


def sum_sq_eigenvalues(lay: SELayout, n_draws: int, block: int, seed: int, X, Y, k: int,
                       settled=None, min_draws: int = 16):
    """tr((S^- Q)^2) per component: the deflated directions exactly, the
    rest by Hutchinson over zeta = X~'xi (leaveout_weakid.py:467-551).
    Returns (estimates, Monte Carlo standard errors, draws)."""
    ctx = lay.ctx
    n = len(ctx.w)
    comps = [c for c in range(3) for _ in range(k)]
    ql, qg = q_block(lay, comps, Y[0], Y[1])
    zl, zg, _ = apply_inverse(ctx, ql, qg)
    exact = ((ql * zl).sum(0) + (qg * zg).sum(0)).reshape(3, k).sum(axis=1)
    draws, drawn = [], 0
    size = max(1, block // 3)
    while drawn < n_draws:
        m = min(size, n_draws - drawn)
        xi = rademacher(0, n, m, seed, drawn)
        zeta = list(se_reduce(lay, xi))
        yl, yg, _ = apply_inverse(ctx, zeta[0], zeta[1])
        y = [yl, yg]
        yd = [np.empty((b.shape[0], 3 * m)) for b in y]
        for c in range(3):
            cols = slice(c * k, (c + 1) * k)
            proj = sum(zb.T @ Yb[:, cols] for zb, Yb in zip(zeta, Y))
            out = slice(c * m, (c + 1) * m)
            for d, yb, Yb in zip(yd, y, Y):
                d[:, out] = yb - Yb[:, cols] @ proj.T
        ql, qg = q_block(lay, [c for c in range(3) for _ in range(m)], yd[0], yd[1])
        bl, bg, _ = apply_inverse(ctx, ql, qg)
        draws.append(((ql * bl).sum(0) + (qg * bg).sum(0)).reshape(3, m))
        drawn += m
        vals = np.hstack(draws)
        if settled is not None and min_draws <= drawn < n_draws and \
                settled(*_summarize(exact, vals)):
            break
    est, se = _summarize(exact, np.hstack(draws))
    return est, se, drawn


def _summarize(exact, values):
    n = values.shape[1]
    se = values.std(axis=1, ddof=1) / np.sqrt(n) if n > 1 else np.full(values.shape[0], np.nan)
    return exact + values.mean(axis=1), se


def fit_smoother_1d(sigma2, leverage, n_bins=None):
    """sigma2 smoothed on the leverage alone (leaveout_se.py:982-1035, one
    axis): (knots, means, n_bins) or None."""
    n = len(sigma2)
    nb = auto_bins(n, 1) if n_bins is None else int(n_bins)
    if nb < 1:
        return None
    kp = smoother_knots(leverage, nb, smoother_sample(n))
    cp = np.minimum((smoother_rank(kp, leverage) * nb).astype(np.int64), nb - 1)
    tot = np.bincount(cp, weights=np.asarray(sigma2, np.float64), minlength=nb)
    cnt = np.bincount(cp, minlength=nb).astype(np.float64)
    overall = float(tot.sum() / max(cnt.sum(), 1.0))
    return kp, np.where(cnt > 0, tot / np.maximum(cnt, 1.0), overall), nb


def smooth_eval_1d(table, leverage):
    kp, means, nb = table
    f = smoother_rank(kp, leverage) * nb - 0.5
    i0 = np.clip(np.floor(f).astype(np.int64), 0, nb - 2)
    a = np.clip(f - i0, 0.0, 1.0)
    return (1.0 - a) * means[i0] + a * means[i0 + 1]


def weak_id_diagnostics(lay: SELayout, y: np.ndarray, sigma2: np.ndarray, leverage,
                        n_draws: int = 64, block: int = 8, seed: int = 0, max_iter: int = 150,
                        tol: float = 1e-4, top: int = 3, smooth_bins=None) -> dict:
    """KSS's eigenvalue diagnostic (leaveout_weakid.py:553-744): Lanczos to
    the leading eigenvalues, deflated sum lambda^2, the q rule, and
    F = b1^2 / V[b1] of the leading direction."""
    ctx = lay.ctx
    caps, reorth = krylov_plan(lay, max_iter)

    def converged_all(alphas, betas, lengths, steps):
        for c in range(3):
            if lengths[c] < steps:
                continue
            theta, bounds, _ = ritz_top(alphas[c], betas[c], top)
            lead = float(np.abs(theta[0])) if len(theta) else 0.0
            if np.any(bounds[:2] > tol * max(lead, 1e-300)):
                return False
        return True

    alphas, betas, lengths, n_iter = lanczos(lay, max_iter, seed, check=converged_all,
                                             every=10, caps=caps, reorth=reorth)
    checked = [ritz_top(alphas[c], betas[c], top) for c in range(3)]
    n_coef = ctx.K + ctx.G
    k = top if 2 * 3 * top * n_coef * 8 <= SCRATCH_BYTES else 1
    ritz = []
    for _t, _b, S in checked:
        coef = np.zeros((S.shape[0], k))
        coef[:, :min(k, S.shape[1])] = S[:, :k]
        ritz.append(coef)
    X, Y = lanczos(lay, n_iter, seed, ritz=ritz, caps=caps, reorth=reorth)
    X, Y = s_orthonormalize(X, Y, k)
    u = [blk[:, ::k].copy() for blk in Y]
    leading = [t for t, _b, _S in checked]

    def settled(est, se):
        for c in range(3):
            if not est[c] > 0 or not np.isfinite(se[c]):
                return False
            r = leading[c] ** 2 / est[c]
            if np.any(np.abs(r - WEAK_ID_THRESHOLD) < 2.0 * (se[c] / est[c]) * r):
                return False
        return True

    sum_sq, sum_sq_se, used = sum_sq_eigenvalues(lay, n_draws, block, seed + WEAKID_SUM_OFFSET,
                                                 X, Y, k, settled=settled)
    table = fit_smoother_1d(sigma2, leverage, smooth_bins)
    sm = sigma2 if table is None else smooth_eval_1d(table, leverage)
    w = ctx.w
    ybar = float(w @ y / w.sum())
    yt = lay.root * (y - ybar)
    x1 = eigen_rows(lay, u[0], u[1])
    num, den = x1.T @ yt, (x1 ** 2).T @ sm
    f_stat = np.where(den > 0, num ** 2 / np.where(den > 0, den, 1.0), np.nan)
    out = {"eigenvalues": [], "bounds": [], "ratios": [], "q": [], "q_capped": [],
           "q_borderline": [], "converged": [], "sum_sq": sum_sq, "sum_sq_se": sum_sq_se,
           "f_stat": f_stat, "steps": n_iter, "lengths": lengths, "deflated": k,
           "n_draws": used, "caps": caps, "reorth": reorth, "lead": u}
    for c in range(3):
        theta, b, _ = checked[c]
        lead = float(np.abs(theta[0])) if len(theta) else 0.0
        out["eigenvalues"].append(theta)
        out["bounds"].append(b)
        out["converged"].append(bool(lengths[c] < n_iter or lengths[c] >= caps[c]
                                     or np.all(b[:2] <= tol * max(lead, 1e-300))))
        r = theta ** 2 / sum_sq[c] if sum_sq[c] > 0 else np.zeros_like(theta)
        out["ratios"].append(r)
        cnt = 0
        for v in r:
            if v >= WEAK_ID_THRESHOLD:
                cnt += 1
            else:
                break
        out["q"].append(cnt)
        out["q_capped"].append(bool(cnt == len(r) and cnt > 0))
        rel_se = sum_sq_se[c] / sum_sq[c] if sum_sq[c] > 0 else np.inf
        out["q_borderline"].append(bool(np.any(np.abs(r - WEAK_ID_THRESHOLD)
                                               < 2.0 * rel_se * r)))
    return out



# ------------------------------------------------ the q = 1 weak interval
#
# leaveout_se.py:786-889 (the deflated moments) and am_interval.py (the
# Andrews-Mikusheva projection), restated.

WEAK_MOMENTS_OFFSET = 8_800_021    # leaveout_se.py:747


def eigen_rows(lay: SELayout, Ul: np.ndarray, Ug: np.ndarray) -> np.ndarray:
    """sqrt(w_i) (A u)_i for coefficient blocks (K, c), (G, c)
    (leaveout_se.py:486-495)."""
    ctx = lay.ctx
    return lay.root[:, None] * (ctx.Do @ Ul[:ctx.L] + ctx.X @ Ul[ctx.L:] + Ug[ctx.g])


def weak_moments(lay: SELayout, y, sigma2, s, leverage, complement, factor, b_ii, lams,
                 Ul, Ug, components, n_draws: int, block: int, seed: int, stayer_rows=None,
                 centering: str = "weighted") -> dict:
    """(b1, raw V[b1], the 2x2 covariance of (b1, theta1)) per weak component:
    b1 = x1'y~ with x1 = X~ u1, and theta1 = y~'C2 y~ for C2 built from
    B2 = B - lambda_1 x1 x1'; `s` (3, n) is the smoothed sigma2, floored at
    zero here; `sigma2` the point estimate's (for V[b1] raw)."""
    ctx = lay.ctx
    n = len(ctx.w)
    x1 = eigen_rows(lay, Ul, Ug)
    w = ctx.w
    ybar = float(w @ y / w.sum())
    yt_c = lay.root * (y - ybar)
    b1 = x1.T @ yt_c
    v_raw = (x1 ** 2).T @ sigma2
    sf = np.maximum(s, 0.0)
    s11 = np.array([(x1[:, k] ** 2) @ sf[k] for k in range(3)])
    inv_m = factor / complement
    b2 = np.stack([(b_ii[k] - lams[k] * x1[:, k] ** 2) * inv_m for k in range(3)])
    if stayer_rows is not None:
        b2[:, stayer_rows] = 0.0
    rank_two = (m_times(lay, b2), 0.5 / n) if centering == "reference" else None
    yt = lay.root * (y - (ybar if centering == "weighted" else 0.0))
    deflate = (np.asarray(lams, np.float64), x1)
    cy = apply_c(lay, yt[:, None], b2, rank_two=rank_two, deflate=deflate)[:, :, 0]
    s12 = np.array([2.0 * (x1[:, k] * sf[k]) @ cy[:, k] for k in range(3)])
    first = np.array([4.0 * sf[k] @ cy[:, k] ** 2 for k in range(3)])
    quad = np.array([yt @ cy[:, k] for k in range(3)])
    total = np.zeros(3)
    drawn = 0
    while drawn < n_draws:
        size = min(block, n_draws - drawn)
        z = rademacher(0, n, size, seed, drawn)
        v = np.zeros((n, 3, size))
        for k in components:
            v[:, k, :] = z * np.sqrt(sf[k])[:, None]
        cv = apply_c(lay, v, b2, per_component=True, rank_two=rank_two, deflate=deflate)
        for k in components:
            total[k] += sf[k] @ (cv[:, k, :] ** 2).sum(axis=1)
        drawn += size
    trace = total / n_draws
    out = {}
    for k in components:
        s22 = first[k] - 2.0 * trace[k]
        conservative = not (s22 > 0 and s12[k] ** 2 < s11[k] * s22)
        if conservative:
            s22 = first[k]
        out[k] = {"lambda_1": float(lams[k]), "b1": float(b1[k]), "v_raw": float(v_raw[k]),
                  "sigma": np.array([[s11[k], s12[k]], [s12[k], s22]]),
                  "conservative": bool(conservative), "trace": float(trace[k]),
                  "quadratic": float(quad[k])}
    return out


def am_cdf(z: float, kappa: float) -> float:
    """P(sqrt(a^2 + (b + c)^2) - c <= z), c = 1/kappa, for independent
    half-normals a, b (am_interval.py:45-62; KSS Appendix C.6.1 for q = 1):
    over b in [0, z] with density 2 phi(b), a reaches
    sqrt((z - b)(z + b + 2c)) -- the circle of radius z + c about (0, -c),
    factored so nothing cancels as c grows -- with probability
    2 Phi(.) - 1."""
    from scipy import integrate, special
    if z <= 0:
        return 0.0
    if kappa <= 0:
        return float(special.erf(z / np.sqrt(2.0)))
    c = 1.0 / kappa

    def inner(b):
        top = np.sqrt(max((z - b) * (z + b + 2.0 * c), 0.0))
        return 2.0 * np.exp(-0.5 * b * b) / np.sqrt(2.0 * np.pi) * special.erf(top / np.sqrt(2.0))

    # for small kappa the erf leaves one only within ~8 kappa of b = z: a
    # breakpoint there keeps quad from stepping over that sliver
    edge = z - 20.0 * kappa
    pts = [edge] if 0.0 < edge < z else None
    val, _ = integrate.quad(inner, 0.0, z, epsabs=1e-14, epsrel=1e-13, limit=200, points=pts)
    return float(val)


def am_critical_value(kappa: float, alpha: float = 0.05) -> float:
    """The (1 - alpha) quantile of the above (am_interval.py:65-79), by
    bisection between the chi-squared(1) and chi-squared(2) quantiles'
    roots (its limits)."""
    from scipy import special
    lo = float(np.sqrt(special.chdtri(1, alpha)))
    if kappa <= 0:
        return lo
    hi = float(np.sqrt(special.chdtri(2, alpha)))
    a, b = lo * (1 - 1e-9), hi * (1 + 1e-9)
    for _ in range(60):
        mid = 0.5 * (a + b)
        if am_cdf(mid, kappa) < 1.0 - alpha:
            a = mid
        else:
            b = mid
    return 0.5 * (a + b)


def am_interval(lam1: float, b1: float, theta1: float, sigma, alpha: float = 0.05):
    """The q = 1 interval (am_interval.py:90-123): the extremes of lambda_1 b^2 + t on the boundary
    of the ellipse of (b, t) at the critical value, from the first-order
    condition in closed form (KSS Appendix C.6.2): with the boundary
    (b1, theta1) + z L (cos p, sin p), L = chol(sigma), and s = tan(p / 2),
    the stationary points are the real roots of a quartic in s."""
    sigma = np.asarray(sigma, np.float64)
    vb, vt, cv = sigma[0, 0], sigma[1, 1], sigma[0, 1]
    kappa = 2.0 * abs(lam1) * vb / np.sqrt(vt * (1.0 - cv * cv / (vb * vt)))
    z = am_critical_value(kappa, alpha)
    Lc = np.linalg.cholesky(sigma)
    A, B, Cc = z * Lc[0, 0], z * Lc[1, 0], z * Lc[1, 1]

    def f(p):
        return lam1 * (b1 + A * np.cos(p)) ** 2 + theta1 + B * np.cos(p) + Cc * np.sin(p)

    # f'(p) (1 + s^2)^2 = -Cc s^4 - (4 lam A (b1 - A) + 2 B) s^3
    #                     - (4 lam A (b1 + A) + 2 B) s + Cc
    coef = [-Cc, -(4 * lam1 * A * (b1 - A) + 2 * B), 0.0, -(4 * lam1 * A * (b1 + A) + 2 * B), Cc]
    roots = np.roots(coef) if np.any(np.abs(coef) > 0) else np.array([])
    cand = [np.pi] + [2.0 * np.arctan(r.real) for r in roots if abs(r.imag) <= 1e-9 * (1 + abs(r))]
    vals = np.array([f(p) for p in cand])
    return float(vals.min()), float(vals.max()), {"kappa": float(kappa), "z": float(z),
                                                  "rho": float(cv / np.sqrt(vb * vt))}
