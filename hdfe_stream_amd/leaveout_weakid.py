"""KSS's weak-identification diagnostic on the device.

What is estimated (the reference's `weak_id_diagnostics`,
leaveout_weakid.py:553-744): for each variance component (var(psi),
var(alpha), cov(psi, alpha)) the leading eigenvalues lambda of S^- Q, their
shares lambda^2 / sum lambda^2, KSS's count q of leading shares of at least
1/10 (the normal interval of the standard errors needs q = 0), and the
first-stage F of the leading direction.

How it is computed here, with the Krylov work resident in HBM:

* Lanczos on T = Q S^-, which is self-adjoint in <x, y> = x'S^- y.  The three
  components advance in lockstep, so every step costs one three-column block
  solve.  A vector x is kept together with its image y = S^- x.
* Full re-orthogonalisation at every step.  Each component's whole Krylov
  basis (x_j and y_j on the component's support: the psi levels, the groups,
  or both) stays on the device -- a few GB at a million groups and 150 steps,
  small against HBM -- and the new vector is orthogonalised against all of it
  by classical Gram-Schmidt applied twice (one tall product and one GEMV per
  pass; two passes restore orthogonality to rounding: Giraud, Langou, Rozloznik
  and van den Eshof, Numer. Math. 101 (2005)).  The tridiagonal is read off
  the projection coefficients.  With the basis orthonormal to rounding, the
  tridiagonal has no spurious copies of converged eigenvalues, so its
  eigenvalues need no ghost filtering.
* Ritz values from the FP64 tridiagonal on the host (scipy), with the standard
  residual bound |beta_m s_mi|; Ritz vectors as one product of the resident
  basis with the tridiagonal's eigenvectors, their S^- images by one solve of
  3k columns, and the k vectors of a component made S^- orthonormal by
  Cholesky QR on their k x k Gram matrix (columns that add nothing dropped).
* sum lambda^2 = tr((S^- Q)^2): the deflated directions exactly, the rest by
  Hutchinson over zeta = X~'xi with an early stop once q is settled.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .leaveout_se import COMPONENTS, SELayout, apply_q, eigen_rows, smoothed_sigma2, tall_dot

WEAK_ID_THRESHOLD = 0.1            # KSS's share threshold (leaveout_weakid.py:67)
SEED_OFFSET = 9_900_019            # the diagnostic's seed offset (leaveout.py:758)
SUM_SEED_OFFSET = 3_300_007        # the sum-of-squares draws' offset (leaveout_weakid.py:651)
SCRATCH_BYTES = 32 << 20           # the reference's scratch_mb (core/estimator.py:77): it
#                                    decides how many directions are deflated, so the
#                                    estimate's draws combine as the reference's do
BASIS_MEMORY_SHARE = 0.25          # at most this share of free device memory for the bases
BREAKDOWN = 1e-8                   # a component stops once beta falls below this share of
#                                    the largest beta it has seen (an invariant subspace)


@dataclass
class WeakIdDiagnostics:
    """The diagnostic for the three components (the fields of
    leaveout_weakid.py:70-152): dicts keyed by component name."""
    eigenvalues: dict
    sum_sq: dict
    sum_sq_se: dict
    ratios: dict
    q: dict
    q_capped: dict
    q_borderline: dict
    f_stat: dict
    bounds: dict
    converged: dict
    threshold: float = WEAK_ID_THRESHOLD
    diagnostics: dict = field(default_factory=dict)
    lead_vectors: tuple = field(default=None, repr=False)

    @property
    def weakly_identified(self) -> list:
        return [name for name in COMPONENTS if self.q[name] >= 1]

    def summary(self) -> str:
        out = [f"KSS weak-identification diagnostic: q counts the leading eigenvalues of "
               f"S^-Q with lambda^2 / sum lambda^2 >= {self.threshold:g}"]
        for name in COMPONENTS:
            shares = ", ".join(f"{r:.4f}" for r in self.ratios[name][:3]) or "-"
            flags = ("+" if self.q_capped[name] else "") + \
                    (" (borderline)" if self.q_borderline[name] else "")
            verdict = ("normal interval justified" if self.q[name] == 0
                       else "normal interval not justified")
            tail = "" if self.converged[name] else "; Lanczos not converged"
            out.append(f"  {name}: q = {self.q[name]}{flags}; shares {shares}; "
                       f"F = {self.f_stat[name]:.2f}; {verdict}{tail}")
        return "\n".join(out)


# ---------------------------------------------------------------- blocks

def _support_rows(lay, c):
    """Rows of component c's support: (level rows, whether the groups)."""
    return (lay.psi if c in (0, 2) else None), c in (1, 2)


def _gather(lay, c, blocks, col):
    """Column `col` of the coefficient blocks on component c's support."""
    import torch
    lv, grp = _support_rows(lay, c)
    parts = ([blocks[0][lv, col]] if lv is not None else []) + \
            ([blocks[1][:, col]] if grp else [])
    return torch.cat(parts).contiguous()


def _put(lay, c, blocks, col, vec):
    """Write a support vector back into column `col` of the blocks."""
    lv, grp = _support_rows(lay, c)
    at = 0
    if lv is not None:
        n = lv.stop - lv.start
        blocks[0][lv, col] = vec[:n]
        at = n
    if grp:
        blocks[1][:, col] = vec[at:]


def _apply_q(lay, comps, levels, groups):
    """Q applied to each column with the component comps[j]."""
    import torch
    comps = np.asarray(comps)
    out_l = torch.zeros((lay.K, levels.shape[1]), dtype=torch.float64, device="cuda")
    out_g = torch.zeros_like(groups)
    for c in np.unique(comps):
        idx = torch.from_numpy(np.flatnonzero(comps == c)).cuda()
        out_l[:, idx], out_g[:, idx] = apply_q(lay, int(c), levels[:, idx], groups[:, idx])
    return out_l, out_g


def _to_range(lay, blocks):
    """Onto range(Q): Q's outputs sum to zero over the psi levels and over the
    groups of the components living there."""
    for c in (0, 2):
        blocks[0][lay.psi, c] -= blocks[0][lay.psi, c].mean()
    for c in (1, 2):
        blocks[1][:, c] -= blocks[1][:, c].mean()


def _inner(a, b):
    """x'y per column over both blocks (host)."""
    return ((a[0] * b[0]).sum(0) + (a[1] * b[1]).sum(0)).cpu().numpy()


def _inverse(lay, blocks):
    levels, groups, _ = lay.inv(blocks[0], blocks[1], want_rows=False)
    return [levels, groups]


# ---------------------------------------------------------------- Lanczos

class _Basis:
    """One component's Krylov basis on its support, resident in HBM: x_j in
    the columns of X, y_j = S^- x_j in the columns of Y, (rows, cap)."""

    def __init__(self, rows, cap):
        import torch
        self.X = torch.zeros((rows, max(cap, 1)), dtype=torch.float64, device="cuda")
        self.Y = torch.zeros_like(self.X)
        self.n = 0

    def append(self, x, y):
        self.X[:, self.n] = x
        self.Y[:, self.n] = y
        self.n += 1

    def orthogonalize(self, w):
        """w minus its S^- projection on the basis, twice (CGS2); returns the
        vector and its total coefficient on the newest basis vector."""
        lead = 0.0
        for _ in range(2):
            h = tall_dot(self.Y[:, :self.n], w)[:, 0]
            w = w - self.X[:, :self.n] @ h
            lead += float(h[self.n - 1])
        return w, lead


def _krylov_caps(lay, max_iter):
    """Steps per component: at most max_iter, and at most the dimension of
    the Krylov space T can span from range(Q) (leaveout_weakid.py:186-215:
    n_psi - 1 for var(psi), K + 1 for var(alpha), 2 (n_psi - 1) for the
    covariance); then fewer if the bases would take more than a quarter of
    the free device memory."""
    n_psi = lay.psi.stop - lay.psi.start
    dims = [max(n_psi - 1, 0), max(lay.K + 1, 1), max(2 * (n_psi - 1), 0)]
    caps = [min(max_iter, d) for d in dims]
    rows = [n_psi, lay.G, n_psi + lay.G]
    free, _total = _lib.device_memory()
    budget = BASIS_MEMORY_SHARE * free
    need = sum(2 * 8 * r * c for r, c in zip(rows, caps))
    if need > budget:
        scale = budget / need
        caps = [max(1, int(c * scale)) for c in caps]
    return caps, rows


def _ritz(alpha, beta, top):
    """The `top` Ritz values of largest magnitude of the Lanczos tridiagonal
    (alpha, beta), their residual bounds |beta_m s_mi| and eigenvectors."""
    from scipy.linalg import eigh_tridiagonal
    m = len(alpha)
    if m == 0:
        return np.zeros(0), np.zeros(0), np.zeros((0, 0))
    theta, vec = eigh_tridiagonal(np.asarray(alpha, float), np.asarray(beta[:m - 1], float))
    order = np.argsort(-np.abs(theta), kind="stable")[:top]
    return theta[order], np.abs(beta[m - 1] * vec[m - 1, order]), vec[:, order]


def _lanczos(lay, max_steps, seed, caps, rows, tol, top, every=10):
    """The lockstep recurrence; returns (alphas, betas, lengths, steps,
    bases).  Every `every` steps it stops once each component still running
    has its two leading Ritz values within tol of the largest."""
    import torch
    lib = _lib.load()
    L, G = lay.L, lay.G
    bases = [_Basis(rows[c], caps[c]) for c in range(3)]
    r_l = torch.empty((L, 3), dtype=torch.float64, device="cuda")
    r_g = torch.empty((G, 3), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    # the start: Q r for r Rademacher over the levels, then the groups
    _lib.check(lib.hdfe_rademacher(0, L, 3, seed, 0, r_l.data_ptr()), "hdfe_rademacher")
    _lib.check(lib.hdfe_rademacher(L, G, 3, seed, 0, r_g.data_ptr()), "hdfe_rademacher")
    x = list(_apply_q(lay, [0, 1, 2], r_l, r_g))
    del r_l, r_g
    _to_range(lay, x)
    y = _inverse(lay, x)
    size = np.sqrt(np.maximum(_inner(x, y), 0.0))
    running = (size > 0) & (np.asarray(caps) > 0)

    def rescale(blocks, by):
        f = torch.from_numpy(np.where(running & (by > 0), 1.0 / np.where(by > 0, by, 1.0),
                                      0.0)).cuda()
        for b in blocks:
            b.mul_(f[None, :])

    rescale(x, size)
    rescale(y, size)
    alphas, betas = [[], [], []], [[], [], []]
    lengths = np.zeros(3, dtype=int)
    beta_top = np.zeros(3)
    steps = 0

    def settled(j):
        for c in range(3):
            if lengths[c] < j:
                continue            # stopped earlier: its space is exhausted
            theta, bound, _ = _ritz(alphas[c], betas[c], top)
            lead = float(np.abs(theta[0])) if len(theta) else 0.0
            if np.any(bound[:2] > tol * max(lead, 1e-300)):
                return False
        return True

    for j in range(max_steps):
        if not running.any() or (j and j % every == 0 and settled(j)):
            break
        steps = j + 1
        live = np.flatnonzero(running)
        for c in live:
            bases[c].append(_gather(lay, c, x, c), _gather(lay, c, y, c))
        w = list(_apply_q(lay, [0, 1, 2], y[0], y[1]))
        alpha = np.zeros(3)
        for c in live:
            wc, alpha[c] = bases[c].orthogonalize(_gather(lay, c, w, c))
            _put(lay, c, w, c, wc)
        _to_range(lay, w)
        z = _inverse(lay, w)
        beta = np.sqrt(np.maximum(_inner(w, z), 0.0))
        for c in live:
            alphas[c].append(float(alpha[c]))
            betas[c].append(float(beta[c]))
            lengths[c] = j + 1
        beta_top = np.maximum(beta_top, beta)
        running &= (beta > BREAKDOWN * np.maximum(beta_top, 1e-300)) & (lengths < caps)
        x, y = w, z
        rescale(x, beta)
        rescale(y, beta)
    return [np.asarray(a) for a in alphas], [np.asarray(b) for b in betas], lengths, steps, bases


def _ritz_vectors(lay, bases, vecs, k):
    """Each component's k leading Ritz vectors U (coefficient blocks, 3k
    columns, component-major), their S^- images Y by one solve, made S^-
    orthonormal by Cholesky QR on the k x k Gram matrix U'S^-U = U'Y."""
    import torch
    Ul = torch.zeros((lay.K, 3 * k), dtype=torch.float64, device="cuda")
    Ug = torch.zeros((lay.G, 3 * k), dtype=torch.float64, device="cuda")
    for c in range(3):
        B, S = bases[c], vecs[c]
        for i in range(min(k, S.shape[1])):
            coef = torch.from_numpy(np.ascontiguousarray(S[:B.n, i])).cuda()
            _put(lay, c, (Ul, Ug), c * k + i, B.X[:, :B.n] @ coef)
    Yl, Yg = _inverse(lay, [Ul, Ug])
    for c in range(3):
        cols = slice(c * k, (c + 1) * k)
        gram = (tall_dot(Ul[:, cols], Yl[:, cols]) + tall_dot(Ug[:, cols], Yg[:, cols]))
        T = torch.from_numpy(_cholesky_qr(gram.cpu().numpy())).cuda()
        for blk in (Ul, Ug, Yl, Yg):
            blk[:, cols] = blk[:, cols] @ T
    return Ul, Ug, Yl, Yg


def _cholesky_qr(gram, drop=1e-20):
    """T with T'GT = I on the columns kept: the upper-triangular inverse of
    G's Cholesky factor, column by column, so the first column is only
    scaled; a column whose remaining square norm is at most `drop` of the
    first's is dropped (a zero column of T)."""
    k = gram.shape[0]
    R = np.zeros((k, k))
    keep = np.zeros(k, bool)
    first = max(float(gram[0, 0]), 1e-300) if k else 1.0
    for a in range(k):
        r = gram[:a, a].copy()
        for p in range(a):
            if keep[p]:
                r[p] = (gram[p, a] - R[:p, p] @ r[:p]) / R[p, p]
            else:
                r[p] = 0.0
        R[:a, a] = r
        d = float(gram[a, a] - r @ r)
        if d > drop * first and d > 0:
            R[a, a] = np.sqrt(d)
            keep[a] = True
    T = np.zeros((k, k))
    # solve R T = I on the kept columns (back substitution)
    for a in range(k):
        if not keep[a]:
            continue
        e = np.zeros(k)
        e[a] = 1.0
        t = np.zeros(k)
        for p in range(a, -1, -1):
            if keep[p]:
                t[p] = (e[p] - R[p, p + 1:a + 1] @ t[p + 1:a + 1]) / R[p, p]
        T[:, a] = t
    return T


def _sum_squares(lay, n_draws, block, seed, U, Y, k, settled=None, min_draws=16):
    """tr((S^- Q)^2) per component: over the k deflated directions exactly
    (sum of y'Q S^- Q y), over the complement by Hutchinson: for draws xi,
    zeta = X~'xi, v = S^- zeta minus its part along the deflated images,
    v'Q S^- Q v.  Draws come in blocks of block // 3; after each block the
    estimate stops early once `settled(mean, se)` (at least min_draws)."""
    import torch
    lib = _lib.load()
    comps_k = np.repeat(np.arange(3), k)
    ql, qg = _apply_q(lay, comps_k, Y[0], Y[1])
    sl, sg = _inverse(lay, [ql, qg])
    exact = _inner((ql, qg), (sl, sg)).reshape(3, k).sum(axis=1)
    per_draw = []
    done = 0
    width = max(1, block // 3)
    while done < n_draws:
        m = min(width, n_draws - done)
        xi = torch.empty((lay.n, m), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        _lib.check(lib.hdfe_rademacher(0, lay.n, m, seed, done, xi.data_ptr()),
                   "hdfe_rademacher")
        zeta = list(lay.reduce(xi))
        del xi
        v = _inverse(lay, zeta)
        # coefficients of zeta on every deflated image: (m, 3k)
        coef = tall_dot(zeta[0], Y[0]) + tall_dot(zeta[1], Y[1])
        vd = [torch.empty((b.shape[0], 3 * m), dtype=torch.float64, device="cuda") for b in v]
        for c in range(3):
            cols = slice(c * k, (c + 1) * k)
            for out, vb, Yb in zip(vd, v, Y):
                out[:, c * m:(c + 1) * m] = vb - Yb[:, cols] @ coef[:, cols].T
        ql, qg = _apply_q(lay, np.repeat(np.arange(3), m), vd[0], vd[1])
        sl, sg = _inverse(lay, [ql, qg])
        per_draw.append(_inner((ql, qg), (sl, sg)).reshape(3, m))
        done += m
        if settled is not None and min_draws <= done < n_draws:
            mean, se = _mean_se(exact, np.hstack(per_draw))
            if settled(mean, se):
                break
    mean, se = _mean_se(exact, np.hstack(per_draw))
    return mean, se, done


def _mean_se(exact, draws):
    n = draws.shape[1]
    se = (draws.std(axis=1, ddof=1) / np.sqrt(n)) if n > 1 else np.full(3, np.nan)
    return exact + draws.mean(axis=1), se


def _near_threshold(shares, rel_se):
    """Shares within two standard errors of the threshold."""
    return np.abs(shares - WEAK_ID_THRESHOLD) < 2.0 * rel_se * shares


def weak_id_diagnostics(inv, y, sigma2, leverages, *, psi: int = 0, rows: int | None = None,
                        n_draws: int = 64, block: int = 8, seed: int = 0, max_iter: int = 150,
                        tol: float = 1e-4, top: int = 3, smooth_bins=None) -> WeakIdDiagnostics:
    """KSS's eigenvalue diagnostic (the reference's weak_id_diagnostics,
    leaveout_weakid.py:553-744): the leading eigenvalues of each component's
    S^- Q by Lanczos, sum lambda^2 by deflated Hutchinson (`n_draws` at
    most, stopping once q is settled), q, and the first-stage F of the
    leading direction with sigma2 smoothed on the leverage.  `inv` is the
    panel's CoefficientInverse, `y` the outcome on the sorted rows, `seed`
    the diagnostic's own (the caller's seed + SEED_OFFSET)."""
    import torch
    if n_draws < 2:
        raise ValueError("n_draws must be at least 2, so the Monte Carlo error of "
                         "sum lambda^2 can be reported")
    lay = SELayout(inv, psi, inv.n if rows is None else rows)
    caps, sup_rows = _krylov_caps(lay, max_iter)
    alphas, betas, lengths, steps, bases = _lanczos(lay, max_iter, seed, caps, sup_rows, tol, top)
    ritz = [_ritz(alphas[c], betas[c], top) for c in range(3)]
    k = top if 2 * 3 * top * (lay.K + lay.G) * 8 <= SCRATCH_BYTES else 1
    Ul, Ug, Yl, Yg = _ritz_vectors(lay, bases, [r[2] for r in ritz], k)
    del bases
    lead_vectors = (Yl[:, ::k].contiguous(), Yg[:, ::k].contiguous())
    leading = [r[0] for r in ritz]

    def settled(mean, se):
        for c in range(3):
            if not (mean[c] > 0 and np.isfinite(se[c])):
                return False
            if np.any(_near_threshold(leading[c] ** 2 / mean[c], se[c] / mean[c])):
                return False
        return True

    sum_sq, sum_sq_se, used = _sum_squares(lay, n_draws, block, seed + SUM_SEED_OFFSET,
                                           (Ul, Ug), (Yl, Yg), k, settled=settled)
    # F = b1^2 / V[b1], b1 = x1'y~ and V[b1] = sum x1^2 sigma2~, x1 = X~ u1
    smooth = smoothed_sigma2(sigma2, leverages.leverage, None, smooth_bins)
    x1 = eigen_rows(lay, *lead_vectors)
    y = y.to(torch.float64)
    centred = lay.root * (y - float((lay.w * y).sum() / lay.w.sum()))
    b1 = (x1 * centred[:, None]).sum(0).cpu().numpy()
    vb1 = (x1 * x1 * smooth[:, None]).sum(0).cpu().numpy()
    res = {key: {} for key in ("eig", "ratios", "q", "capped", "bounds", "conv", "border", "f")}
    for c, name in enumerate(COMPONENTS):
        theta, bound, _ = ritz[c]
        lead = float(np.abs(theta[0])) if len(theta) else 0.0
        shares = theta ** 2 / sum_sq[c] if sum_sq[c] > 0 else np.zeros_like(theta)
        above = shares >= WEAK_ID_THRESHOLD
        q = int(np.argmin(above)) if not above.all() else len(above)
        rel_se = sum_sq_se[c] / sum_sq[c] if sum_sq[c] > 0 else np.inf
        res["eig"][name] = theta.tolist()
        res["bounds"][name] = bound.tolist()
        res["ratios"][name] = shares.tolist()
        res["q"][name] = q
        res["capped"][name] = bool(q > 0 and q == len(shares))
        res["border"][name] = bool(np.any(_near_threshold(shares, rel_se)))
        res["conv"][name] = bool(lengths[c] < steps or lengths[c] >= caps[c]
                                 or np.all(bound[:2] <= tol * max(lead, 1e-300)))
        res["f"][name] = float(b1[c] ** 2 / vb1[c]) if vb1[c] > 0 else float("nan")
    return WeakIdDiagnostics(
        eigenvalues=res["eig"], sum_sq=dict(zip(COMPONENTS, sum_sq.tolist())),
        sum_sq_se=dict(zip(COMPONENTS, sum_sq_se.tolist())), ratios=res["ratios"], q=res["q"],
        q_capped=res["capped"], q_borderline=res["border"], f_stat=res["f"],
        bounds=res["bounds"], converged=res["conv"], lead_vectors=lead_vectors,
        diagnostics={"lanczos_steps": int(steps), "deflated": k,
                     "lanczos_lengths": lengths.tolist(), "krylov_caps": caps,
                     "reorthogonalization": "full, CGS2 against the resident basis",
                     "n_draws": used, "max_draws": n_draws})
