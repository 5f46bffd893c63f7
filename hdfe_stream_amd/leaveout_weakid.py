"""KSS's weak-identification diagnostic on the device (leaveout_weakid.py of
hdfe_stream).

The normal interval of the standard errors rests on lambda_1^2 / sum lambda^2
being small for the eigenvalues of S^- Q (KSS Theorem 2 (ii)); KSS's rule
counts q, the leading eigenvalues with a ratio of at least 1/10.  Mirrored:

  the recurrence        Lanczos on Q S^-, self-adjoint in <x, y> = x'S^- y,
                        three components in lockstep: one solve of three
                        columns per step, local and (small Krylov spaces) full
                        reorthogonalisation, projection onto range(Q)
                                                 leaveout_weakid.py:186-383
  the Ritz values       ghosts merged, Cullum-Willoughby  :385-433
  the Ritz vectors      by retracing the recurrence, S-orthonormalised
                                                 :436-465, 653-668
  sum lambda^2          tr((S^- Q)^2), deflated Hutchinson with an early stop
                                                 :467-551
  F                     b1^2 / V[b1] of the leading direction  :715-744

Solves are the panel's block PCG (`CoefficientInverse`), Q the coefficient-
space kernels of `leaveout_se`; the tridiagonals are FP64 eigenproblems on
the host (scipy), a few hundred entries."""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .leaveout_se import COMPONENTS, SELayout, apply_q, eigen_rows, smoothed_sigma2, tall_dot

WEAK_ID_THRESHOLD = 0.1            # leaveout_weakid.py:67
SEED_OFFSET = 9_900_019            # leaveout.py:758
SUM_SEED_OFFSET = 3_300_007        # leaveout_weakid.py:651
SCRATCH_BYTES = 32 << 20           # core/estimator.py:77 (scratch_mb)


@dataclass
class WeakIdDiagnostics:
    """KSS's weak-identification diagnostic for the three components
    (leaveout_weakid.py:70-152): dicts keyed by component."""
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
        lines = [f"weak-identification diagnostic (KSS threshold "
                 f"{self.threshold:g} on lambda^2 / sum lambda^2)",
                 f"{'component':<18}{'lam1^2/sum':>11}{'lam2^2/sum':>11}"
                 f"{'lam3^2/sum':>11}{'q':>5}{'F':>9}  normal interval",
                 "-" * 82]
        for name in COMPONENTS:
            r = list(self.ratios[name]) + [float("nan")] * 3
            q = (f"{'~' if self.q_borderline[name] else ''}{self.q[name]}"
                 f"{'+' if self.q_capped[name] else ''}")
            verdict = "justified" if self.q[name] == 0 else "NOT justified"
            if not self.converged[name]:
                verdict += " (unconverged)"
            lines.append(f"{name:<18}{r[0]:>11.4f}{r[1]:>11.4f}{r[2]:>11.4f}"
                         f"{q:>5}{self.f_stat[name]:>9.2f}  {verdict}")
        return "\n".join(lines)


def _q_block(lay, comps, Ul, Ug):
    """Q applied column by column, column j with component comps[j]."""
    import torch
    comps = np.asarray(comps)
    ol = torch.zeros((lay.K, Ul.shape[1]), dtype=torch.float64, device="cuda")
    og = torch.zeros_like(Ug)
    for k in range(3):
        cols = np.flatnonzero(comps == k)
        if len(cols):
            ci = torch.from_numpy(cols).cuda()
            a, b = apply_q(lay, k, Ul[:, ci], Ug[:, ci])
            ol[:, ci], og[:, ci] = a, b
    return ol, og


def _dot(a, b):
    """Column-wise inner products over the coefficient blocks, to the host."""
    return sum((x * y).sum(0) for x, y in zip(a, b)).cpu().numpy()


def _solve(lay, blocks):
    U, Ug, _ = lay.inv(blocks[0], blocks[1], want_rows=False)
    return [U, Ug]


def _krylov_plan(lay, max_iter):
    """Steps per component and whether to keep its basis
    (leaveout_weakid.py:186-215)."""
    n_psi = lay.psi.stop - lay.psi.start
    dims = [max(n_psi - 1, 0), max(lay.K + 1, 1), max(2 * (n_psi - 1), 0)]
    support = [n_psi, lay.G, n_psi + lay.G]
    caps = [min(max_iter, d) for d in dims]
    reorth = [cap * 2 * rows * 8 <= SCRATCH_BYTES for cap, rows in zip(caps, support)]
    return caps, reorth


def _lanczos(lay, n_iter, seed, ritz=None, check=None, every=20, caps=None, reorth=None):
    """The recurrence for the three components in lockstep
    (leaveout_weakid.py:217-383); blocks (levels (L, 3), groups (G, 3))."""
    import torch
    lib = _lib.load()
    L, G = lay.L, lay.G
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
        return torch.cat(parts) if len(parts) > 1 else parts[0].clone()

    def project(blocks):
        for c in (0, 2):
            blocks[0][psi, c] -= blocks[0][psi, c].mean()
        for c in (1, 2):
            blocks[1][:, c] -= blocks[1][:, c].mean()

    def scatter(c, flat, blocks):
        at = 0
        if c in (0, 2):
            m = psi.stop - psi.start
            blocks[0][psi, c] -= flat[at:at + m]
            at += m
        if c in (1, 2):
            blocks[1][:, c] -= flat[at:]

    def scale(blocks, factor, live):
        f = torch.from_numpy(np.where(live, factor, 0.0)).cuda()
        for blk in blocks:
            blk *= f[None, :]

    rl = torch.empty((L, 3), dtype=torch.float64, device="cuda")
    rg = torch.empty((G, 3), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    _lib.check(lib.hdfe_rademacher(0, L, 3, seed, 0, rl.data_ptr()), "hdfe_rademacher")
    _lib.check(lib.hdfe_rademacher(L, G, 3, seed, 0, rg.data_ptr()), "hdfe_rademacher")
    x = list(_q_block(lay, comps, rl, rg))
    project(x)
    y = _solve(lay, x)
    norm = np.sqrt(np.maximum(_dot(x, y), 0.0))
    live = (norm > 0) & (np.array(caps) > 0)
    inv_norm = np.where(norm > 0, 1.0 / np.where(norm > 0, norm, 1.0), 0.0)
    scale(x, inv_norm, live)
    scale(y, inv_norm, live)
    x_prev = [torch.zeros((lay.K, 3), dtype=torch.float64, device="cuda"),
              torch.zeros((G, 3), dtype=torch.float64, device="cuda")]
    beta_prev = np.zeros(3)
    basis = [([], []) for _ in comps]
    alphas, betas = [[] for _ in comps], [[] for _ in comps]
    lengths = np.zeros(3, dtype=int)
    acc = None
    if ritz is not None:
        k = ritz[0].shape[1]
        acc = ([torch.zeros((lay.K, 3 * k), dtype=torch.float64, device="cuda"),
                torch.zeros((G, 3 * k), dtype=torch.float64, device="cuda")],
               [torch.zeros((lay.K, 3 * k), dtype=torch.float64, device="cuda"),
                torch.zeros((G, 3 * k), dtype=torch.float64, device="cuda")])
        ritz_d = [torch.from_numpy(np.ascontiguousarray(r)).cuda() for r in ritz]
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
                if j < ritz[c].shape[0]:
                    cols = slice(c * k, (c + 1) * k)
                    for ax, ay, xb, yb in zip(acc[0], acc[1], x, y):
                        ax[:, cols] += torch.outer(xb[:, c], ritz_d[c][j])
                        ay[:, cols] += torch.outer(yb[:, c], ritz_d[c][j])
        for c in np.flatnonzero(live):
            if reorth[c]:
                basis[c][0].append(support(c, x))
                basis[c][1].append(support(c, y))
        w = list(_q_block(lay, comps, y[0], y[1]))
        bp = torch.from_numpy(beta_prev).cuda()
        for wb, xp in zip(w, x_prev):
            wb -= bp[None, :] * xp
        alpha = _dot(w, y)
        ad = torch.from_numpy(alpha).cuda()
        for wb, xb in zip(w, x):
            wb -= ad[None, :] * xb
        fix = _dot(w, y)
        fd = torch.from_numpy(fix).cuda()
        for wb, xb in zip(w, x):
            wb -= fd[None, :] * xb
        alpha = alpha + fix
        project(w)
        for c in np.flatnonzero(live):
            if not reorth[c] or not basis[c][0]:
                continue
            Xs, Ys = torch.stack(basis[c][0], 1), torch.stack(basis[c][1], 1)
            for _ in range(2):
                wc = support(c, w)
                scatter(c, Xs @ tall_dot(Ys, wc)[:, 0], w)
        project(w)
        z = _solve(lay, w)
        beta = np.sqrt(np.maximum(_dot(w, z), 0.0))
        beta_max = np.maximum(beta_max, beta)
        for c in np.flatnonzero(live):
            alphas[c].append(float(alpha[c]))
            betas[c].append(float(beta[c]))
            lengths[c] = j + 1
        live = live & (beta > 1e-8 * np.maximum(beta_max, 1e-300)) & (lengths < np.array(caps))
        x_prev = x
        x = [wb.clone() for wb in w]
        y = z
        inv_beta = np.where(beta > 0, 1.0 / np.where(beta > 0, beta, 1.0), 0.0)
        scale(x, inv_beta, live)
        scale(y, inv_beta, live)
        beta_prev = np.where(live, beta, 0.0)
    if acc is not None:
        return acc
    return [np.array(a) for a in alphas], [np.array(b) for b in betas], lengths, steps


def _ritz_top(alpha, beta, top):
    """Leading genuine Ritz values, error bounds and eigenvectors of the
    tridiagonal, FP64 on the host (leaveout_weakid.py:385-433)."""
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


def _s_orthonormalize(X, Y, k):
    """Modified Gram-Schmidt of each component's k Ritz vectors in
    <x, y> = x'S^- y (leaveout_weakid.py:436-465)."""
    for c in range(3):
        base = c * k
        for a in range(k):
            ca = base + a
            for p in range(a):
                cp = base + p
                coef = float(sum((xb[:, ca] * yb[:, cp]).sum() for xb, yb in zip(X, Y)))
                for xb, yb in zip(X, Y):
                    xb[:, ca] -= coef * xb[:, cp]
                    yb[:, ca] -= coef * yb[:, cp]
            n2 = float(sum((xb[:, ca] * yb[:, ca]).sum() for xb, yb in zip(X, Y)))
            lead = float(sum((xb[:, base] * yb[:, base]).sum() for xb, yb in zip(X, Y)))
            if n2 <= 1e-20 * max(lead, 1e-300) or n2 <= 0:
                for blk in (*X, *Y):
                    blk[:, ca] = 0.0
                continue
            for blk in (*X, *Y):
                blk[:, ca] *= 1.0 / np.sqrt(n2)
    return X, Y


def _summarize(exact, values):
    n = values.shape[1]
    se = values.std(axis=1, ddof=1) / np.sqrt(n) if n > 1 else np.full(values.shape[0], np.nan)
    return exact + values.mean(axis=1), se


def _sum_sq_eigenvalues(lay, n_draws, block, seed, X, Y, k, settled=None, min_draws=16):
    """tr((S^- Q)^2) per component: the k deflated directions exactly, the
    rest by Hutchinson over zeta = X~'xi (leaveout_weakid.py:467-551)."""
    import torch
    lib = _lib.load()
    n = lay.n
    comps = [c for c in range(3) for _ in range(k)]
    ql, qg = _q_block(lay, comps, Y[0], Y[1])
    zl, zg = _solve(lay, [ql, qg])
    exact = ((ql * zl).sum(0) + (qg * zg).sum(0)).cpu().numpy().reshape(3, k).sum(axis=1)
    draws, drawn = [], 0
    size = max(1, block // 3)
    while drawn < n_draws:
        m = min(size, n_draws - drawn)
        xi = torch.empty((n, m), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        _lib.check(lib.hdfe_rademacher(0, n, m, seed, drawn, xi.data_ptr()), "hdfe_rademacher")
        zeta = list(lay.reduce(xi))
        del xi
        y = _solve(lay, zeta)
        yd = [torch.empty((b.shape[0], 3 * m), dtype=torch.float64, device="cuda") for b in y]
        for c in range(3):
            cols = slice(c * k, (c + 1) * k)
            proj = sum(tall_dot(zb, Yb[:, cols]) for zb, Yb in zip(zeta, Y))
            out = slice(c * m, (c + 1) * m)
            for d, yb, Yb in zip(yd, y, Y):
                d[:, out] = yb - Yb[:, cols] @ proj.T
        ql, qg = _q_block(lay, [c for c in range(3) for _ in range(m)], yd[0], yd[1])
        bl, bg = _solve(lay, [ql, qg])
        draws.append(((ql * bl).sum(0) + (qg * bg).sum(0)).cpu().numpy().reshape(3, m))
        drawn += m
        vals = np.hstack(draws)
        if settled is not None and min_draws <= drawn < n_draws and \
                settled(*_summarize(exact, vals)):
            break
    est, se = _summarize(exact, np.hstack(draws))
    return est, se, drawn


def weak_id_diagnostics(inv, y, sigma2, leverages, *, psi: int = 0, rows: int | None = None,
                        n_draws: int = 64, block: int = 8, seed: int = 0, max_iter: int = 150,
                        tol: float = 1e-4, top: int = 3, smooth_bins=None) -> WeakIdDiagnostics:
    """KSS's eigenvalue diagnostic (leaveout_weakid.py:553-713): Lanczos to the
    leading eigenvalues of each component's S^- Q, sum lambda^2 by deflated
    Hutchinson (`n_draws` at most, stopping once q is settled), KSS's q rule
    and the first-stage F of the leading direction from sigma2 smoothed on
    the leverage.  `inv` is the panel's CoefficientInverse, `y` the outcome
    on the sorted rows; `seed` the one the reference's diagnostic receives
    (the caller's seed + 9_900_019)."""
    import torch
    if n_draws < 2:
        raise ValueError("n_draws must be at least 2, so the Monte Carlo error of "
                         "sum lambda^2 can be reported")
    lay = SELayout(inv, psi, inv.n if rows is None else rows)
    caps, reorth = _krylov_plan(lay, max_iter)

    def converged_all(alphas, betas, lengths, steps):
        for c in range(3):
            if lengths[c] < steps:
                continue
            theta, bounds, _ = _ritz_top(alphas[c], betas[c], top)
            lead = float(np.abs(theta[0])) if len(theta) else 0.0
            if np.any(bounds[:2] > tol * max(lead, 1e-300)):
                return False
        return True

    alphas, betas, lengths, n_iter = _lanczos(lay, max_iter, seed, check=converged_all, every=10,
                                              caps=caps, reorth=reorth)
    checked = [_ritz_top(alphas[c], betas[c], top) for c in range(3)]
    k = top if 2 * 3 * top * (lay.K + lay.G) * 8 <= SCRATCH_BYTES else 1
    ritz = []
    for _t, _b, S in checked:
        coef = np.zeros((S.shape[0], k))
        coef[:, :min(k, S.shape[1])] = S[:, :k]
        ritz.append(coef)
    X, Y = _lanczos(lay, n_iter, seed, ritz=ritz, caps=caps, reorth=reorth)
    X, Y = _s_orthonormalize(X, Y, k)
    u = [blk[:, ::k].contiguous() for blk in Y]
    leading = [t for t, _b, _S in checked]

    def settled(est, se):
        for c in range(3):
            if not est[c] > 0 or not np.isfinite(se[c]):
                return False
            r = leading[c] ** 2 / est[c]
            if np.any(np.abs(r - WEAK_ID_THRESHOLD) < 2.0 * (se[c] / est[c]) * r):
                return False
        return True

    sum_sq, sum_sq_se, used = _sum_sq_eigenvalues(lay, n_draws, block, seed + SUM_SEED_OFFSET,
                                                  X, Y, k, settled=settled)
    # F: b1 = x1'y~, V[b1] = sum x1^2 sigma2-tilde, x1 = sqrt(w) (A u1)
    sm = smoothed_sigma2(sigma2, leverages.leverage, None, smooth_bins)
    x1 = eigen_rows(lay, u[0], u[1])
    y = y.to(torch.float64)
    ybar = float((lay.w * y).sum() / lay.w.sum())
    yt = lay.root * (y - ybar)
    num = (x1 * yt[:, None]).sum(0).cpu().numpy()
    den = (x1 * x1 * sm[:, None]).sum(0).cpu().numpy()
    names = list(COMPONENTS)
    f_stat = {nm: float(num[c] ** 2 / den[c]) if den[c] > 0 else float("nan")
              for c, nm in enumerate(names)}
    eig, ratios, q, capped, bounds, converged, borderline = {}, {}, {}, {}, {}, {}, {}
    for c, nm in enumerate(names):
        theta, b, _S = checked[c]
        eig[nm] = theta.tolist()
        bounds[nm] = b.tolist()
        lead = float(np.abs(theta[0])) if len(theta) else 0.0
        converged[nm] = bool(lengths[c] < n_iter or lengths[c] >= caps[c]
                             or np.all(b[:2] <= tol * max(lead, 1e-300)))
        r = theta ** 2 / sum_sq[c] if sum_sq[c] > 0 else np.zeros_like(theta)
        ratios[nm] = r.tolist()
        cnt = 0
        for v in r:
            if v >= WEAK_ID_THRESHOLD:
                cnt += 1
            else:
                break
        q[nm] = cnt
        capped[nm] = bool(cnt == len(r) and cnt > 0)
        rel_se = sum_sq_se[c] / sum_sq[c] if sum_sq[c] > 0 else np.inf
        borderline[nm] = bool(np.any(np.abs(r - WEAK_ID_THRESHOLD) < 2.0 * rel_se * r))
    return WeakIdDiagnostics(
        eigenvalues=eig, sum_sq=dict(zip(names, sum_sq.tolist())),
        sum_sq_se=dict(zip(names, sum_sq_se.tolist())), ratios=ratios, q=q, q_capped=capped,
        q_borderline=borderline, f_stat=f_stat, bounds=bounds, converged=converged,
        lead_vectors=tuple(u),
        diagnostics={"lanczos_steps": int(n_iter), "deflated": k,
                     "lanczos_lengths": lengths.tolist(), "krylov_caps": caps,
                     "full_reorthogonalization": reorth, "n_draws": used,
                     "max_draws": n_draws})
