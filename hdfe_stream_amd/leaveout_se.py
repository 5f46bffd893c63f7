"""Standard errors of the KSS variance components on the device
(leaveout_se.py of hdfe_stream).

In the square-root-weight metric X~ = W^(1/2) A, component k's estimator is
the quadratic form y~'C_k y~ with the zero-diagonal kernel

    C = B - (diag(b) M + M diag(b)) / 2,   B = X~ S^- Q S^- X~',  M = I - P,
    b_i = B_ii / M_ii (times the leverage pass's non-linearity factor),

and its variance is estimated by V-hat = 4 sum_i s_i (C y~)_i^2 -
2 tr(C Om C Om), s the leave-out sigma2 smoothed on (P_ii, B_ii)
(leaveout_se.py:1-49).  Mirrored here:

  bii_rows              B_ii by random projection        leaveout_se.py:228-327
  apply_q               Q_k u in coefficient space       :339-385
  apply_c               C v, two solves, three passes    :426-596
  m_times               M b (the reference centring)     :387-424
  fit / eval smoother   sigma2 on the (P_ii, B_ii) grid  :955-1054
  component_variances   first term, trace, V-hat, SEs    :599-784, 891-937
  eigen_rows            x1 = X~ u of a lead vector       :486-495
  weak_moments          (b1, theta1) moments for q = 1   :786-889

The row passes are hand-written kernels (se.hip, kss.hip, rows.hip); the
solves are the panel's block PCG through `CoefficientInverse`.  The draws
are the counter-based Rademacher draws keyed on the sorted row ordinal, so a
panel in the reference's canonical row order gets the reference's draws."""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib

COMPONENTS = ("var(psi)", "var(alpha)", "cov(psi, alpha)")
BII_SEED_OFFSET = 5_500_011        # leaveout.py:772
TRACE_SEED_OFFSET = 7_700_017      # leaveout_se.py:752
WEAK_SEED_OFFSET = 8_800_021       # leaveout_se.py:747
QUANTILE_SAMPLE = 2_000_000        # leaveout_se.py:62


@dataclass
class ComponentSE:
    """Standard errors of the three components and the parts they came from
    (leaveout_se.py:74-126): dicts keyed by component; NaN where none is
    reported.  `weak` holds, per component the diagnostic flags (q >= 1),
    the moments of the q = 1 interval (see `weak_moments`)."""
    se: dict
    variance: dict
    first_term: dict
    trace_term: dict
    se_conservative: dict
    theta: dict
    n_draws: int
    diagnostics: dict = field(default_factory=dict)
    weak: dict = None

    def summary(self) -> str:
        lines = [f"leave-out standard errors ({self.n_draws} draws)",
                 f"{'component':<20}{'estimate':>12}{'se':>11}"
                 f"{'se (cons.)':>12}{'trace share':>13}",
                 "-" * 68]
        for key in COMPONENTS:
            share = (self.trace_term[key] / self.first_term[key]
                     if self.first_term[key] else float("nan"))
            lines.append(f"{key:<20}{self.theta[key]:>12.6f}{self.se[key]:>11.6f}"
                         f"{self.se_conservative[key]:>12.6f}{share:>12.2%}")
        return "\n".join(lines)


def _moment_divisor(total, rows):
    return total * (rows - 1.0) / rows if rows > 1 else total


class SELayout:
    """The weight totals per level and group, the grand total, the divisor
    T (R - 1) / R and the psi block (leaveout_se.py:135-180), over a
    `CoefficientInverse`."""

    def __init__(self, inv, psi: int, rows: int):
        import torch
        self.inv = inv
        panel = inv.panel
        n = inv.n
        self.n, self.L, self.G = n, inv.L, inv.G
        self.K = inv.K                      # levels, then the design's covariates
        if inv.w is not None:
            w = panel.tensor("w", torch.float64)
        else:
            w = torch.ones(n, dtype=torch.float64, device="cuda")
        self.weighted = inv.w is not None
        self.w = w
        self.root = torch.sqrt(w)
        self.inv_root = torch.where(self.root > 0, 1.0 / torch.where(self.root > 0, self.root,
                                                                     torch.ones_like(w)),
                                    torch.zeros_like(w))
        ones = torch.ones((n, 1), dtype=torch.float64, device="cuda")
        self.level_w = inv.reduce(1, R=ones)[:, 0].contiguous()
        self.group_w = inv.Wg
        self.total = float(self.group_w.sum())
        self.norm = _moment_divisor(self.total, rows)
        o = int(inv.offs[psi])
        self.psi = slice(o, o + int(inv.nlev[psi]))
        self.psi_off = o
        self.psi_codes = inv.codes + 4 * psi * n

    def reduce(self, R):
        """X~'R for rows R (n, k): (levels and covariates (K, k), groups
        (G, k))."""
        x = self.inv_root if self.weighted else None
        R = R.contiguous()
        return self.inv.reduce(R.shape[1], R=R, x=x), self.inv.group_sums(R, x=x)


def tall_dot(A, B, w=None):
    """A' diag(w) B on the device for tall A (K, a) and B (K, b) (vectors count
    as one column), FP64, fixed-order sums (hdfe_tall_dot)."""
    import torch
    A2 = A[:, None] if A.dim() == 1 else A
    B2 = B[:, None] if B.dim() == 1 else B
    A2 = A2.to(torch.float64)
    B2 = B2.to(torch.float64)
    if A2.stride(1) != 1 or A2.stride(0) < A2.shape[1]:
        A2 = A2.contiguous()
    if B2.stride(1) != 1 or B2.stride(0) < B2.shape[1]:
        B2 = B2.contiguous()
    K, a, b = A2.shape[0], A2.shape[1], B2.shape[1]
    out = torch.empty((a, b), dtype=torch.float64, device="cuda")
    if K == 0:
        return out.zero_()
    wd = None if w is None else w.to(torch.float64).contiguous()
    torch.cuda.synchronize()
    _lib.check(_lib.load().hdfe_tall_dot(K, a, b, A2.data_ptr(), A2.stride(0), B2.data_ptr(),
                                         B2.stride(0), None if wd is None else wd.data_ptr(),
                                         out.data_ptr()), "hdfe_tall_dot")
    return out


def bii_rows(lay: SELayout, n_draws: int, block: int, seed: int):
    """B_ii of the three components, (3, n) device, by random projection
    (leaveout_se.py:228-327): per block of draws q, the two solves of the
    weight-centred selectors L_psi'q and L_alpha'q, read back at the rows and
    squared and crossed."""
    import torch
    n, L, G = lay.n, lay.L, lay.G
    lib = _lib.load()
    out = torch.zeros((3, n), dtype=torch.float64, device="cuda")
    lw = lay.level_w[lay.psi]
    sq = np.sqrt(lay.norm)
    drawn = 0
    while drawn < n_draws:
        size = min(block, n_draws - drawn)
        q = torch.empty((n, size), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        _lib.check(lib.hdfe_rademacher(0, n, size, seed, drawn, q.data_ptr()), "hdfe_rademacher")
        lv, gs = lay.reduce(q)
        del q
        grand = lv[lay.psi].sum(0) / lay.total
        rl = torch.zeros((L, 2 * size), dtype=torch.float64, device="cuda")
        rl[lay.psi, :size] = (lv[lay.psi] - torch.outer(lw, grand)) / sq
        rg = torch.zeros((G, 2 * size), dtype=torch.float64, device="cuda")
        rg[:, size:] = (gs - torch.outer(lay.group_w, grand)) / sq
        _, _, rows = lay.inv(rl, rg)
        torch.cuda.synchronize()
        _lib.check(lib.hdfe_se_bii(n, size, rows.data_ptr(),
                                   lay.w.data_ptr() if lay.weighted else None, out.data_ptr()),
                   "hdfe_se_bii")
        del rows
        drawn += size
    return out / n_draws


def apply_q(lay: SELayout, k: int, Ul, Ug):
    """Q_k u (leaveout_se.py:339-385): the variances by weight-centred
    scaling in coefficient space, the covariance by crossing the two
    dimensions over the rows (kernels_se.py:44-72)."""
    import torch
    s = Ul.shape[1]
    ol = torch.zeros((lay.K, s), dtype=torch.float64, device="cuda")
    og = torch.zeros((lay.G, s), dtype=torch.float64, device="cuda")
    lw = lay.level_w[lay.psi]
    psi_mean = tall_dot(lw, Ul[lay.psi])[0] / lay.total
    alpha_mean = tall_dot(lay.group_w, Ug)[0] / lay.total
    if k == 0:
        ol[lay.psi] = lw[:, None] * (Ul[lay.psi] - psi_mean) / lay.norm
    elif k == 1:
        og[:] = lay.group_w[:, None] * (Ug - alpha_mean) / lay.norm
    else:
        inv = lay.inv
        # sum_i w_i Ug[g(i)] per level: the spread reduction divides by W_g
        cross_l = inv.reduce(s, V=(Ug * lay.group_w[:, None]).contiguous())
        cross_g = torch.empty((lay.G, s), dtype=torch.float64, device="cuda")
        Ulc = Ul.contiguous()
        torch.cuda.synchronize()
        _lib.check(_lib.load().hdfe_se_cross_groups(
            lay.G, s, inv.gstart, lay.psi_codes, lay.psi_off, inv.w, Ulc.data_ptr(),
            cross_g.data_ptr()), "hdfe_se_cross_groups")
        ol[lay.psi] = (cross_l[lay.psi] - lw[:, None] * alpha_mean) / (2 * lay.norm)
        og[:] = (cross_g - lay.group_w[:, None] * psi_mean) / (2 * lay.norm)
    return ol, og


def apply_c(lay: SELayout, b, v=None, z=None, s3=None, rank_two=None, yt=None, deflate=None,
            weak: bool = False, mask: int = 7):
    """C_k v for the three components, reduced (leaveout_se.py:426-596).

    With a shared input v (n, s) and `yt`: (first (3), theta (3)) of
    4 sum s_k (C v)_k^2 and yt'(C v)_k (s = 1, v = yt).  Per component,
    v_k = z sqrt(max(s_k, 0)) from draws z (n, s), zero for a component not
    in the bit `mask`: sum_i s_k sum_j (C v)_k^2 (3).  b and s3 are (3, n);
    rank_two = (u (3, n), c) adds c (1 u_k' + u_k 1') v; deflate =
    (lambdas (3,), x1 (3, n)) makes it the C2 of B2 = B - lambda_k x1_k x1_k'.
    `weak` floors s at zero and, for the shared input, reduces to (s12 (3),
    first (3), quad (3)): 2 sum x1_k s_k (C2 v)_k, 4 sum s_k (C2 v)_k^2 and
    yt'(C2 v)_k (leaveout_se.py:843-853)."""
    import torch
    lib = _lib.load()
    n = lay.n
    per = z is not None
    s = (z if per else v).shape[1]
    n_in = 3 if per else 1
    wide = n_in + 3
    if weak and deflate is None:
        raise ValueError("the weak moments need the deflation")
    stack = torch.empty((n, wide * s), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    _lib.check(lib.hdfe_se_stack(n, s, None if per else v.data_ptr(),
                                 z.data_ptr() if per else None, s3.data_ptr(), b.data_ptr(),
                                 int(mask), stack.data_ptr()), "hdfe_se_stack")
    vin = [stack[:, (k if per else 0) * s:((k if per else 0) + 1) * s] for k in range(3)]
    udot = odot = None
    if rank_two is not None:
        u = rank_two[0]
        udot = torch.stack([tall_dot(u[k], vin[k]) for k in range(3)]).reshape(3, s).contiguous()
        odot = torch.stack([vin[k].sum(0) for k in range(3)]).contiguous()
    x1 = defl = None
    if deflate is not None:
        x1 = deflate[1]
        along = torch.stack([tall_dot(x1[k], vin[k]) for k in range(3)]).reshape(3 * s)
        lam = torch.as_tensor(np.asarray(deflate[0], np.float64)).cuda()
        defl = torch.cat([lam, along]).contiguous()
    lv, gs = lay.reduce(stack)
    U0l, U0g, base = lay.inv(lv, gs)
    del lv, gs
    ql = torch.zeros((lay.K, 3 * s), dtype=torch.float64, device="cuda")
    qg = torch.zeros((lay.G, 3 * s), dtype=torch.float64, device="cuda")
    for k in range(3):
        j = k if per else 0
        ol, og = apply_q(lay, k, U0l[:, j * s:(j + 1) * s], U0g[:, j * s:(j + 1) * s])
        ql[:, k * s:(k + 1) * s] = ol
        qg[:, k * s:(k + 1) * s] = og
    del U0l, U0g, stack
    _, _, thq = lay.inv(ql, qg)
    mode = (1 if per else 0) + (2 if weak else 0)
    out = np.empty({0: 6, 2: 9}.get(mode, 3))
    torch.cuda.synchronize()
    _lib.check(lib.hdfe_se_assemble(
        n, s, mode, None if per else v.data_ptr(), z.data_ptr() if per else None,
        s3.data_ptr(), b.data_ptr(), lay.w.data_ptr() if lay.weighted else None,
        base.data_ptr(), thq.data_ptr(), None if rank_two is None else rank_two[0].data_ptr(),
        0.0 if rank_two is None else float(rank_two[1]),
        None if udot is None else udot.data_ptr(), None if odot is None else odot.data_ptr(),
        None if yt is None else yt.data_ptr(), None if x1 is None else x1.data_ptr(),
        None if defl is None else defl.data_ptr(), int(mask), out.ctypes.data),
        "hdfe_se_assemble")
    return out


def m_times(lay: SELayout, b):
    """u_k = M b_k for the three row vectors b (3, n) (leaveout_se.py:387-424)."""
    lv, gs = lay.reduce(b.T.contiguous())
    _, _, base = lay.inv(lv, gs)
    return (b - (lay.root[:, None] * base).T).contiguous()


def auto_bins(n: int, dims: int = 2) -> int:
    """Bins per axis, n^(1/6) in two dimensions (leaveout_se.py:955-964)."""
    power = 1.0 / 6.0 if dims == 2 else 1.0 / 3.0
    return int(np.clip(round(n ** power), 2, 500))


def _sample(n: int) -> np.ndarray:
    """The rows the knots come from (leaveout_se.py:1021-1023)."""
    if n <= QUANTILE_SAMPLE:
        return np.arange(n)
    return np.sort(np.random.default_rng(0).choice(n, QUANTILE_SAMPLE, replace=False))


def _knots(values_host: np.ndarray, n_bins: int) -> np.ndarray:
    """Quantile knots made strictly increasing (leaveout_se.py:967-973)."""
    q = np.quantile(values_host, np.linspace(0.0, 1.0, n_bins + 1))
    q = np.maximum.accumulate(q)
    span = max(float(q[-1] - q[0]), 1e-300)
    return q + np.arange(len(q)) * span * 1e-12


def smoothed_sigma2(sigma2, leverage, b_ii=None, n_bins=None):
    """sigma2 smoothed on the grid of (P_ii, B_ii) quantile ranks -- or of the
    leverage alone without b_ii -- cell means interpolated (bi)linearly
    between cell centres (leaveout_se.py:982-1054); device (n,).
    n_bins < 1: sigma2 itself."""
    import torch
    n = int(sigma2.shape[0])
    dims = 1 if b_ii is None else 2
    nb = auto_bins(n, dims) if n_bins is None else int(n_bins)
    if nb < 1:
        return sigma2.to(torch.float64).contiguous()
    if nb == 1:
        raise ValueError("smooth_bins must be 0 (no smoothing) or at least 2")
    take = torch.from_numpy(_sample(n)).cuda()
    lev = leverage.to(torch.float64).contiguous()
    kp_d = torch.from_numpy(_knots(lev[take].cpu().numpy(), nb)).cuda()
    bii = kb_d = None
    if b_ii is not None:
        bii = b_ii.to(torch.float64).contiguous()
        kb_d = torch.from_numpy(_knots(bii[take].cpu().numpy(), nb)).cuda()
    ncell = nb ** dims
    tot, cnt = np.empty(ncell), np.empty(ncell)
    s2 = sigma2.to(torch.float64).contiguous()
    lib = _lib.load()
    ptr = (lambda t: None if t is None else t.data_ptr())
    torch.cuda.synchronize()
    _lib.check(lib.hdfe_se_cell_sums(n, lev.data_ptr(), ptr(bii), s2.data_ptr(),
                                     kp_d.data_ptr(), ptr(kb_d), nb, tot.ctypes.data,
                                     cnt.ctypes.data), "hdfe_se_cell_sums")
    overall = float(tot.sum() / max(cnt.sum(), 1.0))
    means = np.where(cnt > 0, tot / np.maximum(cnt, 1.0), overall)
    means_d = torch.from_numpy(means).cuda()
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    _lib.check(lib.hdfe_se_smooth(n, lev.data_ptr(), ptr(bii), kp_d.data_ptr(), ptr(kb_d),
                                  means_d.data_ptr(), nb, out.data_ptr()), "hdfe_se_smooth")
    return out


def eigen_rows(lay: SELayout, Ul, Ug):
    """x1 = X~ u: sqrt(w_i) (A u)_i for coefficient blocks Ul (K, c) and
    Ug (G, c) (leaveout_se.py:486-495); device (n, c)."""
    import torch
    inv = lay.inv
    panel = inv.panel
    g = panel.tensor("g", torch.int32).long()
    codes = panel.tensor("codes", torch.int32).view(inv.D, inv.n).long()
    au = Ug[g]
    for d in range(inv.D):
        au = au + Ul[int(inv.offs[d]) + codes[d]]
    if inv.kc:
        au = au + inv.border.X @ Ul[inv.L:]
    return au * lay.root[:, None]


def weak_moments(lay: SELayout, y, sigma2, s3, leverages, b_ii, lams, Ul, Ug, components, *,
                 n_draws: int, block: int, seed: int, stayer_rows=None,
                 centering: str = "weighted") -> dict:
    """The moments of the q = 1 interval per component in `components`
    (leaveout_se.py:786-889): b1 = x1'y~ with x1 = X~ u1 (u1 the lead
    vector, u1'S u1 = 1), V[b1] raw from `sigma2` (the point estimate's),
    and the covariance of (b1, theta1), theta1 = y~'C2 y~ for the C2 of
    B2 = B - lambda_1 x1 x1':

        V[b1] = sum x1^2 s,   Cov = 2 sum x1 s (C2 y~),
        V[theta1] = 4 sum s (C2 y~)^2 - 2 tr(C2 Om C2 Om),

    s the smoothed sigma2 `s3` floored at zero.  Where subtracting the trace
    leaves the matrix indefinite, V[theta1] falls back to its first term
    ("conservative").  The trace's draws use `seed` directly."""
    import torch
    n = lay.n
    x1 = eigen_rows(lay, Ul, Ug).T.contiguous()                 # (3, n)
    w = lay.w
    y = y.to(torch.float64).contiguous()
    ybar = float((w * y).sum() / w.sum())
    yt_c = lay.root * (y - ybar)
    b1 = tall_dot(yt_c, x1.T)[0].cpu().numpy()
    v_raw = tall_dot(sigma2, (x1 * x1).T)[0].cpu().numpy()
    sf = s3.clamp_min(0.0)
    s11 = (x1 * x1 * sf).sum(1).cpu().numpy()
    lam_d = torch.as_tensor(np.asarray(lams, np.float64)).cuda()
    b2 = ((b_ii - lam_d[:, None] * x1 * x1)
          * (leverages.factor / leverages.complement)[None, :]).contiguous()
    if stayer_rows is not None:
        b2[:, stayer_rows] = 0.0
    rank_two = (m_times(lay, b2), 0.5 / n) if centering == "reference" else None
    yt = (lay.root * (y - (ybar if centering == "weighted" else 0.0))).contiguous()
    deflate = (lams, x1)
    fy = apply_c(lay, b2, v=yt[:, None].contiguous(), s3=s3, rank_two=rank_two, yt=yt,
                 deflate=deflate, weak=True)
    s12, first, quad = fy[:3], fy[3:6], fy[6:]
    mask = sum(1 << k for k in components)
    total = np.zeros(3)
    lib = _lib.load()
    drawn = 0
    while drawn < n_draws:
        size = min(block, n_draws - drawn)
        z = torch.empty((n, size), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        _lib.check(lib.hdfe_rademacher(0, n, size, seed, drawn, z.data_ptr()),
                   "hdfe_rademacher")
        total += apply_c(lay, b2, z=z, s3=s3, rank_two=rank_two, deflate=deflate, weak=True,
                         mask=mask)
        drawn += size
    trace = total / n_draws
    out = {}
    for k in components:
        s22 = first[k] - 2.0 * trace[k]
        conservative = not (s22 > 0 and s12[k] ** 2 < s11[k] * s22)
        if conservative:
            s22 = first[k]
        out[COMPONENTS[k]] = {
            "lambda_1": float(lams[k]), "b1": float(b1[k]), "v_raw": float(v_raw[k]),
            "sigma": np.array([[s11[k], s12[k]], [s12[k], s22]]),
            "conservative": bool(conservative), "trace": float(trace[k]),
            "quadratic": float(quad[k])}
    return out


def component_variances(inv, y, sigma2, leverages, *, psi: int = 0, rows: int | None = None,
                        n_draws: int = 250, block: int | None = None, seed: int = 0,
                        trace: bool = True, smooth_bins=None, b_ii=None, stayer_rows=None,
                        centering: str = "weighted", reported=(0, 1, 2), weak_id=None,
                        weak_components=None, sigma2_point=None) -> ComponentSE:
    """V-hat and the standard errors of the three components
    (leaveout_se.py:599-784).  `inv` is the panel's CoefficientInverse, `y`
    the outcome on the sorted rows (device), `sigma2` the leave-out variances
    (device), `leverages` the JLA result.  `seed` is the one the reference's
    component_variances receives (the caller's seed + 5_500_011); the B_ii
    draws use it, the trace's seed + 7_700_017.  At match level
    `stayer_rows` (device bool) zeroes b on stayers' matches, `centering`
    "reference" adds the rank-two term, and `reported` lists the components
    with a standard error.  With `weak_id` (the diagnostic's result), every
    reported component it flags (q >= 1) -- at match level var(psi) alone,
    since stayers load on the others' weak direction -- gets the moments of
    the q = 1 interval, from `sigma2_point` (default sigma2) and draws with
    seed + 8_800_021 (leaveout_se.py:736-756)."""
    import torch
    if centering not in ("reference", "weighted"):
        raise ValueError("centering must be 'reference' or 'weighted'")
    block = 32 if block is None else int(block)
    if block < 1 or n_draws < 1:
        raise ValueError("n_draws and block must be at least 1")
    n = inv.n
    lay = SELayout(inv, psi, n if rows is None else rows)
    if b_ii is None:
        b_ii = bii_rows(lay, n_draws, block, seed)
    b = (b_ii * (leverages.factor / leverages.complement)[None, :]).contiguous()
    if stayer_rows is not None:
        b[:, stayer_rows] = 0.0
    rank_two = (m_times(lay, b), 0.5 / n) if centering == "reference" else None
    s3 = torch.stack([smoothed_sigma2(sigma2, leverages.leverage, b_ii[k], smooth_bins)
                      for k in range(3)]).contiguous()
    w = lay.w
    y = y.to(torch.float64).contiguous()
    ybar = float((w * y).sum() / w.sum())
    yt = (lay.root * (y - (ybar if centering == "weighted" else 0.0))).contiguous()
    fy = apply_c(lay, b, v=yt[:, None].contiguous(), s3=s3, rank_two=rank_two, yt=yt)
    first, theta = fy[:3], fy[3:]
    total = np.zeros(3)
    blocks = 0
    if trace:
        lib = _lib.load()
        drawn = 0
        while drawn < n_draws:
            size = min(block, n_draws - drawn)
            z = torch.empty((n, size), dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            _lib.check(lib.hdfe_rademacher(0, n, size, seed + TRACE_SEED_OFFSET, drawn,
                                           z.data_ptr()), "hdfe_rademacher")
            total += apply_c(lay, b, z=z, s3=s3, rank_two=rank_two)
            drawn += size
            blocks += 1
        total /= n_draws
    weak = None
    if weak_id is not None:
        if weak_components is None:
            weak_components = [k for k, nm in enumerate(COMPONENTS) if weak_id.q[nm] >= 1]
        weak_components = [k for k in weak_components
                           if k in reported and (stayer_rows is None or k == 0)]
        if weak_components:
            lams = np.array([float(weak_id.eigenvalues[nm][0]) if weak_id.eigenvalues[nm]
                             else 0.0 for nm in COMPONENTS])
            Ul, Ug = weak_id.lead_vectors
            weak = weak_moments(lay, y, sigma2 if sigma2_point is None else sigma2_point, s3,
                                leverages, b_ii, lams, Ul, Ug, weak_components,
                                n_draws=n_draws, block=block, seed=seed + WEAK_SEED_OFFSET,
                                stayer_rows=stayer_rows, centering=centering)
    variance = first - 2.0 * total
    names = list(COMPONENTS)
    for k in range(3):
        if k not in reported:
            variance[k] = first[k] = total[k] = theta[k] = np.nan
    return ComponentSE(
        weak=weak,
        se={nm: float(np.sqrt(v)) if v > 0 else float("nan") for nm, v in zip(names, variance)},
        variance=dict(zip(names, variance.tolist())),
        first_term=dict(zip(names, first.tolist())),
        trace_term=dict(zip(names, (2.0 * total).tolist())),
        se_conservative={nm: float(np.sqrt(f)) if f > 0 else float("nan")
                         for nm, f in zip(names, first)},
        theta=dict(zip(names, theta.tolist())),
        n_draws=n_draws if trace else 0,
        diagnostics={"blocks": blocks, "trace_included": bool(trace),
                     "smooth_bins": smooth_bins, "centering": centering,
                     "reported": [names[k] for k in reported],
                     "max_b_ii": {nm: float(b_ii[k].max()) for k, nm in enumerate(names)},
                     "stayer_rows": 0 if stayer_rows is None else int(stayer_rows.sum()),
                     "negative_trace": {nm: bool(t < 0) for nm, t in zip(names, total)},
                     "b_ii": b_ii, "smoothed_sigma2": s3})
