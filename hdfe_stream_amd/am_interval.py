"""The q = 1 interval of a weakly identified variance component
(am_interval.py of hdfe_stream; Kline, Saggio and Solvsten 2020, Section 6
and Appendix C.6).

Given the leading eigenvalue lambda_1 of S^- Q and the estimate (b1, theta1)
with its 2x2 covariance Sigma, the interval is the range of

    lambda_1 b^2 + t   over the ellipse  (b - b1, t - theta1) Sigma^-1 (.)' <= z^2,

with z Andrews and Mikusheva's curvature-adjusted critical value: the
(1 - alpha) quantile of sqrt(a^2 + (b + c)^2) - c for independent
half-normals a, b and c = 1 / kappa, where

    kappa = 2 |lambda_1| V[b1] / sqrt(V[theta1] (1 - rho^2)).

Host arithmetic on a handful of scalars; the moments come from the device
(`leaveout_se.weak_moments`).

  am_cdf             the distribution of the statistic     am_interval.py:45-62
  am_critical_value  its quantile                          :65-79
  curvature          kappa                                 :82-87
  am_interval        the projection of the ellipse         :90-123
"""

from __future__ import annotations

import math

import numpy as np

# Composite Gauss-Legendre rule for the one-dimensional form of am_cdf.
_PANELS = 48
_NODES, _WEIGHTS = np.polynomial.legendre.leggauss(20)
# Beyond this the half-normal density is below exp(-800): no mass.
_A_CAP = 40.0


def _erf(x):
    from scipy.special import erf
    return erf(x)


def am_cdf(z: float, kappa: float) -> float:
    """P(sqrt(a^2 + (b + c)^2) - c <= z), c = 1 / kappa, a and b independent
    half-normals: the quarter plane inside the circle of radius z + c about
    (0, -c).  Integrated over a, the inner half-normal probability in b is
    erf(r(a) / sqrt(2)) with r(a) = sqrt((z + c)^2 - a^2) - c, up to
    a_max = sqrt(z (z + 2c)); the substitution a = a_max sin(t) removes the
    square-root end point, and r is written as
    a_max^2 cos^2 t / (sqrt(c^2 + a_max^2 cos^2 t) + c), so nothing cancels
    as kappa -> 0 (c -> infinity).  kappa <= 0: erf(z / sqrt(2)), the
    chi-squared(1) law of |N(0, 1)|."""
    if not z > 0:
        return 0.0
    if kappa <= 0:
        return float(_erf(z / math.sqrt(2.0)))
    c = 1.0 / kappa
    amax = math.sqrt(z * (z + 2.0 * c))
    tmax = math.asin(min(1.0, _A_CAP / amax))
    edges = np.linspace(0.0, tmax, _PANELS + 1)
    half = 0.5 * (edges[1:] - edges[:-1])
    mid = 0.5 * (edges[1:] + edges[:-1])
    t = (mid[:, None] + half[:, None] * _NODES[None, :]).ravel()
    wt = (half[:, None] * _WEIGHTS[None, :]).ravel()
    a = amax * np.sin(t)
    r2 = (amax * np.cos(t)) ** 2
    inner = r2 / (np.sqrt(c * c + r2) + c)
    dens = math.sqrt(2.0 / math.pi) * np.exp(-0.5 * a * a)
    val = float(np.sum(wt * dens * _erf(inner / math.sqrt(2.0)) * amax * np.cos(t)))
    return min(max(val, 0.0), 1.0)


def am_critical_value(kappa: float, alpha: float = 0.05) -> float:
    """z with am_cdf(z, kappa) = 1 - alpha.  The statistic lies between
    |N(0, 1)| (kappa = 0) and the chi(2) law (kappa -> infinity), so z lies
    between their quantiles, which bracket the root."""
    from scipy import optimize
    from scipy.special import chdtri
    if not 0.0 < alpha < 1.0:
        raise ValueError("alpha must lie in (0, 1)")
    lo = math.sqrt(chdtri(1, alpha))
    if kappa <= 0:
        return lo
    hi = math.sqrt(chdtri(2, alpha))
    target = 1.0 - alpha
    f = (lambda x: am_cdf(x, kappa) - target)
    a, b = lo * (1.0 - 1e-9), hi * (1.0 + 1e-9)
    fa, fb = f(a), f(b)
    if fa >= 0.0:
        return a
    if fb <= 0.0:
        return b
    return float(optimize.brentq(f, a, b, xtol=1e-14, rtol=4 * np.finfo(float).eps))


def curvature(lam1: float, sigma) -> float:
    """kappa = 2 |lambda_1| V[b1] / sqrt(V[theta1] - Cov^2 / V[b1])."""
    sigma = np.asarray(sigma, np.float64)
    vb, vt, cv = float(sigma[0, 0]), float(sigma[1, 1]), float(sigma[0, 1])
    return 2.0 * abs(lam1) * vb / math.sqrt(vt - cv * cv / vb)


def am_interval(lam1: float, b1: float, theta1: float, sigma, alpha: float = 0.05,
                grid: int = 720):
    """(lower, upper, {"kappa", "z", "rho"}) of the q = 1 interval.

    The map is lambda_1 b^2 + t, increasing in t, so its extremes over the
    ellipse lie on the boundary (b1, theta1) + z L (cos p, sin p), L the
    Cholesky factor of `sigma` (positive definite, else LinAlgError).  On
    the boundary the objective is g(p) = lambda_1 (b1 + A cos p)^2 + theta1
    + B cos p + C sin p: its local extremes on a grid of `grid` angles are
    polished by Newton's method on g'(p) = 0 with the exact g''."""
    sigma = np.asarray(sigma, np.float64)
    L = np.linalg.cholesky(sigma)
    kappa = curvature(lam1, sigma)
    z = am_critical_value(kappa, alpha)
    A, B, Cc = z * L[0, 0], z * L[1, 0], z * L[1, 1]

    def g(p):
        return lam1 * (b1 + A * np.cos(p)) ** 2 + theta1 + B * np.cos(p) + Cc * np.sin(p)

    def dg(p):
        c, s = math.cos(p), math.sin(p)
        d1 = -2.0 * lam1 * A * s * (b1 + A * c) - B * s + Cc * c
        d2 = -2.0 * lam1 * A * c * (b1 + A * c) + 2.0 * lam1 * A * A * s * s - B * c - Cc * s
        return d1, d2

    p = np.arange(grid) * (2.0 * math.pi / grid)
    v = g(p)
    prev, nxt = np.roll(v, 1), np.roll(v, -1)
    lows = np.flatnonzero((v <= prev) & (v <= nxt))
    highs = np.flatnonzero((v >= prev) & (v >= nxt))
    lower, upper = float(v.min()), float(v.max())
    step = 2.0 * math.pi / grid
    for idx, sign in [(i, 1.0) for i in lows] + [(i, -1.0) for i in highs]:
        x = float(p[idx])
        for _ in range(60):
            d1, d2 = dg(x)
            if sign * d2 <= 0.0:          # not curving the right way: keep the grid point
                break
            dx = -d1 / d2
            x += max(-step, min(step, dx))
            if abs(dx) <= 1e-15 * (1.0 + abs(x)):
                break
        val = float(g(x))
        lower, upper = min(lower, val), max(upper, val)
    rho = float(sigma[0, 1] / math.sqrt(sigma[0, 0] * sigma[1, 1]))
    return lower, upper, {"kappa": float(kappa), "z": float(z), "rho": rho}
