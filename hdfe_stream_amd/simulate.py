"""Seeded synthetic AKM panels of the shapes the benchmarks and tests use.

The reference ships its own simulator (hdfe_stream/simulate.py:42-143,
`simulate_akm`, on Polars).  These generators are the numpy panels of the
workloads measured here: workers observed for `n_years` consecutive years, a
uniform first firm, and each year a move with probability `p_move` to a
uniform random firm.  Effects: worker alpha ~ N(0, 0.5^2), firm psi ~ N(0,
0.3^2), year gamma ~ N(0, 0.1^2); covariates x1, x2 ~ N(0, 1);
y = alpha + psi + gamma + 0.3 x1 - 0.2 x2 + eps, eps ~ N(0, 0.4^2).
Ids are int64.  With `corr` > 0 the first firm is drawn by a Gaussian copula
between the worker's alpha and the firms' psi ranks (sorting), which gives a
positive row-level corr(alpha, psi).
"""

from __future__ import annotations

import numpy as np

__all__ = ["simulate_akm", "akm_panel", "kss_panel", "weighted_panel", "bottleneck_panel", "poisson_models",
           "poisson_models_device", "ols_models"]

# Row-level corr(alpha, psi) is about 0.6 times the copula parameter with
# p_move 0.15 and 8 years (movers' later firms are uniform), so a copula of
# 1/3 gives a row-level correlation of about 0.20.
KSS_COPULA = 1.0 / 3.0


def _first_firm(rng, alpha, psi, corr):
    n_workers, n_firms = len(alpha), len(psi)
    if corr <= 0:
        return rng.integers(0, n_firms, n_workers)
    # Gaussian copula: latent = rho * z(alpha) + sqrt(1 - rho^2) * noise,
    # mapped to the firm whose psi has the same rank among firms.
    za = (alpha - alpha.mean()) / alpha.std()
    lat = corr * za + np.sqrt(1.0 - corr * corr) * rng.standard_normal(n_workers)
    q = np.argsort(np.argsort(lat)) / max(n_workers - 1, 1)
    order = np.argsort(psi)
    return order[np.minimum((q * n_firms).astype(np.int64), n_firms - 1)]


def akm_panel(n_workers: int, n_firms: int, n_years: int = 8, p_move: float = 0.10,
              seed: int = 1, corr: float = 0.0, copula: float | None = None) -> dict:
    """A balanced AKM panel as a dict of numpy columns, rows ordered by
    (worker, year): worker_id, firm_id, year (int64), x1, x2, y (float64),
    plus the true effects alpha (per row), psi (per row) and gamma (per row).

    `corr` is the copula correlation used for the first firm (0: uniform);
    `copula` overrides it directly when given."""
    rng = np.random.default_rng(seed)
    n = n_workers * n_years
    alpha_w = rng.normal(0.0, 0.5, n_workers)
    psi_f = rng.normal(0.0, 0.3, n_firms)
    gamma_t = rng.normal(0.0, 0.1, n_years)
    rho = corr if copula is None else copula
    firm = np.empty((n_workers, n_years), np.int64)
    firm[:, 0] = _first_firm(rng, alpha_w, psi_f, rho)
    for t in range(1, n_years):
        move = rng.random(n_workers) < p_move
        dest = rng.integers(0, n_firms, n_workers)
        firm[:, t] = np.where(move, dest, firm[:, t - 1])
    worker = np.repeat(np.arange(n_workers, dtype=np.int64), n_years)
    year = np.tile(np.arange(n_years, dtype=np.int64) + 2000, n_workers)
    firm = firm.reshape(-1)
    x1 = rng.standard_normal(n)
    x2 = rng.standard_normal(n)
    eps = rng.normal(0.0, 0.4, n)
    a = alpha_w[worker]
    p = psi_f[firm]
    g = gamma_t[year - 2000]
    y = a + p + g + 0.3 * x1 - 0.2 * x2 + eps
    return {"worker_id": worker, "firm_id": firm, "year": year, "x1": x1, "x2": x2, "y": y,
            "alpha": a, "psi": p, "gamma": g}


def simulate_akm(n_workers: int = 50_000, n_firms: int | None = None, years=range(2005, 2015),
                 p_move: float = 0.08, p_obs: float = 0.85, sd_worker: float = 0.4,
                 sd_firm: float = 0.25, sd_noise: float = 0.3, seed: int = 0,
                 keep_effects: bool = False) -> dict:
    """The reference's AKM panel (hdfe_stream/simulate.py:42-143,
    `simulate_akm`) in numpy, draw for draw: the same generator calls in the
    same order, so the same seed gives the same rows.  Kept alike on purpose:
    the reference's benchmark coefficients (benchmarks/results/akm.csv) are
    quoted on these panels.

    Firms are ranked by effect; a worker's first firm and every mover's new
    firm are, with probability 1/2 each, the firm at the worker's noisy rank
    or a size-weighted (Pareto 1.2 sizes) draw; each year a worker moves with
    p_move and is observed with p_obs; ages run from a birth year in
    [22, 55).  log_earn = 10 + worker + firm + 0.08 age^2/100 - 0.012
    age^3/1000 + N(0, sd_noise^2) noise drawn from seed + 1 over the rows in
    year order.

    Returns a dict of numpy columns in generation order (year by year, workers
    ascending within a year); the reference then shuffles the rows with
    Polars, which changes no estimate.  `firm_id` is the firm's index (int64;
    the reference's string ids are "F" + this index), `worker_id` is
    7 * worker + 1,000,000 as there."""
    if n_firms is None:
        n_firms = max(n_workers // 15, 50)
    rng = np.random.default_rng(seed)
    years = list(years)
    worker_effect = rng.normal(0, sd_worker, n_workers)
    firm_effect = rng.normal(0, sd_firm, n_firms)
    ranked = np.argsort(firm_effect)
    size = rng.pareto(1.2, n_firms) + 1
    weights = size / size.sum()
    birth = rng.integers(22, 55, n_workers)

    def new_firm(effect):
        pos = (effect / max(sd_worker, 1e-12) + rng.normal(0, 1.5, effect.size)) / 6 + 0.5
        by_rank = ranked[(np.clip(pos, 0, 0.999) * n_firms).astype(int)]
        by_size = rng.choice(n_firms, effect.size, p=weights)
        return np.where(rng.random(effect.size) < 0.5, by_rank, by_size)

    firm = new_firm(worker_effect)
    cols = {k: [] for k in ("worker_id", "firm_id", "year", "age", "w", "f")}
    for t, yr in enumerate(years):
        if t:
            firm = np.where(rng.random(n_workers) < p_move, new_firm(worker_effect), firm)
        seen = np.flatnonzero(rng.random(n_workers) < p_obs)
        cols["worker_id"].append(seen.astype(np.int64) * 7 + 1_000_000)
        cols["firm_id"].append(firm[seen].astype(np.int64))
        cols["year"].append(np.full(seen.size, yr, np.int64))
        cols["age"].append((birth[seen] + t).astype(float))
        cols["w"].append(worker_effect[seen])
        cols["f"].append(firm_effect[firm[seen]])
    out = {k: np.concatenate(v) for k, v in cols.items()}
    noise = np.random.default_rng(seed + 1).normal(0, sd_noise, out["age"].size)
    age = out["age"]
    out["age_squared"] = age ** 2 / 100
    out["age_cubed"] = age ** 3 / 1000
    out["log_earn"] = (10 + out["w"] + out["f"] + 0.08 * out["age_squared"]
                       - 0.012 * out["age_cubed"] + noise)
    w, f = out.pop("w"), out.pop("f")
    if keep_effects:
        out["true_worker_effect"], out["true_firm_effect"] = w, f
    return out


def kss_panel(n_workers: int = 1_000_000, n_firms: int = 100_000, seed: int = 3,
              n_years: int = 8) -> dict:
    """The leave-out panel: `akm_panel` with p_move 0.15 and sorting, a
    row-level corr(alpha, psi) of about 0.2."""
    return akm_panel(n_workers, n_firms, n_years=n_years, p_move=0.15, seed=seed,
                     copula=KSS_COPULA)


def weighted_panel(n_workers: int, n_firms: int, seed: int = 4, n_years: int = 8,
                   p_move: float = 0.10) -> dict:
    """`akm_panel` with row weights w ~ Uniform(0.5, 2.0) as float32."""
    d = akm_panel(n_workers, n_firms, n_years=n_years, p_move=p_move, seed=seed)
    rng = np.random.default_rng(seed + 1000)
    d["w"] = rng.uniform(0.5, 2.0, len(d["y"])).astype(np.float32)
    return d


def bottleneck_panel(n_bridge: int = 6, firms_per_side: int = 8, movers_per_side: int = 150,
                     stayers_per_firm: int = 8, gap: float = 0.5, seed: int = 0) -> dict:
    """A panel with weakly identified components (the design of
    hdfe_stream/simulate.py:279-360, `simulate_bottleneck`): two sides of
    `firms_per_side` firms, each with `movers_per_side` workers who move
    between two firms of their side (two years at each) and
    `stayers_per_firm` three-year stayers per firm, joined only by
    `n_bridge` workers who move from one side to the other.  The gap
    between the sides' mean firm effects (`gap` is added to the second
    side's) is then estimated far worse than any other contrast: one
    eigenvalue of S^- Q dominates.  Effects alpha ~ N(0, 0.4^2), psi ~
    N(0, 0.3^2), noise N(0, 0.15^2); columns worker_id, firm_id, y."""
    rng = np.random.default_rng(seed)
    F = firms_per_side
    psi = rng.normal(0.0, 0.3, 2 * F)
    psi[F:] += gap
    worker, firm = [], []
    wid = 0

    def spell(f, years):
        nonlocal wid
        worker.extend([wid] * years)
        firm.extend([f] * years)

    for side in range(2):
        base = side * F
        for _ in range(movers_per_side):
            a, b = rng.choice(F, 2, replace=False)
            spell(base + a, 2)
            spell(base + b, 2)
            wid += 1
        for f in range(F):
            for _ in range(stayers_per_firm):
                spell(base + f, 3)
                wid += 1
    for _ in range(n_bridge):
        spell(int(rng.integers(0, F)), 2)
        spell(F + int(rng.integers(0, F)), 2)
        wid += 1
    worker = np.asarray(worker, np.int64)
    firm = np.asarray(firm, np.int64)
    alpha_w = rng.normal(0.0, 0.4, wid)
    y = alpha_w[worker] + psi[firm] + rng.normal(0.0, 0.15, len(worker))
    return {"worker_id": worker, "firm_id": firm, "y": y}


def poisson_models(n_models: int = 32, lo: int = 50_000, hi: int = 2_000_000,
                   n_cov: int = 300, seed: int = 5, scale: float = 1.0) -> list:
    """Uneven Poisson models: rows log-uniform in [lo, hi] (times `scale`),
    worker rows/8 levels, firm rows/80, year 10, `n_cov` float32 covariates;
    y ~ Poisson(exp(0.5 + 0.1 x1 + alpha + psi)), alpha, psi ~ N(0, 0.2^2)."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), n_models)) * scale
    out = []
    for n in sizes.astype(np.int64):
        n = max(int(n), 160)
        nw, nf = max(n // 8, 2), max(n // 80, 2)
        worker = np.sort(rng.integers(0, nw, n))
        firm = rng.integers(0, nf, n)
        year = rng.integers(0, 10, n)
        X = rng.standard_normal((n, n_cov)).astype(np.float32)
        a = rng.normal(0.0, 0.2, nw)[worker]
        p = rng.normal(0.0, 0.2, nf)[firm]
        mu = np.exp(0.5 + 0.1 * X[:, 0].astype(np.float64) + a + p)
        y = rng.poisson(mu).astype(np.float64)
        out.append({"worker_id": worker.astype(np.int64), "firm_id": firm.astype(np.int64),
                    "year": year.astype(np.int64), "X": X, "y": y})
    return out


def poisson_models_device(n_models: int = 32, lo: int = 50_000, hi: int = 2_000_000,
                          n_cov: int = 300, seed: int = 5) -> list:
    """`poisson_models`' shapes generated on the GPU (torch's seeded CUDA
    generator), for the sizes whose covariates would take minutes to draw on
    the host: rows log-uniform in [lo, hi] (numpy, `seed`), worker rows/8
    levels (rows sorted by worker), firm rows/80, year 10, `n_cov` float32
    covariates N(0, 1), y ~ Poisson(exp(0.5 + 0.1 x1 + alpha + psi)) with
    alpha, psi ~ N(0, 0.2^2).  Device tensors."""
    import torch
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), n_models)).astype(np.int64)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    out = []
    for n in sizes:
        n = max(int(n), 160)
        nw, nf = max(n // 8, 2), max(n // 80, 2)
        worker = torch.sort(torch.randint(0, nw, (n,), generator=gen, device="cuda")).values
        firm = torch.randint(0, nf, (n,), generator=gen, device="cuda")
        year = torch.randint(0, 10, (n,), generator=gen, device="cuda")
        X = torch.randn((n, n_cov), generator=gen, device="cuda", dtype=torch.float32)
        a = 0.2 * torch.randn(nw, generator=gen, device="cuda", dtype=torch.float64)
        p = 0.2 * torch.randn(nf, generator=gen, device="cuda", dtype=torch.float64)
        mu = torch.exp(0.5 + 0.1 * X[:, 0].double() + a[worker] + p[firm])
        y = torch.poisson(mu, generator=gen)
        out.append({"worker_id": worker, "firm_id": firm, "year": year, "X": X, "y": y})
    return out


def ols_models(n_models: int = 32, lo: int = 50_000, hi: int = 2_000_000, n_cov: int = 300,
               seed: int = 6) -> list:
    """Uneven linear models of `poisson_models`' shapes: rows log-uniform in
    [lo, hi], worker rows/8 levels (rows sorted by worker), firm rows/80,
    year 10, `n_cov` float32 covariates N(0, 1), and y = 0.3 x1 - 0.2 x2 +
    alpha + psi + gamma + eps with alpha ~ N(0, 0.5^2), psi ~ N(0, 0.3^2),
    gamma ~ N(0, 0.1^2), eps ~ N(0, 0.4^2); float64."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), n_models))
    out = []
    for n in sizes.astype(np.int64):
        n = max(int(n), 160)
        nw, nf = max(n // 8, 2), max(n // 80, 2)
        worker = np.sort(rng.integers(0, nw, n))
        firm = rng.integers(0, nf, n)
        year = rng.integers(0, 10, n)
        X = rng.standard_normal((n, n_cov)).astype(np.float32)
        y = (0.3 * X[:, 0].astype(np.float64) - 0.2 * X[:, 1].astype(np.float64)
             + rng.normal(0.0, 0.5, nw)[worker] + rng.normal(0.0, 0.3, nf)[firm]
             + rng.normal(0.0, 0.1, 10)[year] + rng.normal(0.0, 0.4, n))
        out.append({"worker_id": worker.astype(np.int64), "firm_id": firm.astype(np.int64),
                    "year": year.astype(np.int64), "X": X, "y": y})
    return out
