"""Variance-covariance matrices of a fit's coefficients on the device (the
reference's step 4: the meats of `_pass2`, core/inference.py:430-668, and
their small-sample factors, :880-920).

With X~ the covariates partialled on the fixed effects (M_F x, per row),
e the residuals, w the weights and B = X~'W X~ (the fit's V'M_D V block):

    iid     rss / (N - K) B^-1
    hetero  N / (N - K) B^-1 M B^-1,          M = sum_i w_i^2 e_i^2 x~_i x~_i'
    CRV1    G/(G-1) (N-1)/(N-Kc) B^-1 M_c B^-1, M_c = sum_c s_c s_c',
            s_c = sum_{i in c} w_i e_i x~_i

K = k + k_fe, k_fe the fixed effects' free parameters: every level less the
redundant ones (core/estimator.py:446-464).  fe_dof="exact" counts the
connected components of the pair of dimensions with the most (every pair's
components on the device), plus one per further dimension; "pyfixest" one
per dimension beyond the first.  For CRV1, Kc drops the levels of fixed
effects nested in the clusters (every level inside one cluster) and adds one
back per nested effect, and under "exact" a pair of two nested dimensions is
not counted; a multi-way request "a+b" combines the terms a, b, a^b by
inclusion-exclusion with G the smallest one-way count (pyfixest's
conventions, which the reference follows).  Scores, cluster sums and the
meats' tall products run on the device in a fixed order (the same result
every run); the k x k algebra on the host.
"""

from __future__ import annotations

import ctypes as C
from itertools import combinations

import numpy as np

from . import _lib

FE_DOF = ("exact", "pyfixest")
SCORES_MAX_K = 8        # covariates the one-pass score kernel takes


def _parse(vcov):
    """('iid' | 'hetero' | 'CRV1', [cluster terms]) from 'iid', 'hetero',
    'HC1', 'CRV1:a+b' or {'CRV1': 'a+b'}."""
    if isinstance(vcov, dict):
        if len(vcov) != 1:
            raise ValueError("vcov dict must have one key")
        (kind, req), = vcov.items()
        vcov = f"{kind}:{req}"
    v = str(vcov).strip()
    if v in ("iid",):
        return "iid", None
    if v in ("hetero", "HC1"):
        return "hetero", None
    if v.startswith("CRV1:"):
        req = v.split(":", 1)[1]
        parts = [[c.strip() for c in p.split("^")] for p in req.split("+") if p.strip()]
        if not parts:
            raise ValueError(f"no cluster variable in {vcov!r}")
        return "CRV1", parts
    raise ValueError(f"unknown vcov {vcov!r}: use 'iid', 'hetero', 'HC1' or 'CRV1:<cols>'")


class VcovEngine:
    """The pieces every vcov of one fit shares: X~, the scores w e x~ and B
    (device), the fixed effects' dof."""

    @property
    def res(self):
        r = self._res()
        if r is None:
            raise RuntimeError("the fit this vcov engine belongs to is gone")
        return r

    def __init__(self, result, data=None):
        import torch
        import weakref
        # the result owns the engine: a weak reference back, so that the
        # pair is freed (device arrays included) as soon as the result is,
        # without waiting for the cycle collector
        self._res = weakref.ref(result)
        panel = result.panel
        info = panel.info
        self.n, self.k = info["n"], len(result.coefnames)
        m = info["m"]
        self.data = data
        self.w = panel.tensor("w", torch.float64) if info["weighted"] else None
        self._e = result.resid_sorted.to(torch.float64).contiguous()
        self.N = self.n
        xj = list(range(1, m))
        self.B = np.asarray(result.A)[np.ix_(xj, xj)] if self.k else np.zeros((0, 0))
        self._cols = np.arange(1, m, dtype=np.int32)
        self._U = None
        self.meat_g = None           # the streamed dimension's clusters' meat
        if 0 < self.k <= SCORES_MAX_K:
            # one pass: rss, the hetero meat and the streamed clusters' meat
            self.rss, self.meat_hc, self.meat_g = self._scores(None)
        else:
            e = self._e
            self.rss = float(((e * e) * (self.w if self.w is not None else 1.0)).sum())
            self.meat_hc = None
        # the fixed effects' parameters: every level, less the redundant
        G = info["G"]
        nlev = [int(v) for v in panel.n_levels]
        self.levels = {result.fe_names[0]: G}
        for t, name in enumerate(result.fe_names[1:]):
            self.levels[name] = nlev[t]
        self.ncomp = max(int(result.n_components), 1)
        self.D = len(result.fe_names)

    def _scores(self, U):
        """(rss, sum_i u_i u_i', sum_g s_g s_g') by hdfe_panel_scores; U (n,
        k) on the device receives the scores u_i when given."""
        import torch
        k = self.k
        out = np.zeros(1 + 2 * k * k)
        torch.cuda.current_stream().synchronize()
        _lib.check(_lib.load().hdfe_panel_scores(
            self.res.panel._h, C.c_void_p(self.res.Gamma.data_ptr()), self._cols.ctypes.data, k,
            C.c_void_p(self._e.data_ptr()), C.c_void_p(U.data_ptr() if U is not None else None),
            out.ctypes.data), "hdfe_panel_scores")
        return float(out[0]), out[1:1 + k * k].reshape(k, k), out[1 + k * k:].reshape(k, k)

    @property
    def U(self):
        """The scores w_i e_i x~_i (n, k) on the device, made on first use."""
        import torch
        if self._U is None and self.k:
            U = torch.empty((self.n, self.k), dtype=torch.float64, device="cuda")
            if self.k <= SCORES_MAX_K:
                self._scores(U)
            else:
                torch.cuda.current_stream().synchronize()
                _lib.check(_lib.load().hdfe_panel_partial(
                    self.res.panel._h, C.c_void_p(self.res.Gamma.data_ptr()),
                    self._cols.ctypes.data, self.k, C.c_void_p(U.data_ptr())),
                    "hdfe_panel_partial")
                f = self._e if self.w is None else self.w * self._e
                U *= f[:, None]
            self._U = U
        return self._U

    def pair_components(self):
        """{(dimension a, dimension b): connected components} for every pair
        of the fit's fixed effects (the first pair's from the fit)."""
        if getattr(self, "_pair_comps", None) is None:
            names = self.res.fe_names
            out = {}
            for a in range(self.D):
                for b in range(a + 1, self.D):
                    if (a, b) == (0, 1):
                        c = self.ncomp
                    else:
                        v = C.c_int64()
                        _lib.check(_lib.load().hdfe_panel_pair_components(
                            self.res.panel._h, a, b, C.byref(v)), "hdfe_panel_pair_components")
                        c = int(v.value)
                    out[(names[a], names[b])] = c
            self._pair_comps = out
        return self._pair_comps

    def k_fe(self, nested=(), fe_dof="exact"):
        total = sum(self.levels.values())
        if self.D < 2:
            return total
        if fe_dof == "pyfixest":
            return total - (self.D - 1)
        # the pair with the most components, a pair of two nested dimensions
        # aside (its redundant levels are among those not counted)
        comps = max((c for (a, b), c in self.pair_components().items()
                     if a not in nested or b not in nested), default=1)
        return total - (comps + self.D - 2)

    def _codes(self, cols):
        """Each sorted row's code of the cluster formed by `cols` (fixed
        effects of the fit, or columns of the data the fit was given); codes
        are dense, 0 .. clusters - 1.  Kept per column set."""
        key = tuple(cols)
        cache = self.__dict__.setdefault("_code_cache", {})
        if key not in cache:
            cache[key] = self._make_codes(cols)
        return cache[key]

    def _make_codes(self, cols):
        import torch
        res, panel = self.res, self.res.panel
        perm = panel.tensor("perm", torch.int32).long()[:self.n]
        keys = []
        for c in cols:
            if c == res.fe_names[0]:
                keys.append(panel.tensor("g", torch.int32).long()[:self.n])
            elif c in res.fe_names:
                t = res.fe_names.index(c) - 1
                keys.append(panel.tensor("codes", torch.int32).view(-1, self.n)[t].long())
            else:
                if self.data is None:
                    raise ValueError(f"cluster column {c!r} is not a fixed effect: pass the "
                                     "fit's data to vcov()")
                from .estimator import _column
                col = _column(self.data, c)
                v = (col if torch.is_tensor(col) else torch.from_numpy(np.asarray(col))).to(
                    "cuda")
                keys.append(torch.unique(v[perm], return_inverse=True)[1])
        code = keys[0]
        for kk in keys[1:]:
            code = torch.unique(torch.stack([code, kk], 1), dim=0, return_inverse=True)[1]
        if len(keys) == 1:
            code = torch.unique(code, return_inverse=True)[1]
        return code

    def _meat(self, cols):
        """sum_c s_c s_c' and the cluster count, s_c the scores' sum over the
        cluster's rows in a fixed order (no float atomics): the streamed
        dimension's groups are the panel's contiguous row ranges; any other
        clustering is sorted stably by cluster first."""
        import torch
        from .leaveout_se import tall_dot
        res, panel = self.res, self.res.panel
        if list(cols) == [res.fe_names[0]] and self.meat_g is not None:
            return self.meat_g, panel.info["G"]
        if list(cols) == [res.fe_names[0]]:
            G = panel.info["G"]
            S = torch.empty((G, self.k), dtype=torch.float64, device="cuda")
            torch.cuda.current_stream().synchronize()
            _lib.check(_lib.load().hdfe_kss_group_sums(G, self.k, panel.ptr("gstart")[0], None,
                                                       None, self.U.data_ptr(), S.data_ptr()),
                       "hdfe_kss_group_sums")
        else:
            code = self._codes(cols)
            G = int(code.max()) + 1 if code.numel() else 0
            order = torch.argsort(code, stable=True)
            lengths = torch.bincount(code, minlength=G)
            S = torch.segment_reduce(self.U[order], "sum", lengths=lengths, axis=0)
        return tall_dot(S, S).cpu().numpy(), G

    def _n_clusters(self, cols):
        if len(cols) == 1 and cols[0] in self.levels:
            return self.levels[cols[0]]
        code = self._codes(cols)
        return int(code.max()) + 1 if code.numel() else 0

    def _nested(self, cols):
        """Fixed effects all of whose levels lie in one cluster of `cols`: as
        many distinct (level, cluster) pairs over the rows as levels.  Every
        cluster holds a level and a nested level lies in one cluster, so a
        fixed effect with fewer levels than there are clusters is not nested
        (no pass over the rows)."""
        import torch
        out = []
        nc = None
        for name in self.res.fe_names:
            if set(cols) <= {name}:
                out.append(name)
                continue
            if nc is None:
                nc = self._n_clusters(cols)
            if self.levels[name] < nc:
                continue
            code = self._codes(cols)
            lv = self._codes([name])
            pairs = torch.unique(lv * nc + code).numel()
            if pairs == int(lv.max()) + 1:
                out.append(name)
        return out

    def vcov(self, spec, fe_dof="exact"):
        if fe_dof not in FE_DOF:
            raise ValueError(f"fe_dof must be one of {FE_DOF}")
        kind, parts = _parse(spec)
        k, N = self.k, self.N
        if not k:
            return np.zeros((0, 0)), {"kind": kind}
        Binv = np.linalg.inv(self.B)
        K = k + self.k_fe(fe_dof=fe_dof)
        ratio = (lambda a, b: a / b if b else np.nan)
        if kind == "iid":
            return ratio(self.rss, N - K) * Binv, {"kind": kind, "K": K, "df": N - K}
        if kind == "hetero":
            from .leaveout_se import tall_dot
            if self.meat_hc is None:
                self.meat_hc = tall_dot(self.U, self.U).cpu().numpy()
            M = self.meat_hc
            return ratio(N, N - K) * Binv @ M @ Binv, {"kind": kind, "K": K, "df": N - K}
        terms = []
        for r in range(1, len(parts) + 1):
            for combo in combinations(range(len(parts)), r):
                cols = list(dict.fromkeys(c for i in combo for c in parts[i]))
                terms.append((cols, (-1) ** (r + 1), r == 1))
        meats, Gs = [], []
        for cols, sign, single in terms:
            M, G = self._meat(cols)
            meats.append((sign, M))
            if single:
                Gs.append(G)
        Gm = min(Gs)
        nest = list(dict.fromkeys(d for cols, _, _ in terms for d in self._nested(cols)))
        Kc = k + self.k_fe(nest, fe_dof) - sum(self.levels[d] for d in nest) + len(nest)
        adj = ratio(Gm, Gm - 1) * ratio(N - 1, N - Kc)
        V = sum(sign * adj * (Binv @ M @ Binv) for sign, M in meats)
        return V, {"kind": "CRV1", "clusters": Gs, "Kc": Kc, "adj": adj, "df": Gm - 1,
                   "nested": nest}
