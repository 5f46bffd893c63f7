"""The leave-one-out connected set on the device (core/connected.py:378-562).

`leave_one_out(workers, firms)` returns which rows survive the reference's
pruning -- the largest connected component by rows, then workers with a
single row dropped, then rounds of dropping workers whose removal would
disconnect the network (articulation points), each followed by keeping the
largest component again -- and the same counts the reference records
(`components_before`, `single_observation`, `cut`, `rounds`)."""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib


def leave_one_out(workers, firms):
    """(keep: device bool tensor per row, info dict)."""
    import torch
    from .estimator import _to_device
    a = _to_device(workers, torch.int64)
    b = _to_device(firms, torch.int64)
    n = int(a.shape[0])
    keep = torch.empty(n, dtype=torch.uint8, device="cuda")
    info = (C.c_int64 * 8)()
    torch.cuda.synchronize()
    _lib.check(_lib.load().hdfe_loo_prune(n, a.data_ptr(), b.data_ptr(), keep.data_ptr(), info),
               "hdfe_loo_prune")
    names = ["components_before", "single_observation", "cut", "rounds", "workers", "firms",
             "pairs", "vertices"]
    return keep.bool(), dict(zip(names, [int(v) for v in info]))


COMPONENT = "component"
IN_LARGEST = "in_largest"


def _parse_pair(fe):
    """The two dimensions and the columns of each, from "worker_id + firm_id^year"
    or a pair of names (core/connected.py:341-364)."""
    if isinstance(fe, str):
        if "~" in fe or "|" in fe:
            raise ValueError(f"fe={fe!r}: pass only the fixed effects, e.g. "
                             "'worker_id + firm_id'")
        terms = [t.strip() for t in fe.split("+") if t.strip()]
    else:
        terms = [str(t).strip() for t in fe]
    if len(terms) != 2:
        raise ValueError(f"fe must name exactly two fixed-effect dimensions, got {terms}. "
                         "Components are those of a pair; with three or more "
                         "dimensions, pass the pair whose connectedness matters "
                         "(e.g. workers and firms)")
    for t in terms:
        if "[" in t or "/" in t:
            raise ValueError(f"{t!r}: varying slopes have no bearing on which levels are "
                             "connected; pass the dimension alone")
    if terms[0] == terms[1]:
        raise ValueError(f"the two dimensions are the same: {terms[0]!r}")
    return tuple(terms), {d: [c.strip() for c in d.split("^")] for d in terms}


def _min_degree(min_degree):
    """(first, second) from an int or a pair (core/connected.py:367-375)."""
    k = ((min_degree, min_degree) if isinstance(min_degree, (int, np.integer))
         else tuple(min_degree))
    if (len(k) != 2 or not all(isinstance(v, (int, np.integer)) and not isinstance(v, bool)
                               and v >= 1 for v in k)):
        raise ValueError("min_degree must be a positive integer or a pair of them, "
                         f"got {min_degree!r}")
    return int(k[0]), int(k[1])


def _host_column(data, name):
    from .estimator import _column
    col = _column(data, name)
    if hasattr(col, "detach"):
        col = col.detach().cpu().numpy()
    return np.asarray(col)


def _complete_mask(cols):
    """Rows with no null or NaN in any of the columns (they join nothing)."""
    import pandas as pd
    ok = np.ones(len(cols[0]), bool)
    for c in cols:
        if c.dtype.kind in "fcOUSM":
            ok &= ~pd.isna(c)
    return ok


class _Coding:
    """A dimension's level keys: one column's integer ids as they are, or the
    rank of each column's value among its distinct values, combined in the
    columns' order (so keys sort as the levels do, lexicographically)."""

    def __init__(self, cols):
        self.uniques, codes, self.radix = [], [], []
        direct = len(cols) == 1 and cols[0].dtype.kind in "iu"
        self.direct = direct
        if direct:
            self.key = cols[0].astype(np.int64)
            return
        key = np.zeros(len(cols[0]), np.int64)
        for c in cols:
            u, inv = np.unique(c, return_inverse=True)
            self.uniques.append(u)
            key = key * len(u) + inv
        self.key = key

    def keys_of(self, cols):
        """The keys of other rows (-1 where a value was never seen)."""
        if self.direct:
            return cols[0].astype(np.int64)
        key = np.zeros(len(cols[0]), np.int64)
        bad = np.zeros(len(cols[0]), bool)
        for c, u in zip(cols, self.uniques):
            pos = np.clip(np.searchsorted(u, c), 0, len(u) - 1)
            bad |= u[pos] != c
            key = key * len(u) + pos
        return np.where(bad, -1, key)

    def columns(self, keys, names):
        """The level table's columns for sorted keys."""
        if self.direct:
            return {names[0]: keys}
        out = {}
        rest = keys.copy()
        for name, u in reversed(list(zip(names, self.uniques))):
            out[name] = u[rest % len(u)]
            rest //= len(u)
        return {n: out[n] for n in names}


@dataclass
class ConnectedComponents:
    """The connected components of the graph of two fixed-effect dimensions
    (core/connected.py:45-208): per-level tables, not rows.  `levels[d]` is a
    DataFrame of the dimension's columns and `component` (nullable UInt32,
    0 the largest; missing for levels removed by `min_degree` or
    `leave_one_out`); `sizes` one row per component, in order: `component`,
    `rows`, and the levels of each dimension."""

    fe: tuple
    levels: dict
    sizes: object
    min_degree: tuple = (1, 1)
    leave_one_out: bool = False
    diagnostics: dict = field(default_factory=dict)
    _codings: dict = field(default=None, repr=False)
    _cols: dict = field(default=None, repr=False)
    _level_keys: dict = field(default=None, repr=False)
    _source: object = field(default=None, repr=False)

    @property
    def _pruned(self):
        return self.min_degree != (1, 1) or self.leave_one_out

    @property
    def n_components(self) -> int:
        return len(self.sizes)

    def _data(self, data):
        if data is not None:
            return data
        if self._source is None:
            raise ValueError("these components hold no reference to the data they were "
                             "found in; pass the data")
        return self._source

    def _row_components(self, data):
        """Each row's component (-1: none) and the data's columns."""
        import pandas as pd
        names = list(dict.fromkeys(c for d in self.fe for c in self._cols[d]))
        cols = {}
        for c in names:
            try:
                cols[c] = _host_column(data, c)
            except (KeyError, ValueError):
                raise ValueError(f"columns not found in data: {[c]}") from None
        ok = _complete_mask([cols[c] for c in names])
        comp = np.where(ok, 0.0, np.nan)
        for d in self.fe:
            keys = self._codings[d].keys_of([cols[c] for c in self._cols[d]])
            lk = self._level_keys[d]
            pos = np.clip(np.searchsorted(lk, keys), 0, len(lk) - 1)
            found = (lk[pos] == keys) & (keys >= 0)
            cd = self.levels[d][COMPONENT].to_numpy(dtype="float64", na_value=np.nan)
            c = np.where(found, cd[pos], np.nan)
            # both levels must be in a component (they are then in the same)
            comp = np.where(np.isnan(c), np.nan, np.where(np.isnan(comp), np.nan, c))
        return pd.array(np.where(np.isnan(comp), None, comp), dtype="UInt32"), names

    def _frame(self, data):
        import pandas as pd
        if isinstance(data, pd.DataFrame):
            return data.reset_index(drop=True)
        return pd.DataFrame({k: _host_column(data, k) for k in data})

    def label(self, data=None):
        """The data, in its order, with `component` (null for a row in no
        component) and `in_largest` added (core/connected.py:150-177)."""
        data = self._data(data)
        comp, _ = self._row_components(data)
        df = self._frame(data)
        clash = [c for c in (COMPONENT, IN_LARGEST) if c in df.columns]
        if clash:
            raise ValueError(f"the data has columns named {clash}, which label() adds; "
                             "rename them")
        df[COMPONENT] = comp
        df[IN_LARGEST] = (comp == 0).fillna(False).to_numpy(dtype=bool)
        return df

    def largest(self, data=None):
        """The rows in the largest component, in their order: the largest
        connected set (core/connected.py:122-148)."""
        data = self._data(data)
        comp, _ = self._row_components(data)
        keep = (comp == 0).fillna(False).to_numpy(dtype=bool)
        return self._frame(data)[keep].reset_index(drop=True)

    def summary(self) -> str:
        """A short report: the components, the largest, and what pruning
        removed (core/connected.py:179-208)."""
        d, s = self.diagnostics, self.sizes
        d0, d1 = self.fe
        lines = [f"connected components of {d0} x {d1}"]
        if self.min_degree != (1, 1):
            lines.append(f"  min_degree={self.min_degree} removed "
                         f"{d['pruned_levels'][d0]:,} {d0}, {d['pruned_levels'][d1]:,} {d1} "
                         f"and {d['pruned_rows']:,} rows in {d['pruning_rounds']} round(s)")
        if self.leave_one_out:
            lo = d["leave_one_out"]
            lines.append(f"  leave-one-out connected set: from {lo['components_before']:,} "
                         f"components, kept the largest; removed "
                         f"{lo['single_observation']:,} {d0} observed once and "
                         f"{lo['cut']:,} {d0} whose removal would disconnect it, in "
                         f"{lo['rounds']} round(s); {lo['levels'][d0]:,} {d0}, "
                         f"{lo['levels'][d1]:,} {d1} and {lo['rows']:,} rows removed")
        lines.append(f"  components: {self.n_components:,}")
        rows = d["rows"]
        if self.n_components:
            top = s.iloc[0]
            lines.append(f"  largest: {int(top['rows']):,} of {rows:,} rows "
                         f"({top['rows'] / max(rows, 1):.1%}), "
                         f"{int(top[d0]):,} of {d['levels'][d0]:,} {d0}, "
                         f"{int(top[d1]):,} of {d['levels'][d1]:,} {d1}")
        if self.n_components > 1:
            lines.append(f"  next largest: {int(s.iloc[1]['rows']):,} rows")
        return "\n".join(lines)


def connected_components(data, fe, *, min_degree=1, leave_one_out=False, rank_by="rows",
                         batch_rows=None, verbose=False, logger=None, log_level=None):
    """Connected components of the graph of two fixed-effect dimensions on the
    device (core/connected.py:211-338): a vertex per level, an edge per row;
    `min_degree` prunes to the (k0, k1)-core of the distinct pairs first,
    `leave_one_out` then keeps the leave-one-out connected set (the first
    dimension plays the workers).  Components are numbered by size, rows or
    the levels of `rank_by`, ties going to the component whose smallest
    first-dimension level sorts first.  `batch_rows`, `verbose`, `logger`
    and `log_level` are accepted for the reference's signature; the rows are
    on the device at once."""
    import time

    import pandas as pd
    import torch
    t0 = time.time()
    dims, fe_cols = _parse_pair(fe)
    k = _min_degree(min_degree)
    if rank_by != "rows" and rank_by not in dims:
        raise ValueError(f"rank_by must be 'rows' or one of {list(dims)}, got {rank_by!r}")
    names = list(dict.fromkeys(c for d in dims for c in fe_cols[d]))
    cols = {}
    for c in names:
        try:
            cols[c] = _host_column(data, c)
        except (KeyError, ValueError):
            raise ValueError(f"columns not found in data: {[c]}") from None
    ok = _complete_mask([cols[c] for c in names])
    if not ok.all():
        cols = {c: v[ok] for c, v in cols.items()}
    codings = {d: _Coding([cols[c] for c in fe_cols[d]]) for d in dims}
    n = int(ok.sum())
    if n == 0:
        raise ValueError("no complete rows: every row has a null or NaN level")
    a = torch.from_numpy(codings[dims[0]].key).cuda()
    b = torch.from_numpy(codings[dims[1]].key).cuda()
    lib = _lib.load()
    h = C.c_void_p()
    torch.cuda.synchronize()
    rb = -1 if rank_by == "rows" else dims.index(rank_by)
    _lib.check(lib.hdfe_components(n, a.data_ptr(), b.data_ptr(), k[0], k[1],
                                   int(bool(leave_one_out)), rb, C.byref(h)), "hdfe_components")
    try:
        info = (C.c_int64 * 16)()
        _lib.check(lib.hdfe_components_info(h, info), "hdfe_components_info")
        info = [int(v) for v in info]
        n0, n1, nc = info[0], info[1], info[2]
        level_keys, comps = {}, {}
        for t, (d, nd) in enumerate(zip(dims, (n0, n1))):
            lv = np.empty(nd, np.int64)
            cp = np.empty(nd, np.int64)
            _lib.check(lib.hdfe_components_levels(h, t, lv.ctypes.data, cp.ctypes.data),
                       "hdfe_components_levels")
            level_keys[d], comps[d] = lv, cp
        sz = np.empty((max(nc, 1), 3), np.int64)
        _lib.check(lib.hdfe_components_sizes(h, sz.ctypes.data), "hdfe_components_sizes")
        sz = sz[:nc]
    finally:
        lib.hdfe_components_destroy(h)
    levels = {}
    for d in dims:
        table = pd.DataFrame(codings[d].columns(level_keys[d], fe_cols[d]))
        cp = comps[d]
        table[COMPONENT] = pd.array(np.where(cp >= 0, cp, None), dtype="UInt32")
        levels[d] = table
    sizes = pd.DataFrame({COMPONENT: np.arange(nc, dtype=np.uint32), "rows": sz[:, 0],
                          dims[0]: sz[:, 1], dims[1]: sz[:, 2]})
    diag = {"levels": {dims[0]: n0, dims[1]: n1}, "min_degree": list(k), "rows": info[3]}
    if k != (1, 1) or leave_one_out:
        diag["distinct_pairs"] = info[4]
    if k != (1, 1):
        diag["pruning_rounds"] = info[5]
        diag["pruned_levels"] = {dims[0]: info[6], dims[1]: info[7]}
        diag["pruned_rows"] = info[8]
    if leave_one_out:
        diag["leave_one_out"] = {"components_before": info[9], "single_observation": info[10],
                                 "cut": info[11], "rounds": info[12],
                                 "levels": {dims[0]: info[13], dims[1]: info[14]},
                                 "rows": info[15]}
    diag["seconds"] = round(time.time() - t0, 2)
    cc = ConnectedComponents(fe=dims, levels=levels, sizes=sizes, min_degree=k,
                             leave_one_out=bool(leave_one_out), diagnostics=diag,
                             _codings=codings, _cols=fe_cols, _level_keys=level_keys,
                             _source=data)
    if verbose and nc:
        print(f"connected components: {nc:,}; the largest has {int(sz[0, 0]):,} of "
              f"{info[3]:,} rows")
    return cc
