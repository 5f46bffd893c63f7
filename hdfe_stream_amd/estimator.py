"""`device="rocm"` fits: the reference's estimator interface over the HIP path.

Mirrors the parts of hdfe_stream's public API that reach the solve path:
`feols_stream` (api.py:22-107) and `StreamingHDFE(...).fit` (core/estimator.py:
69-246, 467-546) with the same options, validation and solver-info keys, and
the same FP64 algorithm -- pass 0 (singletons, codes, row sort), pass 1
(cells), step 2 (explicit S, block Jacobi-PCG, normalisation), step 3 (V' M_D V
and beta) and pass 2 (streamed effects, residuals) -- all of it on the device
(libhdfe_hip.so), with the panel's rows resident in HBM.

Host columns and Parquet files reach HBM through the batched ingest (`ingest.py`).
Not mirrored (outside the solve path): Polars buckets and hash partitions, the
pyfixest formula language beyond `y ~ x1 + x2 | fe1 + fe2 + ...`, standard
errors, IV, collinearity checks.
"""

from __future__ import annotations

import ctypes as C
import re
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .rocm import DeviceCSR, PRECISIONS, rocm_backend

RHS_ZERO_TOL = 1e-11      # core/solve.py:44


def parse_formula(fml: str):
    """`y ~ x1 + x2 | fe1 + fe2` -> ("y", ["x1", "x2"], ["fe1", "fe2"])."""
    m = re.fullmatch(r"\s*([^~|]+?)\s*~\s*([^|]*?)\s*(?:\|\s*(.*?)\s*)?", fml)
    if not m:
        raise ValueError(f"cannot parse formula {fml!r}")
    y = m.group(1).strip()
    rhs = [t.strip() for t in m.group(2).split("+") if t.strip()]
    xs = [t for t in rhs if t not in ("1", "0")]
    fe = [t.strip() for t in (m.group(3) or "").split("+") if t.strip()]
    return y, xs, fe


def _column(data, name):
    """A column of `data` (dict, pandas DataFrame, or Parquet path) as a
    numpy array or torch tensor, unchanged."""
    if isinstance(data, (str, bytes)) or hasattr(data, "__fspath__"):
        import pyarrow.parquet as pq
        return pq.read_table(data, columns=[name]).column(name).to_numpy()
    col = data[name]
    if hasattr(col, "to_numpy") and not type(col).__module__.startswith("torch"):
        return col.to_numpy()
    return col


def _to_device(a, dtype):
    import torch
    if isinstance(a, torch.Tensor):
        return a.to(device="cuda", dtype=dtype).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)


def _pinned(n, dtype):
    """A numpy array of n elements in page-locked host memory (it keeps its
    torch buffer alive)."""
    import torch
    t = torch.empty(n, dtype=torch.from_numpy(np.empty(0, dtype)).dtype, pin_memory=True)
    return t.numpy()


_side_stream = None


def _host_async(t):
    """Start a device tensor's copy to page-locked host memory on a side
    stream, after the work queued so far on the current stream; returns the
    array and an event (None for an empty tensor) that must complete before
    the array is read."""
    import torch
    global _side_stream
    if _side_stream is None:
        _side_stream = torch.cuda.Stream()
    pin = torch.empty(t.numel(), dtype=t.dtype, pin_memory=True)
    ev = None
    if t.numel():
        cur = torch.cuda.current_stream()
        _side_stream.wait_stream(cur)
        with torch.cuda.stream(_side_stream):
            pin.copy_(t.reshape(-1), non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(_side_stream)
        t.record_stream(_side_stream)
    return pin.numpy().reshape(tuple(t.shape)), ev


def _host(t):
    """A device tensor's host copy through page-locked memory."""
    out = _pinned(t.numel(), np.dtype(str(t.dtype).replace("torch.", "")))
    if t.numel():
        import torch
        torch.from_numpy(out).copy_(t.reshape(-1))
    return out.reshape(tuple(t.shape))


class Panel:
    """A panel's rows on the device after passes 0 and 1 (hdfe_panel_*)."""

    INFO = ["n_in", "n", "G", "C", "M", "Gi", "L", "D", "gdim", "singletons",
            "single_rounds", "E", "nnz", "ncomp", "m", "weighted"]

    def __init__(self, fe_cols, var_cols, weights=None, fixef_rm="singleton", stream_dim=-1,
                 sort_vars=False):
        import torch
        lib = _lib.load()
        self._keep = []
        fe_t = [_to_device(c, torch.int64) for c in fe_cols]
        var_t = [_to_device(c, torch.float64) for c in var_cols]
        w_t = None
        w32 = 0
        if weights is not None:
            if getattr(weights, "dtype", None) in (np.float32, torch.float32):
                w_t, w32 = _to_device(weights, torch.float32), 1
            else:
                w_t = _to_device(weights, torch.float64)
        self._keep += fe_t + var_t + ([w_t] if w_t is not None else [])
        n = int(fe_t[0].shape[0])
        fe_p = (C.c_void_p * len(fe_t))(*[t.data_ptr() for t in fe_t])
        var_p = (C.c_void_p * len(var_t))(*[t.data_ptr() for t in var_t])
        h = C.c_void_p()
        torch.cuda.synchronize()
        lib.hdfe_panel_create.restype = C.c_int
        _lib.check(lib.hdfe_panel_create(
            C.c_int64(n), C.c_int(len(fe_t)), fe_p, C.c_int(len(var_t)), var_p,
            C.c_void_p(w_t.data_ptr() if w_t is not None else None), C.c_int(w32),
            C.c_int(1 if fixef_rm == "singleton" else 0), C.c_int(stream_dim),
            C.c_int(1 if sort_vars else 0), C.byref(h)), "hdfe_panel_create")
        self._h = h
        self._keep = []      # the panel holds its own sorted copies
        self.matrix = None
        self.refresh()

    def refresh(self):
        info = (C.c_int64 * 16)()
        nlev = (C.c_int64 * 8)()
        dims = (C.c_int * 8)()
        _lib.check(_lib.load().hdfe_panel_info(self._h, info, nlev, dims), "hdfe_panel_info")
        self.info = dict(zip(self.INFO, [int(v) for v in info]))
        D = self.info["D"]
        self.n_levels = [int(nlev[t]) for t in range(D)]
        self.dims = [int(dims[t]) for t in range(D)]
        self.offs = np.concatenate(([0], np.cumsum(self.n_levels)[:-1])).astype(np.int64)

    def ptr(self, name):
        p = C.c_void_p()
        n = C.c_int64()
        _lib.check(_lib.load().hdfe_panel_array(self._h, name.encode(), C.byref(p), C.byref(n)),
                   "hdfe_panel_array")
        return p.value, int(n.value)

    def array(self, name, dtype, pinned=False):
        """A copy of a panel array on the host; with `pinned`, in page-locked
        memory from torch's caching host allocator (a DMA copy at full link
        speed, where pageable memory goes through staging buffers)."""
        p, n = self.ptr(name)
        if pinned and n:
            out = _pinned(n, dtype)
        else:
            out = np.empty(n, dtype)
        if n:
            _lib.check(_lib.load().hdfe_memcpy(out.ctypes.data, C.c_void_p(p), out.nbytes, 0),
                       "hdfe_memcpy")
        return out

    def array_async(self, name, dtype):
        """A panel array's copy to page-locked host memory, started now on a
        stream beside the panel's own (ready after copy_wait)."""
        p, n = self.ptr(name)
        out = _pinned(n, dtype) if n else np.empty(0, dtype)
        # held with the panel, whose release waits for the device first, so
        # the buffer outlives the copy however the fit ends
        self._keep.append(out)
        if n:
            _lib.check(_lib.load().hdfe_panel_copy_async(self._h, C.c_void_p(p),
                                                         out.ctypes.data, out.nbytes),
                       "hdfe_panel_copy_async")
        return out

    def copy_wait(self):
        _lib.check(_lib.load().hdfe_panel_copy_wait(self._h), "hdfe_panel_copy_wait")

    def tensor(self, name, dtype):
        """A copy of a panel array on the device (torch)."""
        import torch
        p, n = self.ptr(name)
        out = torch.empty(n, dtype=dtype, device="cuda")
        if n:
            torch.cuda.current_stream().synchronize()
            _lib.check(_lib.load().hdfe_memcpy(C.c_void_p(out.data_ptr()), C.c_void_p(p),
                                               out.numel() * out.element_size(), 2), "hdfe_memcpy")
        return out

    def system(self, with_rhs=True, budget=None):
        """diag, rhs and the explicit S (`matrix`).  With `budget` (bytes) an
        S whose build would take more is not built: `matrix` is None and
        `s_skipped` its estimate (the caller's fallback is stream_matrix)."""
        h = C.c_void_p()
        secs = C.c_double()
        lib = _lib.load()
        _lib.check(lib.hdfe_panel_budget(self._h, -1 if budget is None else int(budget), None,
                                         None), "hdfe_panel_budget")
        _lib.check(lib.hdfe_panel_system(self._h, int(with_rhs), C.byref(h), C.byref(secs)),
                   "hdfe_panel_system")
        skipped = C.c_int64()
        _lib.check(lib.hdfe_panel_budget(self._h, -2, None, C.byref(skipped)), "hdfe_panel_budget")
        self.s_skipped = int(skipped.value)
        self.refresh()
        if h.value:
            self.matrix = DeviceCSR(h, self.info["L"], 0, owned=False)
            self.matrix.nnz = self.info["nnz"]
        else:
            self.matrix = None
        return float(secs.value)

    def stream_matrix(self, reuse=False):
        """The matrix-free S over the identifying cells (hdfe_stream_create):
        the same solves as the explicit S, for an S too large to build (the
        reference's stream_cg, core/solve.py:251-271).  Kept as `matrix`.
        With `reuse`, the panel's previous operator is rebound to the current
        cells when their structure is the same (hdfe_stream_rebind: a
        reweighting changes only the weights), keeping its plan."""
        import torch
        info = self.info
        L, Gi, Mc, D = info["L"], info["Gi"], info["M"], info["D"]
        diag = self.tensor("diag", torch.float64)
        self._mf_dinv = torch.where(diag > 1e-12, 1.0 / torch.where(diag > 1e-12, diag,
                                                                     torch.ones_like(diag)),
                                    torch.zeros_like(diag)).contiguous()
        torch.cuda.synchronize()
        prev = getattr(self, "_stream", None)
        if reuse and prev is not None and prev._h:
            ok = C.c_int(0)
            _lib.check(_lib.load().hdfe_stream_rebind(
                prev._h, L, Gi, Mc, D, self.ptr("igstart")[0], self.ptr("icodes")[0], Mc,
                self.ptr("in")[0], self._mf_dinv.data_ptr(), C.byref(ok)), "hdfe_stream_rebind")
            if ok.value:
                self.matrix = prev
                return self.matrix
        offs = np.ascontiguousarray(self.offs, np.int64)
        h = C.c_void_p()
        _lib.check(_lib.load().hdfe_stream_create(
            L, Gi, Mc, D, self.ptr("igstart")[0], self.ptr("icodes")[0], Mc, offs.ctypes.data,
            self.ptr("in")[0], self._mf_dinv.data_ptr(), C.byref(h)), "hdfe_stream_create")
        self.matrix = DeviceCSR(h, L, 0, owned=True)
        self._stream = self.matrix
        return self.matrix

    def components(self):
        n = C.c_int64()
        _lib.check(_lib.load().hdfe_panel_components(self._h, C.byref(n)), "hdfe_panel_components")
        self.refresh()
        return int(n.value)

    def normalize(self, G):
        import torch
        torch.cuda.synchronize()
        _lib.check(_lib.load().hdfe_panel_normalize(self._h, C.c_void_p(G.data_ptr()),
                                                    C.c_int(G.shape[1]), None),
                   "hdfe_panel_normalize")

    def assemble(self, G):
        import torch
        m = self.info["m"]
        A = np.empty((m, m))
        torch.cuda.synchronize()
        _lib.check(_lib.load().hdfe_panel_assemble(self._h, C.c_void_p(G.data_ptr()),
                                                   A.ctypes.data), "hdfe_panel_assemble")
        return A

    def pass2(self, G, beta, xcols):
        import torch
        L, Gn, n = self.info["L"], self.info["G"], self.info["n"]
        gam_y = torch.empty(L, dtype=torch.float64, device="cuda")
        geff = torch.empty(Gn, dtype=torch.float64, device="cuda")
        resid = torch.empty(n, dtype=torch.float64, device="cuda")
        b = np.ascontiguousarray(beta, np.float64)
        xc = np.ascontiguousarray(xcols, np.int32)
        torch.cuda.synchronize()
        _lib.check(_lib.load().hdfe_panel_pass2(
            self._h, C.c_void_p(G.data_ptr()), b.ctypes.data, xc.ctypes.data, C.c_int(len(xc)),
            C.c_void_p(gam_y.data_ptr()), C.c_void_p(geff.data_ptr()),
            C.c_void_p(resid.data_ptr())), "hdfe_panel_pass2")
        return gam_y, geff, resid

    def close(self):
        if getattr(self, "_h", None):
            if self.matrix is not None:
                self.matrix.close()         # the handle itself is the panel's
            prev = getattr(self, "_stream", None)
            if prev is not None and prev is not self.matrix:
                prev.close()
            self._stream = None
            _lib.load().hdfe_panel_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class HDFEResult:
    """What a fit returns: coefficients, effects, solver facts."""
    coefnames: list
    beta: np.ndarray
    solver_info: dict
    fe_names: list
    fe_levels: dict
    fe_values: dict
    n_obs: int
    n_components: int
    A: np.ndarray
    diagnostics: dict = field(default_factory=dict)
    panel: object = None
    Gamma: object = None
    resid_sorted: object = None
    gam_y: object = None
    geff: object = None
    vcov_matrix: object = None      # the fit's default vcov (its vcov= option)
    vcov_details: dict = None
    vcovs: dict = field(default_factory=dict)   # spec -> (V, details) made at fit time
    fe_dof: str = "exact"

    def coef(self):
        return dict(zip(self.coefnames, self.beta.tolist()))

    def fixef(self, dim):
        """(level values, effects) of one fixed-effect dimension."""
        return self.fe_levels[dim], self.fe_values[dim]

    def vcov(self, vcov="iid", data=None, fe_dof=None):
        """The coefficients' variance matrix: 'iid', 'hetero' ('HC1'), or
        'CRV1:<cols>' / {'CRV1': '<cols>'} with '+' for multi-way clustering
        and '^' for intersections (core/inference.py:430-668, 880-920).
        Cluster columns that are not fixed effects are read from `data` (the
        data the fit was given); fe_dof ('exact' or 'pyfixest', the fit's by
        default) counts the fixed effects' parameters.  Returns (V, details)."""
        from .vcov import VcovEngine
        eng = getattr(self, "_vcov_engine", None)
        if eng is None or (data is not None and eng.data is not data):
            eng = self._vcov_engine = VcovEngine(self, data)
        return eng.vcov(vcov, fe_dof or self.fe_dof)

    def se(self, vcov="iid", data=None, fe_dof=None):
        """Standard errors under `vcov` (see `vcov`)."""
        V, _ = self.vcov(vcov, data, fe_dof)
        return np.sqrt(np.diag(V))

    def solve_normal_equations(self, data, columns, block=16, **options):
        """u_j = (Z'WZ)^- Z'v_j for the columns v_j of `data`
        (normal_equations.py:115-231; see `normal_equations`)."""
        from .normal_equations import solve_normal_equations
        return solve_normal_equations(self, data, columns, block=block, **options)

    def resid(self):
        """Residuals in the source's row order (NaN for dropped rows)."""
        perm = self.panel.array("perm", np.int32)[: self.n_obs]
        r = self.resid_sorted.cpu().numpy()
        out = np.full(self.panel.info["n_in"], np.nan)
        out[perm] = r
        return out


class StreamingHDFE:
    """The estimator with `device="rocm"` (core/estimator.py:69-246)."""

    def __init__(self, y, x, fe, *, tol=1e-10, maxiter=5000, rhs_block=8, solver="auto",
                 precond="jacobi", fixef_rm="singleton", weights=None, stream=None,
                 device="rocm", precision="fp64", inner_tol=1e-5, keep_intermediates=False,
                 verbose=False, spmm_timing=False, max_memory_gb=None, vcov="iid", fe_dof="exact",
                 **ignored):
        if device != "rocm":
            raise ValueError("device must be 'rocm' (this package is the MI355X backend)")
        if solver not in ("auto", "explicit", "stream_cg", "within"):
            raise ValueError("solver must be 'auto', 'explicit', 'stream_cg' or 'within'")
        if precond not in ("jacobi", "amg"):
            raise ValueError("precond must be 'jacobi' or 'amg'")
        if solver not in ("auto", "explicit", "stream_cg") or precond != "jacobi":
            raise ValueError("device='rocm' runs the explicit or the matrix-free (stream_cg) "
                             "solver with the Jacobi preconditioner: use solver='auto', "
                             "'explicit' or 'stream_cg' and precond='jacobi'")
        if max_memory_gb is not None and not max_memory_gb > 0:
            raise ValueError("max_memory_gb must be positive")
        if fixef_rm not in ("none", "singleton"):
            raise ValueError("fixef_rm must be 'none' or 'singleton'")
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        self.backend = rocm_backend()       # fail now, without the library or a GPU
        self.y = y
        self.x = list(x or [])
        self.fe = list(fe)
        if len(set(self.fe)) < len(self.fe):
            raise ValueError(f"duplicate fixed-effect dimensions in {self.fe}")
        if stream is not None and stream not in self.fe:
            raise ValueError(f"stream={stream!r} is not one of the fixed effects {self.fe}")
        self.tol, self.maxiter = tol, maxiter
        self.solver = solver
        self.max_memory_gb = max_memory_gb
        self.vcov = vcov
        if vcov is not None:
            from .vcov import _parse
            _parse(vcov)
        if fe_dof not in ("exact", "pyfixest"):
            raise ValueError("fe_dof must be 'exact' or 'pyfixest'")
        self.fe_dof = fe_dof
        self.rhs_block = max(1, int(rhs_block))
        self.fixef_rm = fixef_rm
        self.weights = weights
        self.stream = stream
        self.precision = precision
        self.inner_tol = inner_tol
        self.keep_intermediates = keep_intermediates
        self.verbose = verbose
        self.spmm_timing = spmm_timing

    def _s_budget(self, panel):
        """Bytes the explicit S build may take: `max_memory_gb` if given,
        else 80% of the device's free memory less the block PCG's work
        arrays (the reference's headroom rule, core/solve.py:172-176);
        None (no limit) for solver='explicit', 0 for 'stream_cg'."""
        if self.solver == "stream_cg":
            return 0
        work = 7 * panel.info["L"] * min(panel.info["m"], self.rhs_block) * 8
        if self.max_memory_gb is not None:
            return max(int(self.max_memory_gb * 1e9) - work, 0)
        if self.solver == "explicit":
            return None
        return max(int(0.8 * _lib.device_memory()[0]) - work, 0)

    def _ingest(self, data, names):
        """The model's columns on the device: host arrays streamed to HBM in
        batches of rows, a Parquet file read and streamed batch by batch,
        device tensors as they are (`ingest`)."""
        import torch
        from .ingest import parquet_to_device, to_device
        dtypes = {f: torch.int64 for f in self.fe}
        dtypes.update({v: torch.float64 for v in names})
        needed = list(dict.fromkeys(list(self.fe) + names +
                                    ([self.weights] if self.weights is not None else [])))
        stats = {}
        if isinstance(data, (str, bytes)) or hasattr(data, "__fspath__"):
            cols = parquet_to_device(data, needed, dtypes, stats=stats)
        else:
            cols = to_device({k: _column(data, k) for k in needed}, dtypes, stats=stats)
        return cols, {k: (round(v, 4) if isinstance(v, float) else v) for k, v in stats.items()}

    def fit(self, data) -> HDFEResult:
        import torch
        t0 = time.time()
        names = [self.y] + [v for v in self.x if v != self.y]
        if len(self.fe) < 2:
            raise ValueError("device='rocm' fits need at least two fixed effects "
                             "(one streamed, one solved for)")
        cols, ingest = self._ingest(data, names)
        t_in = time.time()
        panel = Panel([cols[f] for f in self.fe], [cols[v] for v in names],
                      None if self.weights is None else cols[self.weights],
                      self.fixef_rm, -1 if self.stream is None else self.fe.index(self.stream),
                      sort_vars=self.keep_intermediates)
        info = {"solver": "explicit", "device": self.backend.name}
        # the streamed dimension's level values go to the host beside the
        # system build and the solve
        glevels = panel.array_async("glevels", np.int64)
        t1 = time.time()
        s_secs = panel.system(with_rhs=True, budget=self._s_budget(panel))
        if panel.matrix is None:
            # S would not fit the headroom: the matrix-free operator over the
            # cells (core/solve.py:205-218, 251-271), on the device
            budget = self._s_budget(panel)
            if budget is not None and panel.s_skipped > budget:
                why = (f"S would take about {panel.s_skipped / 1e9:.3g} GB > the "
                       f"{budget / 1e9:.3g} GB of device memory headroom")
            else:
                why = "S would hold 2**31 entries or more (the device build counts in 32 bits)"
            if self.solver == "explicit":
                raise MemoryError(f"{why}; raise max_memory_gb or use solver='stream_cg'")
            panel.stream_matrix()
            info["solver"] = "stream_cg"
            if self.solver == "auto":
                info["fallback"] = f"{why}: matrix-free stream_cg on the device"
            info.update(S_nnz=0, S_mb=0.0, S_build_seconds=0.0)
        else:
            info.update(S_nnz=panel.info["nnz"],
                        S_mb=round(panel.info["nnz"] * 12 / 1e6
                                   + (panel.info["L"] + 1) * 4 / 1e6, 1),
                        S_build_seconds=round(s_secs, 4))
        info["device_setup_seconds"] = 0.0
        ncomp = panel.components()
        L, m = panel.info["L"], panel.info["m"]
        rhs = panel.tensor("rhs", torch.float64).view(L, m)
        raw_ss = panel.array("raw_ss", np.float64)
        if self.spmm_timing:
            panel.matrix.timing(True)
        t2 = time.time()
        Gamma = torch.empty((L, m), dtype=torch.float64, device="cuda")
        its, conv, rel, fe_spanned, steps = [], True, 0.0, [], []
        for j0 in range(0, m, self.rhs_block):
            j1 = min(m, j0 + self.rhs_block)
            b = rhs[:, j0:j1].contiguous()
            bn = np.empty(j1 - j0)
            torch.cuda.synchronize()
            _lib.check(_lib.load().hdfe_col_norms(C.c_void_p(b.data_ptr()), L, j1 - j0,
                                                  bn.ctypes.data), "hdfe_col_norms")
            zero = bn <= RHS_ZERO_TOL * np.sqrt(np.maximum(raw_ss[j0:j1], 0.0))
            fe_spanned += [names[j0 + k] for k in np.flatnonzero(zero)]
            sinfo = {}
            X, it, c, r = panel.matrix.solve(b, zero, None, tol=self.tol, maxiter=self.maxiter,
                                             precision=self.precision,
                                             inner_tol=self.inner_tol, info=sinfo)
            Gamma[:, j0:j1] = X
            its.append(it)
            steps.append(sinfo.get("refinement_steps", 0))
            conv &= c
            rel = max(rel, r)
        t3 = time.time()
        panel.normalize(Gamma)
        A = panel.assemble(Gamma)
        xj = list(range(1, m))
        beta = np.linalg.solve(A[np.ix_(xj, xj)], A[xj, 0]) if xj else np.zeros(0)
        gam_y, geff, resid = panel.pass2(Gamma, beta, xj)
        torch.cuda.synchronize()
        t4 = time.time()
        info.update({"iterations": its if len(its) > 1 else its[0], "converged": bool(conv),
                     "blocks": len(its), "seconds": round(t3 - t2, 4),
                     "max_rel_residual": rel, "precision": self.precision})
        if self.precision == "mixed":
            info["refinement_steps"] = steps if len(steps) > 1 else steps[0]
        if fe_spanned:
            info["fe_spanned"] = fe_spanned
        if self.spmm_timing:
            info["spmm"] = panel.matrix.timing_read()
        # effects per dimension, with their level values
        fe_order = [self.fe[panel.info["gdim"]]] + [self.fe[d] for d in panel.dims]
        gy = _host(gam_y)
        panel.copy_wait()
        levels = {fe_order[0]: glevels}
        values = {}
        for t, d in enumerate(fe_order[1:]):
            o, nl = int(panel.offs[t]), panel.n_levels[t]
            levels[d] = panel.array(f"levels{t}", np.int64, pinned=True)
            values[d] = gy[o:o + nl]
        # the streamed dimension's effects (one per group: the largest host
        # copy of the fit, after the smaller ones, which would queue behind it
        # on the bus) travel while the variances are computed
        geff_h, geff_ev = _host_async(geff)
        values = {fe_order[0]: geff_h, **values}
        diag = {"seconds_ingest": round(t_in - t0, 4), "ingest": ingest,
                "seconds_pass01": round(t1 - t_in, 4), "seconds_system": round(t2 - t1, 4),
                "seconds_solve": round(t3 - t2, 4), "seconds_pass2": round(t4 - t3, 4),
                "seconds_total": round(t4 - t0, 4), "singletons": panel.info["singletons"],
                "cells": panel.info["C"], "identifying_cells": panel.info["M"],
                "identifying_groups": panel.info["Gi"]}
        res = HDFEResult(coefnames=names[1:], beta=beta, solver_info=info, fe_names=fe_order,
                         fe_levels=levels, fe_values=values, n_obs=panel.info["n"],
                         n_components=ncomp, A=A, diagnostics=diag, panel=panel, Gamma=Gamma,
                         resid_sorted=resid, gam_y=gam_y, geff=geff, fe_dof=self.fe_dof)
        try:
            if self.vcov is not None and res.coefnames:
                # the coefficients' variance at fit time, as the reference's
                # fit (core/estimator.py:467-481): iid and hetero always, and
                # the default `vcov` (cluster columns that are not fixed
                # effects come from the fit's data); vcov=None skips them
                t5 = time.time()
                res.vcovs = {spec: res.vcov(spec, data) for spec in ("iid", "hetero")}
                if not isinstance(self.vcov, str) or self.vcov not in res.vcovs:
                    res.vcovs[str(self.vcov)] = res.vcov(self.vcov, data)
                res.vcov_matrix, res.vcov_details = res.vcovs[str(self.vcov)]
                res.diagnostics["seconds_vcov"] = round(time.time() - t5, 4)
        finally:
            if geff_ev is not None:
                geff_ev.synchronize()
        if self.verbose:
            print(f"fit: {diag}")
        return res


def feols_stream(fml: str, data, **options) -> HDFEResult:
    """OLS with high-dimensional fixed effects on the device
    (hdfe_stream/api.py:22-107, with device='rocm')."""
    y, xs, fe = parse_formula(fml)
    options.setdefault("device", "rocm")
    return StreamingHDFE(y, xs, fe, **options).fit(data)
