"""The `device="rocm"` solve backend: the MI355X counterpart of the reference's
CuPy plugin (hdfe_stream/core/gpu.py:33-51 `cuda_backend`, :186-196
`explicit_solver`, :199-253 `pcg`).

Same contract as the reference:
  explicit_solver(S, dinv, tol, maxiter, debug=None) -> (solve, device name)
  solve(b, zero=None, x0=None) -> (X, iterations, converged, max_rel)
S (a scipy CSR, or device arrays) and dinv go to the device once; every call
of `solve` reuses them.  b, x0 and X are host float64 (L, k) arrays, or torch
tensors already on the device (then X comes back on the device, ordered on
torch's current stream).  Non-convergence is reported, not raised.

Beyond the reference: `precision="mixed"` runs FP32 CG inside FP64 iterative
refinement (the residual b - S x recomputed in FP64), stopping on the true
FP64 residual, so the tolerance means the same thing in both modes.
"""

from __future__ import annotations

import ctypes as C
from types import SimpleNamespace

import numpy as np

from . import _lib

PRECISIONS = {"fp64": 0, "mixed": 1}


def rocm_backend():
    """What the device solver runs on: the HIP library and the device.
    ImportError without the built library, RuntimeError without a device
    (the reference's cuda_backend contract, core/gpu.py:33-51)."""
    lib = _lib.load()
    name = _lib.require_device()
    return SimpleNamespace(
        lib=lib, name=name,
        free_bytes=lambda: _lib.device_memory()[0],
        release=_lib.release_memory,
        csr=lambda S: DeviceCSR.from_scipy(S, None),
    )


def _is_torch(a) -> bool:
    return type(a).__module__.startswith("torch")


class DeviceCSR:
    """An explicit S (and its Jacobi inverse diagonal) resident on the device."""

    def __init__(self, handle, L: int, nnz: int, owned: bool = True):
        self._h = handle
        self.L = L
        self.nnz = nnz
        self._owned = owned      # False: the handle belongs to a Panel

    @classmethod
    def from_scipy(cls, S, dinv):
        S = S.tocsr()
        L = S.shape[0]
        indptr = np.ascontiguousarray(S.indptr, np.int64)
        indices = np.ascontiguousarray(S.indices, np.int32)
        data = np.ascontiguousarray(S.data, np.float64)
        d = np.ascontiguousarray(np.zeros(L) if dinv is None else dinv, np.float64)
        return cls._create(L, indptr, indices, data, d, 0)

    @classmethod
    def from_device(cls, L, indptr, indices, data, dinv):
        """From torch device tensors: indptr int64 (L+1), indices int32,
        data float64, dinv float64 (L)."""
        return cls._create(L, indptr, indices, data, dinv, 1)

    @classmethod
    def _create(cls, L, indptr, indices, data, dinv, on_device):
        lib = _lib.load()
        h = C.c_void_p()
        ptr = (lambda a: a.data_ptr()) if on_device else (lambda a: a.ctypes.data)
        nnz = int(indices.shape[0]) if on_device else int(indptr[-1])
        _lib.check(lib.hdfe_csr_create(L, nnz, ptr(indptr), ptr(indices), ptr(data), ptr(dinv),
                                       on_device, C.byref(h)), "hdfe_csr_create")
        return cls(h, L, nnz)

    def nlong(self) -> int:
        n = C.c_int(0)
        _lib.check(_lib.load().hdfe_csr_info(self._h, None, None, C.byref(n)), "hdfe_csr_info")
        return int(n.value)

    def matmat(self, X):
        """S X on the device (FP64)."""
        lib = _lib.load()
        if _is_torch(X):
            import torch
            X = X.contiguous()
            k = X.shape[1] if X.dim() == 2 else 1
            Y = torch.empty_like(X)
            torch.cuda.current_stream().synchronize()
            _lib.check(lib.hdfe_csr_matmat(self._h, k, X.data_ptr(), Y.data_ptr(), 1),
                       "hdfe_csr_matmat")
            return Y
        X = np.ascontiguousarray(X, np.float64)
        k = X.shape[1] if X.ndim == 2 else 1
        Y = np.empty_like(X)
        _lib.check(lib.hdfe_csr_matmat(self._h, k, X.ctypes.data, Y.ctypes.data, 0),
                   "hdfe_csr_matmat")
        return Y

    def matmat_f32(self, X):
        """S X in FP32 on the device (host float32 in and out)."""
        X = np.ascontiguousarray(X, np.float32)
        k = X.shape[1] if X.ndim == 2 else 1
        Y = np.empty_like(X)
        _lib.check(_lib.load().hdfe_csr_matmat_f32(self._h, k, X.ctypes.data, Y.ctypes.data, 0),
                   "hdfe_csr_matmat_f32")
        return Y

    def solve(self, b, zero=None, x0=None, tol=1e-10, maxiter=5000, precision="fp64",
              inner_tol=1e-5, info=None):
        """(X, iterations, converged, max_rel); see the module docstring."""
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        lib = _lib.load()
        dev = _is_torch(b)
        if dev:
            import torch
            B = b.contiguous().to(torch.float64)
            L, k = B.shape[0], (B.shape[1] if B.dim() == 2 else 1)
            X = torch.empty_like(B)
            X0 = None if x0 is None else x0.contiguous().to(torch.float64)
            bp, xp, x0p = B.data_ptr(), X.data_ptr(), (None if X0 is None else X0.data_ptr())
            stream = torch.cuda.current_stream().cuda_stream
        else:
            B = np.ascontiguousarray(np.asarray(b, np.float64))
            L, k = B.shape[0], (B.shape[1] if B.ndim == 2 else 1)
            X = np.empty_like(B)
            X0 = None if x0 is None else np.ascontiguousarray(np.asarray(x0, np.float64))
            bp, xp, x0p = B.ctypes.data, X.ctypes.data, (None if X0 is None else X0.ctypes.data)
            stream = None
        if L != self.L:
            raise ValueError(f"b has {L} rows; S is {self.L} x {self.L}")
        z = None if zero is None else np.ascontiguousarray(np.asarray(zero, np.uint8).reshape(-1))
        if z is not None and len(z) != k:
            raise ValueError("zero must have one entry per column of b")
        it = C.c_int64(0)
        conv = C.c_int(0)
        rel = C.c_double(0.0)
        outer = C.c_int64(0)
        _lib.check(lib.hdfe_pcg_solve(self._h, k, bp, None if z is None else z.ctypes.data,
                                      x0p, float(tol), int(maxiter), PRECISIONS[precision],
                                      float(inner_tol), xp, C.byref(it), C.byref(conv),
                                      C.byref(rel), C.byref(outer), int(dev), stream),
                   "hdfe_pcg_solve")
        if info is not None:
            info["refinement_steps"] = int(outer.value)
        return X, int(it.value), bool(conv.value), float(rel.value)

    def timing(self, enable: bool = True) -> None:
        """HIP-event timing of the SpMM launches of later solves (resets)."""
        _lib.check(_lib.load().hdfe_csr_timing(self._h, int(enable)), "hdfe_csr_timing")

    def spmm_time(self, k: int, reps: int = 50, precision: str = "fp64") -> float:
        """Average microseconds of the solve's SpMM kernel (k columns), from
        HIP events around `reps` back-to-back launches on the solver's stream."""
        us = C.c_double()
        _lib.check(_lib.load().hdfe_csr_spmm_time(self._h, k, reps, PRECISIONS[precision],
                                                  C.byref(us)), "hdfe_csr_spmm_time")
        return float(us.value)

    def blk_parts(self, k: int, precision: str = "fp64", set: int = 0) -> int:
        """The blocked SpMM's tuned row parts per XCD for k columns (0: not
        tuned); `set` overrides it.  Results do not depend on it."""
        out = C.c_int()
        _lib.check(_lib.load().hdfe_csr_blk_parts(self._h, int(k), PRECISIONS[precision],
                                                  int(set), C.byref(out)), "hdfe_csr_blk_parts")
        return int(out.value)

    def wide_mode(self, k: int, precision: str = "fp64", set: int = -1) -> int:
        """The wide SpMM's form for k > 8 columns (-1 not tuned yet, 0 a wave
        per row, 1 and 2 column slabs per XCD); `set` overrides it."""
        out = C.c_int()
        _lib.check(_lib.load().hdfe_csr_wide_mode(self._h, int(k), PRECISIONS[precision],
                                                  int(set), C.byref(out)), "hdfe_csr_wide_mode")
        return int(out.value)

    def small_form(self, k: int, precision: str = "fp64", set: int = -1) -> int:
        """The column-blocked SpMM's form for k <= 8 columns (0 the default:
        tiled where it applies; 1 groups of column blocks with partial
        products and a combine; 2 tiled; 3 tiled with 32-bit indices); `set`
        overrides it."""
        out = C.c_int()
        _lib.check(_lib.load().hdfe_csr_small_form(self._h, int(k), PRECISIONS[precision],
                                                   int(set), C.byref(out)), "hdfe_csr_small_form")
        return int(out.value)

    def row_pointer_bits(self, set: int = -1) -> int:
        """The row pointers' width on the device (32 or 64); set=64 widens
        them (only the plain CSR products run then, as for a matrix of 2**31
        nonzeros or more, which is created with 64-bit ones)."""
        out = C.c_int()
        _lib.check(_lib.load().hdfe_csr_row_pointer_bits(self._h, int(set), C.byref(out)),
                   "hdfe_csr_row_pointer_bits")
        return int(out.value)

    def graphs(self) -> int:
        """CG batches held as HIP graphs (captured on small systems only)."""
        n = C.c_int64()
        _lib.check(_lib.load().hdfe_csr_graphs(self._h, C.byref(n)), "hdfe_csr_graphs")
        return int(n.value)

    def timing_read(self) -> dict:
        ms = C.c_double()
        n = C.c_int64()
        _lib.check(_lib.load().hdfe_csr_timing_read(self._h, C.byref(ms), C.byref(n)),
                   "hdfe_csr_timing_read")
        return {"ms": float(ms.value), "launches": int(n.value)}

    def export(self):
        """(scipy CSR S, dinv) copied back from the device."""
        import scipy.sparse as sp
        L = self.L
        nnz = C.c_int64()
        _lib.check(_lib.load().hdfe_csr_info(self._h, None, C.byref(nnz), None), "hdfe_csr_info")
        nnz = int(nnz.value)
        indptr = np.empty(L + 1, np.int64)
        indices = np.empty(max(nnz, 1), np.int32)
        data = np.empty(max(nnz, 1))
        dinv = np.empty(max(L, 1))
        _lib.check(_lib.load().hdfe_csr_export(self._h, indptr.ctypes.data, indices.ctypes.data,
                                               data.ctypes.data, dinv.ctypes.data),
                   "hdfe_csr_export")
        return sp.csr_matrix((data[:nnz], indices[:nnz], indptr), shape=(L, L)), dinv[:L]

    def close(self):
        if self._h and self._owned:
            _lib.load().hdfe_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def explicit_solver(S, dinv, tol, maxiter, debug=None, precision="fp64", inner_tol=1e-5):
    """(solve(b, zero, x0) -> (X, iterations, converged, rel), device name):
    S and the preconditioner are copied to the device here, once, and stay
    there for every right-hand side (core/gpu.py:186-196)."""
    be = rocm_backend()
    A = DeviceCSR.from_scipy(S, dinv)

    def solve(b, zero=None, x0=None):
        out = A.solve(b, zero, x0, tol=tol, maxiter=maxiter, precision=precision,
                      inner_tol=inner_tol)
        if debug is not None:
            debug(f"  CG: {out[1]} iterations, max rel. residual {out[3]:.2e}")
        return out
    solve.matrix = A
    return solve, be.name


def pcg(A: DeviceCSR, b, zero=None, x0=None, tol=1e-10, maxiter=5000, debug=None,
        precision="fp64"):
    """Jacobi block PCG on the device (core/gpu.py:199-253)."""
    out = A.solve(b, zero, x0, tol=tol, maxiter=maxiter, precision=precision)
    if debug is not None:
        debug(f"  CG: {out[1]} iterations, max rel. residual {out[3]:.2e}")
    return out
