"""Host -> HBM ingest of a panel's columns in row batches (hdfe_stage_*).

The reference streams its input from Parquet in row batches
(core/utils.py:26-60, `iter_group_chunks`) and works on each batch on the
CPU.  Here the batches go to HBM, where the whole fit runs: each batch of
rows is packed by host threads into a page-locked slot and copied to the
device while the next batch is packed.  Two sources:

* host arrays (numpy, in ordinary pageable memory): `to_device`;
* a Parquet file, read batch by batch with pyarrow (`parquet_to_device`):
  the decode of batch b + 1 overlaps the DMA of batch b, and the file is
  never held whole in host memory.

Columns already on the device pass through untouched.
"""

from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np

from . import _lib

DEFAULT_BATCH_ROWS = 1_000_000      # rows per batch (48 MB of the AKM panel's columns)
DEFAULT_DEPTH = 3                   # page-locked slots in the ring


def default_threads() -> int:
    """Host threads for the packs: the process's CPUs, at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


class Stager:
    """A ring of page-locked slots for columns of the given element sizes."""

    def __init__(self, elem_bytes, batch_rows: int = DEFAULT_BATCH_ROWS,
                 depth: int = DEFAULT_DEPTH, threads: int | None = None):
        self.elem_bytes = [int(b) for b in elem_bytes]
        eb = (C.c_int * len(self.elem_bytes))(*self.elem_bytes)
        h = C.c_void_p()
        _lib.check(_lib.load().hdfe_stage_create(len(self.elem_bytes), eb, int(batch_rows),
                                                 int(depth), int(threads or default_threads()),
                                                 C.byref(h)), "hdfe_stage_create")
        self._h = h

    def stage(self, host_cols, dev_cols, dev_row: int = 0, wait: bool = True) -> dict:
        """host_cols[c] (contiguous numpy, the element size of column c) into
        dev_cols[c] (device tensors) from row dev_row on."""
        n = len(host_cols[0])
        for h, d, eb in zip(host_cols, dev_cols, self.elem_bytes):
            if len(h) != n or h.itemsize != eb or not h.flags["C_CONTIGUOUS"]:
                raise ValueError("host columns must be contiguous, of equal length and of "
                                 "the stager's element sizes")
            if d.element_size() != eb or d.numel() < dev_row + n or not d.is_contiguous():
                raise ValueError("a device column is too short or of another element size")
        hp = (C.c_void_p * len(host_cols))(*[h.ctypes.data for h in host_cols])
        dp = (C.c_void_p * len(dev_cols))(*[d.data_ptr() for d in dev_cols])
        stats = np.zeros(4)
        _lib.check(_lib.load().hdfe_stage_columns(self._h, n, hp, dp, int(dev_row), int(wait),
                                                  stats.ctypes.data), "hdfe_stage_columns")
        return {"seconds": float(stats[0]), "pack_seconds": float(stats[1]),
                "bytes": int(stats[2]), "batches": int(stats[3])}

    def wait(self) -> None:
        _lib.check(_lib.load().hdfe_stage_wait(self._h), "hdfe_stage_wait")

    def close(self) -> None:
        if self._h:
            _lib.load().hdfe_stage_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_CACHE: dict = {}


def cached_stager(elem_bytes, batch_rows: int = DEFAULT_BATCH_ROWS, depth: int = DEFAULT_DEPTH,
                  threads: int | None = None) -> Stager:
    """A process-wide stager per (element sizes, batch, depth, threads): its
    page-locked slots and host threads are made once (hipHostMalloc of the
    slots costs milliseconds per call), then reused by every ingest of that
    shape.  Calls from one thread only, as the fits."""
    key = (tuple(int(b) for b in elem_bytes), int(batch_rows), int(depth),
           int(threads or default_threads()))
    st = _CACHE.get(key)
    if st is None or st._h is None:
        st = _CACHE[key] = Stager(key[0], key[1], key[2], key[3])
    return st


def _torch_dtype(np_dtype):
    import torch
    return torch.from_numpy(np.empty(0, np_dtype)).dtype


def to_device(columns: dict, dtypes: dict | None = None, batch_rows: int = DEFAULT_BATCH_ROWS,
              threads: int | None = None, stats: dict | None = None) -> dict:
    """{name: device tensor} for {name: host array or tensor}: host columns
    (cast to dtypes[name] first when given and different) streamed to HBM
    together in batches of `batch_rows` rows; device tensors pass through
    (cast on the device).  `stats` receives seconds, bytes and GB/s."""
    import torch
    dtypes = dtypes or {}
    out, host = {}, {}
    for name, col in columns.items():
        want = dtypes.get(name)
        if isinstance(col, torch.Tensor):
            t = col if col.is_cuda else None
            if t is not None:
                out[name] = t.to(dtype=want) if want is not None else t
                out[name] = out[name].contiguous()
                continue
            col = col.numpy()
        a = np.asarray(col)
        if want is not None and a.dtype != np.dtype(str(want).replace("torch.", "")):
            a = a.astype(str(want).replace("torch.", ""))
        host[name] = np.ascontiguousarray(a)
    if host:
        names = list(host)
        n = len(host[names[0]])
        if any(len(host[k]) != n for k in names):
            raise ValueError("columns of different lengths")
        dev = [torch.empty(n, dtype=_torch_dtype(host[k].dtype), device="cuda") for k in names]
        torch.cuda.synchronize()
        st = cached_stager([host[k].itemsize for k in names], batch_rows, threads=threads)
        s = st.stage([host[k] for k in names], dev) if n else \
            {"seconds": 0.0, "pack_seconds": 0.0, "bytes": 0, "batches": 0}
        out.update(zip(names, dev))
        if stats is not None:
            stats.update(s)
            stats["GBps"] = s["bytes"] / s["seconds"] / 1e9 if s["seconds"] > 0 else 0.0
    elif stats is not None:
        stats.update(seconds=0.0, pack_seconds=0.0, bytes=0, batches=0, GBps=0.0)
    return {k: out[k] for k in columns}


def parquet_to_device(path, columns, dtypes: dict | None = None,
                      batch_rows: int = DEFAULT_BATCH_ROWS, threads: int | None = None,
                      stats: dict | None = None) -> dict:
    """{name: device tensor} of Parquet columns, read in batches of
    `batch_rows` rows (pyarrow) and streamed to HBM batch by batch: the
    decode of a batch overlaps the copy of the previous one."""
    import pyarrow.parquet as pq
    import torch
    dtypes = dtypes or {}
    columns = list(columns)
    t0 = time.perf_counter()
    pf = pq.ParquetFile(path)
    n = pf.metadata.num_rows
    schema = pf.schema_arrow
    np_types = []
    for c in columns:
        want = dtypes.get(c)
        np_types.append(np.dtype(str(want).replace("torch.", "")) if want is not None
                        else np.dtype(schema.field(c).type.to_pandas_dtype()))
    dev = [torch.empty(n, dtype=_torch_dtype(t), device="cuda") for t in np_types]
    torch.cuda.synchronize()
    st = cached_stager([t.itemsize for t in np_types], batch_rows, threads=threads)
    row, nbytes, pack = 0, 0, 0.0
    try:
        for batch in pf.iter_batches(batch_size=batch_rows, columns=columns):
            arrs = [np.ascontiguousarray(batch.column(i).to_numpy(zero_copy_only=False),
                                         dtype=t) for i, t in enumerate(np_types)]
            s = st.stage(arrs, dev, dev_row=row, wait=False)
            row += len(arrs[0])
            nbytes += s["bytes"]
            pack += s["pack_seconds"]
    finally:
        st.wait()
    if row != n:
        raise RuntimeError(f"read {row} rows of {n}")
    if stats is not None:
        secs = time.perf_counter() - t0
        stats.update(seconds=secs, pack_seconds=pack, bytes=nbytes,
                     GBps=nbytes / secs / 1e9 if secs > 0 else 0.0)
    return dict(zip(columns, dev))
