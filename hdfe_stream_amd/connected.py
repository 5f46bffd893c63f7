"""The leave-one-out connected set on the device (core/connected.py:378-562).

`leave_one_out(workers, firms)` returns which rows survive the reference's
pruning -- the largest connected component by rows, then workers with a
single row dropped, then rounds of dropping workers whose removal would
disconnect the network (articulation points), each followed by keeping the
largest component again -- and the same counts the reference records
(`components_before`, `single_observation`, `cut`, `rounds`)."""

from __future__ import annotations

import ctypes as C

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
