"""hdfe-stream's solve path on AMD Instinct MI355X (gfx950).

The device backend (`rocm`) and the estimators that reach it through
`device="rocm"`.  All compute runs in libhdfe_hip.so (hand-written HIP);
there is no CPU fallback.
"""

from __future__ import annotations

__version__ = "0.1.0"

from . import rocm  # noqa: F401
from .rocm import DeviceCSR, explicit_solver, rocm_backend  # noqa: F401
