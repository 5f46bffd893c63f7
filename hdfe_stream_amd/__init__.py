"""hdfe-stream's solve path on AMD Instinct MI355X (gfx950).

The device backend (`rocm`) and the estimators that reach it through
`device="rocm"`.  All compute runs in libhdfe_hip.so (hand-written HIP);
there is no CPU fallback.
"""

from __future__ import annotations

__version__ = "0.1.0"

from . import rocm  # noqa: F401
from .rocm import DeviceCSR, explicit_solver, rocm_backend  # noqa: F401
from .connected import ConnectedComponents, connected_components  # noqa: F401,E402
from .estimator import feols_stream  # noqa: F401,E402
from .glm import fepois_stream  # noqa: F401,E402
from .leaveout import leave_out_kss  # noqa: F401,E402
