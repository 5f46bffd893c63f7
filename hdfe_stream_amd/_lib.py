"""ctypes binding of libhdfe_hip.so, the C ABI declared in include/hdfe_hip.h.

The library is the only compute path of this package: there is no CPU
fallback.  Loading fails loudly (ImportError) when the library has not been
built, and `require_device()` raises RuntimeError without a HIP device.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libhdfe_hip.so"
BUILD_HINT = ("device='rocm' needs the HIP library built for gfx950: run "
              "`python -m hdfe_stream_amd._build` (hipcc from ROCm) and use an AMD Instinct "
              "MI355X")

STATUS = {0: "ok", 1: "invalid argument", 2: "HIP error", 3: "out of device memory",
          4: "no device"}

_lib = None

# (name, restype, argtypes) of every exported function; the header declares
# exactly these.
P = C.c_void_p
I64 = C.c_int64
I32 = C.c_int
D = C.c_double
SIGNATURES = [
    ("hdfe_last_error", C.c_char_p, []),
    ("hdfe_version", I32, []),
    ("hdfe_device_info", I32, [C.c_char_p, I32, P, P]),
    ("hdfe_csr_create", I32, [I64, I64, P, P, P, P, I32, P]),
    ("hdfe_csr_info", I32, [P, P, P, P]),
    ("hdfe_pcg_solve", I32, [P, I64, P, P, P, D, I64, I32, D, P, P, P, P, P, I32, P]),
    ("hdfe_csr_matmat", I32, [P, I64, P, P, I32]),
    ("hdfe_stream_create", I32, [I64, I64, I64, I32, P, P, I64, P, P, P, P]),
    ("hdfe_stream_rebind", I32, [P, I64, I64, I64, I32, P, P, I64, P, P, P]),
    ("hdfe_csr_matmat_f32", I32, [P, I64, P, P, I32]),
    ("hdfe_csr_export", I32, [P, P, P, P, P]),
    ("hdfe_col_norms", I32, [P, I64, I64, P]),
    ("hdfe_csr_timing", I32, [P, I32]),
    ("hdfe_csr_spmm_time", I32, [P, I64, I32, I32, P]),
    ("hdfe_csr_timing_read", I32, [P, P, P]),
    ("hdfe_memcpy", I32, [P, P, I64, I32]),
    ("hdfe_cache_release", I32, [P]),
    ("hdfe_destroy", None, [P]),
    ("hdfe_panel_create", I32, [I64, I32, P, I32, P, P, I32, I32, I32, I32, P]),
    ("hdfe_panel_info", I32, [P, P, P, P]),
    ("hdfe_panel_array", I32, [P, C.c_char_p, P, P]),
    ("hdfe_panel_copy_async", I32, [P, P, P, I64]),
    ("hdfe_panel_copy_wait", I32, [P]),
    ("hdfe_panel_system", I32, [P, I32, P, P]),
    ("hdfe_panel_budget", I32, [P, I64, P, P]),
    ("hdfe_panel_components", I32, [P, P]),
    ("hdfe_panel_pair_components", I32, [P, I32, I32, P]),
    ("hdfe_panel_scores", I32, [P, P, P, I32, P, P, P]),
    ("hdfe_panel_normalize", I32, [P, P, I32, P]),
    ("hdfe_panel_assemble", I32, [P, P, P]),
    ("hdfe_panel_pass2", I32, [P, P, P, P, I32, P, P, P]),
    ("hdfe_panel_destroy", None, [P]),
    ("hdfe_rademacher", I32, [I64, I64, I32, C.c_uint64, I64, P]),
    ("hdfe_rows_create", I32, [I64, I64, I32, P, P, P, P]),
    ("hdfe_rows_center", I32, [P, P, P, P, I32, C.c_uint64, I64, I32, P]),
    ("hdfe_rows_project_moments", I32, [P, P, P, P, I32, C.c_uint64, I64, I32, P, P, I32, P,
                                         P, P]),
    ("hdfe_rows_cov_center", I32, [P, I32, P, P, I32, C.c_uint64, I64, I32, P]),
    ("hdfe_jla_combine", I32, [I64, I64, P, P]),
    ("hdfe_rows_destroy", None, [P]),
    ("hdfe_loo_prune", I32, [I64, P, P, P, P]),
    ("hdfe_components", I32, [I64, P, P, I32, I32, I32, I32, P]),
    ("hdfe_csr_blk_parts", I32, [P, I32, I32, I32, P]),
    ("hdfe_csr_graphs", I32, [P, P]),
    ("hdfe_csr_wide_mode", I32, [P, I32, I32, I32, P]),
    ("hdfe_csr_small_form", I32, [P, I32, I32, I32, P]),
    ("hdfe_csr_row_pointer_bits", I32, [P, I32, P]),
    ("hdfe_tall_dot", I32, [I64, I32, I32, P, I64, P, I64, P, P]),
    ("hdfe_tall_dot_noalloc", I32, [I64, I32, I32, P, I64, P, I64, P, P, P]),
    ("hdfe_se_stack", I32, [I64, I32, P, P, P, P, I32, P]),
    ("hdfe_se_assemble", I32, [I64, I32, I32, P, P, P, P, P, P, P, P, D, P, P, P, P, P, I32, P]),
    ("hdfe_se_bii", I32, [I64, I32, P, P, P]),
    ("hdfe_se_cross_groups", I32, [I64, I32, P, P, I64, P, P, P]),
    ("hdfe_se_cell_sums", I32, [I64, P, P, P, P, P, I32, P, P]),
    ("hdfe_se_smooth", I32, [I64, P, P, P, P, P, I32, P]),
    ("hdfe_components_info", I32, [P, P]),
    ("hdfe_components_levels", I32, [P, I32, P, P]),
    ("hdfe_components_sizes", I32, [P, P]),
    ("hdfe_components_destroy", None, [P]),
    ("hdfe_rows_csr", I32, [P, P, P, P, P, P, P]),
    ("hdfe_rows_chunk_sum", I32, [P, I32, P, P]),
    ("hdfe_kss_group_weight", I32, [I64, P, P, P]),
    ("hdfe_kss_reduce", I32, [P, I32, P, P, P, P, P, P, P]),
    ("hdfe_kss_project", I32, [I64, I32, I32, P, P, I64, P, P, P, P, P, P, P]),
    ("hdfe_kss_group_sums", I32, [I64, I32, P, P, P, P, P]),
    ("hdfe_kss_trace_moments", I32, [I64, I32, P, P, I64, P, P, P, P, P, P]),
    ("hdfe_kss_drop_null", I32, [I64, I64, I32, I64, I32, P, P, P, P, P, P, P]),
    ("hdfe_kss_sigma2", I32, [I64, I32, P, P, I32, P, P, P, P, P]),
    ("hdfe_kss_plugin", I32, [I64, P, P, I64, P, P, P, P]),
    ("hdfe_copy_calibrate", I32, [I64, I32, P]),
    ("hdfe_panel_partial", I32, [P, P, P, I32, P]),
    ("hdfe_adj_component_rows", I32, [I64, P, P, I64, P, P, P, P, P, P]),
    ("hdfe_adj_xwr", I32, [I64, I32, P, P, I32, P, P, P]),
    ("hdfe_adj_gradients", I32, [I64, I32, P, P, P, P, P, P, P, P]),
    ("hdfe_panel_reweight", I32, [P, P, P]),
    ("hdfe_glm_rows", I32, [I32, P, P, P, I32, I32, P, P, P, I32, P, P, I64, I32, P, P, P, P, P,
                            P, P, P]),
    ("hdfe_glm_rhs", I32, [P, I64, P, P, P, P, I32, P, P, P]),
    ("hdfe_glm_partial", I32, [I32, P, P, P, P, P, P, I32, I32, P, P, I32, P, P, I64, I32, P, P,
                               P, P, P, P]),
    ("hdfe_glm_syrk", I32, [I32, P, I32, P, P, I64, P]),
    ("hdfe_glm_coefficients", I32, [I64, I64, I32, P, P, P, P, P, P, P, P, I64, I32, P, P, P,
                                    P]),
    ("hdfe_glm_toward", I32, [I64, P, I32, P, P, P, P]),
    ("hdfe_stage_create", I32, [I32, P, I64, I32, I32, P]),
    ("hdfe_stage_columns", I32, [P, I64, P, P, I64, I32, P]),
    ("hdfe_stage_wait", I32, [P]),
    ("hdfe_stage_destroy", None, [P]),
]


def load():
    """The loaded library (ctypes.CDLL), loading it on first use."""
    global _lib
    if _lib is None:
        # torch first: its HIP runtime (same SONAMEs) is then the one the
        # library binds to, so the process holds a single HIP runtime
        import torch  # noqa: F401
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH.name} not found. {BUILD_HINT}")
        lib = C.CDLL(str(LIB_PATH))
        for name, res, args in SIGNATURES:
            try:
                fn = getattr(lib, name)
            except AttributeError:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def exported(name: str) -> bool:
    return hasattr(load(), name)


class HipError(RuntimeError):
    pass


def check(status: int, what: str) -> None:
    """Raise for a non-zero status: ValueError for bad arguments, MemoryError
    for device OOM, RuntimeError otherwise."""
    if status == 0:
        return
    msg = load().hdfe_last_error().decode(errors="replace")
    text = f"{what}: {STATUS.get(status, status)}: {msg}"
    if status == 1:
        raise ValueError(text)
    if status == 3:
        raise MemoryError(text)
    raise HipError(text)


def require_device() -> str:
    """The device name; RuntimeError when no HIP device is visible."""
    lib = load()
    name = C.create_string_buffer(256)
    free = C.c_int64(0)
    total = C.c_int64(0)
    st = lib.hdfe_device_info(name, 256, C.byref(free), C.byref(total))
    if st == 4:
        raise RuntimeError("device='rocm': no HIP device visible")
    check(st, "hdfe_device_info")
    return name.value.decode()


def release_memory() -> int:
    """Device memory the library and torch hold for reuse, returned to the
    driver (torch's cache emptied, the library's free blocks freed); the
    bytes the library returned."""
    import torch
    freed = C.c_int64(0)
    check(load().hdfe_cache_release(C.byref(freed)), "hdfe_cache_release")
    torch.cuda.empty_cache()
    return int(freed.value)


def device_memory() -> tuple[int, int]:
    lib = load()
    free = C.c_int64(0)
    total = C.c_int64(0)
    check(lib.hdfe_device_info(None, 0, C.byref(free), C.byref(total)), "hdfe_device_info")
    return int(free.value), int(total.value)
