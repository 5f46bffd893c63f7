"""Build libhdfe_hip.so in-tree with hipcc for gfx950.

Usage: python -m hdfe_stream_amd._build   (or hdfe_stream_amd._build.build())
"""

from __future__ import annotations

import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libhdfe_hip.so"
SOURCES = ["util.hip", "pcg.hip", "build.hip", "rows.hip", "kss.hip", "graph.hip", "adjoint.hip", "se.hip",
           "glm.hip"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         # packed FP32 instructions stay off: hipcc would otherwise emit them
         # from plain float code
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


def torch_lib_dir() -> str | None:
    """torch's lib directory: the library resolves the HIP runtime there
    first, so that a process holds one HIP runtime (torch's)."""
    try:
        import importlib.util
        spec = importlib.util.find_spec("torch")
        if spec and spec.origin:
            d = Path(spec.origin).parent / "lib"
            if (d / "libamdhip64.so").exists():
                return str(d)
    except Exception:
        pass
    return None


def sources() -> list[Path]:
    return [CSRC / s for s in SOURCES if (CSRC / s).exists()]


def build(verbose: bool = False) -> Path:
    """Compile every HIP source into one shared library (always rebuilds)."""
    tmp = LIB.with_suffix(".so.tmp")
    rpath = torch_lib_dir()
    link = [f"-Wl,-rpath,{rpath}"] if rpath else []
    cmd = ["hipcc", *FLAGS, *link, "-I", str(CSRC), *map(str, sources()), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode}) building {LIB.name}")
    if verbose:
        sys.stderr.write(res.stderr)
    tmp.replace(LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
