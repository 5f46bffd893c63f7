"""Build libhdfe_hip.so in-tree with hipcc for gfx950.

Every source is compiled to its own object in parallel (the kernels of one
file share no device symbols with another), then the objects are linked
into one shared library.

Usage: python -m hdfe_stream_amd._build   (or hdfe_stream_amd._build.build())
"""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libhdfe_hip.so"
OBJ = PKG / "_build"
SOURCES = ["util.hip", "pcg.hip", "build.hip", "rows.hip", "kss.hip", "graph.hip", "adjoint.hip", "se.hip",
           "glm.hip", "ingest.hip"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
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


def _run(cmd: list[str]) -> subprocess.CompletedProcess:
    return subprocess.run(cmd, capture_output=True, text=True)


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile every HIP source and link one shared library (always
    rebuilds)."""
    OBJ.mkdir(exist_ok=True)
    srcs = sources()
    objs = [OBJ / (s.stem + ".o") for s in srcs]
    cmds = [["hipcc", *FLAGS, "-I", str(CSRC), "-c", str(s), "-o", str(o)]
            for s, o in zip(srcs, objs)]
    jobs = jobs or min(len(cmds), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_run, cmds))
    for s, res in zip(srcs, results):
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"hipcc failed ({res.returncode}) compiling {s.name}")
        if verbose and res.stderr:
            sys.stderr.write(res.stderr)
    tmp = LIB.with_suffix(".so.tmp")
    rpath = torch_lib_dir()
    link = [f"-Wl,-rpath,{rpath}"] if rpath else []
    res = _run(["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", *link,
                *map(str, objs), "-o", str(tmp)])
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode}) linking {LIB.name}")
    tmp.replace(LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
