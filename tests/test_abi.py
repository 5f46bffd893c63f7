"""The C ABI: the built library loads and exports every function the header
declares (no GPU needed: nothing here launches a kernel)."""

import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared():
    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names += re.findall(r"\b(hdfe_\w+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_functions():
    assert "hdfe_pcg_solve" in declared()
    assert "hdfe_csr_create" in declared()


def test_library_exports_every_declared_symbol():
    from hdfe_stream_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.fail("libhdfe_hip.so is not built: run python -m hdfe_stream_amd._build")
    lib = _lib.load()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_the_header():
    from hdfe_stream_amd import _lib
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert set(declared()) <= bound


def test_missing_library_raises_import_error(monkeypatch, tmp_path):
    from hdfe_stream_amd import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "libhdfe_hip.so")
    with pytest.raises(ImportError, match="hdfe_stream_amd._build"):
        _lib.load()
