"""Host-side logic that needs no GPU: every module of the package imports,
and the vcov specifications parse as the reference's fit accepts them
(core/estimator.py:475-481)."""

import importlib
import pathlib

import pytest


def test_every_module_imports():
    import hdfe_stream_amd
    for f in sorted(pathlib.Path(hdfe_stream_amd.__file__).parent.glob("*.py")):
        importlib.import_module(f"hdfe_stream_amd.{f.stem}")


@pytest.mark.parametrize("spec, want", [
    ("iid", ("iid", None)),
    ("hetero", ("hetero", None)),
    ("HC1", ("hetero", None)),
    ("CRV1:worker_id", ("CRV1", [["worker_id"]])),
    ({"CRV1": "worker_id+firm_id"}, ("CRV1", [["worker_id"], ["firm_id"]])),
    ("CRV1:firm_id^year", ("CRV1", [["firm_id", "year"]])),
])
def test_vcov_specs(spec, want):
    from hdfe_stream_amd.vcov import _parse
    assert _parse(spec) == want


@pytest.mark.parametrize("spec", ["CRV3:worker_id", "CRV1:", "robust", {"a": 1, "b": 2}])
def test_vcov_specs_refused(spec):
    from hdfe_stream_amd.vcov import _parse
    with pytest.raises(ValueError):
        _parse(spec)
