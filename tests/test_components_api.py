"""connected_components' argument checks, which run before anything reaches
the device (core/connected.py:320-375)."""

import numpy as np
import pytest


@pytest.mark.parametrize("fe,kw", [
    ("worker_id", {}),
    ("y ~ 1 | worker_id + firm_id", {}),
    ("worker_id + firm_id", {"min_degree": 0}),
    ("worker_id + firm_id", {"min_degree": (1, 2, 3)}),
    ("worker_id + firm_id", {"rank_by": "year"}),
    ("worker_id + worker_id", {}),
    ("worker_id + firm_id + year", {}),
])
def test_bad_arguments_are_refused(fe, kw):
    from hdfe_stream_amd import connected_components
    d = {"worker_id": np.arange(4), "firm_id": np.arange(4), "year": np.arange(4)}
    with pytest.raises(ValueError):
        connected_components(d, fe, **kw)


def test_missing_columns_are_named():
    from hdfe_stream_amd import connected_components
    with pytest.raises(ValueError, match="not found"):
        connected_components({"worker_id": np.arange(3)}, "worker_id + firm_id")
