"""leave_out_kss's argument checks, which run before anything reaches the
device (leaveout.py:1196-1215, _check_leave_out :1017-1038)."""

import pytest


@pytest.mark.parametrize("fml,kw,err", [
    ("y ~ x1", {}, ValueError),                                   # no fixed effects
    ("y ~ 1 | worker_id", {}, ValueError),                        # one fixed effect
    ("y ~ 1 | worker_id + firm_id", {"psi": "worker_id"}, ValueError),
    ("y ~ 1 | worker_id + firm_id", {"psi": "year"}, ValueError),
    ("y ~ 1 | worker_id + firm_id", {"leave_out": "spell"}, ValueError),
    ("y ~ 1 | worker_id + firm_id", {"stayers": "own"}, ValueError),     # match level
    ("y ~ 1 | worker_id + firm_id", {"centering": "mean"}, ValueError),
    ("y ~ 1 | worker_id + firm_id", {"leave_out": "observation", "stayers": "drop"},
     NotImplementedError),
    ("y ~ 1 | worker_id + firm_id", {"se": True, "se_variance": "row"}, ValueError),
    ("y ~ 1 | worker_id + firm_id", {"se": True, "weights": "w"}, ValueError),  # person_year
])
def test_bad_arguments_are_refused_before_any_work(fml, kw, err):
    from hdfe_stream_amd.leaveout import leave_out_kss
    with pytest.raises(err):
        leave_out_kss(fml, {}, **kw)


def test_moment_divisor():
    from hdfe_stream_amd.leaveout import _moment_divisor
    assert _moment_divisor(10.0, 10) == 9.0
    assert _moment_divisor(20.0, 10) == 18.0
    assert _moment_divisor(3.0, 1) == 3.0
