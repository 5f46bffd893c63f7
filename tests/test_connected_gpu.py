"""Leave-one-out pruning on the device against the oracle's restatement of
the reference's rule (which is itself pinned against brute force): the same
rows, bit for bit, and the same counts."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n_workers,n_firms,p_move,seed", [
    (400, 60, 0.15, 4), (3000, 400, 0.05, 9), (2000, 80, 0.3, 2), (20000, 2000, 0.15, 3)])
def test_leave_one_out_matches_the_oracle(oracle_mod, n_workers, n_firms, p_move, seed):
    from hdfe_stream_amd.connected import leave_one_out
    from hdfe_stream_amd.simulate import akm_panel
    O = oracle_mod
    d = akm_panel(n_workers, n_firms, p_move=p_move, seed=seed)
    rng = np.random.default_rng(seed)
    keep_in = rng.random(len(d["y"])) > 0.3       # ragged: workers with 1..8 rows
    wid, fid = d["worker_id"][keep_in], d["firm_id"][keep_in]
    keep, info = leave_one_out(wid, fid)
    a = np.unique(wid, return_inverse=True)[1]
    b = np.unique(fid, return_inverse=True)[1]
    n0, n1 = a.max() + 1, b.max() + 1
    alive, ref = O.leave_one_out(a, b, n0, n1)
    np.testing.assert_array_equal(keep.cpu().numpy(), alive[a] & alive[n0 + b])
    for key in ("components_before", "single_observation", "cut", "rounds"):
        assert info[key] == ref[key], key
    assert info["workers"] == int(alive[:n0].sum()) and info["firms"] == int(alive[n0:].sum())
