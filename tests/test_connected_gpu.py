"""Leave-one-out pruning on the device against the oracle's restatement of
the reference's rule (which is itself pinned against brute force): the same
rows, bit for bit, and the same counts."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n_workers,n_firms,p_move,seed", [
    (400, 60, 0.15, 4), (3000, 400, 0.05, 9), (2000, 80, 0.3, 2), (20000, 2000, 0.15, 3),
    (200000, 20000, 0.05, 7)])
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


def _panel(seed, n_workers=2000, n_firms=700, p_move=0.03):
    from hdfe_stream_amd.simulate import akm_panel
    d = akm_panel(n_workers, n_firms, p_move=p_move, seed=seed)
    keep = np.random.default_rng(seed).random(len(d["y"])) > 0.4
    return {k: np.asarray(v)[keep] for k, v in d.items()}


def _oracle_cc(O, a_ids, b_ids, min_degree, loo, rank_by):
    ua, a = np.unique(a_ids, return_inverse=True)
    ub, b = np.unique(b_ids, return_inverse=True)
    comp, sizes, diag = O.connected_components(a, b, len(ua), len(ub), min_degree, loo, rank_by)
    return ua, ub, a, b, comp, sizes, diag


@pytest.mark.parametrize("min_degree,loo,rank_by", [
    ((1, 1), False, "rows"), ((2, 1), False, "rows"), ((2, 2), False, "firm_id"),
    ((1, 1), False, "worker_id"), ((1, 1), True, "rows"), ((1, 2), True, "worker_id"),
    ((2, 1), True, "firm_id")])
def test_connected_components_match_the_oracle(oracle_mod, min_degree, loo, rank_by):
    from hdfe_stream_amd import connected_components
    d = _panel(7)
    print(f"components min_degree={min_degree} loo={loo} rank_by={rank_by}", flush=True)
    cc = connected_components(d, "worker_id + firm_id", min_degree=min_degree,
                              leave_one_out=loo, rank_by=rank_by)
    rb = None if rank_by == "rows" else ["worker_id", "firm_id"].index(rank_by)
    ua, ub, a, b, comp, sizes, diag = _oracle_cc(oracle_mod, d["worker_id"], d["firm_id"],
                                                 min_degree, loo, rb)
    n0 = len(ua)
    lw, lf = cc.levels["worker_id"], cc.levels["firm_id"]
    np.testing.assert_array_equal(lw["worker_id"].to_numpy(), ua)
    np.testing.assert_array_equal(lf["firm_id"].to_numpy(), ub)
    got_w = lw["component"].to_numpy(dtype="float64", na_value=-1).astype(np.int64)
    got_f = lf["component"].to_numpy(dtype="float64", na_value=-1).astype(np.int64)
    np.testing.assert_array_equal(got_w, comp[:n0])
    np.testing.assert_array_equal(got_f, comp[n0:])
    s = cc.sizes
    np.testing.assert_array_equal(s[["rows", "worker_id", "firm_id"]].to_numpy(), sizes)
    assert cc.n_components == len(sizes)
    assert cc.diagnostics["rows"] == len(a)
    if min_degree != (1, 1):
        assert cc.diagnostics["pruning_rounds"] == diag["pruning_rounds"]
        assert (cc.diagnostics["pruned_levels"]["worker_id"],
                cc.diagnostics["pruned_levels"]["firm_id"]) == diag["pruned_levels"]
        assert cc.diagnostics["pruned_rows"] == diag["pruned_rows"]
    if loo:
        lo, rlo = cc.diagnostics["leave_one_out"], diag["leave_one_out"]
        for k in ("components_before", "single_observation", "cut", "rounds", "rows"):
            assert lo[k] == rlo[k], k
        assert (lo["levels"]["worker_id"], lo["levels"]["firm_id"]) == rlo["levels"]
    # rows: each row's component, null unless both of its levels are in one
    lab = cc.label()
    want = np.where((comp[a] >= 0) & (comp[n0 + b] >= 0), comp[a], -1)
    got = lab["component"].to_numpy(dtype="float64", na_value=-1).astype(np.int64)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(lab["in_largest"].to_numpy(), want == 0)
    assert len(cc.largest()) == (int(sizes[0, 0]) if len(sizes) else 0)
    assert "components" in cc.summary()


def test_components_of_an_interaction_and_string_ids(oracle_mod):
    from hdfe_stream_amd import connected_components
    d = _panel(3, 1500, 300, 0.05)
    d["worker_id"] = np.array([f"w{v:06d}" for v in d["worker_id"]])
    d["firm_id"] = d["firm_id"].astype(np.float64)
    d["firm_id"][:7] = np.nan          # rows with a NaN level are in no component
    cc = connected_components(d, "worker_id + firm_id^year")
    ok = ~np.isnan(d["firm_id"])
    fy = np.unique(np.stack([d["firm_id"][ok], d["year"][ok]], 1), axis=0, return_inverse=True)[1]
    ua, ub, a, b, comp, sizes, _ = _oracle_cc(oracle_mod, d["worker_id"][ok], fy.ravel(),
                                              (1, 1), False, None)
    np.testing.assert_array_equal(cc.sizes[["rows", "worker_id", "firm_id^year"]].to_numpy(),
                                  sizes)
    t = cc.levels["firm_id^year"]
    assert list(t.columns) == ["firm_id", "year", "component"]
    got = t["component"].to_numpy(dtype="float64", na_value=-1).astype(np.int64)
    np.testing.assert_array_equal(got, comp[len(ua):])
    lab = cc.label()
    assert lab["component"][:7].isna().all()
    assert int(lab["in_largest"].sum()) == int(sizes[0, 0])

