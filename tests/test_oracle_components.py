"""The oracle's standalone components (core/connected.py:378-493) pinned by
what fixes them independently: scipy's connected components for the
partition, a brute-force search over vertex subsets for the (k0, k1)-core,
and hand-built graphs for the labelling and ranking rules."""

import itertools

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components as scipy_cc


def _random_rows(seed, n0=40, n1=15, n=90):
    rng = np.random.default_rng(seed)
    return rng.integers(0, n0, n), rng.integers(0, n1, n), n0, n1


def _used(a, b, n0, n1):
    """Re-code to the levels present (the device sees only those)."""
    ua, a2 = np.unique(a, return_inverse=True)
    ub, b2 = np.unique(b, return_inverse=True)
    return a2, b2, len(ua), len(ub)


@pytest.mark.parametrize("seed", range(6))
def test_partition_matches_scipy(oracle_mod, seed):
    a, b, n0, n1 = _used(*_random_rows(seed))
    comp, sizes, diag = oracle_mod.connected_components(a, b, n0, n1)
    A = sp.coo_matrix((np.ones(len(a)), (a, n0 + b)), shape=(n0 + n1, n0 + n1))
    nc, lab = scipy_cc(A, directed=False)
    assert len(sizes) == nc
    # the same partition: a bijection between the two labellings
    pairs = set(zip(comp.tolist(), lab.tolist()))
    assert len(pairs) == nc == len({p[0] for p in pairs}) == len({p[1] for p in pairs})
    # sizes: rows and levels per component, largest first (rows)
    assert sizes[:, 0].sum() == len(a)
    assert list(sizes[:, 0]) == sorted(sizes[:, 0], reverse=True)
    for c in range(nc):
        assert sizes[c, 1] == (comp[:n0] == c).sum() and sizes[c, 2] == (comp[n0:] == c).sum()
        assert sizes[c, 0] == (comp[a] == c).sum()


def test_ties_go_to_the_smallest_first_dimension_level(oracle_mod):
    # three components of two rows each: {w0,w3 | f1}, {w1 | f0}, {w2 | f2}
    a = np.array([3, 0, 1, 1, 2, 2])
    b = np.array([1, 1, 0, 0, 2, 2])
    comp, sizes, _ = oracle_mod.connected_components(a, b, 4, 3)
    # equal rows: ranked in the order of their smallest worker (0, 1, 2)
    assert comp[:4].tolist() == [0, 1, 2, 0]
    assert comp[4:].tolist() == [1, 0, 2]
    # by the workers' count the first has 2 and goes first; the others tie
    comp, sizes, _ = oracle_mod.connected_components(a, b, 4, 3, rank_by=0)
    assert sizes[:, 1].tolist() == [2, 1, 1] and comp[:4].tolist() == [0, 1, 2, 0]


def _core_brute(keys, n0, n1, k0, k1):
    """The largest vertex set in which every vertex has at least k partners
    inside it (the union of all such sets is one), by brute force."""
    ka, kb = keys // n1, n0 + keys % n1
    nv = n0 + n1
    best = np.zeros(nv, bool)
    for r in range(nv, 0, -1):
        for sub in itertools.combinations(range(nv), r):
            al = np.zeros(nv, bool)
            al[list(sub)] = True
            ok = al[ka] & al[kb]
            deg = np.bincount(ka[ok], minlength=nv) + np.bincount(kb[ok], minlength=nv)
            thr = np.where(np.arange(nv) < n0, k0, k1)
            if np.all(deg[al] >= thr[al]):
                best |= al
        if best.any():
            return best
    return best


@pytest.mark.parametrize("seed,k", [(0, (2, 1)), (1, (2, 2)), (2, (1, 3)), (3, (2, 2)),
                                    (4, (3, 2))])
def test_peel_is_the_core(oracle_mod, seed, k):
    rng = np.random.default_rng(seed)
    a, b = rng.integers(0, 6, 14), rng.integers(0, 5, 14)
    a, b, n0, n1 = _used(a, b, 6, 5)
    keys = np.unique(a * n1 + b)
    alive = np.ones(n0 + n1, bool)
    oracle_mod.peel(keys, n0, n1, k[0], k[1], alive)
    np.testing.assert_array_equal(alive, _core_brute(keys, n0, n1, *k))


def test_peel_rounds_on_a_chain(oracle_mod):
    # a path w0-f0-w1-f1-w2-f2: min_degree 2 peels from both ends inward
    a = np.array([0, 1, 1, 2, 2])
    b = np.array([0, 0, 1, 1, 2])
    keys = np.unique(a * 3 + b)
    alive = np.ones(6, bool)
    rounds = oracle_mod.peel(keys, 3, 3, 2, 2, alive)
    assert not alive.any() and rounds == 3


def test_pruned_levels_have_no_component(oracle_mod):
    a, b, n0, n1 = _used(*_random_rows(7, 30, 10, 70))
    comp, sizes, diag = oracle_mod.connected_components(a, b, n0, n1, min_degree=(2, 1))
    # workers seen at a single firm are gone; what is left forms components
    firms_of = [set(b[a == w]) for w in range(n0)]
    for w in range(n0):
        if len(firms_of[w]) < 2:
            assert comp[w] == -1
    assert diag["pruned_levels"][0] >= sum(len(f) < 2 for f in firms_of)
    assert sizes[:, 0].sum() + diag["pruned_rows"] == len(a)


def test_leave_one_out_after_min_degree(oracle_mod):
    a, b, n0, n1 = _used(*_random_rows(9, 60, 12, 200))
    comp, sizes, diag = oracle_mod.connected_components(a, b, n0, n1, min_degree=(1, 2),
                                                        leave_one_out_set=True)
    assert len(sizes) <= 1 and set(np.unique(comp)) <= {-1, 0}
    lo = diag["leave_one_out"]
    assert sizes[0, 0] + lo["rows"] + diag.get("pruned_rows", 0) == len(a)
