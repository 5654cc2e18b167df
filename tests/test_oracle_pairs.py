"""Pins the oracle's pair search against scipy.spatial.cKDTree.query_pairs,
the library call the reference makes (interaction_simulator.py:93,98)."""
import numpy as np
import pytest

import oracle
from oracle.pairs import pairs_numpy, sorted_kdtree_pairs

R = 1000.0 / (1852.0 * 60.0)


def cloud(seed, n, spread=0.2):
    rng = np.random.default_rng(seed)
    x = (210.0 + rng.uniform(0, spread, n)).astype(np.float32)
    y = (30.0 + rng.uniform(0, spread, n)).astype(np.float32)
    return x, y


@pytest.mark.parametrize("p", [2.0, 1.0, np.inf])
@pytest.mark.parametrize("seed", range(3))
def test_bruteforce_matches_kdtree(seed, p):
    x, y = cloud(seed, 3000)
    ki, kj = sorted_kdtree_pairs(x, y, R, p)
    bi, bj = oracle.pairs_bruteforce(x, y, R, p)
    ni, nj = pairs_numpy(x, y, R, p)
    assert len(ki) > 100
    assert np.array_equal(ki, bi) and np.array_equal(kj, bj)
    assert np.array_equal(ki, ni) and np.array_equal(kj, nj)


def test_closed_ball_ties_and_duplicates():
    # points exactly at distance r along an axis: the pair is kept (<=)
    r = np.float32(0.0078125)  # 2^-7: exact in float32 and its square exact
    x = np.array([210.0, 210.0 + r, 210.0 + 2 * r + 1e-5, 210.0, 210.0], dtype=np.float32)
    y = np.array([30.0, 30.0, 30.0, 30.0 + r, 30.0], dtype=np.float32)
    ki, kj = sorted_kdtree_pairs(x, y, float(r))
    bi, bj = oracle.pairs_bruteforce(x, y, float(r))
    assert np.array_equal(ki, bi) and np.array_equal(kj, bj)
    got = set(zip(bi.tolist(), bj.tolist()))
    assert (0, 1) in got and (0, 3) in got and (0, 4) in got and (3, 4) in got
    assert (1, 2) not in got  # just beyond r


def test_lattice_at_exact_spacing():
    # a float32 lattice whose spacing is exactly r: each point pairs with its 4-neighbours
    r = 2.0 ** -7
    g = np.arange(30, dtype=np.float64) * r
    X, Y = np.meshgrid(210.0 + g, 30.0 + g, indexing="ij")
    x, y = X.ravel().astype(np.float32), Y.ravel().astype(np.float32)
    bi, bj = oracle.pairs_bruteforce(x, y, r)
    ki, kj = sorted_kdtree_pairs(x, y, r)
    assert len(bi) == 2 * 30 * 29
    assert np.array_equal(ki, bi) and np.array_equal(kj, bj)


def test_empty_and_single():
    for n in (0, 1):
        x = np.zeros(n, np.float32)
        bi, bj = oracle.pairs_bruteforce(x, x, R)
        assert len(bi) == 0
