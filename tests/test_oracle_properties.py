"""Property tests (hypothesis) that pin the oracle on the same generated
cases as the GPU property tests (tests/test_gpu_properties.py), with no GPU:

- the brute-force pair search against scipy's cKDTree.query_pairs (the call
  the reference makes, interaction_simulator.py:93,98) and the blocked numpy
  restatement, on lattice point sets with pairs at exactly r, cell edges,
  duplicates and one-ulp jitter, for p = 1, 2 and infinity;
- the C sequential RPS update against the pure-Python restatement of
  interactions.py:13-40 applied in sorted pair order, on chains, stars,
  cliques and random pair lists with invalid species and p of 0, 1 or drawn.
"""
import numpy as np
from hypothesis import given, settings, strategies as st

import oracle
from helpers import PROP_R as R, lattice_points, pair_lists
from oracle.pairs import pairs_numpy, sorted_kdtree_pairs
from oracle.rps import rps_sequential_py

SETTINGS = settings(max_examples=60, deadline=None, derandomize=True)


@SETTINGS
@given(pts=lattice_points(), p=st.sampled_from([1.0, 2.0, np.inf]))
def test_pair_oracles_agree(pts, p):
    x, y = pts
    bi, bj = oracle.pairs_bruteforce(x, y, R, p)
    ni, nj = pairs_numpy(x, y, R, p)
    assert np.array_equal(bi, ni) and np.array_equal(bj, nj)
    if x.size > 1:
        ki, kj = sorted_kdtree_pairs(x, y, R, p)
        assert np.array_equal(bi, ki) and np.array_equal(bj, kj)


@SETTINGS
@given(case=pair_lists(), probs=st.sampled_from(["zero", "one", "drawn"]),
       p3=st.tuples(st.floats(0.0, 1.0), st.floats(0.0, 1.0), st.floats(0.0, 1.0)),
       key=st.tuples(st.integers(0, 2 ** 32 - 1), st.integers(0, 2 ** 32 - 1)), step=st.integers(0, 10 ** 6))
def test_rps_oracles_agree(case, probs, p3, key, step):
    species, pi, pj = case
    pRS, pPR, pSP = {"zero": (0.0, 0.0, 0.0), "one": (1.0, 1.0, 1.0), "drawn": p3}[probs]
    c = oracle.rps_sequential(species, pi, pj, pRS, pPR, pSP, key, step)
    py = rps_sequential_py(species, pi, pj, pRS, pPR, pSP, key, step)
    assert np.array_equal(c, py)
    # invalid species never change; valid ones stay valid
    valid = (species >= 1) & (species <= 3)
    assert np.array_equal(c[~valid], species[~valid])
    assert np.all((c[valid] >= 1) & (c[valid] <= 3))
