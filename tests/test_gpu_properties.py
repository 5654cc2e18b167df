"""Property tests (hypothesis) of the two exact kernels against the oracle,
through the C ABI on the GPU.

- Pair search (interaction_simulator.py:88-99, cKDTree.query_pairs
  semantics): random point sets drawn on a lattice of r/4 with r = 2^-7
  degrees, so that axis pairs lie exactly at distance r (the closed ball
  keeps them) and points sit on cell edges; plus exact duplicates, jittered
  copies one float32 step off the lattice, and sets of 0-2 points.  Norms
  p = 1, 2 and infinity.  Bit-exact against brute force.
- RPS update (interactions.py:13-40 applied in sorted pair order): random
  sorted pair lists over small particle sets (chains, stars, cliques and
  isolated pairs mixed), species including values outside {1, 2, 3}, and
  p in {0, 1, drawn}.  Bit-exact against the sequential oracle.
"""
import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings, strategies as st

import oracle
from helpers import PROP_R as R, lattice_points, pair_lists
from test_gpu_kernels import gpu_pairs, gpu_rps

pytestmark = pytest.mark.gpu

SETTINGS = settings(max_examples=60, deadline=None, derandomize=True,
                    suppress_health_check=[HealthCheck.function_scoped_fixture, HealthCheck.too_slow])


@SETTINGS
@given(pts=lattice_points(), p=st.sampled_from([1.0, 2.0, np.inf]))
def test_pair_search_property(gpu, pts, p):
    x, y = pts
    bbox = (209.9, 210.4, 29.9, 30.4)
    gi, gj = gpu_pairs(x, y, R, p, bbox=bbox)
    bi, bj = oracle.pairs_bruteforce(x, y, R, p)
    assert np.array_equal(gi, bi) and np.array_equal(gj, bj)


@SETTINGS
@given(case=pair_lists(), probs=st.sampled_from(["zero", "one", "drawn"]),
       p3=st.tuples(st.floats(0.0, 1.0), st.floats(0.0, 1.0), st.floats(0.0, 1.0)),
       key=st.tuples(st.integers(0, 2 ** 32 - 1), st.integers(0, 2 ** 32 - 1)), step=st.integers(0, 10 ** 6))
def test_rps_property(gpu, case, probs, p3, key, step):
    species, pi, pj = case
    pRS, pPR, pSP = {"zero": (0.0, 0.0, 0.0), "one": (1.0, 1.0, 1.0), "drawn": p3}[probs]
    got, _ = gpu_rps(species, pi, pj, pRS, pPR, pSP, key, step)
    want = oracle.rps_sequential(species, pi, pj, pRS, pPR, pSP, key, step)
    assert np.array_equal(got, want)
