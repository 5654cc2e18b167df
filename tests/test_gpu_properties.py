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
from test_gpu_kernels import gpu_pairs, gpu_rps

pytestmark = pytest.mark.gpu

SETTINGS = settings(max_examples=60, deadline=None, derandomize=True,
                    suppress_health_check=[HealthCheck.function_scoped_fixture, HealthCheck.too_slow])

R = 2.0 ** -7  # exact in float32: lattice multiples of R / 4 give exact distances


@st.composite
def lattice_points(draw):
    n = draw(st.integers(0, 160))
    cells = draw(st.lists(st.tuples(st.integers(0, 48), st.integers(0, 48)), min_size=n, max_size=n))
    x = np.array([210.0 + c[0] * R / 4 for c in cells], dtype=np.float32)
    y = np.array([30.0 + c[1] * R / 4 for c in cells], dtype=np.float32)
    extra = draw(st.sampled_from(["none", "duplicates", "jitter"]))
    if extra != "none" and n:
        k = draw(st.integers(1, min(n, 20)))
        dx, dy = x[:k].copy(), y[:k].copy()
        if extra == "jitter":  # one float32 step off the lattice, either way
            sgn = np.array(draw(st.lists(st.sampled_from([-1.0, 1.0]), min_size=k, max_size=k)), np.float32)
            dx = np.nextafter(dx, dx + sgn * np.float32(1.0)).astype(np.float32)
        x, y = np.concatenate([x, dx]), np.concatenate([y, dy])
    return x, y


@SETTINGS
@given(pts=lattice_points(), p=st.sampled_from([1.0, 2.0, np.inf]))
def test_pair_search_property(gpu, pts, p):
    x, y = pts
    bbox = (209.9, 210.4, 29.9, 30.4)
    gi, gj = gpu_pairs(x, y, R, p, bbox=bbox)
    bi, bj = oracle.pairs_bruteforce(x, y, R, p)
    assert np.array_equal(gi, bi) and np.array_equal(gj, bj)


@st.composite
def pair_lists(draw):
    n = draw(st.integers(2, 120))
    kind = draw(st.sampled_from(["random", "chain", "star", "clique", "mixed"]))
    pairs = set()
    if kind in ("chain", "mixed"):
        a0 = draw(st.integers(0, n - 2))
        length = draw(st.integers(1, n - 1 - a0))
        pairs |= {(a, a + 1) for a in range(a0, a0 + length)}
    if kind in ("star", "mixed"):
        c = draw(st.integers(0, n - 1))
        leaves = draw(st.lists(st.integers(0, n - 1), max_size=30))
        pairs |= {(min(c, v), max(c, v)) for v in leaves if v != c}
    if kind == "clique":
        members = sorted(set(draw(st.lists(st.integers(0, n - 1), min_size=2, max_size=14))))
        pairs |= {(a, b) for k, a in enumerate(members) for b in members[k + 1:]}
    if kind in ("random", "mixed"):
        extra = draw(st.lists(st.tuples(st.integers(0, n - 1), st.integers(0, n - 1)), max_size=200))
        pairs |= {(min(a, b), max(a, b)) for a, b in extra if a != b}
    arr = np.array(sorted(pairs), dtype=np.int32).reshape(-1, 2)
    species = np.array(draw(st.lists(st.integers(0, 4), min_size=n, max_size=n)), dtype=np.int8)
    return species, arr[:, 0].copy(), arr[:, 1].copy()


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
