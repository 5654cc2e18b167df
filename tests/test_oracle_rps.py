"""Pins the oracle's rock-paper-scissors rule (interactions.py:13-40) and its
sequential sorted-order loop (interaction_simulator.py:104-105)."""
import itertools

import numpy as np
import pytest

import oracle
from oracle.philox import pair_uniforms
from oracle.rps import rps_rule, rps_sequential_py

P = {"pRS": 0.3, "pPR": 0.6, "pSP": 0.8}


def expected_winner_species(s1, s2, r):
    """Game definition written independently of the oracle's branch chain:
    rock beats scissors with pRS, paper beats rock with pPR, scissors beat
    paper with pSP; the dominant species wins iff r < p, else the other."""
    dominant = {frozenset((1, 3)): 1, frozenset((1, 2)): 2, frozenset((2, 3)): 3}[frozenset((s1, s2))]
    prob = {frozenset((1, 3)): P["pRS"], frozenset((1, 2)): P["pPR"], frozenset((2, 3)): P["pSP"]}[
        frozenset((s1, s2))]
    other = s1 if dominant == s2 else s2
    return dominant if r < prob else other


@pytest.mark.parametrize("s1,s2", list(itertools.product([1, 2, 3], repeat=2)))
@pytest.mark.parametrize("r", [0.0, 0.29999, 0.3, 0.59999, 0.6, 0.79999, 0.8, 0.999])
def test_rule_truth_table(s1, s2, r):
    sp = np.array([s1, s2], dtype=np.int8)
    rps_rule(sp, 0, 1, r, P["pRS"], P["pPR"], P["pSP"])
    if s1 == s2:
        assert list(sp) == [s1, s2]
    else:
        w = expected_winner_species(s1, s2, r)
        assert list(sp) == [w, w]


def test_invalid_species_never_interact():
    for a, b in [(0, 1), (4, 2), (-1, 3), (0, 5)]:
        sp = np.array([a, b], dtype=np.int8)
        rps_rule(sp, 0, 1, 0.1, 0.5, 0.5, 0.5)
        assert list(sp) == [a, b]


def test_p_limits_deterministic():
    # p = 1: the dominant species always wins; p = 0: never
    pi = np.array([0, 2, 4], np.int32)
    pj = np.array([1, 3, 5], np.int32)
    sp = np.array([1, 3, 2, 1, 3, 2], np.int8)  # R-S, P-R, S-P
    assert list(oracle.rps_sequential(sp, pi, pj, 1.0, 1.0, 1.0, (0, 0), 0)) == [1, 1, 2, 2, 3, 3]
    assert list(oracle.rps_sequential(sp, pi, pj, 0.0, 0.0, 0.0, (0, 0), 0)) == [3, 3, 1, 1, 2, 2]


def test_order_dependent_chain_hand_worked():
    # pairs (0,1), (0,2), (1,2) with p = 1 for every type.
    # start R P S: (0,1) R-P -> paper wins -> P P S; (0,2) P-S -> scissors -> S P S;
    # (1,2) P-S -> scissors -> S S S.  The reverse order gives R R R.
    sp = np.array([1, 2, 3], np.int8)
    pi, pj = np.array([0, 0, 1], np.int32), np.array([1, 2, 2], np.int32)
    assert list(oracle.rps_sequential(sp, pi, pj, 1, 1, 1, (0, 0), 0)) == [3, 3, 3]
    assert list(rps_sequential_py(sp, pi, pj, 1, 1, 1, (0, 0), 0)) == [3, 3, 3]
    assert list(rps_sequential_py(sp, pi[::-1], pj[::-1], 1, 1, 1, (0, 0), 0)) == [1, 1, 1]


@pytest.mark.parametrize("seed", range(5))
def test_c_and_python_sequential_agree(seed):
    rng = np.random.default_rng(seed)
    n = 300
    pairs = set()
    while len(pairs) < 900:
        a, b = sorted(rng.integers(0, n, 2))
        if a != b:
            pairs.add((int(a), int(b)))
    arr = np.array(sorted(pairs), dtype=np.int32)
    sp = rng.integers(0, 5, n).astype(np.int8)  # includes invalid codes 0 and 4
    args = (0.55, 0.45, 0.65, (seed, 2), 11)
    c = oracle.rps_sequential(sp, arr[:, 0], arr[:, 1], *args)
    py = rps_sequential_py(sp, arr[:, 0], arr[:, 1], *args)
    assert np.array_equal(c, py)
    valid = (sp >= 1) & (sp <= 3)
    assert np.array_equal(c[~valid], sp[~valid])
    assert np.all((c[valid] >= 1) & (c[valid] <= 3))


def test_uniforms_consumed_per_pair_rank():
    # the pair of rank k uses uniform k of (k, step), whether or not earlier pairs drew
    sp = np.array([1, 3, 2, 2, 1, 3], np.int8)
    pi, pj = np.array([0, 2, 4], np.int32), np.array([1, 3, 5], np.int32)
    u = pair_uniforms(3, 5, (1, 1))
    out = oracle.rps_sequential(sp, pi, pj, 0.5, 0.5, 0.5, (1, 1), 5)
    exp0 = 1 if u[0] < 0.5 else 3  # R vs S
    exp2 = 1 if u[2] < 0.5 else 3  # R vs S (pair 1 is a like pair)
    assert list(out) == [exp0, exp0, 2, 2, exp2, exp2]
