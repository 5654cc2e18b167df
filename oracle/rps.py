"""ORACLE (test infrastructure): the rock-paper-scissors rule in pure Python.

``rps_rule`` restates rock_paper_scissors_interaction (interactions.py:13-40)
with the uniform passed in instead of drawn from numpy's global generator
(interactions.py:20).  ``rps_sequential_py`` is the loop of
interaction_simulator.py:104-105 over the sorted pair list, with pair k using
the keyed Philox uniform of (k, step).  Pure-Python loops: small cases only.
"""
import numpy as np

from .philox import pair_uniforms

ROCK, PAPER, SCISSORS = 1, 2, 3


def rps_rule(species, p1, p2, r, pRS, pPR, pSP):
    """One pair interaction, mutating ``species`` in place (interactions.py:13-40)."""
    if species[p1] == species[p2]:
        return
    s1, s2 = species[p1], species[p2]
    winner = None
    if s1 == ROCK and s2 == SCISSORS:
        winner = p1 if r < pRS else p2
    elif s1 == ROCK and s2 == PAPER:
        winner = p2 if r < pPR else p1
    elif s1 == PAPER and s2 == ROCK:
        winner = p1 if r < pPR else p2
    elif s1 == PAPER and s2 == SCISSORS:
        winner = p2 if r < pSP else p1
    elif s1 == SCISSORS and s2 == ROCK:
        winner = p2 if r < pRS else p1
    elif s1 == SCISSORS and s2 == PAPER:
        winner = p1 if r < pSP else p2
    if winner == p1:
        species[p2] = species[p1]
    elif winner == p2:
        species[p1] = species[p2]


def rps_sequential_py(species, pair_i, pair_j, pRS, pPR, pSP, key, step, uniforms=None):
    """Visit the sorted pairs in order; returns the updated copy of species."""
    sp = np.array(species, dtype=np.int8, copy=True)
    u = pair_uniforms(len(pair_i), step, key) if uniforms is None else uniforms
    for k, (a, b) in enumerate(zip(np.asarray(pair_i).tolist(), np.asarray(pair_j).tolist())):
        rps_rule(sp, a, b, float(u[k]), pRS, pPR, pSP)
    return sp
