"""The rock-paper-scissors pair interaction (interactions.py of the reference).

``rock_paper_scissors(N, pRS, pPR, pSP)`` returns the plugin tuple
``(rock_paper_scissors_interaction, parameters, microbe_properties)`` that
InteractionSimulator takes (interactions.py:43-54).  When the simulator sees
this function it runs the GPU kernel; ``rock_paper_scissors_interaction``
itself is the host form of the rule for callers that invoke it directly.
"""
import numpy as np
from numpy import int8

# Species codes and plotting colours (interactions.py:5-10).
ROCK, PAPER, SCISSORS = int8(1), int8(2), int8(3)
ROCK_COLOR = "red"
PAPER_COLOR = "limegreen"
SCISSORS_COLOR = "blue"


def _beats(a, b):
    """Cyclic dominance on {1, 2, 3}: rock > scissors, paper > rock, scissors > paper."""
    return a == b % 3 + 1


def _forward_probability(parameters, a, b):
    """Probability that the dominant species of the unordered pair {a, b} wins."""
    s = int(a) + int(b)
    return parameters["pPR"] if s == 3 else parameters["pRS"] if s == 4 else parameters["pSP"]


def rock_paper_scissors_interaction(parameters, microbe_properties, p1, p2):
    """One interaction between microbes p1 < p2, mutating species in place.

    Same rule and random-number use as interactions.py:13-40: a uniform is
    drawn from numpy's global generator only when the species differ; the
    dominant species wins with its forward probability, else the other one
    wins; the loser takes the winner's species.  Codes outside {1, 2, 3}
    never win.
    """
    species = microbe_properties["species"]
    s1, s2 = species[p1], species[p2]
    if s1 == s2:
        return
    r = np.random.rand()
    if not (1 <= s1 <= 3 and 1 <= s2 <= 3):
        return
    forward = r < _forward_probability(parameters, s1, s2)
    if forward == _beats(s1, s2):
        species[p2] = s1
    else:
        species[p1] = s2


def rock_paper_scissors(N_microbes, pRS, pPR, pSP, seed=None):
    """Plugin tuple for InteractionSimulator (interactions.py:43-54).

    Species are drawn uniformly from {rock, paper, scissors} with numpy's
    global generator as in the reference, or from RandomState(seed) when a
    seed is given.
    """
    rng = np.random if seed is None else np.random.RandomState(seed)
    microbe_properties = {"species": rng.choice([ROCK, PAPER, SCISSORS], N_microbes)}
    interaction_parameters = {"pRS": pRS, "pPR": pPR, "pSP": pSP}
    return rock_paper_scissors_interaction, interaction_parameters, microbe_properties
