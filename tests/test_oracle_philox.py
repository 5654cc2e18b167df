"""Pins the oracle's Philox4x64-10 (C and numpy) against numpy.random.Philox
and the Random123 known-answer vector."""
import numpy as np
import pytest

import oracle
from oracle import philox as nphil

KAT_ZERO = [0x16554D9ECA36314C, 0xDB20FE9D672D0FDC, 0xD7E772CEE186176B, 0x7E68B68AEC7BA23B]
MASK256 = (1 << 256) - 1


def numpy_block(counter, key):
    """numpy's Philox pre-increments its 256-bit counter before each block."""
    c = (counter[0] | counter[1] << 64 | counter[2] << 128 | counter[3] << 192)
    c = (c - 1) & MASK256
    words = [(c >> (64 * w)) & (2 ** 64 - 1) for w in range(4)]
    bg = np.random.Philox(counter=words, key=list(key))
    return [int(v) for v in bg.random_raw(4)]


def test_known_answer_zero():
    assert [int(v) for v in oracle.philox4x64_10([0, 0, 0, 0], [0, 0])] == KAT_ZERO
    w = nphil.philox4x64_10(0, 0, 0, 0, 0, 0)
    assert [int(v[()]) for v in w] == KAT_ZERO
    assert numpy_block([0, 0, 0, 0], (0, 0)) == KAT_ZERO


@pytest.mark.parametrize("seed", range(8))
def test_against_numpy_philox(seed):
    rng = np.random.default_rng(seed)
    ctr = [int(v) for v in rng.integers(0, 2 ** 63, 4)]
    key = [int(v) for v in rng.integers(0, 2 ** 63, 2)]
    ref = numpy_block(ctr, key)
    assert [int(v) for v in oracle.philox4x64_10(ctr, key)] == ref
    w = nphil.philox4x64_10(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1])
    assert [int(v[()]) for v in w] == ref


def test_uniform_mapping_matches_numpy_generator():
    # Generator(Philox).random() maps each 64-bit word to (x >> 11) * 2^-53
    key = (12345, 7)
    bg = np.random.Philox(counter=[2 ** 64 - 1, 2 ** 64 - 1, 2 ** 64 - 1, 2 ** 64 - 1], key=list(key))
    g = np.random.Generator(bg)
    ref = g.random(4)  # the four words of block counter 0
    w = oracle.philox4x64_10([0, 0, 0, 0], key)
    assert np.array_equal(nphil.u01(np.array(w, dtype=np.uint64)), ref)


def test_pair_and_walk_streams():
    key = (3, 9)
    u = oracle.philox_uniforms(100, 0, 17, 0, 0, key[0], key[1], 0)
    assert np.array_equal(u, nphil.pair_uniforms(100, 17, key))
    ul, uo = nphil.walk_uniforms(50, 4, key)
    assert np.array_equal(ul, oracle.philox_uniforms(50, 0, 4, 1, 0, key[0], key[1], 0))
    assert np.array_equal(uo, oracle.philox_uniforms(50, 0, 4, 1, 0, key[0], key[1], 1))
    assert np.all((u >= 0) & (u < 1))
    # distinct streams
    assert not np.array_equal(u[:50], ul)
