"""ORACLE (test infrastructure): numpy restatement of Philox4x64-10.

Second, independent implementation beside lm_oracle.c; both are pinned
against numpy.random.Philox and the Random123 known-answer vector in
tests/test_oracle_philox.py.  The counter layout is the one the GPU path
uses: pair interactions draw word 0 of counter (pair rank, step, 0, 0); the
random walk draws words 0 (lat) and 1 (lon) of (particle, step, 1, 0).
"""
import numpy as np

M0 = np.uint64(0xD2E7470EE14C6C93)
M1 = np.uint64(0xCA5A826395121157)
W0 = np.uint64(0x9E3779B97F4A7C15)
W1 = np.uint64(0xBB67AE8584CAA73B)
_MASK32 = np.uint64(0xFFFFFFFF)
_S32 = np.uint64(32)


def _mulhilo(a, b):
    """128-bit product of uint64 arrays via 32-bit limbs -> (hi, lo)."""
    a = np.asarray(a, dtype=np.uint64)
    b = np.asarray(b, dtype=np.uint64)
    al, ah = a & _MASK32, a >> _S32
    bl, bh = b & _MASK32, b >> _S32
    p0, p1, p2, p3 = al * bl, al * bh, ah * bl, ah * bh
    mid = (p0 >> _S32) + (p1 & _MASK32) + (p2 & _MASK32)
    hi = p3 + (p1 >> _S32) + (p2 >> _S32) + (mid >> _S32)
    with np.errstate(over="ignore"):
        lo = a * b
    return hi, lo


def philox4x64_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x64-10; returns four uint64 arrays."""
    c0, c1, c2, c3 = (np.array(c, dtype=np.uint64, copy=True) for c in np.broadcast_arrays(c0, c1, c2, c3))
    k0 = np.uint64(k0)
    k1 = np.uint64(k1)
    with np.errstate(over="ignore"):
        for _ in range(10):
            hi0, lo0 = _mulhilo(M0, c0)
            hi1, lo1 = _mulhilo(M1, c2)
            c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
            k0 = k0 + W0
            k1 = k1 + W1
    return c0, c1, c2, c3


def u01(x):
    """Top 53 bits of a uint64 word as a double in [0, 1) (numpy's mapping)."""
    return (np.asarray(x, dtype=np.uint64) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def pair_uniforms(npairs, step, key):
    """Uniform of pair rank k = 0..npairs-1 at a step (word 0 of (k, step, 0, 0))."""
    k = np.arange(int(npairs), dtype=np.uint64)
    w = philox4x64_10(k, np.uint64(step), np.uint64(0), np.uint64(0), key[0], key[1])
    return u01(w[0])


def walk_uniforms(n, step, key):
    """(u_lat, u_lon) of particles 0..n-1 at a step (words 0, 1 of (i, step, 1, 0))."""
    i = np.arange(int(n), dtype=np.uint64)
    w = philox4x64_10(i, np.uint64(step), np.uint64(1), np.uint64(0), key[0], key[1])
    return u01(w[0]), u01(w[1])
