"""The oracle against the committed golden fixtures (tests/golden/).

The reference ships no golden vectors, so the fixtures were written by the
oracle (tests/golden/make_fixtures.py); these tests pin the oracle to them and
cross-check the parts a library can confirm (numpy's Philox, scipy's cKDTree).
"""
import hashlib
import os

import numpy as np
from scipy.spatial import cKDTree

import oracle
from oracle.philox import philox4x64_10 as philox_np
from oracle.rps import rps_rule
from lagrangian_microbes_amd import configs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def test_philox_words_and_uniforms():
    g = load("philox_kat")
    for c, k, words, uni in zip(g["counter"], g["key"], g["words"], g["uniforms"]):
        assert np.array_equal(oracle.philox4x64_10(c, k), words)
        got_np = np.array([int(v) for v in philox_np(*[int(v) for v in c], int(k[0]), int(k[1]))], dtype=np.uint64)
        assert np.array_equal(got_np, words)
        got = oracle.philox_uniforms(len(uni), *[int(v) for v in c], int(k[0]), int(k[1]), 0)
        assert np.array_equal(got, uni)


def test_rps_rule_table():
    g = load("rps_rule")
    pRS, pPR, pSP = (float(v) for v in g["probs"])
    for a, b, u, o1, o2 in zip(g["s1"], g["s2"], g["u"], g["out1"], g["out2"]):
        sp = np.array([a, b], dtype=np.int8)
        rps_rule(sp, 0, 1, float(u), pRS, pPR, pSP)
        assert (sp[0], sp[1]) == (o1, o2)


def test_pairs_ties_fixture_matches_oracle_and_ckdtree():
    g = load("pairs_ties")
    x, y, r = g["x"], g["y"], float(g["r"])
    pts = np.stack([x.astype(np.float64), y.astype(np.float64)], axis=1)
    for name, p in (("p1", 1.0), ("p2", 2.0), ("pinf", np.inf)):
        pi, pj = oracle.pairs_bruteforce(x, y, r, p)
        assert np.array_equal(pi, g[name + "_i"]) and np.array_equal(pj, g[name + "_j"])
        ref = np.array(sorted(cKDTree(pts).query_pairs(r, p=p)), dtype=np.int64).reshape(-1, 2)
        assert np.array_equal(ref[:, 0], g[name + "_i"]) and np.array_equal(ref[:, 1], g[name + "_j"])


def test_c1_oracle_reproduces_fixture():
    g = load("c1_24h")
    w = configs.rps_10k_steady()
    f = w.velocity_field()
    lon, lat = w.particles()
    sp = w.species()
    assert np.array_equal(lon, g["lon0"]) and np.array_equal(lat, g["lat0"]) and np.array_equal(sp, g["species0"])
    steps = len(g["npairs"])
    grid = (f.nx, f.ny, f.x0, f.y0, 1.0 / f.dx, 1.0 / f.dy)
    lo, la, rl, rt, oob = oracle.advect_rk4_f64(lon, lat, f.interleaved()[:1], grid, steps, w.dt, steady=True)
    assert oob == int(g["oob"])
    assert np.array_equal(lo, g["lon"]) and np.array_equal(la, g["lat"])
    for s in range(steps):
        assert digest(rl[s], rt[s]) == g["digest_records"][s]
        pi, pj = oracle.pairs_bruteforce(rl[s], rt[s], w.radius, w.norm)
        assert len(pi) == g["npairs"][s]
        assert digest(pi.astype(np.int32), pj.astype(np.int32)) == g["digest_pairs"][s]
        sp = oracle.rps_sequential(sp, pi, pj, w.p, w.p, w.p, w.key, s)
        assert digest(sp) == g["digest_species"][s]
    assert np.array_equal(sp, g["species"])
