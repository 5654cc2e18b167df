"""Write the golden fixtures under tests/golden/ from the oracle.

    python tests/golden/make_fixtures.py

The reference ships no golden vectors, known-answer tests or fixture files
for this path, so these are written by the oracle (oracle/) once and kept:
the CPU tests check that the oracle still reproduces them and the GPU tests
check the HIP path against them.  Contents:

- philox_kat.npz     Philox4x64-10 words and uniforms for a few (counter, key)
- rps_rule.npz       every ordered species pair in {0..4}^2 x u below / above
                     each probability: the species after one interaction
                     (interactions.py:13-40)
- pairs_ties.npz     float32 points on a lattice of exactly-representable
                     offsets, duplicates and points at distance r (closed
                     ball), with the sorted pair list for p = 1, 2, inf
- c1_24h.npz         the 10k-particle steady workload (rps_10k_steady): start
                     state, end state after 24 hourly steps, per-step pair
                     counts and SHA-256 digests of the float32 records, the
                     pair list and the species after every step
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from oracle.rps import rps_rule  # noqa: E402
from lagrangian_microbes_amd import configs  # noqa: E402


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def philox_kat():
    cases = [((0, 0, 0, 0), (0, 0)), ((1, 2, 3, 4), (5, 6)), ((2 ** 64 - 1, 2 ** 64 - 1, 2 ** 64 - 1, 2 ** 64 - 1),
                                                             (2 ** 64 - 1, 2 ** 64 - 1)),
             ((123456789, 7, 0, 0), (1, 3)), ((0x243F6A8885A308D3, 0x13198A2E03707344, 0xA4093822299F31D0,
                                               0x082EFA98EC4E6C89), (0x452821E638D01377, 0xBE5466CF34E90C6C))]
    ctr = np.array([c for c, _ in cases], dtype=np.uint64)
    key = np.array([k for _, k in cases], dtype=np.uint64)
    out = np.array([oracle.philox4x64_10(c, k) for c, k in cases], dtype=np.uint64)
    uni = np.stack([oracle.philox_uniforms(16, *c, k[0], k[1], 0) for c, k in cases])
    return dict(counter=ctr, key=key, words=out, uniforms=uni)


def rps_table():
    probs = (0.3, 0.5, 0.7)  # pRS, pPR, pSP
    us = np.array([0.0, 0.29, 0.31, 0.49, 0.51, 0.69, 0.71, 0.999])
    rows = []
    for a in range(5):
        for b in range(5):
            for u in us:
                sp = np.array([a, b], dtype=np.int8)
                rps_rule(sp, 0, 1, float(u), *probs)
                rows.append((a, b, u, sp[0], sp[1]))
    t = np.array(rows, dtype=np.float64)
    return dict(probs=np.array(probs), s1=t[:, 0].astype(np.int8), s2=t[:, 1].astype(np.int8), u=t[:, 2],
                out1=t[:, 3].astype(np.int8), out2=t[:, 4].astype(np.int8))


def pairs_ties():
    r = np.float32(0.25)
    g = np.arange(0, 8, dtype=np.float32) * np.float32(0.25)  # multiples of r: exact in float32
    x, y = np.meshgrid(g + np.float32(200.0), g + np.float32(30.0), indexing="ij")
    x, y = x.ravel(), y.ravel()
    # duplicates and diagonal points (distance r*sqrt(2) > r for p = 2 and inf-norm ties at r)
    x = np.concatenate([x, x[:5], np.float32([200.125, 200.375])])
    y = np.concatenate([y, y[:5], np.float32([30.125, 30.375])])
    out = dict(x=x.astype(np.float32), y=y.astype(np.float32), r=np.float64(r))
    for name, p in (("p1", 1.0), ("p2", 2.0), ("pinf", np.inf)):
        pi, pj = oracle.pairs_bruteforce(x, y, float(r), p)
        out[name + "_i"], out[name + "_j"] = pi, pj
    return out


def c1_run(steps=24):
    w = configs.rps_10k_steady()
    f = w.velocity_field()
    lon, lat = w.particles()
    sp = w.species()
    lon0, lat0, sp0 = lon.copy(), lat.copy(), sp.copy()
    grid = (f.nx, f.ny, f.x0, f.y0, 1.0 / f.dx, 1.0 / f.dy)
    if f.nt == 1:
        stage, steady = f.interleaved()[:1], True
    else:
        lev0, wts = f.stage_schedule(0.0, w.dt, steps)
        stage, steady = oracle.field_blend_f64(f.interleaved(), lev0, wts), False
    lo, la, rl, rt, oob = oracle.advect_rk4_f64(lon, lat, stage, grid, steps, w.dt, steady=steady)
    counts, d_rec, d_pairs, d_sp = [], [], [], []
    for s in range(steps):
        pi, pj = oracle.pairs_bruteforce(rl[s], rt[s], w.radius, w.norm)
        sp = oracle.rps_sequential(sp, pi, pj, w.p, w.p, w.p, w.key, s)
        counts.append(len(pi))
        d_rec.append(digest(rl[s], rt[s]))
        d_pairs.append(digest(pi.astype(np.int32), pj.astype(np.int32)))
        d_sp.append(digest(sp))
    return dict(lon0=lon0, lat0=lat0, species0=sp0, lon=lo, lat=la, species=sp, oob=np.int64(oob),
                npairs=np.array(counts, dtype=np.int64), digest_records=np.array(d_rec),
                digest_pairs=np.array(d_pairs), digest_species=np.array(d_sp))


def main():
    for name, fn in (("philox_kat", philox_kat), ("rps_rule", rps_table), ("pairs_ties", pairs_ties),
                     ("c1_24h", c1_run)):
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **fn())
        print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
