"""Shared test helpers: oracle runs of the named workloads."""
import numpy as np

import oracle


def grid_of(f):
    return (f.nx, f.ny, f.x0, f.y0, 1.0 / f.dx, 1.0 / f.dy)


def oracle_stage_fields(f, t0, dt, nsteps):
    if f.nt == 1:
        return f.interleaved()[:1], True
    lev0, w = f.stage_schedule(t0, dt, nsteps)
    return oracle.field_blend_f64(f.interleaved(), lev0, w), False


def oracle_advect(f, lon, lat, dt, nsteps, t0=0.0, walk_amp=0.0, key=(0, 0), step0=0):
    sf, steady = oracle_stage_fields(f, t0, dt, nsteps)
    return oracle.advect_rk4_f64(lon, lat, sf, grid_of(f), nsteps, dt, steady=steady, walk_amp=walk_amp,
                                 key=key, step0=step0)


def oracle_two_pass(w, nsteps=None, lon=None, lat=None, species=None):
    """Advection records, then per-step brute-force pairs + sequential RPS."""
    f = w.velocity_field()
    if lon is None:
        lon, lat = w.particles()
    sp = w.species() if species is None else species
    nsteps = w.steps if nsteps is None else nsteps
    lo, la, rl, rt, oob = oracle_advect(f, lon, lat, w.dt, nsteps)
    pairs, species_hist = [], []
    for s in range(nsteps):
        pi, pj = oracle.pairs_bruteforce(rl[s], rt[s], w.radius, w.norm)
        sp = oracle.rps_sequential(sp, pi, pj, w.p, w.p, w.p, w.key, s)
        pairs.append((pi, pj))
        species_hist.append(sp.copy())
    return dict(lon=lo, lat=la, rec_lon=rl, rec_lat=rt, oob=oob, pairs=pairs, species=np.array(species_hist))


# ---------------------------------------------------------------- hypothesis strategies
# Shared by the oracle's and the GPU's property tests.
from hypothesis import strategies as st  # noqa: E402

PROP_R = 2.0 ** -7  # exact in float32: lattice multiples of PROP_R / 4 give exact distances


@st.composite
def lattice_points(draw):
    """Point sets on a lattice of PROP_R / 4 (axis pairs at exactly PROP_R,
    points on cell edges), optionally with exact duplicates or copies one
    float32 step off the lattice; 0 to 180 points."""
    n = draw(st.integers(0, 160))
    cells = draw(st.lists(st.tuples(st.integers(0, 48), st.integers(0, 48)), min_size=n, max_size=n))
    x = np.array([210.0 + c[0] * PROP_R / 4 for c in cells], dtype=np.float32)
    y = np.array([30.0 + c[1] * PROP_R / 4 for c in cells], dtype=np.float32)
    extra = draw(st.sampled_from(["none", "duplicates", "jitter"]))
    if extra != "none" and n:
        k = draw(st.integers(1, min(n, 20)))
        dx, dy = x[:k].copy(), y[:k].copy()
        if extra == "jitter":  # one float32 step off the lattice, either way
            sgn = np.array(draw(st.lists(st.sampled_from([-1.0, 1.0]), min_size=k, max_size=k)), np.float32)
            dx = np.nextafter(dx, dx + sgn * np.float32(1.0)).astype(np.float32)
        x, y = np.concatenate([x, dx]), np.concatenate([y, dy])
    return x, y


@st.composite
def pair_lists(draw):
    """(species int8 [n], pair_i, pair_j): a sorted list of distinct i < j
    pairs mixing chains, stars, cliques and random pairs over n <= 120
    particles; species drawn from 0-4 (0 and 4 never interact)."""
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
