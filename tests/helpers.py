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
