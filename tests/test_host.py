"""Host-side mirror of the reference API: lattice, tiles, field schedule,
output formats and the host rule.  No GPU."""
import os
from datetime import datetime, timedelta

import numpy as np
import pytest

import oracle
from oracle.rps import rps_rule
from lagrangian_microbes_amd import interactions, utils
from lagrangian_microbes_amd.netcdf_io import (decode_times, encode_times, read_particle_dataset,
                                               write_particle_dataset)
from lagrangian_microbes_amd.particle_advecter import (_parse_pickle_name, _pickle_name,
                                                       distribute_particles_across_tiles,
                                                       uniform_particle_locations)
from lagrangian_microbes_amd.velocity_fields import VelocityField, synthetic_fourier_field


@pytest.mark.parametrize("N", [1, 2, 12, 97, 10000, 490000, 4000000, 123456])
def test_factorization_brute_force(N):
    divs = [d for d in range(1, N + 1) if N % d == 0] if N < 200000 else utils.factor(N)
    assert utils.factor(N) == divs
    a, b = utils.most_symmetric_integer_factorization(N)
    best = min(divs, key=lambda d: (abs(d - N ** 0.5), d))
    assert (a, b) == (best, N // best)


def test_lattice_closed_form():
    lons, lats = uniform_particle_locations(490000, 25, 35, 205, 215)
    # particle k = i_lon * N_lat + i_lat, latitude fastest, end points included
    assert np.array_equal(lons, np.repeat(np.linspace(205, 215, 700), 700))
    assert np.array_equal(lats, np.tile(np.linspace(25, 35, 700), 700))


def test_tiles():
    a = np.arange(12.0)
    tl, tt = distribute_particles_across_tiles(a, a + 1, 4)
    assert [list(t) for t in tl] == [[0, 1, 2], [3, 4, 5], [6, 7, 8], [9, 10, 11]]
    with pytest.raises(AssertionError):
        distribute_particles_across_tiles(a, a, 5)


def test_pickle_names_round_trip():
    name = _pickle_name(100, 200, 3)
    assert name == "particle_locations_00100_00200_tile03.pickle"
    assert _parse_pickle_name("/x/" + name) == (100, 200, 3)


def test_stage_schedule():
    f = synthetic_fourier_field(200, 210, 20, 30, 0.5, 4, 0.2, 0, n_levels=3, level_interval=86400.0)
    lev, w = f.stage_schedule(0.0, 3600.0, 48)
    assert len(lev) == 97 and lev[0] == 0 and w[0] == 0.0
    assert lev[48] == 1 and w[48] == 0.0  # t = 24 h lands on level 1
    assert lev[-1] == 1 and w[-1] == 1.0  # t = 48 h is the last level: last interval, weight 1
    assert np.allclose(w[1], 1800.0 / 86400.0)
    with pytest.raises(ValueError):
        f.stage_schedule(0.0, 3600.0, 49)


def test_synthetic_field_properties():
    f = synthetic_fourier_field(200, 220, 20, 40, 0.25, 16, 0.25, 1, n_levels=2, level_interval=86400.0)
    assert f.u.shape == (2, 81, 81)
    for lev in range(2):
        assert np.isclose(np.sqrt(np.mean(f.u[lev] ** 2 + f.v[lev] ** 2)), 0.25)
    # non-divergent on the tangent plane: on a grid fine enough for finite
    # differences, du/dx + dv/dy vanishes relative to the vorticity
    fine = synthetic_fourier_field(205, 215, 25, 35, 0.01, 16, 0.25, 1)
    dx = 0.01 * 111120.0 * np.cos(np.deg2rad(30.0))
    dy = 0.01 * 111120.0
    div = np.gradient(fine.u[0], dx, axis=1) + np.gradient(fine.v[0], dy, axis=0)
    vort = np.gradient(fine.v[0], dx, axis=1) - np.gradient(fine.u[0], dy, axis=0)
    assert np.abs(div).mean() < 1e-3 * np.abs(vort).mean()
    g = synthetic_fourier_field(200, 220, 20, 40, 0.25, 16, 0.25, 1, n_levels=2, level_interval=86400.0)
    assert np.array_equal(f.u, g.u)


def test_velocity_field_flips_descending_latitude():
    lons = np.arange(0.0, 3.0)
    lats = np.array([2.0, 1.0, 0.0])
    u = np.arange(9.0).reshape(3, 3)
    f = VelocityField(lons, lats, u, u * 2)
    assert f.y0 == 0.0 and f.u[0, 0, 0] == 6.0
    with pytest.raises(ValueError):
        VelocityField(np.array([0.0, 1.0, 3.0]), lats, u, u)


def test_netcdf_round_trip(tmp_path):
    times = [datetime(2017, 1, 1) + k * timedelta(hours=1) for k in range(5)]
    vals, units = encode_times(times)
    assert units == "hours since 2017-01-01 00:00:00" and list(vals) == [0, 1, 2, 3, 4]
    assert decode_times(vals, units) == times
    lon = np.random.default_rng(0).normal(size=(7, 5)).astype(np.float32)
    sp = np.random.default_rng(1).integers(1, 4, (7, 5)).astype(np.int8)
    path = os.path.join(tmp_path, "microbe_data.nc")
    write_particle_dataset(path, {"longitude": lon, "species": sp}, times, 7)
    data, t2, n = read_particle_dataset(path)
    assert n == 7 and t2 == times
    assert np.array_equal(data["longitude"], lon) and np.array_equal(data["species"], sp)
    assert data["species"].dtype == np.int8 and data["longitude"].dtype == np.float32


def test_host_rule_matches_oracle_with_global_rng():
    # the host form draws numpy's global generator exactly when species differ
    rng = np.random.default_rng(5)
    sp0 = rng.integers(0, 5, 200).astype(np.int8)
    pairs = sorted({tuple(sorted(map(int, rng.integers(0, 200, 2)))) for _ in range(400)})
    pairs = [p for p in pairs if p[0] != p[1]]
    params = {"pRS": 0.4, "pPR": 0.55, "pSP": 0.7}
    props = {"species": sp0.copy()}
    np.random.seed(99)
    for a, b in pairs:
        interactions.rock_paper_scissors_interaction(params, props, a, b)
    ref = sp0.copy()
    np.random.seed(99)
    for a, b in pairs:
        if ref[a] != ref[b]:
            rps_rule(ref, a, b, np.random.rand(), params["pRS"], params["pPR"], params["pSP"])
    assert np.array_equal(props["species"], ref)


def test_factory_species():
    fn, params, props = interactions.rock_paper_scissors(1000, 0.5, 0.6, 0.7, seed=3)
    assert fn is interactions.rock_paper_scissors_interaction
    assert params == {"pRS": 0.5, "pPR": 0.6, "pSP": 0.7}
    assert props["species"].dtype == np.int8 and set(np.unique(props["species"])) == {1, 2, 3}
    assert np.array_equal(props["species"], np.random.RandomState(3).choice([1, 2, 3], 1000))


def test_pretty():
    assert utils.pretty_time(2e-9) == "2 ns" and utils.pretty_time(90) == "1.5 mins"
    assert utils.pretty_filesize(2048) == "2.0 KiB"


def test_ensemble_runs_layout():
    from lagrangian_microbes_amd.device import EnsembleRuns, RUN_PARAMS
    r = EnsembleRuns([3, 0, 2], [{"pRS": 0.5, "pPR": 0.6, "pSP": 0.7}] * 3, [(100, 0), (101, 1), (102, 2)],
                     device="cpu")
    assert r.n == 5 and list(r.start) == [0, 3, 3, 5]
    assert r.run.tolist() == [0, 0, 0, 2, 2]
    # lm_run_params: two uint64 keys and three doubles, 40 bytes per run
    assert RUN_PARAMS.itemsize == 40 and r.params.numel() == 3 * 40
    t = r.params.numpy().view(RUN_PARAMS)
    assert t["key0"].tolist() == [100, 101, 102] and t["pSP"].tolist() == [0.7] * 3
    assert r.slice(2) == slice(3, 5)


def test_autotune_picks_fastest_candidate():
    from lagrangian_microbes_amd import autotune
    cands = autotune.candidates()
    assert len(cands) == 27 and all(c["cell_scale"] >= 1.0 for c in cands)
    seen = []

    def fake_timer(make_runner, c, steps):
        seen.append(c)
        # fastest at cell scale 1.5, chunk 12, replay (1024, 1024)
        return abs(c["cell_scale"] - 1.5) + abs(c["chunk"] - 12) / 10 + (c["play_blocks"] != (1024, 1024))

    best, table = autotune.tune(None, steps=4, cands=cands, timer=fake_timer)
    assert best == {"cell_scale": 1.5, "chunk": 12, "play_blocks": (1024, 1024)}
    assert len(table) == 27 and seen == cands


def test_chunk_schedule_tapers_the_end():
    from lagrangian_microbes_amd.simulator import chunk_schedule
    assert chunk_schedule(720, 100) == [100] * 6 + [60, 30, 15, 8, 7]
    assert chunk_schedule(24, 5) == [5, 5, 5, 5, 4]  # chunks at or below the taper floor are kept
    assert chunk_schedule(0, 100) == []
    for total in range(0, 400, 7):
        for chunk in (1, 8, 24, 100):
            s = chunk_schedule(total, chunk)
            assert sum(s) == total and all(1 <= k <= chunk for k in s)
