"""End-to-end GPU runs against the oracle's two-pass run (bit-exact pairs,
species and float32 records; float64 positions within 1e-9 degrees)."""
import os
from datetime import datetime, timedelta

import joblib
import numpy as np
import pytest
import torch

import oracle
from oracle.pairs import sorted_kdtree_pairs
from helpers import oracle_two_pass
from lagrangian_microbes_amd import (InteractionSimulator, MicrobeSimulator, ParticleAdvecter, configs,
                                     rock_paper_scissors)
from lagrangian_microbes_amd.netcdf_io import read_particle_dataset
from lagrangian_microbes_amd.simulator import FusedRunner

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1_oracle():
    return oracle_two_pass(configs.rps_10k_steady())


def test_fused_runner_c1(gpu, c1_oracle):
    w = configs.rps_10k_steady()
    lon, lat = w.particles()
    r = FusedRunner(lon, lat, w.species(), w.velocity_field(), w.radius, w.params, w.key, w.dt, chunk=10)
    done = 0
    while done < w.steps:
        k = r.advance(w.steps - done)
        for s in range(k):
            S = done + s
            assert np.array_equal(r.rec_x[s].cpu().numpy(), c1_oracle["rec_lon"][S])
            assert int(r.npairs[s].item()) == len(c1_oracle["pairs"][S][0]), S
            assert np.array_equal(r.rec_s[s].cpu().numpy(), c1_oracle["species"][S]), S
        done += k
    r.check_domain()
    assert np.max(np.abs(r.lon.cpu().numpy() - c1_oracle["lon"])) <= 1e-9
    assert np.max(np.abs(r.lat.cpu().numpy() - c1_oracle["lat"])) <= 1e-9


def test_pair_buffer_growth_replays_chunk(gpu, c1_oracle):
    w = configs.rps_10k_steady()
    lon, lat = w.particles()
    r = FusedRunner(lon, lat, w.species(), w.velocity_field(), w.radius, w.params, w.key, w.dt, chunk=24)
    r.engine.capacity = 0
    r.engine._grow(64)  # far too small: every step with > 64 pairs overflows
    r.advance(24)
    assert r.engine.capacity >= max(len(p[0]) for p in c1_oracle["pairs"])
    assert np.array_equal(r.rec_s[23].cpu().numpy(), c1_oracle["species"][23])


def test_dropin_two_pass_c1(gpu, c1_oracle, tmp_path):
    w = configs.rps_10k_steady()
    lon, lat = w.particles()
    start = datetime(2017, 1, 1)
    dt = timedelta(hours=1)
    end = start + w.steps * dt
    adv_dir = str(tmp_path / "adv")
    pa = ParticleAdvecter(lon, lat, N_procs=4, velocity_field=w.velocity_field(), output_dir=adv_dir,
                          output_chunk_iters=10)
    pa.time_step(start, end, dt)
    pk = sorted(os.listdir(adv_dir))
    assert pk[0] == "particle_locations_00000_00010_tile00.pickle" and len(pk) == 12
    chunk = joblib.load(os.path.join(adv_dir, pk[0]))
    assert chunk["lat"].shape == (10, 2500) and chunk["lat"].dtype == np.float32
    assert chunk["time"][0] == start + dt
    pa.create_netcdf_file(start, end, dt)
    assert not any(f.endswith(".pickle") for f in os.listdir(adv_dir))
    data, times, N = read_particle_dataset(os.path.join(adv_dir, "particle_data.nc"))
    assert N == 10000 and times[0] == start and len(times) == 24
    assert np.array_equal(data["longitude"], c1_oracle["rec_lon"].T)
    assert np.array_equal(data["latitude"], c1_oracle["rec_lat"].T)

    rps = rock_paper_scissors(w.n, w.p, w.p, w.p, seed=w.species_seed)
    isim = InteractionSimulator(rps, w.radius, advection_dir=adv_dir, output_dir=str(tmp_path / "int"),
                                philox_key=w.key)
    isim.time_step(start, end, dt)
    out, _, _ = read_particle_dataset(os.path.join(str(tmp_path / "int"), "microbe_data.nc"))
    assert np.array_equal(out["species"], c1_oracle["species"].T)
    assert np.array_equal(rps[2]["species"], c1_oracle["species"][-1])
    assert np.array_equal(out["longitude"], data["longitude"])
    assert list(isim.pairs_per_step) == [len(p[0]) for p in c1_oracle["pairs"]]


def test_oscar_file_dropin_c1(gpu, c1_oracle, tmp_path, monkeypatch):
    # velocity_field="OSCAR" reads oscar_vel<year>.nc from the working
    # directory: the C1 field written in OSCAR's layout (latitude decreasing,
    # a depth axis, CF time) advects exactly as the field object does
    from scipy.io import netcdf_file
    w = configs.rps_10k_steady()
    f = w.velocity_field()
    with netcdf_file(str(tmp_path / "oscar_vel2017.nc"), "w", version=2) as nc:
        for name, size in (("time", 1), ("depth", 1), ("latitude", f.ny), ("longitude", f.nx)):
            nc.createDimension(name, size)
        t = nc.createVariable("time", "i4", ("time",))
        t[:] = [8858]
        t.units = "day since 1992-10-05 00:00:00"
        nc.createVariable("depth", "f4", ("depth",))[:] = [15.0]
        nc.createVariable("latitude", "f8", ("latitude",))[:] = f.lats[::-1]
        nc.createVariable("longitude", "f8", ("longitude",))[:] = f.lons
        for name, data in (("u", f.u), ("v", f.v)):
            nc.createVariable(name, "f8", ("time", "depth", "latitude", "longitude"))[:] = \
                data[:, None, ::-1, :]
    monkeypatch.chdir(tmp_path)
    lon, lat = w.particles()
    start, dt = datetime(2017, 1, 1), timedelta(hours=1)
    pa = ParticleAdvecter(lon, lat, N_procs=1, velocity_field="OSCAR", output_dir=str(tmp_path / "adv"),
                          output_chunk_iters=24)
    pa.time_step(start, start + w.steps * dt, dt)
    pa.create_netcdf_file(start, start + w.steps * dt, dt)
    data, _, _ = read_particle_dataset(str(tmp_path / "adv" / "particle_data.nc"))
    assert np.array_equal(data["longitude"], c1_oracle["rec_lon"].T)
    assert np.array_equal(data["latitude"], c1_oracle["rec_lat"].T)


def test_generic_callback_path(gpu, c1_oracle, tmp_path):
    # a non-RPS callback sees every pair once, in sorted (i, j) order
    w = configs.rps_10k_steady()
    lon, lat = w.particles()
    start, dt = datetime(2017, 1, 1), timedelta(hours=1)
    end = start + 12 * dt
    pa = ParticleAdvecter(lon, lat, N_procs=1, velocity_field=w.velocity_field(), output_dir=str(tmp_path))
    pa.time_step(start, end, dt)
    pa.create_netcdf_file(start, end, dt)
    seen = []

    def count_pairs(params, props, p1, p2):
        seen.append((p1, p2))
        props["species"][p1] = props["species"][p1]

    isim = InteractionSimulator((count_pairs, {}, {"species": w.species()}), w.radius, advection_dir=str(tmp_path),
                                output_dir=str(tmp_path))
    isim.time_step(start, end, dt)
    expect = [p for s in range(12) for p in zip(*[a.tolist() for a in c1_oracle["pairs"][s]])]
    assert seen == expect


@pytest.mark.parametrize("mode,kw", [("device", {}), ("stream", {"device_record_bytes": 0}),
                                     ("stream", {"device_record_bytes": 0, "ring_bytes": 1})])
def test_microbe_simulator_outputs(gpu, c1_oracle, tmp_path, mode, kw):
    # device: the records stay in HBM during the run and are written afterwards (in blocks of 3,000
    # particles here, so the transposed blocks alternate between the two pinned buffers); stream:
    # they go to the files chunk by chunk, and ring_bytes=1 leaves two pinned slots, so the run
    # waits for the writer thread before most chunks
    from lagrangian_microbes_amd import records
    w = configs.rps_10k_steady()
    lon, lat = w.particles()
    rps = rock_paper_scissors(w.n, w.p, w.p, w.p, seed=w.species_seed)
    sim = MicrobeSimulator(lon, lat, w.velocity_field(), rps, w.radius, output_dir=str(tmp_path),
                           output_chunk_iters=7, philox_key=w.key, **kw)
    start, dt = datetime(2017, 1, 1), timedelta(hours=1)
    block = records.BLOCK_BYTES
    records.BLOCK_BYTES = 3000 * 24 * 4
    try:
        sim.time_step(start, start + 24 * dt, dt)
    finally:
        records.BLOCK_BYTES = block
    assert sim.timings["records"] == mode
    if "ring_bytes" in kw:
        assert sim.timings["ring_bytes"] == 2 * 7 * (w.n * 9 + 8)
    out, times, N = read_particle_dataset(os.path.join(str(tmp_path), "microbe_data.nc"))
    assert N == w.n and times == [start + k * dt for k in range(24)]
    assert np.array_equal(out["species"], c1_oracle["species"].T)
    assert np.array_equal(out["latitude"], c1_oracle["rec_lat"].T)
    assert np.array_equal(out["longitude"], c1_oracle["rec_lon"].T)
    part, _, _ = read_particle_dataset(os.path.join(str(tmp_path), "particle_data.nc"))
    assert set(part) == {"longitude", "latitude"}
    assert np.array_equal(part["longitude"], out["longitude"]) and np.array_equal(part["latitude"], out["latitude"])
    assert list(sim.pairs_per_step) == [len(p[0]) for p in c1_oracle["pairs"]]
    assert np.array_equal(rps[2]["species"], c1_oracle["species"][-1])


def test_c2_full_size_properties(gpu):
    """490k particles: advect to mid-run, then pairs vs cKDTree and species vs
    the sequential oracle at full size, per step on identical inputs."""
    w = configs.rps_490k_30day()
    lon, lat = w.particles()
    r = FusedRunner(lon, lat, w.species(), w.velocity_field(), w.radius, w.params, w.key, w.dt, chunk=100)
    for _ in range(3):
        r.advance(100)
    r.check_domain()
    sp_before = r.species.clone()
    x, y = r.rec_x[99].clone(), r.rec_y[99].clone()
    # one more interaction step on the recorded positions of step 299, done by hand
    ki, kj = sorted_kdtree_pairs(x.cpu().numpy(), y.cpu().numpy(), w.radius)
    pi, pj = r.engine.search.query_sync(x, y, w.radius)
    assert np.array_equal(pi.cpu().numpy(), ki) and np.array_equal(pj.cpu().numpy(), kj)
    assert len(ki) > 100000
    npairs = torch.tensor([len(ki)], dtype=torch.int64, device="cuda")
    sp = sp_before.clone()
    r.engine.rps.update(sp, pi, pj, npairs, w.p, w.p, w.p, w.key, 300)
    o = oracle.rps_sequential(sp_before.cpu().numpy(), ki, kj, w.p, w.p, w.p, w.key, 300)
    assert np.array_equal(sp.cpu().numpy(), o)
    counts = np.bincount(sp.cpu().numpy(), minlength=4)
    assert counts[0] == 0 and counts.sum() == w.n
