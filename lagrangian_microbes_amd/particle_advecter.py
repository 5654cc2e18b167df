"""Particle advection: the reference's ParticleAdvecter API on the GPU.

Same constructor, ``time_step`` and ``create_netcdf_file`` as
particle_advecter.py:69-307, same chunk pickles
(particle_locations_SSSSS_EEEEE_tileTT.pickle holding {"time", "lat", "lon"}
with float32 [iters, particles_per_tile], zlib level 3) and the same
particle_data.nc variables.  Inside, each output chunk is one launch of the
HIP RK4 kernel (lm_advect_rk4_f64) that advances all particles through every
step of the chunk and records the post-step float32 positions.

Differences, all deliberate:
- ``velocity_field`` may be a VelocityField; "OSCAR" reads the local
  oscar_vel<year>.nc in the working directory (NetCDF-3 only) and, where the
  reference would download it (velocity_fields.py:21-32), raises.
- Particle state is float64 on the device (parcels keeps float32,
  particle_advecter.py:208-209); records stay float32 as in the reference.
- The Kh random walk draws keyed Philox uniforms (particle, step) instead of
  numpy's global generator (particle_advecter.py:240-242); ``seed`` fixes the
  key, otherwise it is drawn from numpy's global generator.
- A span that is not a whole number of dt stops after the last whole step
  (the reference loop would not advance).
Kept as in the reference: each ``time_step`` call starts the particle clock at
the field's first time level (ParticleSet.from_list without time=,
particle_advecter.py:186-187), and a later call restarts from the float32
positions in the last N_procs chunk pickles (particle_advecter.py:124-138).
"""
import logging
import math
import os
import pickle
from glob import glob
from time import time

import joblib
import numpy as np
import torch

from . import device as dev
from .netcdf_io import write_particle_dataset
from .utils import most_symmetric_integer_factorization, pretty_filesize, pretty_time
from .velocity_fields import VelocityField, oscar_dataset

logger = logging.getLogger(__name__)


def uniform_particle_locations(N_particles, lat_min, lat_max, lon_min, lon_max):
    """Regular lattice of N_particles (particle_advecter.py:26-35).

    (N_lat, N_lon) is the most symmetric factorisation of N; particle
    k = i_lon * N_lat + i_lat (latitude varies fastest); both axes include
    their end points.
    """
    n_lat, n_lon = most_symmetric_integer_factorization(N_particles)
    logger.info("Generating ({:d}, {:d}) particles along each (lat, lon).".format(n_lat, n_lon))
    lon_axis = np.linspace(lon_min, lon_max, n_lon)
    lat_axis = np.linspace(lat_min, lat_max, n_lat)
    lon_grid, lat_grid = np.meshgrid(lon_axis, lat_axis, indexing="ij")  # [n_lon, n_lat]
    return lon_grid.ravel(), lat_grid.ravel()


def distribute_particles_across_tiles(particle_lons, particle_lats, tiles):
    """Split into `tiles` equal contiguous slices (particle_advecter.py:38-66)."""
    assert particle_lons.size == particle_lats.size
    assert (particle_lons.size / tiles).is_integer()
    return list(np.split(particle_lons, tiles)), list(np.split(particle_lats, tiles))


def _pickle_name(start_iter, end_iter, tile_id):
    return "particle_locations_{:05d}_{:05d}_tile{:02d}.pickle".format(start_iter, end_iter, tile_id)


def _parse_pickle_name(path):
    parts = os.path.basename(path).split("_")
    return int(parts[2]), int(parts[3]), int(parts[4][4:6])


class ParticleAdvecter:
    def __init__(self, particle_lons, particle_lats, N_procs=-1, velocity_field="OSCAR", output_dir=".",
                 output_chunk_iters=100, Kh=0, seed=None, device=None):
        if isinstance(velocity_field, str):
            assert velocity_field == "OSCAR", "OSCAR is the only supported velocity field name."
        else:
            assert isinstance(velocity_field, VelocityField), "velocity_field must be 'OSCAR' or a VelocityField"

        assert 1 <= N_procs or N_procs == -1, \
            "Number of processors N_procs must be a positive integer or -1 (use all processors)."
        max_procs = joblib.cpu_count()
        N_procs = N_procs if 1 <= N_procs <= max_procs else max_procs
        logger.info("Requested {:d} tiles. Found {:d} processors. Using {:d}.".format(N_procs, max_procs, N_procs))

        assert particle_lons.size == particle_lats.size
        N_particles = particle_lons.size
        assert (N_particles / N_procs).is_integer()

        output_dir = os.path.abspath(output_dir)
        if not os.path.exists(output_dir):
            logger.info("Creating directory: {:s}".format(output_dir))
            os.makedirs(output_dir)

        particle_lons, particle_lats = distribute_particles_across_tiles(particle_lons, particle_lats, N_procs)
        assert output_chunk_iters >= 1

        self.iteration = 0
        self.particle_lons = particle_lons
        self.particle_lats = particle_lats
        self.velocity_field = velocity_field
        self.N_particles = N_particles
        self.N_procs = N_procs
        self.particles_per_tile = N_particles // N_procs
        self.output_dir = output_dir
        self.output_chunk_iters = output_chunk_iters
        self.Kh = Kh / 1e10  # m^2/s -> deg^2/s with 1 deg = 100 km, as particle_advecter.py:121
        if seed is None:
            seed = int(np.random.randint(0, 2 ** 31 - 1))
        self.walk_key = (int(seed), 0x6B68)  # key word 1 separates the walk from other streams
        self.device = torch.device(device or "cuda")
        self._dev_field = None
        self._oscar = None  # (year, VelocityField) read from the local OSCAR file

    # -- helpers -----------------------------------------------------------
    def _field(self, start_time):
        field = self.velocity_field
        if isinstance(field, str):  # "OSCAR": the local oscar_vel<year>.nc, read once per year
            if self._oscar is None or self._oscar[0] != start_time.year:
                self._oscar = (start_time.year, oscar_dataset(start_time.year))
            field = self._oscar[1]
        if self._dev_field is None or self._dev_field.field is not field:
            self._dev_field = dev.DeviceField(field, self.device)
        return self._dev_field

    def _restore_from_pickles(self):
        logger.info("Restoring particle locations from disk...")
        pkl_files = sorted(glob(os.path.join(self.output_dir, "particle_locations_*.pickle")))
        for path in pkl_files[-self.N_procs:]:
            _, _, tile_id = _parse_pickle_name(path)
            data = joblib.load(path)
            self.particle_lons[tile_id] = data["lon"][-1, :]
            self.particle_lats[tile_id] = data["lat"][-1, :]

    # -- API ---------------------------------------------------------------
    def time_step(self, start_time, end_time, dt):
        if self.iteration != 0:
            self._restore_from_pickles()
        logger.info("Starting time stepping: {:} -> {:} (dt={:}) on the GPU, {:d} output tile(s)."
                    .format(start_time, end_time, dt, self.N_procs))

        dfield = self._field(start_time)
        dt_s = dt.total_seconds()
        n = self.N_particles
        lon = torch.from_numpy(np.concatenate(self.particle_lons).astype(np.float64)).to(self.device)
        lat = torch.from_numpy(np.concatenate(self.particle_lats).astype(np.float64)).to(self.device)
        oob = torch.zeros(1, dtype=torch.int32, device=self.device)
        walk_amp = math.sqrt(6.0 * abs(dt_s) * self.Kh) if self.Kh > 0 else 0.0

        t = start_time
        t_rel = 0.0  # particle clock starts at the field's first level on every call
        iteration = self.iteration
        ppt = self.particles_per_tile
        while t < end_time:
            iters_to_do = min(self.output_chunk_iters, (end_time - t) // dt)
            if iters_to_do <= 0:
                break
            tic = time()
            stage, stride = dfield.stage_fields(t_rel, dt_s, iters_to_do)
            rec_lon = torch.empty((iters_to_do, n), dtype=torch.float32, device=self.device)
            rec_lat = torch.empty((iters_to_do, n), dtype=torch.float32, device=self.device)
            dev.advect_rk4_f64(lon, lat, stage, stride, dfield.grid, iters_to_do, dt_s, rec_lon, rec_lat, n,
                               walk_amp, self.walk_key, iteration, oob)
            rl = rec_lon.cpu().numpy()
            rt = rec_lat.cpu().numpy()
            n_oob = int(oob.item())
            advection_time = time() - tic
            if n_oob:
                raise RuntimeError(f"{n_oob} particle-step(s) left the velocity field's domain "
                                   f"{dfield.field.bounds} (parcels raises OutOfBounds here)")

            times = [t + (k + 1) * dt for k in range(iters_to_do)]
            tic = time()
            size = 0
            for tile_id in range(self.N_procs):
                path = os.path.join(self.output_dir, _pickle_name(iteration, iteration + iters_to_do, tile_id))
                out = {"time": times,
                       "lat": np.ascontiguousarray(rt[:, tile_id * ppt:(tile_id + 1) * ppt]),
                       "lon": np.ascontiguousarray(rl[:, tile_id * ppt:(tile_id + 1) * ppt])}
                with open(path, "wb") as f:
                    joblib.dump(out, f, compress=("zlib", 3), protocol=pickle.HIGHEST_PROTOCOL)
                size += os.path.getsize(path)
            pickling_time = time() - tic
            logger.info("Advecting particles (iteration {:05d} -> {:05d}): {:s}; pickling {:s} ({:s}, {:s} per "
                        "particle per iteration)".format(iteration, iteration + iters_to_do,
                                                         pretty_time(advection_time), pretty_time(pickling_time),
                                                         pretty_filesize(size),
                                                         pretty_filesize(size / (iters_to_do * n))))
            t = t + iters_to_do * dt
            t_rel += iters_to_do * dt_s
            iteration += iters_to_do

        self.final_lons = lon.cpu().numpy()
        self.final_lats = lat.cpu().numpy()
        self.iteration += (end_time - start_time) // dt

    def create_netcdf_file(self, start_time, end_time, dt):
        iters = (end_time - start_time) // dt
        times = [start_time + k * dt for k in range(iters)]
        plons = np.zeros((self.N_particles, iters), dtype=np.float32)
        plats = np.zeros((self.N_particles, iters), dtype=np.float32)

        pkl_files = sorted(glob(os.path.join(self.output_dir, "particle_locations_*.pickle")))
        for path in pkl_files:
            logger.info("Collecting particle locations from {:s}...".format(path))
            t1, t2, tile_id = _parse_pickle_name(path)
            data = joblib.load(path)
            i1, i2 = tile_id * self.particles_per_tile, (tile_id + 1) * self.particles_per_tile
            plons[i1:i2, t1:t2] = data["lon"].T
            plats[i1:i2, t1:t2] = data["lat"].T

        nc_path = os.path.join(self.output_dir, "particle_data.nc")
        logger.info("Writing particle locations to {:s}...".format(nc_path))
        write_particle_dataset(nc_path, {"longitude": plons, "latitude": plats}, times, self.N_particles)
        for path in pkl_files:
            os.remove(path)
