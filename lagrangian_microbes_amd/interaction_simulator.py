"""Pair interactions: the reference's InteractionSimulator API on the GPU.

Same constructor and ``time_step`` as interaction_simulator.py:22-122: read
particle_data.nc from ``advection_dir``, for each step find all pairs within
``interaction_radius`` (degree space, Minkowski ``interaction_norm``) of the
stored float32 positions, apply the pair interaction, record lon/lat/species
per step and write microbe_data.nc to ``output_dir``.  The final species are
left in ``microbe_properties["species"]`` as the reference's in-place
callback leaves them.

Routing: for ``rock_paper_scissors_interaction`` the whole step runs on the
GPU (pair search kernel, then the rock-paper-scissors kernel).  Any other
callback gets the GPU pair list, copied back and visited on the host in sorted
(i, j) order with the reference's call signature fn(params, props, i, j).

Differences, all deliberate: pairs are visited in sorted (i, j) order instead
of Python set order (interaction_simulator.py:104), and the
rock-paper-scissors uniforms are Philox4x64-10 keyed by (step, pair rank)
instead of numpy's global generator (interactions.py:20).  ``philox_key``
fixes the key (seed, stream); by default it is drawn from numpy's global
generator.
"""
import logging
import os
from time import time

import numpy as np
import torch

from . import device as dev
from .engine import InteractionEngine, new_key
from .interactions import rock_paper_scissors_interaction
from .netcdf_io import read_particle_dataset, write_particle_dataset
from .utils import pretty_time

logger = logging.getLogger(__name__)


class InteractionSimulator:
    def __init__(self, pair_interaction, interaction_radius, interaction_norm=2, self_interaction=None,
                 advection_dir=".", output_dir=".", philox_key=None, device=None):
        output_dir = os.path.abspath(output_dir)
        logger.info("Microbe interactions output directory: {:s}".format(output_dir))
        if not os.path.exists(output_dir):
            logger.info("Creating directory: {:s}".format(output_dir))
            os.makedirs(output_dir)

        pair_interaction_function, pair_interaction_parameters, microbe_properties = pair_interaction
        self.microbe_properties = microbe_properties
        self.pair_interaction = pair_interaction_function
        self.pair_interaction_parameters = pair_interaction_parameters
        self.interaction_radius = interaction_radius
        self.interaction_norm = interaction_norm
        self.self_interaction = self_interaction  # stored, unused (as interaction_simulator.py:28,49)
        self.advection_dir = advection_dir
        self.output_dir = output_dir
        self.iteration = 0
        self.philox_key = tuple(philox_key) if philox_key is not None else new_key()
        self.device = torch.device(device or "cuda")
        self.batch_steps = 100  # steps whose pair search and grouping share one set of launches

    def time_step(self, start_time, end_time, dt):
        path = os.path.join(self.advection_dir, "particle_data.nc")
        data, _, N = read_particle_dataset(path, ["longitude", "latitude"])
        plon, plat = data["longitude"], data["latitude"]
        Nt = plon.shape[1]
        times = [start_time + k * dt for k in range(Nt)]
        mspecies = np.zeros((N, Nt), dtype=np.int8)

        steps = []
        t = start_time
        while t < end_time:
            steps.append(self.iteration + len(steps))
            t = t + dt
        if steps and steps[-1] >= Nt:
            raise IndexError(f"step {steps[-1]} is beyond the {Nt} recorded steps of {path}")
        logger.info("Simulating interactions for {:d} microbes over {:d} time steps ({:} -> {:})."
                    .format(N, len(steps), start_time, end_time))

        tic = time()
        if steps:
            cols = np.asarray(steps)
            rec_x = torch.from_numpy(np.ascontiguousarray(plon[:, cols].T)).to(self.device)
            rec_y = torch.from_numpy(np.ascontiguousarray(plat[:, cols].T)).to(self.device)
            finite = np.isfinite(plon[:, cols]) & np.isfinite(plat[:, cols])
            bbox = (float(plon[:, cols][finite].min(initial=0.0)), float(plon[:, cols][finite].max(initial=0.0)),
                    float(plat[:, cols][finite].min(initial=0.0)), float(plat[:, cols][finite].max(initial=0.0)))
            engine = InteractionEngine(N, bbox, self.interaction_radius, self.interaction_norm, self.device,
                                       steps_max=min(self.batch_steps, len(steps)))
            if self.pair_interaction is rock_paper_scissors_interaction:
                self._run_rps(engine, rec_x, rec_y, steps, mspecies)
            else:
                self._run_generic(engine, rec_x, rec_y, steps, mspecies)
        logger.info("Simulating interactions: {:s}.".format(pretty_time(time() - tic)))
        self.iteration += len(steps)

        nc_path = os.path.join(self.output_dir, "microbe_data.nc")
        logger.info("Writing microbe data to {:s}...".format(nc_path))
        mlon = np.zeros((N, Nt), dtype=np.float32)
        mlat = np.zeros((N, Nt), dtype=np.float32)
        if steps:
            mlon[:, steps] = plon[:, steps]
            mlat[:, steps] = plat[:, steps]
        write_particle_dataset(nc_path, {"longitude": mlon, "latitude": mlat, "species": mspecies}, times, N)

    def _run_rps(self, engine, rec_x, rec_y, steps, mspecies):
        props = self.microbe_properties
        species = torch.from_numpy(np.ascontiguousarray(props["species"], dtype=np.int8)).to(self.device)
        rec_s = torch.empty((len(steps), species.numel()), dtype=torch.int8, device=self.device)
        counts = []
        B = engine.steps_max
        for a in range(0, len(steps), B):
            b = min(a + B, len(steps))
            cnt = engine.run_chunk(rec_x[a:b], rec_y[a:b], species, self.pair_interaction_parameters,
                                   self.philox_key, steps[a], rec_species=rec_s[a:b])
            counts.append(cnt.cpu().numpy().copy())
        mspecies[:, steps] = rec_s.cpu().numpy().T
        props["species"][:] = species.cpu().numpy()
        self.pairs_per_step = np.concatenate(counts)

    def _run_generic(self, engine, rec_x, rec_y, steps, mspecies):
        props = self.microbe_properties
        fn, params = self.pair_interaction, self.pair_interaction_parameters
        counts = []
        for s, col in enumerate(steps):
            pi, pj = engine.search.query_sync(rec_x[s], rec_y[s], engine.radius, engine.norm)
            a = pi.cpu().numpy().tolist()
            b = pj.cpu().numpy().tolist()
            counts.append(len(a))
            for p1, p2 in zip(a, b):
                fn(params, props, p1, p2)
            mspecies[:, col] = props["species"]
        self.pairs_per_step = np.asarray(counts, dtype=np.int64)
