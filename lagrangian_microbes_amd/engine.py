"""The per-step device pipeline shared by the simulators and the benchmark.

A chunk of steps runs without host synchronisation: the pair count of every
step lands in its own device slot, and the host looks at them once at the end
of the chunk.  If a step found more pairs than the pair buffers hold, the
chunk's interactions are replayed from its starting species with larger
buffers; the pair search, the pair ranks and the keyed uniforms are
deterministic, so the replay gives the same result as an unbounded buffer.

Advection does not depend on species, so a chunk is advected first (one
launch, recording the float32 positions of every step) and the pair search
and rock-paper-scissors update then run step by step over those records,
exactly as the reference's two passes do through particle_data.nc
(particle_advecter.py:220-235, interaction_simulator.py:83-110).
"""
import math

import numpy as np
import torch

from . import device as dev


class InteractionEngine:
    """Pair search + rock-paper-scissors for n particles on the device."""

    def __init__(self, n, bbox, radius, norm=2.0, device=None, initial_capacity=None):
        self.n = int(n)
        self.radius = float(radius)
        self.norm = float(norm)
        self.device = torch.device(device or "cuda")
        self.search = dev.PairSearch(self.n, bbox, self.radius, self.device)
        self.rps = dev.RPSUpdater(self.n, self.device)
        self.capacity = 0
        self._grow(initial_capacity or max(2 * self.n, 1024))

    def _grow(self, cap):
        cap = int(cap)
        if cap > self.capacity:
            self.capacity = cap
            self.pair_i = torch.empty(cap, dtype=torch.int32, device=self.device)
            self.pair_j = torch.empty(cap, dtype=torch.int32, device=self.device)

    def pairs_step(self, x32, y32, npairs_slot):
        self.search.query(x32, y32, self.radius, self.norm, self.pair_i, self.pair_j, npairs_slot)

    def rps_step(self, species, npairs_slot, params, key, step):
        self.rps.update(species, self.pair_i, self.pair_j, npairs_slot, params["pRS"], params["pPR"],
                        params["pSP"], key, step)

    def run_chunk(self, rec_x, rec_y, species, params, key, step0, rec_species=None, npairs=None):
        """Steps s = 0..len(rec_x)-1 over recorded float32 positions rec_x[s], rec_y[s].

        species (int8, device) is updated in place; rec_species[s] receives the
        species after step s.  Returns the per-step pair counts (device int64).
        """
        steps = rec_x.shape[0]
        if npairs is None:
            npairs = torch.zeros(steps, dtype=torch.int64, device=self.device)
        start = species.clone()
        while True:
            for s in range(steps):
                slot = npairs[s:s + 1]
                self.pairs_step(rec_x[s], rec_y[s], slot)
                self.rps_step(species, slot, params, key, step0 + s)
                if rec_species is not None:
                    rec_species[s].copy_(species)
            worst = int(npairs.max().item()) if steps else 0
            if worst <= self.capacity:
                return npairs
            self._grow(worst + worst // 4 + 1024)
            species.copy_(start)


def bbox_with_margin(lon_min, lon_max, lat_min, lat_max, margin):
    return (lon_min - margin, lon_max + margin, lat_min - margin, lat_max + margin)


def new_key(seed=None, stream=0):
    """Philox key (seed, stream); seed drawn from numpy's global generator if None."""
    if seed is None:
        seed = int(np.random.randint(0, 2 ** 31 - 1))
    return (int(seed), int(stream))


def walk_amplitude(Kh_deg2_per_s, dt_s):
    return math.sqrt(6.0 * abs(dt_s) * Kh_deg2_per_s) if Kh_deg2_per_s > 0 else 0.0
