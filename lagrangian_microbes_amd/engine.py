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
    """Pair search + rock-paper-scissors for n particles on the device, run
    over batches of up to steps_max steps."""

    def __init__(self, n, bbox, radius, norm=2.0, device=None, steps_max=1, initial_capacity=None):
        self.n = int(n)
        self.radius = float(radius)
        self.norm = float(norm)
        self.steps_max = int(steps_max)
        self.device = torch.device(device or "cuda")
        self.search = dev.PairSearch(self.n, bbox, self.radius, self.device, self.steps_max)
        self.rps = dev.RPSUpdater(self.n, self.device, self.steps_max)
        self.npairs = torch.zeros(self.steps_max, dtype=torch.int64, device=self.device)
        self.capacity = 0
        self._grow(initial_capacity or max(2 * self.n, 1024))

    def _grow(self, cap):
        cap = int(cap)
        if cap > self.capacity:
            self.capacity = cap
            self.pair_i = torch.empty((self.steps_max, cap), dtype=torch.int32, device=self.device)
            self.pair_j = torch.empty((self.steps_max, cap), dtype=torch.int32, device=self.device)

    def run_chunk(self, rec_x, rec_y, species, params, key, step0, rec_species=None, npairs=None):
        """Steps s = 0..S-1 over recorded float32 positions rec_x[s], rec_y[s] ([S, n]).

        The pair search and the grouping of every step run once for the whole
        batch; the replay runs step by step.  species (int8, device) ends as
        the state after the last step; rec_species[s] (if given) receives the
        state after step s.  One host synchronisation per call, to check the
        pair counts against the buffers (the batch is redone with larger
        buffers if a step overflowed).  Returns the per-step pair counts.
        """
        S = rec_x.shape[0]
        if S == 0:
            return self.npairs[:0]
        if S > self.steps_max:
            raise ValueError(f"{S} steps exceed the engine's batch of {self.steps_max}")
        start = species.clone() if rec_species is None else None
        while True:
            pi, pj, cnt = self.pair_i[:S], self.pair_j[:S], self.npairs[:S]
            self.search.query_batch(rec_x, rec_y, self.radius, self.norm, pi, pj, cnt)
            self.rps.prepare(self.n, pi, pj, cnt, params["pRS"], params["pPR"], params["pSP"], key, step0)
            sp_in = species
            for s in range(S):
                out = rec_species[s] if rec_species is not None else species
                self.rps.play(s, sp_in, out)
                sp_in = out
            worst = int(cnt.max().item())
            if worst <= self.capacity:
                if rec_species is not None:
                    species.copy_(rec_species[S - 1])
                if npairs is not None:
                    npairs.copy_(cnt)
                return cnt
            self._grow(worst + worst // 4 + 1024)
            if start is not None:
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
