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
    over batches of up to steps_max steps.

    ``nsets`` independent sets of pair buffers and grouping workspaces let a
    caller prepare one batch (pair search and grouping, which need positions
    only) while another batch replays (which needs the species): set b is
    addressed by the ``*_set`` methods; ``run_chunk`` uses set 0.
    """

    def __init__(self, n, bbox, radius, norm=2.0, device=None, steps_max=1, initial_capacity=None, nsets=1,
                 runs=None, cell_scale=1.0, play_blocks=None, graphs=False):
        self.n = int(n)
        self.radius = float(radius)
        self.norm = float(norm)
        self.steps_max = int(steps_max)
        self.device = torch.device(device or "cuda")
        self.runs = runs  # EnsembleRuns: independent runs, each with its own key and p
        self.search = dev.PairSearch(self.n, bbox, self.radius, self.device, self.steps_max, cell_scale)
        self.nsets = int(nsets)
        self.rps_sets = [dev.RPSUpdater(self.n, self.device, self.steps_max) for _ in range(self.nsets)]
        self.npairs_sets = [torch.zeros(self.steps_max, dtype=torch.int64, device=self.device)
                            for _ in range(self.nsets)]
        self.pair_sets = [None] * self.nsets
        self.capacity = 0
        if play_blocks is not None:
            for r in self.rps_sets:
                r.set_play_blocks(*play_blocks)
        # the per-step replay launches of a batch, captured once per (set, steps, buffers)
        # into a HIP graph and replayed (graphs=True)
        self.graphs = bool(graphs)
        self._graphs = {}
        self._grow(initial_capacity or max(2 * self.n, 1024))

    # set 0 under the single-batch names
    @property
    def rps(self):
        return self.rps_sets[0]

    @property
    def npairs(self):
        return self.npairs_sets[0]

    @property
    def pair_i(self):
        return self.pair_sets[0][0]

    @property
    def pair_j(self):
        return self.pair_sets[0][1]

    def _grow(self, cap):
        cap = int(cap)
        if cap > self.capacity:
            self.capacity = cap
            self._graphs = {}  # captured launches address the old buffers
            self.pair_sets = [(torch.empty((self.steps_max, cap), dtype=torch.int32, device=self.device),
                               torch.empty((self.steps_max, cap), dtype=torch.int32, device=self.device))
                              for _ in range(self.nsets)]
            # every set's per-pair RPS arrays too: growing them inside a pipelined
            # run would stop the stream that prepares the next chunk
            for r in self.rps_sets:
                r.reserve(cap)

    def prepare_set(self, b, rec_x, rec_y, params, key, step0):
        """Pair search and grouping of S steps into set b (asynchronous)."""
        S = rec_x.shape[0]
        if S > self.steps_max:
            raise ValueError(f"{S} steps exceed the engine's batch of {self.steps_max}")
        pi, pj = (t[:S] for t in self.pair_sets[b])
        cnt = self.npairs_sets[b][:S]
        self.search.query_batch(rec_x, rec_y, self.radius, self.norm, pi, pj, cnt, self.runs)
        self.rps_sets[b].prepare(self.n, pi, pj, cnt, params["pRS"], params["pPR"], params["pSP"], key, step0,
                                 self.runs)

    def worst_count(self, b, S):
        """Largest pair count of set b's S steps (synchronises the stream)."""
        return int(self.npairs_sets[b][:S].max().item()) if S > 0 else 0

    def fits(self, worst):
        """True if the pair buffers held every step; otherwise they are grown
        (the caller prepares the affected sets again)."""
        if worst <= self.capacity:
            return True
        self._grow(worst + worst // 4 + 1024)
        return False

    def play_set(self, b, S, species, rec_species):
        """Replay set b's S steps: rec_species[s] = species after step s, and
        species ends as the state after the last step (asynchronous)."""
        if S <= 0:
            return
        if self.graphs and not dev.KernelTimer.active:
            # The launches' arguments are the workspace's arrays, the species
            # buffers and the pair capacity, fixed for a key; only their contents
            # change between replays.  The capture stream holds no memset node.
            key = (b, S, species.data_ptr(), rec_species.data_ptr(), self.capacity)
            g = self._graphs.get(key)
            if g is None:
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    self._play_steps(b, S, species, rec_species)
                self._graphs[key] = g
            g.replay()
        else:
            self._play_steps(b, S, species, rec_species)
        species.copy_(rec_species[S - 1])

    def _play_steps(self, b, S, species, rec_species):
        sp_in = species
        for s in range(S):
            self.rps_sets[b].play(s, sp_in, rec_species[s])
            sp_in = rec_species[s]

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
            self.prepare_set(0, rec_x, rec_y, params, key, step0)
            cnt = self.npairs[:S]
            sp_in = species
            for s in range(S):
                out = rec_species[s] if rec_species is not None else species
                self.rps.play(s, sp_in, out)
                sp_in = out
            if self.fits(self.worst_count(0, S)):
                if rec_species is not None:
                    species.copy_(rec_species[S - 1])
                if npairs is not None:
                    npairs.copy_(cnt)
                return cnt
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
