"""Launch-shape autotuner for fused runs (ensembles above all).

The pair search's cell side (a multiple of the radius: larger cells mean
fewer table entries but more candidates per particle), the chunk length
(steps prepared per batch: launch count against memory and pipeline depth)
and the replay's launch shape (workgroups on wide and on small components)
change the speed of a run but never its result.  ``tune`` builds a runner for
every candidate, runs a warm-up chunk, times a few steps from the initial
state, and returns the fastest candidate with the whole table.  Runs that
keep the same workload can reuse the choice.
"""
import itertools
import time

import torch

CELL_SCALES = (1.0, 1.5, 2.0)
CHUNKS = (8, 12, 24)
PLAY_BLOCKS = ((2048, 2048), (1024, 1024), (4096, 2048))


def candidates(cell_scales=CELL_SCALES, chunks=CHUNKS, play_blocks=PLAY_BLOCKS):
    return [{"cell_scale": c, "chunk": k, "play_blocks": tuple(pb)}
            for c, k, pb in itertools.product(cell_scales, chunks, play_blocks)]


def time_runner(make_runner, cand, steps, reps=2):
    """Seconds per step of a runner built with ``cand`` (best of reps)."""
    r = make_runner(**cand)
    r.run(cand["chunk"])  # warm-up: sizes buffers, captures graphs
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        r.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r.run(steps)
        torch.cuda.synchronize()
        best = min(best, (time.perf_counter() - t0) / steps)
    del r
    torch.cuda.empty_cache()
    return best


def tune(make_runner, steps=24, cands=None, timer=time_runner):
    """make_runner(cell_scale=, chunk=, play_blocks=) -> a FusedRunner.

    Returns (best candidate, [(candidate, seconds per step), ...])."""
    cands = candidates() if cands is None else list(cands)
    if not cands:
        raise ValueError("no candidates")
    table = [(c, timer(make_runner, c, steps)) for c in cands]
    best = min(table, key=lambda t: t[1])[0]
    return best, table
