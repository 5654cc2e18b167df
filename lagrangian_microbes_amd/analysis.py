"""Reductions of a run's outputs on the device: the species count time
series that the reference plots (analysis.py:20-57 counts species == ROCK,
PAPER, SCISSORS at every time of microbe_data.nc) and the histogram of all
pairwise great-circle distances (sandbox/pairwise_distance_histogram_*.jl:
bin = round(10 log10(max(1, d / 1 m))), d the haversine distance on a sphere
of radius 6371.228 km).  Plotting stays out: these return the numbers a
figure is drawn from.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr, require, stream_ptr

EARTH_RADIUS_M = 6371.228e3  # the radius of the reference's distance histogram


def species_counts(rec_species):
    """rec_species: int8 device tensor [steps, n] (or [n]); returns int64
    [steps, 3] counts of rock, paper and scissors per step (codes outside
    1..3 are not counted)."""
    r = rec_species if rec_species.dim() == 2 else rec_species[None]
    require(r, torch.int8, name="rec_species")
    S, n = r.shape
    out = torch.zeros((S, 4), dtype=torch.int64, device=r.device)
    for a in range(0, S, 65535):
        b = min(S, a + 65535)
        check(_lib.load().lm_species_counts(ctypes.c_void_p(r.data_ptr() + a * n), n, b - a, n,
                                            ctypes.c_void_p(out.data_ptr() + 32 * a), stream_ptr()))
    return out[:, 1:]


def distance_bin_edges_m(bins):
    """Distance where bin b (b = 1 .. bins-1) of round(10 log10(max(1, d))) starts."""
    return 10.0 ** ((np.arange(1, bins) - 0.5) / 10.0)


def haversine_thresholds(bins, radius=EARTH_RADIUS_M):
    """The haversine terms a = sin^2(d / (2 R)) of the bin edges."""
    return np.sin(distance_bin_edges_m(bins) / (2.0 * radius)) ** 2


def distance_histogram(lon, lat, bins=70, radius=EARTH_RADIUS_M):
    """Histogram (numpy uint64 [bins]) of the great-circle distances of all
    unordered pairs of float32 device positions: bin b holds the pairs with
    round(10 log10(max(1, d))) == b, the last bin also everything farther."""
    n = lon.numel()
    require(lon, torch.float32, name="lon")
    require(lat, torch.float32, numel=n, name="lat")
    edges = torch.from_numpy(haversine_thresholds(bins, radius)).to(lon.device)
    hist = torch.zeros(bins, dtype=torch.int64, device=lon.device)
    check(_lib.load().lm_distance_histogram(n, ptr(lon), ptr(lat), int(bins), ptr(edges), ptr(hist), stream_ptr()))
    return hist.cpu().numpy().astype(np.uint64)
