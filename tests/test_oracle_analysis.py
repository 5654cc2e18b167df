"""Pins the oracle's pairwise distance histogram (the restatement of the
device reduction) with an independent numpy haversine and round(10 log10)
binning, the form of sandbox/pairwise_distance_histogram_distributed.jl:49-55."""
import numpy as np
import pytest

import oracle
from lagrangian_microbes_amd.analysis import EARTH_RADIUS_M, haversine_thresholds


def numpy_histogram(lon, lat, bins, radius):
    lon = lon.astype(np.float64)
    lat = lat.astype(np.float64)
    i, j = np.triu_indices(lon.size, 1)
    p1, p2 = np.radians(lat[i]), np.radians(lat[j])
    a = np.sin((p2 - p1) / 2) ** 2 + np.cos(p1) * np.cos(p2) * np.sin(np.radians(lon[j] - lon[i]) / 2) ** 2
    d = 2 * radius * np.arcsin(np.minimum(1.0, np.sqrt(a)))
    b = np.minimum(np.rint(10 * np.log10(np.maximum(1.0, d))), bins - 1).astype(np.int64)
    return np.bincount(b, minlength=bins), d


def points(n, seed):
    rng = np.random.default_rng(seed)
    lon = (200 + 20 * rng.random(n)).astype(np.float32)
    lat = (20 + 20 * rng.random(n)).astype(np.float32)
    # a cluster at metre scale, exact duplicates (d = 0) and a far point (clamped into the last bin)
    lon[:40] = np.float32(210.0) + np.float32(1e-5) * np.arange(40, dtype=np.float32)
    lat[:40] = np.float32(30.0)
    lon[40:45], lat[40:45] = lon[0], lat[0]
    lon[45], lat[45] = np.float32(30.0), np.float32(-60.0)
    return lon, lat


def test_sine_polynomial_against_numpy():
    x = np.linspace(-90, 90, 20001)
    ours = np.array([oracle.sin_deg(v) for v in x])
    assert np.max(np.abs(ours - np.sin(np.radians(x)))) < 1e-13


@pytest.mark.parametrize("bins", [70, 8])
def test_histogram_against_numpy(bins):
    lon, lat = points(1500, bins)
    got = oracle.distance_histogram(lon, lat, haversine_thresholds(bins, EARTH_RADIUS_M)).astype(np.int64)
    ref, d = numpy_histogram(lon, lat, bins, EARTH_RADIUS_M)
    assert got.sum() == lon.size * (lon.size - 1) // 2
    # only pairs within rounding of a bin edge may fall on the other side
    edges = np.concatenate([[0.0], 10.0 ** ((np.arange(1, bins) - 0.5) / 10.0)])
    near = np.sum(np.min(np.abs(d[:, None] / edges[None, 1:] - 1.0), axis=1) < 1e-9)
    assert np.abs(got - ref).sum() <= 2 * near
    assert got[0] >= 5 * 6 // 2  # the duplicates
    assert got[-1] > 0  # the far point


def test_planted_radius_mistake_is_caught():
    lon, lat = points(400, 1)
    got = oracle.distance_histogram(lon, lat, haversine_thresholds(70, EARTH_RADIUS_M)).astype(np.int64)
    wrong, _ = numpy_histogram(lon, lat, 70, 2 * EARTH_RADIUS_M)
    assert np.abs(got - wrong).sum() > 1000
