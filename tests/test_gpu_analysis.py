"""Device reductions (lm_species_counts, lm_distance_histogram) against numpy
and the oracle: exact integer counts."""
import numpy as np
import pytest
import torch

import oracle
from lagrangian_microbes_amd import analysis, configs
from lagrangian_microbes_amd.simulator import FusedRunner
from test_oracle_analysis import points

pytestmark = pytest.mark.gpu


def test_species_counts(gpu):
    rng = np.random.default_rng(3)
    sp = rng.integers(1, 4, size=(7, 100001)).astype(np.int8)
    sp[2, :10] = 0  # codes outside 1..3 are not counted
    sp[5, 5] = 9
    got = analysis.species_counts(torch.from_numpy(sp).cuda()).cpu().numpy()
    ref = np.stack([[np.sum(row == c) for c in (1, 2, 3)] for row in sp])
    assert np.array_equal(got, ref)
    assert analysis.species_counts(torch.from_numpy(sp[0]).cuda()).shape == (1, 3)


@pytest.mark.parametrize("n,bins", [(3000, 70), (257, 8), (1, 70), (0, 70)])
def test_distance_histogram_bit_exact(gpu, n, bins):
    lon, lat = points(max(n, 46), n)
    lon, lat = lon[:n], lat[:n]
    got = analysis.distance_histogram(torch.from_numpy(lon).cuda(), torch.from_numpy(lat).cuda(), bins)
    ref = oracle.distance_histogram(lon, lat, analysis.haversine_thresholds(bins))
    assert np.array_equal(got, ref)
    assert int(got.sum()) == n * (n - 1) // 2


def test_full_size_reductions(gpu):
    # the 490k run after 24 steps: every pair counted once, and a subsample bit-exact with the oracle
    w = configs.rps_490k_30day()
    lon, lat = w.particles()
    r = FusedRunner(lon, lat, w.species(), w.velocity_field(), w.radius, w.params, w.key, w.dt, chunk=24)
    r.advance(24)
    counts = analysis.species_counts(r.rec_s[:24]).cpu().numpy()
    assert np.all(counts.sum(axis=1) == w.n)
    x, y = r.rec_x[23].contiguous(), r.rec_y[23].contiguous()
    h = analysis.distance_histogram(x, y, 70)
    assert int(h.sum()) == w.n * (w.n - 1) // 2
    sub = slice(0, w.n, 97)
    hs = analysis.distance_histogram(x[sub].contiguous(), y[sub].contiguous(), 70)
    assert np.array_equal(hs, oracle.distance_histogram(x[sub].cpu().numpy(), y[sub].cpu().numpy(),
                                                        analysis.haversine_thresholds(70)))
