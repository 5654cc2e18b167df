"""The HIP path against the committed golden fixtures (tests/golden/)."""
import hashlib
import os

import numpy as np
import pytest
import torch

from lagrangian_microbes_amd import configs, device as dev
from lagrangian_microbes_amd.simulator import FusedRunner

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def test_philox_uniforms_fixture(gpu):
    g = load("philox_kat")
    for c, k, uni in zip(g["counter"], g["key"], g["uniforms"]):
        got = dev.philox_uniforms(len(uni), [int(v) for v in c], (int(k[0]), int(k[1]))).cpu().numpy()
        assert np.array_equal(got, uni)


@pytest.mark.parametrize("name,p", [("p1", 1.0), ("p2", 2.0), ("pinf", np.inf)])
def test_pairs_ties_fixture(gpu, name, p):
    g = load("pairs_ties")
    x, y, r = g["x"], g["y"], float(g["r"])
    ps = dev.PairSearch(len(x), (float(x.min()), float(x.max()), float(y.min()), float(y.max())), r)
    pi, pj = ps.query_sync(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(), r, p)
    assert np.array_equal(pi.cpu().numpy(), g[name + "_i"]) and np.array_equal(pj.cpu().numpy(), g[name + "_j"])


def test_c1_fused_runner_fixture(gpu):
    g = load("c1_24h")
    w = configs.rps_10k_steady()
    lon, lat = w.particles()
    steps = len(g["npairs"])
    r = FusedRunner(lon, lat, w.species(), w.velocity_field(), w.radius, w.params, w.key, w.dt, chunk=10)
    done = 0
    while done < steps:
        k = r.advance(steps - done)
        for s in range(k):
            S = done + s
            assert digest(r.rec_x[s].cpu().numpy(), r.rec_y[s].cpu().numpy()) == g["digest_records"][S], S
            P = int(r.npairs[s].item())
            assert P == g["npairs"][S], S
            pi = r.engine.pair_i[s, :P].cpu().numpy()
            pj = r.engine.pair_j[s, :P].cpu().numpy()
            assert digest(pi, pj) == g["digest_pairs"][S], S
            assert digest(r.rec_s[s].cpu().numpy()) == g["digest_species"][S], S
        done += k
    r.check_domain()
    assert np.array_equal(r.lon.cpu().numpy(), g["lon"]) and np.array_equal(r.lat.cpu().numpy(), g["lat"])
    assert np.array_equal(r.species.cpu().numpy(), g["species"])


@pytest.mark.parametrize("chunk,grow", [(7, False), (10, True), (24, False)])
def test_c1_pipelined_run_fixture(gpu, chunk, grow):
    """Chunks overlapped on three streams give the same records and end
    state, also when the pair buffers must grow in the middle of the run."""
    g = load("c1_24h")
    w = configs.rps_10k_steady()
    lon, lat = w.particles()
    steps = len(g["npairs"])
    r = FusedRunner(lon, lat, w.species(), w.velocity_field(), w.radius, w.params, w.key, w.dt, chunk=chunk)
    if grow:
        r.engine.capacity = 0
        r.engine._grow(64)  # far too small: most steps after the first hours overflow
    seen = []

    def check(k, rx, ry, rs, cnt):
        base = len(seen)
        for s in range(k):
            S = base + s
            assert digest(rx[s].cpu().numpy(), ry[s].cpu().numpy()) == g["digest_records"][S], S
            assert int(cnt[s].item()) == g["npairs"][S], S
            assert digest(rs[s].cpu().numpy()) == g["digest_species"][S], S
            seen.append(S)

    assert r.run(steps, on_chunk=check) == steps
    assert seen == list(range(steps)) and r.step == steps
    r.check_domain()
    assert np.array_equal(r.lon.cpu().numpy(), g["lon"]) and np.array_equal(r.lat.cpu().numpy(), g["lat"])
    assert np.array_equal(r.species.cpu().numpy(), g["species"])
    if grow:
        assert r.engine.capacity >= int(g["npairs"].max())
