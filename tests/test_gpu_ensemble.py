"""Ensembles on the GPU: independent runs of uneven size in one index range
(lm_query_pairs_runs, lm_rps_prepare_runs, FusedRunner(runs=...)).

Every run of the batched ensemble must evolve exactly as that run alone
through the plain FP32 path: the same float32 records, the same pair counts
and the same species at every step (each run keeps its own p and Philox key,
and the uniforms are ranked within the run).  Pairs never cross runs, also
when different runs occupy the same cells.
"""
import numpy as np
import pytest
import torch

import oracle
from lagrangian_microbes_amd import configs, device as dev
from lagrangian_microbes_amd.simulator import FusedRunner
from oracle.pairs import sorted_kdtree_pairs

pytestmark = pytest.mark.gpu


def run_alone(w, field, chunk, steps=None):
    lon, lat = w.particles()
    r = FusedRunner(lon, lat, w.species(), field, w.radius, w.params, w.key, w.dt, chunk=chunk, precision="f32")
    recs = []

    def keep(k, rx, ry, rs, cnt):
        recs.append((rx.cpu().numpy().copy(), rs.cpu().numpy().copy(), cnt.cpu().numpy().copy()))

    r.run(w.steps if steps is None else steps, keep)
    r.check_domain()
    rx = np.concatenate([a for a, _, _ in recs])
    rs = np.concatenate([b for _, b, _ in recs])
    cnt = np.concatenate([c for _, _, c in recs])
    return rx, rs, cnt


@pytest.mark.parametrize("cell_scale,chunk", [(1.0, 10), (2.0, 24)])
def test_ensemble_matches_runs_alone(gpu, cell_scale, chunk):
    e = configs.ensemble_small()
    field = e.velocity_field()
    lon, lat = e.particles()
    runs = dev.EnsembleRuns([w.n for w in e.runs], [w.params for w in e.runs], [w.key for w in e.runs])
    r = FusedRunner(lon, lat, e.species(), field, e.runs[0].radius, {"pRS": 0, "pPR": 0, "pSP": 0}, (0, 0),
                    e.runs[0].dt, chunk=chunk, precision="f32", runs=runs, cell_scale=cell_scale)
    recs = []

    def keep(k, rx, ry, rs, cnt):
        recs.append((rx.cpu().numpy().copy(), rs.cpu().numpy().copy(), cnt.cpu().numpy().copy()))

    r.run(e.steps, keep)
    r.check_domain()
    ex = np.concatenate([a for a, _, _ in recs])
    es = np.concatenate([b for _, b, _ in recs])
    ecnt = np.concatenate([c for _, _, c in recs])
    total = np.zeros(e.steps, dtype=np.int64)
    for k, w in enumerate(e.runs):
        rx, rs, cnt = run_alone(w, field, chunk)
        sl = runs.slice(k)
        assert np.array_equal(ex[:, sl], rx), k
        assert np.array_equal(es[:, sl], rs), k
        total += cnt
    assert np.array_equal(ecnt, total)
    assert total.sum() > 1000  # the runs do interact


def test_pairs_never_cross_runs(gpu):
    # two runs on the same lattice: every particle has a twin at the same position in the other run
    w = configs.rps_10k_steady()
    lon, lat = w.particles()
    lon2, lat2 = np.concatenate([lon, lon]), np.concatenate([lat, lat])
    x = torch.from_numpy(lon2.astype(np.float32)).cuda()[None]
    y = torch.from_numpy(lat2.astype(np.float32)).cuda()[None]
    runs = dev.EnsembleRuns([w.n, w.n], [w.params, w.params], [w.key, w.key])
    s = dev.PairSearch(2 * w.n, w.velocity_field().bounds, 0.02)
    pi = torch.empty((1, 400000), dtype=torch.int32, device="cuda")
    pj = torch.empty_like(pi)
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    s.query_batch(x, y, 0.02, 2.0, pi, pj, cnt, runs)
    P = int(cnt.item())
    a, b = pi[0, :P].cpu().numpy(), pj[0, :P].cpu().numpy()
    assert P > 0 and np.all((a < w.n) == (b < w.n))
    # each run finds the pairs it finds alone; the second run's are the first's shifted by n
    one, _ = dev.PairSearch(w.n, w.velocity_field().bounds, 0.02).query_sync(x[0, :w.n].contiguous(),
                                                                             y[0, :w.n].contiguous(), 0.02)
    m = a < w.n
    assert np.array_equal(a[m], one.cpu().numpy()) and np.array_equal(a[~m], one.cpu().numpy() + w.n)


def test_graph_replay_matches_eager(gpu):
    # the same ensemble with its per-step replay launches captured once per chunk shape and replayed:
    # 24 steps in chunks of 5 replay the full-chunk graph four times (with host reads in between) and
    # the last, shorter chunk's once; every step's species must equal the eager run's
    e = configs.ensemble_small()
    field = e.velocity_field()
    lon, lat = e.particles()
    out = {}
    for graphs in (False, True):
        runs = dev.EnsembleRuns([w.n for w in e.runs], [w.params for w in e.runs], [w.key for w in e.runs])
        r = FusedRunner(lon, lat, e.species(), field, e.runs[0].radius, {"pRS": 0, "pPR": 0, "pSP": 0}, (0, 0),
                        e.runs[0].dt, chunk=5, precision="f32", runs=runs, graphs=graphs, play_blocks=(64, 32))
        recs = []
        r.run(e.steps, lambda k, rx, ry, rs, cnt: recs.append(rs.cpu().numpy().copy()))
        out[graphs] = np.concatenate(recs)
        if graphs:
            # the three buffer sets' 5-step chunks and the last 4-step chunk
            assert len(r.engine._graphs) == 4
    assert np.array_equal(out[True], out[False])


def test_full_size_ensemble_matches_runs_alone(gpu):
    # The 64-run ensemble at full size (14.7M particles) takes the paths only
    # sizes above 2^23 particles reach: 2^23 buckets with 13 bits counted in
    # LDS per binning block, and the grouping by root with 14 fine bits per
    # block.  Its smallest, median and largest runs, each alone (2^19
    # particles or fewer: the small-size paths), must match it step by step
    # over the first two days (the lattices start wider apart than the
    # radius; pairs form as the flow converges).
    e = configs.ensemble_64runs()
    field = e.velocity_field()
    lon, lat = e.particles()
    steps, chunk = 48, 24
    print(f"[full-size ensemble] {e.n} particles, {len(e.runs)} runs, {steps} steps", flush=True)
    assert e.n > 1 << 23  # bucket table capped at 2^23; 24 root bits
    runs = dev.EnsembleRuns([w.n for w in e.runs], [w.params for w in e.runs], [w.key for w in e.runs])
    r = FusedRunner(lon, lat, e.species(), field, e.runs[0].radius, {"pRS": 0, "pPR": 0, "pSP": 0}, (0, 0),
                    e.runs[0].dt, chunk=chunk, precision="f32", runs=runs)
    sizes = np.array([w.n for w in e.runs])
    order = np.argsort(sizes, kind="stable")
    picks = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    sel = {q: ([], [], []) for q in picks}

    def keep(k, rx, ry, rs, cnt):
        for q in picks:
            sl = runs.slice(q)
            sel[q][0].append(rx[:, sl].cpu().numpy())
            sel[q][1].append(rs[:, sl].cpu().numpy())
            sel[q][2].append(ry[:, sl].cpu().numpy())

    r.run(steps, keep)
    r.check_domain()
    del r
    torch.cuda.empty_cache()
    pairs = changed = 0
    for q in picks:
        w = e.runs[q]
        rx, rs, cnt = run_alone(w, field, chunk, steps)
        assert np.array_equal(np.concatenate(sel[q][0]), rx), q
        assert np.array_equal(np.concatenate(sel[q][1]), rs), q
        pairs += int(cnt.sum())
        changed += int(np.sum(rs[-1] != w.species()))
    print(f"[full-size ensemble] runs {picks}: {pairs} pairs, {changed} species changed", flush=True)
    assert pairs > 0 and changed > 0  # the runs do interact
    # and the batched run's last step of its smallest and largest runs against the oracle, on the
    # run's own records: pairs of the float32 positions against cKDTree, the species against the
    # sequential update of the step before
    s, checked = steps - 1, 0
    for q in (picks[0], picks[-1]):
        w = e.runs[q]
        X, S, Y = (np.concatenate(a) for a in sel[q])
        pi, pj = sorted_kdtree_pairs(X[s], Y[s], w.radius, w.norm)
        want = oracle.rps_sequential(S[s - 1], pi, pj, w.p, w.p, w.p, w.key, s)
        assert np.array_equal(S[s], want), q
        checked += len(pi)
    assert checked > 0
