"""GPU parity of each liblm entry point against the oracle (bit-exact)."""
import numpy as np
import pytest
import torch

import oracle
from oracle.pairs import sorted_kdtree_pairs
from helpers import oracle_advect
from lagrangian_microbes_amd import configs, device as dev
from lagrangian_microbes_amd.velocity_fields import VelocityField

pytestmark = pytest.mark.gpu
R = configs.KM_DEG


def T(a, device="cuda"):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


# ---------------------------------------------------------------- philox
def test_philox_uniforms_bit_exact(gpu):
    for (c, key, word) in [((0, 7, 0, 0), (1, 0), 0), ((123456789, 3, 1, 0), (99, 5), 1),
                           ((2 ** 63, 1, 2, 3), (2 ** 62, 2 ** 61), 3)]:
        g = dev.philox_uniforms(5000, c, key, word).cpu().numpy()
        o = oracle.philox_uniforms(5000, *c, key[0], key[1], word)
        assert np.array_equal(g, o)


# ---------------------------------------------------------------- advection
def gpu_advect(f, lon, lat, dt, nsteps, walk_amp=0.0, key=(0, 0), step0=0):
    df = dev.DeviceField(f)
    stage, stride = df.stage_fields(0.0, dt, nsteps)
    x, y = T(lon.astype(np.float64)), T(lat.astype(np.float64))
    n = x.numel()
    rl = torch.empty((nsteps, n), dtype=torch.float32, device="cuda")
    rt = torch.empty((nsteps, n), dtype=torch.float32, device="cuda")
    oob = dev.advect_rk4_f64(x, y, stage, stride, df.grid, nsteps, dt, rl, rt, n, walk_amp, key, step0)
    return x.cpu().numpy(), y.cpu().numpy(), rl.cpu().numpy(), rt.cpu().numpy(), int(oob.item())


@pytest.mark.parametrize("name,steps,stride", [("rps_10k_steady_24h", 24, 1), ("rps_490k_30day", 720, 97)])
def test_advect_bit_exact(gpu, name, steps, stride):
    w = configs.WORKLOADS[name]()
    f = w.velocity_field()
    lon, lat = w.particles()
    lon, lat = lon[::stride], lat[::stride]
    g = gpu_advect(f, lon, lat, w.dt, steps)
    o = oracle_advect(f, lon, lat, w.dt, steps)
    assert g[4] == 0 and o[4] == 0
    err = max(np.max(np.abs(g[0] - o[0])), np.max(np.abs(g[1] - o[1])))
    assert err <= 1e-9, err  # stated tolerance; the operation order is shared, so expect 0
    assert err == 0.0, f"not bit-exact: {err}"
    assert np.array_equal(g[2], o[2]) and np.array_equal(g[3], o[3])


def test_advect_random_walk_and_step_offset(gpu):
    w = configs.rps_10k_steady()
    f = w.velocity_field()
    lon, lat = w.particles()
    g = gpu_advect(f, lon, lat, w.dt, 5, walk_amp=1e-3, key=(4, 0x6B68), step0=17)
    o = oracle_advect(f, lon, lat, w.dt, 5, walk_amp=1e-3, key=(4, 0x6B68), step0=17)
    assert np.array_equal(g[0], o[0]) and np.array_equal(g[1], o[1]) and np.array_equal(g[2], o[2])


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_advect_steps_matches_advect(gpu, precision):
    """advect_steps (the stage schedule sliced from the span's, on the device)
    against advect (schedule computed and uploaded per call), bit for bit, at
    a clock offset; a step whose 2 dt is not whole takes the advect path; a
    window past the span raises."""
    w = configs.rps_490k_30day()
    f = w.velocity_field()
    lon, lat = w.particles()
    lon, lat = lon[::977], lat[::977]
    df = dev.DeviceField(f, precision=precision)
    n = lon.size
    for dt, clock, k in [(w.dt, 0, 10), (w.dt, 331, 17), (w.dt, 719 - 2, 2), (1234.25, 5, 6)]:
        out = []
        for fn, arg in [(df.advect, clock * dt), (df.advect_steps, clock)]:
            x, y = T(lon.astype(np.float64)), T(lat.astype(np.float64))
            rl = torch.empty((k, n), dtype=torch.float32, device="cuda")
            rt = torch.empty((k, n), dtype=torch.float32, device="cuda")
            oob = fn(x, y, arg, dt, k, rl, rt, n, 0.0, (0, 0), 3)
            out.append([a.cpu().numpy() for a in (x, y, rl, rt)] + [int(oob.item())])
        for a, b in zip(out[0], out[1]):
            assert np.array_equal(a, b), (dt, clock, k)
    span_steps = int((f.times[-1] - f.times[0]) // w.dt)
    x, y = T(lon.astype(np.float64)), T(lat.astype(np.float64))
    with pytest.raises(ValueError):
        df.advect_steps(x, y, span_steps - 1, w.dt, 2)


def test_advect_out_of_grid(gpu):
    lons = np.arange(200.0, 220.01, 0.5)
    lats = np.arange(20.0, 40.01, 0.5)
    f = VelocityField(lons, lats, np.zeros((lats.size, lons.size)), np.ones((lats.size, lons.size)))
    g = gpu_advect(f, np.array([210.0, 210.0]), np.array([30.0, 39.99]), 3600.0, 2)
    assert g[4] == 2 and g[1][1] == 39.99


def test_field_blend_bit_exact(gpu):
    rng = np.random.default_rng(0)
    levels = rng.normal(size=(5, 33, 47, 2))
    lev0 = np.array([0, 1, 3, 3, 2], np.int32)
    wts = rng.uniform(size=5)
    out = torch.empty(5 * 33 * 47 * 2, dtype=torch.float64, device="cuda")
    dev.field_blend_f64(T(levels), 33 * 47 * 2, T(lev0), T(wts), out)
    assert np.array_equal(out.cpu().numpy().reshape(5, 33, 47, 2), oracle.field_blend_f64(levels, lev0, wts))


# ---------------------------------------------------------------- pairs
def gpu_pairs(x, y, r, p=2.0, bbox=None, max_radius=None):
    x32, y32 = np.asarray(x, np.float32), np.asarray(y, np.float32)
    if bbox is None:
        bbox = (float(x32.min()), float(x32.max()), float(y32.min()), float(y32.max())) if x32.size else (0, 1, 0, 1)
    ps = dev.PairSearch(max(x32.size, 1), bbox, max_radius or max(r, 1e-6))
    pi, pj = ps.query_sync(T(x32), T(y32), r, p)
    return pi.cpu().numpy(), pj.cpu().numpy()


@pytest.mark.parametrize("p", [2.0, 1.0, np.inf, 3.0])
@pytest.mark.parametrize("seed,n,spread", [(0, 5000, 0.3), (1, 20000, 0.5), (2, 3000, 0.02)])
def test_pairs_match_bruteforce(gpu, seed, n, spread, p):
    rng = np.random.default_rng(seed)
    x = (210.0 + rng.uniform(0, spread, n)).astype(np.float32)
    y = (30.0 + rng.uniform(0, spread, n)).astype(np.float32)
    gi, gj = gpu_pairs(x, y, R, p)
    bi, bj = oracle.pairs_bruteforce(x, y, R, p)
    assert len(bi) > 0
    assert np.array_equal(gi, bi) and np.array_equal(gj, bj)
    if p != 3.0:
        ki, kj = sorted_kdtree_pairs(x, y, R, p)
        assert np.array_equal(gi, ki) and np.array_equal(gj, kj)


def band_pairs(lo, rng, ncen=8):
    """ncen^2 isolated pairs on the float32 grid of the binade [lo, 2 lo) (lo a power of two, or its
    negative): centre and partner are exact float32 values u apart, so a pair's squared distance is
    exactly (a^2 + b^2) u^2.  The radius is sqrt(N) u with N near (|lo| / 40 / u)^2 and the offsets
    (a, b) put (a^2 + b^2) / N - 1 across -5e-6 .. 5e-6: inside the band around r^2 where the p = 2
    test takes float64, and on both sides just outside it."""
    u = float(np.spacing(np.float32(abs(lo))))
    r = abs(lo) / 40
    N = round((r / u) ** 2)
    span = abs(lo)
    cen = abs(lo) + span * (0.05 + 0.9 * (np.arange(ncen) + 0.5) / ncen)  # spacing > 3 r: pairs stay apart
    assert span * 0.9 / ncen > 3 * r
    cx, cy = [c.ravel() for c in np.meshgrid(cen, cen)]
    sgn = 1.0 if lo > 0 else -1.0
    cx = (sgn * cx / u).round() * u
    cy = (sgn * cy / u).round() * u
    m = cx.size
    target = rng.uniform(-5e-6, 5e-6, m)
    a = np.floor(rng.uniform(0.2, 0.7, m) * np.sqrt(N))
    b = np.round(np.sqrt(N * (1 + target) - a * a))
    rel = (a * a + b * b) / N - 1
    x = np.concatenate([cx, cx + a * u]).astype(np.float32)
    y = np.concatenate([cy, cy + b * u]).astype(np.float32)
    assert np.array_equal(x[m:].astype(np.float64) - x[:m], a * u)  # exact float32 grid points
    return x, y, np.sqrt(N) * u, rel


@pytest.mark.parametrize("lo", [0.25, -0.125, 2.0 ** -10])
def test_pairs_near_the_radius(gpu, lo):
    rng = np.random.default_rng(7)
    x, y, r, rel = band_pairs(lo, rng)
    assert np.sum(np.abs(rel) < 1e-6) >= 5 and np.sum(rel > 1e-6) >= 5 and np.sum(rel < -1e-6) >= 5
    gi, gj = gpu_pairs(x, y, r)
    bi, bj = oracle.pairs_bruteforce(x, y, r)
    ki, kj = sorted_kdtree_pairs(x, y, r)
    assert np.array_equal(gi, bi) and np.array_equal(gj, bj)
    assert np.array_equal(gi, ki) and np.array_equal(gj, kj)
    m = rel.size
    assert np.array_equal(np.sort(bi), np.flatnonzero(rel <= 0)) and np.all(bj == bi + m)


def test_pairs_dense_cluster_selection_path(gpu):
    # 400 points inside one radius: every particle has far more than 8 neighbours
    rng = np.random.default_rng(3)
    x = (210.0 + rng.uniform(0, R / 3, 400)).astype(np.float32)
    y = (30.0 + rng.uniform(0, R / 3, 400)).astype(np.float32)
    gi, gj = gpu_pairs(x, y, R)
    assert len(gi) == 400 * 399 // 2
    bi, bj = oracle.pairs_bruteforce(x, y, R)
    assert np.array_equal(gi, bi) and np.array_equal(gj, bj)


def test_pairs_ties_duplicates_and_outside_box(gpu):
    r = 2.0 ** -7
    g = np.arange(40, dtype=np.float64) * r
    X, Y = np.meshgrid(210.0 + g, 30.0 + g, indexing="ij")
    x, y = X.ravel().astype(np.float32), Y.ravel().astype(np.float32)
    x = np.concatenate([x, x[:50], np.float32([150.0, 150.0, 260.0])])  # duplicates and far outliers
    y = np.concatenate([y, y[:50], np.float32([10.0, 10.0, 60.0])])
    gi, gj = gpu_pairs(x, y, r, bbox=(210.0, 210.3, 30.0, 30.3))
    bi, bj = oracle.pairs_bruteforce(x, y, r)
    assert np.array_equal(gi, bi) and np.array_equal(gj, bj)
    assert (len(x) - 3, len(x) - 2) in set(zip(gi.tolist(), gj.tolist()))  # coincident outliers pair


def test_pairs_empty_and_tiny(gpu):
    for n in (1, 2):
        x = np.full(n, 210.0, np.float32)
        gi, gj = gpu_pairs(x, x, R, bbox=(209, 211, 209, 211))
        assert len(gi) == n * (n - 1) // 2
    ps = dev.PairSearch(16, (0, 1, 0, 1), R)
    e = torch.empty(0, dtype=torch.float32, device="cuda")
    pi, pj = ps.query_sync(e, e, R)
    assert pi.numel() == 0


def test_pairs_capacity_overflow_reports_count(gpu):
    rng = np.random.default_rng(4)
    x = T((210.0 + rng.uniform(0, R, 300)).astype(np.float32))
    y = T((30.0 + rng.uniform(0, R, 300)).astype(np.float32))
    ps = dev.PairSearch(300, (210, 210.01, 30, 30.01), R)
    pi = torch.full((10,), -7, dtype=torch.int32, device="cuda")
    pj = torch.full((10,), -7, dtype=torch.int32, device="cuda")
    npairs = torch.zeros(1, dtype=torch.int64, device="cuda")
    ps.query(x, y, R, 2.0, pi, pj, npairs)
    bi, _ = oracle.pairs_bruteforce(x.cpu().numpy(), y.cpu().numpy(), R)
    assert int(npairs.item()) == len(bi) > 10
    assert torch.all(pi == -7)  # nothing written past the capacity check


# ---------------------------------------------------------------- rps
def random_pairs(rng, n, m, clustered=False):
    # clustered pairs (a, a+1..a+3) exist only 3n-6 times: ask for fewer
    limit = 3 * n - 6 if clustered else n * (n - 1) // 2
    assert m <= limit, f"{m} distinct pairs requested, only {limit} possible"
    pairs = set()
    while len(pairs) < m:
        if clustered:
            a = int(rng.integers(0, n))
            b = int(min(n - 1, a + rng.integers(1, 4)))
        else:
            a, b = (int(v) for v in rng.integers(0, n, 2))
        if a != b:
            pairs.add((min(a, b), max(a, b)))
    arr = np.array(sorted(pairs), dtype=np.int32)
    return arr[:, 0].copy(), arr[:, 1].copy()


def gpu_rps(species, pi, pj, pRS, pPR, pSP, key, step, updater=None):
    n = species.size
    up = updater or dev.RPSUpdater(n)
    sp = T(species.astype(np.int8))
    npairs = torch.tensor([len(pi)], dtype=torch.int64, device="cuda")
    gi = T(pi.astype(np.int32)) if len(pi) else torch.zeros(1, dtype=torch.int32, device="cuda")
    gj = T(pj.astype(np.int32)) if len(pj) else torch.zeros(1, dtype=torch.int32, device="cuda")
    up.update(sp, gi, gj, npairs, pRS, pPR, pSP, key, step)
    return sp.cpu().numpy(), up


@pytest.mark.parametrize("seed,n,m,clustered", [(0, 1000, 800, False), (1, 20000, 30000, False),
                                                (2, 5000, 12000, True), (3, 300, 4000, False)])
def test_rps_matches_sequential_oracle(gpu, seed, n, m, clustered):
    rng = np.random.default_rng(seed)
    pi, pj = random_pairs(rng, n, m, clustered)
    sp = rng.integers(0, 5, n).astype(np.int8)
    args = (0.55, 0.45, 0.65, (seed, 3), 7 + seed)
    g, up = gpu_rps(sp, pi, pj, *args)
    o = oracle.rps_sequential(sp, pi, pj, *args)
    assert np.array_equal(g, o)
    # the union-find's partition against scipy's connected components: species
    # alone would not show two components merged (their pairs replay in order)
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    _, lab = connected_components(coo_matrix((np.ones(m), (pi, pj)), shape=(n, n)), directed=False)
    touched = np.unique(np.concatenate([pi, pj]))
    comp_pairs = np.bincount(lab[pi], minlength=n)
    comp_parts = np.bincount(lab[touched], minlength=n)
    st = up.component_stats()
    assert st.pairs == m and st.touched == touched.size
    assert st.components == np.count_nonzero(comp_pairs)
    assert st.max_pairs == comp_pairs.max() and st.max_particles == comp_parts.max()


@pytest.mark.parametrize("seed", [0, 1])
def test_rps_union_across_tiles(gpu, seed):
    """More particles than one union-find tile (2^15): local pairs (some
    across tile edges), long-range pairs, a chain that zigzags through every
    tile from a small root, and pairs at the end of the partial last tile,
    so most components are joined in a tile's LDS forest and the rest
    through the global one.  Species (0-4, some outside {1, 2, 3}) bit-exact
    against the sequential oracle; the partition against scipy."""
    rng = np.random.default_rng(40 + seed)
    n = 3 * 2 ** 15 + 4321
    pairs = set()
    a = rng.integers(0, n - 41, 30000)
    for x, d in zip(a.tolist(), rng.integers(1, 41, a.size).tolist()):
        pairs.add((x, x + d))
    for x, y in rng.integers(0, n, (15000, 2)).tolist():
        if x != y:
            pairs.add((min(x, y), max(x, y)))
    zig = [5 + 2 * k if k % 2 == 0 else n - 1 - 3 * k for k in range(60)]
    for x, y in zip(zig[:-1], zig[1:]):
        pairs.add((min(x, y), max(x, y)))
    for x in range(n - 200, n - 1, 3):
        pairs.add((x, n - 1))
        pairs.add((x - 2 ** 15, x))
    arr = np.array(sorted(pairs), dtype=np.int32)
    pi, pj = arr[:, 0].copy(), arr[:, 1].copy()
    m = pi.size
    assert np.count_nonzero((pi >> 15) != (pj >> 15)) > 10000  # plenty of pairs cross tiles
    sp = rng.integers(0, 5, n).astype(np.int8)
    args = (0.55, 0.45, 0.65, (seed, 11), 123 + seed)
    g, up = gpu_rps(sp, pi, pj, *args)
    o = oracle.rps_sequential(sp, pi, pj, *args)
    assert np.array_equal(g, o)
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    _, lab = connected_components(coo_matrix((np.ones(m), (pi, pj)), shape=(n, n)), directed=False)
    touched = np.unique(np.concatenate([pi, pj]))
    comp_pairs = np.bincount(lab[pi], minlength=n)
    comp_parts = np.bincount(lab[touched], minlength=n)
    st = up.component_stats()
    assert st.pairs == m and st.touched == touched.size
    assert st.components == np.count_nonzero(comp_pairs)
    assert st.max_pairs == comp_pairs.max() and st.max_particles == comp_parts.max()


@pytest.mark.parametrize("n", [10000, 20000])
def test_rps_giant_chain_component(gpu, n):
    # one component of 20,000 pairs: a chain plus random chords (n = 20000:
    # more particles than the replay keeps in LDS, so it runs in global memory)
    rng = np.random.default_rng(9)
    pairs = {(k, k + 1) for k in range(n - 1)}
    while len(pairs) < 20000:
        a, b = sorted(int(v) for v in rng.integers(0, n, 2))
        if a != b:
            pairs.add((a, b))
    arr = np.array(sorted(pairs), dtype=np.int32)
    sp = rng.integers(1, 4, n).astype(np.int8)
    g, up = gpu_rps(sp, arr[:, 0], arr[:, 1], 0.5, 0.5, 0.5, (1, 1), 3)
    o = oracle.rps_sequential(sp, arr[:, 0], arr[:, 1], 0.5, 0.5, 0.5, (1, 1), 3)
    assert np.array_equal(g, o)
    st = up.component_stats()
    assert st.max_pairs == 20000 and st.wide_components == 1 and st.wide_pairs == 20000 and st.touched == n


@pytest.mark.parametrize("seed", [0, 1])
def test_rps_every_replay_tier(gpu, seed):
    """Components of controlled sizes, indices interleaved across components:
    narrow (<= 8 heads), mid (<= 128 heads, <= 512 pairs), mid spilled to large
    (<= 128 heads, > 512 pairs), large (<= 2048 heads) and the serial fallback."""
    rng = np.random.default_rng(seed)
    sizes = [2, 3, 5, 9] * 50 + [12, 40, 100] * 10 + [60] * 3 + [400, 1500] + [3000]
    n = sum(sizes)
    perm = rng.permutation(n)
    pairs = set()
    at = 0
    for s in sizes:
        ids = perm[at:at + s]
        at += s
        for a in range(1, s):  # spanning tree keeps the component connected
            b = int(rng.integers(0, a))
            pairs.add((int(min(ids[a], ids[b])), int(max(ids[a], ids[b]))))
        extra = 12 * s if s == 60 else s  # dense 60-particle groups exceed 512 pairs
        for _ in range(extra):
            a, b = (int(v) for v in rng.choice(ids, 2, replace=False))
            pairs.add((min(a, b), max(a, b)))
    arr = np.array(sorted(pairs), dtype=np.int32)
    sp = rng.integers(1, 4, n).astype(np.int8)
    args = (0.5, 0.6, 0.4, (seed, 9), 21)
    g, up = gpu_rps(sp, arr[:, 0], arr[:, 1], *args)
    o = oracle.rps_sequential(sp, arr[:, 0], arr[:, 1], *args)
    assert np.array_equal(g, o)
    st = up.component_stats()
    assert st.components == len(sizes) and st.pairs == len(arr) and st.touched == n
    assert st.max_particles == max(sizes)


def test_rps_empty_and_overflow(gpu):
    sp = np.array([1, 2, 3, 1], np.int8)
    g, up = gpu_rps(sp, np.zeros(0, np.int32), np.zeros(0, np.int32), 0.5, 0.5, 0.5, (0, 0), 0)
    assert np.array_equal(g, sp)
    # npairs beyond the capacity: no change
    s = T(sp)
    pi = T(np.array([0], np.int32)); pj = T(np.array([1], np.int32))
    up.update(s, pi, pj, torch.tensor([5], dtype=torch.int64, device="cuda"), 1, 1, 1, (0, 0), 0)
    assert np.array_equal(s.cpu().numpy(), sp)


def test_rps_workspace_reuse(gpu):
    rng = np.random.default_rng(11)
    up = dev.RPSUpdater(5000)
    for step in range(4):
        n = [5000, 1200, 5000, 17][step]
        clustered = step % 2 == 0
        pi, pj = random_pairs(rng, n, 2 * n if clustered else min(3 * n, n * (n - 1) // 2), clustered=clustered)
        sp = rng.integers(1, 4, n).astype(np.int8)
        g, _ = gpu_rps(sp, pi, pj, 0.6, 0.5, 0.4, (2, 2), step, updater=up)
        assert np.array_equal(g, oracle.rps_sequential(sp, pi, pj, 0.6, 0.5, 0.4, (2, 2), step))


# ---------------------------------------------------------------- radix sort
@pytest.mark.parametrize("L,bits,maxkey", [(1, 32, 2 ** 32), (2047, 8, 7), (2048, 16, 2 ** 16), (2049, 24, 300),
                                           (70001, 32, 2 ** 32), (300000, 20, 2 ** 20), (5000, 19, 2 ** 32),
                                           (100000, 10, 2 ** 32), (9000, 9, 2 ** 9)])
def test_radix_sort_stable_segments(gpu, L, bits, maxkey):
    rng = np.random.default_rng(L + bits)
    nseg = 3
    keys = rng.integers(0, maxkey, size=(nseg, L), dtype=np.uint64).astype(np.uint32)
    vals = np.tile(np.arange(L, dtype=np.uint32), (nseg, 1))
    lens = np.array([L, max(L // 3, 0), L - 1 if L > 1 else 0], dtype=np.int64)
    k = T(keys.view(np.int32))
    v = T(vals.view(np.int32))
    dev.sort_pairs_u32(k, v, key_bits=bits, seg_len=T(lens))
    gk = k.cpu().numpy().view(np.uint32)
    gv = v.cpu().numpy().view(np.uint32)
    mask = np.uint32((1 << bits) - 1) if bits < 32 else np.uint32(0xFFFFFFFF)
    for g in range(nseg):
        n = int(lens[g])
        order = np.argsort(keys[g, :n] & mask, kind="stable")
        assert np.array_equal(gk[g, :n], keys[g, :n][order])
        assert np.array_equal(gv[g, :n], vals[g, :n][order])
        # entries past the segment length are left alone
        assert np.array_equal(gk[g, n:], keys[g, n:])
