"""The hand-written RK4 adjoint on the GPU, through lm_adjoint_forward /
lm_adjoint_backward behind lagrangian_microbes_amd.adjoint.Rk4Adjoint.

- forward positions bit-exact with the C oracle on the float32 field widened
  to double (stated tolerance 1e-9 degrees);
- gradients of the great-circle loss against the float64 torch-autograd
  reference (oracle/adjoint_ref.py) at 10,000 particles / 24 steps:
  positions within 1e-9 and u, v within 1e-5 of the largest entry (float32
  accumulation);
- central finite differences (eps = 1e-3 m/s, float32 perturbations) on 32
  random sampled grid nodes: median relative error <= 1e-5 and every node
  within 5e-3 (a difference that straddles a trajectory's cell crossing sees
  the interpolation's kink);
- closed forms of a uniform zonal flow; checkpoint intervals 1..64 give the
  same gradients; the global-memory fallback for blocks whose LDS window
  does not fit; steps leaving the grid; a steady field; ragged and empty
  sizes; finite differences on 32 nodes at the full 1M-particle, 240-step
  size (eps = 1e-3 and 1e-4 m/s); and the forward-only RPS update on that
  run's final positions against cKDTree pairs and the sequential oracle.
"""
import numpy as np
import pytest
import torch

from oracle import adjoint_ref
from helpers import grid_of, oracle_advect
from lagrangian_microbes_amd import adjoint, configs
from lagrangian_microbes_amd.velocity_fields import VelocityField

pytestmark = pytest.mark.gpu
M = 1852.0 * 60.0


def widened(f):
    return VelocityField(f.lons, f.lats, f.u.astype(np.float32).astype(np.float64),
                         f.v.astype(np.float32).astype(np.float64), f.times if f.nt > 1 else None)


def f32_tensors(f, requires_grad=False):
    u = torch.tensor(f.u, dtype=torch.float32, device="cuda", requires_grad=requires_grad)
    v = torch.tensor(f.v, dtype=torch.float32, device="cuda", requires_grad=requires_grad)
    return u, v


def targets_of(w, lon, lat, steps):
    """Target positions: the same particles through the second seeded field (C oracle)."""
    lo, la, _, _, _ = oracle_advect(w.target_velocity_field(), lon, lat, w.dt, steps)
    return lo, la


def gpu_grads(f, lon, lat, tlon, tlat, dt, steps, ckpt=16, loss="great_circle"):
    plan = adjoint.AdjointPlan(f, 0.0, dt, steps, ckpt)
    u, v = f32_tensors(f, requires_grad=True)
    x = torch.tensor(lon, dtype=torch.float64, device="cuda", requires_grad=True)
    y = torch.tensor(lat, dtype=torch.float64, device="cuda", requires_grad=True)
    X, Y = adjoint.advect(x, y, u, v, plan)
    tx = torch.tensor(tlon, device="cuda")
    ty = torch.tensor(tlat, device="cuda")
    L = adjoint.trajectory_loss(X, Y, tx, ty)
    L.backward()
    torch.cuda.synchronize()
    return dict(X=X.detach().cpu().numpy(), Y=Y.detach().cpu().numpy(), L=float(L.detach()),
                gx=x.grad.cpu().numpy(), gy=y.grad.cpu().numpy(), gu=u.grad.cpu().numpy(), gv=v.grad.cpu().numpy(),
                oob=int(plan.oob.item()))


def ref_grads(f, lon, lat, tlon, tlat, dt, steps):
    fw = widened(f)
    lev0, w = fw.stage_schedule(0.0, dt, steps)
    u = torch.tensor(fw.u, requires_grad=True)
    v = torch.tensor(fw.v, requires_grad=True)
    x = torch.tensor(lon, dtype=torch.float64, requires_grad=True)
    y = torch.tensor(lat, dtype=torch.float64, requires_grad=True)
    X, Y, oob = adjoint_ref.rk4(x, y, u, v, grid_of(fw), lev0, w, dt, steps)
    L = adjoint.trajectory_loss(X, Y, torch.tensor(tlon), torch.tensor(tlat))
    L.backward()
    return dict(X=X.detach().numpy(), Y=Y.detach().numpy(), L=float(L.detach()), gx=x.grad.numpy(), gy=y.grad.numpy(),
                gu=u.grad.numpy(), gv=v.grad.numpy(), oob=oob)


def normwise(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def assert_close_to_ref(g, r, pos_tol=1e-9, field_tol=1e-5):
    assert g["oob"] == r["oob"]
    assert normwise(g["gx"], r["gx"]) <= pos_tol, normwise(g["gx"], r["gx"])
    assert normwise(g["gy"], r["gy"]) <= pos_tol, normwise(g["gy"], r["gy"])
    assert normwise(g["gu"], r["gu"]) <= field_tol, normwise(g["gu"], r["gu"])
    assert normwise(g["gv"], r["gv"]) <= field_tol, normwise(g["gv"], r["gv"])


@pytest.fixture(scope="module")
def case10k():
    w = configs.adjoint_10k_24h()
    f = w.velocity_field()
    lon, lat = w.particles()
    tlon, tlat = targets_of(w, lon, lat, w.steps)
    return w, f, lon, lat, tlon, tlat


def test_forward_bit_exact_with_oracle(gpu, case10k):
    w, f, lon, lat, _, _ = case10k
    plan = adjoint.AdjointPlan(f, 0.0, w.dt, w.steps)
    u, v = f32_tensors(f)
    X, Y = adjoint.advect(torch.tensor(lon, device="cuda"), torch.tensor(lat, device="cuda"), u, v, plan)
    lo, la, _, _, oob = oracle_advect(widened(f), lon, lat, w.dt, w.steps)
    err = max(np.max(np.abs(X.cpu().numpy() - lo)), np.max(np.abs(Y.cpu().numpy() - la)))
    assert oob == 0 and int(plan.oob.item()) == 0
    assert err <= 1e-9, err  # stated tolerance; the operation order is shared, so expect 0
    assert err == 0.0, f"not bit-exact: {err}"


def test_gradients_match_autograd_reference(gpu, case10k):
    w, f, lon, lat, tlon, tlat = case10k
    g = gpu_grads(f, lon, lat, tlon, tlat, w.dt, w.steps)
    r = ref_grads(f, lon, lat, tlon, tlat, w.dt, w.steps)
    assert abs(g["L"] - r["L"]) <= 1e-12 * r["L"]
    assert_close_to_ref(g, r)
    assert np.count_nonzero(g["gu"]) > 1000  # the particles sample many nodes


def fd_loss(f, plan, lon_t, lat_t, tx, ty, u, v):
    with torch.no_grad():
        X, Y = adjoint.advect(lon_t, lat_t, u, v, plan)
        return float(adjoint.trajectory_loss(X, Y, tx, ty))


def check_fd(f, lon, lat, tlon, tlat, dt, steps, grads, nodes, eps=1e-3, median_tol=1e-5, max_tol=5e-3):
    """Central differences of the loss at each node against the gradient.  The loss is smooth except
    where a perturbed trajectory crosses a cell edge, whose kink the +-eps difference straddles:
    most nodes agree to ~1e-7 (eps^2 truncation), a few to ~1e-3 (kinks), hence the median bound and
    the looser bound on every node."""
    plan = adjoint.AdjointPlan(f, 0.0, dt, steps)
    lon_t = torch.tensor(lon, device="cuda")
    lat_t = torch.tensor(lat, device="cuda")
    tx, ty = torch.tensor(tlon, device="cuda"), torch.tensor(tlat, device="cuda")
    u0, v0 = f32_tensors(f)
    scale = max(np.abs(grads["gu"]).max(), np.abs(grads["gv"]).max())
    errs = []
    for comp, node in nodes:
        base = (u0 if comp == 0 else v0)
        val = base[node].item()
        hi = np.float32(val + eps)
        lo = np.float32(val - eps)
        delta = float(hi) - float(lo)  # the float32 perturbation actually applied
        pert = base.clone()
        pert[node] = float(hi)
        Lp = fd_loss(f, plan, lon_t, lat_t, tx, ty, *((pert, v0) if comp == 0 else (u0, pert)))
        pert[node] = float(lo)
        Lm = fd_loss(f, plan, lon_t, lat_t, tx, ty, *((pert, v0) if comp == 0 else (u0, pert)))
        fd = (Lp - Lm) / delta
        an = float((grads["gu"] if comp == 0 else grads["gv"])[node])
        err = abs(fd - an) / max(abs(an), 1e-3 * scale)
        errs.append(err)
        assert err <= max_tol, (comp, node, fd, an)
    assert float(np.median(errs)) <= median_tol, sorted(errs)
    return errs


def test_finite_differences_32_nodes(gpu, case10k):
    w, f, lon, lat, tlon, tlat = case10k
    g = gpu_grads(f, lon, lat, tlon, tlat, w.dt, w.steps)
    rng = np.random.default_rng(2024)
    nodes = []
    for comp, arr in ((0, g["gu"]), (1, g["gv"])):
        touched = np.argwhere(np.abs(arr) > 1e-2 * np.abs(arr).max())
        pick = rng.choice(len(touched), 16, replace=False)
        nodes += [(comp, tuple(int(c) for c in touched[k])) for k in pick]
    assert len(nodes) == 32
    check_fd(f, lon, lat, tlon, tlat, w.dt, w.steps, g, nodes)


def uniform_field(U, V, nlev):
    lons = np.arange(200.0, 220.01, 0.25)
    lats = np.arange(20.0, 40.01, 0.25)
    u = np.full((nlev, lats.size, lons.size), U)
    v = np.full((nlev, lats.size, lons.size), V)
    return VelocityField(lons, lats, u, v, None if nlev == 1 else np.arange(nlev) * 13 * 3600.0)


@pytest.mark.parametrize("nlev", [1, 20])
def test_uniform_zonal_flow_closed_form(gpu, nlev):
    U = float(np.float32(0.3))
    steps, dt, T = 24, 3600.0, 24 * 3600.0
    f = uniform_field(U, 0.0, nlev)
    lon = np.linspace(205.05, 214.95, 1000)
    lat = np.linspace(25.05, 34.95, 1000)
    K = 1.0 / M
    c, s = np.cos(np.deg2rad(lat)), np.sin(np.deg2rad(lat))
    plan = adjoint.AdjointPlan(f, 0.0, dt, steps)
    for which in ("lon", "lat"):
        u, v = f32_tensors(f, requires_grad=True)
        x = torch.tensor(lon, device="cuda", requires_grad=True)
        y = torch.tensor(lat, device="cuda", requires_grad=True)
        X, Y = adjoint.advect(x, y, u, v, plan)
        (X.sum() if which == "lon" else Y.sum()).backward()
        if which == "lon":
            assert np.allclose(X.detach().cpu().numpy(), lon + T * U * K / c, rtol=0, atol=1e-11)
            assert np.allclose(x.grad.cpu().numpy(), 1.0, rtol=0, atol=1e-12)
            assert np.allclose(y.grad.cpu().numpy(), T * U * K * (np.pi / 180) * s / c ** 2, rtol=1e-10, atol=0)
            assert abs(float(u.grad.double().sum()) - np.sum(T * K / c)) <= 1e-5 * np.sum(T * K / c)
        else:
            assert abs(float(v.grad.double().sum()) - lon.size * T * K) <= 1e-5 * lon.size * T * K
            assert np.allclose(y.grad.cpu().numpy(), 1.0, rtol=0, atol=1e-12)


def test_checkpoint_interval_does_not_change_gradients(gpu, case10k):
    w, f, lon, lat, tlon, tlat = case10k
    base = gpu_grads(f, lon, lat, tlon, tlat, w.dt, w.steps, ckpt=16)
    for ckpt in (1, 7, 24, 64):
        g = gpu_grads(f, lon, lat, tlon, tlat, w.dt, w.steps, ckpt=ckpt)
        # the recompute restarts from the forward's exact positions at each checkpoint and runs within
        # rounding of them in between, so segmentation moves the position adjoint by rounding only
        assert normwise(g["gx"], base["gx"]) <= 1e-12 and normwise(g["gy"], base["gy"]) <= 1e-12, ckpt
        # field gradients differ only in float32 summation order
        assert normwise(g["gu"], base["gu"]) <= 1e-5 and normwise(g["gv"], base["gv"]) <= 1e-5, ckpt


def test_shuffled_particles_use_global_fallback(gpu, case10k):
    # Shuffled, each block of 256 particles spans the whole patch: its window of 41x41 nodes over the
    # ~6 levels a 64-step segment reads (41*41*6*8 bytes = 80 KB) exceeds the 32 KB of LDS, so blocks
    # add to global memory directly; with 16-step segments windows of 2 levels still fit.
    w, f, lon, lat, tlon, tlat = case10k
    perm = np.random.default_rng(7).permutation(lon.size)
    lon, lat, tlon, tlat = lon[perm], lat[perm], tlon[perm], tlat[perm]
    r = ref_grads(f, lon, lat, tlon, tlat, w.dt, w.steps)
    for ckpt in (16, 64):
        g = gpu_grads(f, lon, lat, tlon, tlat, w.dt, w.steps, ckpt=ckpt)
        assert_close_to_ref(g, r)


def test_steps_leaving_the_grid(gpu):
    # northward flow carries the northern rows of particles across the field's edge mid-run
    w = configs.adjoint_10k_24h()
    f0 = w.velocity_field()
    f = VelocityField(f0.lons, f0.lats, f0.u, f0.v + 0.8, f0.times)
    lon, lat = w.particles()
    lat = lat + (f.lats[-1] - 0.3 - lat.max())
    tlon, tlat = lon + 0.1, lat - 0.1
    g = gpu_grads(f, lon, lat, tlon, tlat, w.dt, w.steps)
    r = ref_grads(f, lon, lat, tlon, tlat, w.dt, w.steps)
    assert r["oob"] > 1000
    assert np.max(np.abs(g["X"] - r["X"])) < 1e-11 and np.max(np.abs(g["Y"] - r["Y"])) < 1e-11
    assert_close_to_ref(g, r)


def test_steady_field(gpu):
    w = configs.rps_10k_steady()
    f = w.velocity_field()
    lon, lat = w.particles()
    tlon, tlat = lon + 0.01, lat + 0.02
    g = gpu_grads(f, lon, lat, tlon, tlat, w.dt, w.steps, ckpt=5)
    r = ref_grads(f, lon, lat, tlon, tlat, w.dt, w.steps)
    assert_close_to_ref(g, r)


@pytest.mark.parametrize("n", [1, 255, 1000])
def test_ragged_sizes(gpu, case10k, n):
    w, f, lon, lat, tlon, tlat = case10k
    sel = np.random.default_rng(n).choice(lon.size, n, replace=False)
    g = gpu_grads(f, lon[sel], lat[sel], tlon[sel], tlat[sel], w.dt, 9, ckpt=4)
    r = ref_grads(f, lon[sel], lat[sel], tlon[sel], tlat[sel], w.dt, 9)
    assert_close_to_ref(g, r)


def test_empty_and_zero_steps(gpu, case10k):
    w, f, lon, lat, _, _ = case10k
    u, v = f32_tensors(f, requires_grad=True)
    plan = adjoint.AdjointPlan(f, 0.0, w.dt, 0)
    x = torch.tensor(lon[:10], device="cuda", requires_grad=True)
    X, Y = adjoint.advect(x, torch.tensor(lat[:10], device="cuda"), u, v, plan)
    X.sum().backward()
    assert torch.equal(X.detach(), x.detach()) and torch.all(x.grad == 1) and float(u.grad.abs().sum()) == 0
    plan = adjoint.AdjointPlan(f, 0.0, w.dt, 5)
    e = torch.zeros(0, dtype=torch.float64, device="cuda", requires_grad=True)
    X, Y = adjoint.advect(e, e, u, v, plan)
    (X.sum() + Y.sum()).backward()
    assert X.numel() == 0


def test_only_longitude_used(gpu, case10k):
    # a loss that ignores lat hands the backward a None for it
    w, f, lon, lat, _, _ = case10k
    plan = adjoint.AdjointPlan(f, 0.0, w.dt, 8)
    u, v = f32_tensors(f, requires_grad=True)
    X, _ = adjoint.advect(torch.tensor(lon, device="cuda"), torch.tensor(lat, device="cuda"), u, v, plan)
    X.sum().backward()
    fw = widened(f)
    lev0, wt = fw.stage_schedule(0.0, w.dt, 8)
    ur = torch.tensor(fw.u, requires_grad=True)
    vr = torch.tensor(fw.v, requires_grad=True)
    Xr, _, _ = adjoint_ref.rk4(torch.tensor(lon), torch.tensor(lat), ur, vr, grid_of(fw), lev0, wt, w.dt, 8)
    Xr.sum().backward()
    assert normwise(u.grad.cpu().numpy(), ur.grad.numpy()) <= 1e-5
    assert normwise(v.grad.cpu().numpy(), vr.grad.numpy()) <= 1e-5


@pytest.fixture(scope="module")
def full_size():
    """The 1,000,000-particle, 240-step workload: its gradients against targets advected through the
    second seeded field."""
    w = configs.adjoint_1m_10day()
    f = w.velocity_field()
    lon, lat = w.particles()
    tf = w.target_velocity_field()
    tplan = adjoint.AdjointPlan(tf, 0.0, w.dt, w.steps)
    tu, tv = f32_tensors(tf)
    with torch.no_grad():
        TX, TY = adjoint.advect(torch.tensor(lon, device="cuda"), torch.tensor(lat, device="cuda"), tu, tv, tplan)
    tlon, tlat = TX.cpu().numpy(), TY.cpu().numpy()
    g = gpu_grads(f, lon, lat, tlon, tlat, w.dt, w.steps)
    assert g["oob"] == 0
    rng = np.random.default_rng(5)
    nodes = []
    for comp, arr in ((0, g["gu"]), (1, g["gv"])):
        touched = np.argwhere(np.abs(arr) > 1e-2 * np.abs(arr).max())
        nodes += [(comp, tuple(int(c) for c in touched[k])) for k in rng.choice(len(touched), 16, replace=False)]
    return w, f, lon, lat, tlon, tlat, g, nodes


def test_full_size_finite_differences(gpu, full_size):
    # 1,000,000 particles, 240 steps, 15 checkpoint segments: the gradient on 32 random sampled nodes
    # (16 of u, 16 of v) against central differences of the same forward with eps = 1e-3 m/s.  Over
    # 240 steps far more perturbed trajectories straddle a cell crossing than at 24 steps, and a node's
    # float32 gradient sums ~1e5 samples: measured medians 1e-4 .. 2e-4
    w, f, lon, lat, tlon, tlat, g, nodes = full_size
    check_fd(f, lon, lat, tlon, tlat, w.dt, w.steps, g, nodes, median_tol=5e-4)


def test_full_size_finite_differences_small_eps(gpu, full_size):
    # the same 32 nodes with eps = 1e-4 m/s: ten times fewer perturbed trajectories straddle a cell
    # crossing, so the differences come closer to the gradient; a wrong term in the reverse step
    # (a dropped stage, a sign) would be off by O(1) at either eps
    w, f, lon, lat, tlon, tlat, g, nodes = full_size
    check_fd(f, lon, lat, tlon, tlat, w.dt, w.steps, g, nodes, eps=1e-4, median_tol=1e-4)


def test_full_size_final_interaction(gpu, full_size):
    # the RPS update runs forward only on the final state: pairs on the float32-rounded final positions
    # against scipy's cKDTree, species against the sequential oracle with the same keyed uniforms
    import oracle
    from oracle.pairs import sorted_kdtree_pairs
    w, f, lon, lat, _, _, g, _ = full_size
    sp0 = w.species()
    sp = torch.tensor(sp0, device="cuda")
    X = torch.tensor(g["X"], device="cuda")
    Y = torch.tensor(g["Y"], device="cuda")
    npairs = adjoint.interact_final(X, Y, sp, w.radius, w.params, w.key, w.steps, f.bounds)
    rx, ry = g["X"].astype(np.float32), g["Y"].astype(np.float32)
    pi, pj = sorted_kdtree_pairs(rx, ry, w.radius, 2.0)
    assert npairs == len(pi) > 1000
    want = oracle.rps_sequential(sp0, pi, pj, w.p, w.p, w.p, w.key, w.steps)
    got = sp.cpu().numpy()
    assert not np.array_equal(want, sp0)
    assert np.array_equal(got, want)
