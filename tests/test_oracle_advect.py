"""Pins the oracle's RK4 advection (parcels AdvectionRK4 semantics, call site
particle_advecter.py:222-223) with closed forms and an independent
restatement built on scipy's RegularGridInterpolator."""
import numpy as np
import pytest

import oracle
from oracle import advect_ref
from lagrangian_microbes_amd import configs
from lagrangian_microbes_amd.velocity_fields import VelocityField

M = 1852.0 * 60.0


def grid_of(f):
    return (f.nx, f.ny, f.x0, f.y0, 1.0 / f.dx, 1.0 / f.dy)


def run(f, lon, lat, dt, nsteps, t0=0.0):
    if f.nt == 1:
        return oracle.advect_rk4_f64(lon, lat, f.interleaved()[:1], grid_of(f), nsteps, dt, steady=True)
    lev0, w = f.stage_schedule(t0, dt, nsteps)
    return oracle.advect_rk4_f64(lon, lat, oracle.field_blend_f64(f.interleaved(), lev0, w), grid_of(f),
                                 nsteps, dt)


def const_field(u, v, nt=1):
    lons = np.arange(200.0, 220.01, 0.5)
    lats = np.arange(20.0, 40.01, 0.5)
    U = np.full((nt, lats.size, lons.size), u)
    V = np.full((nt, lats.size, lons.size), v)
    return VelocityField(lons, lats, U, V, None if nt == 1 else np.arange(nt) * 86400.0)


def test_cos_polynomial_against_numpy():
    lat = np.linspace(-90, 90, 20001)
    ours = np.array([oracle.cos_deg(v) for v in lat])
    ref = np.cos(np.deg2rad(lat))
    assert np.max(np.abs(ours - ref)) < 4e-16


def test_meridional_uniform_flow_exact():
    f = const_field(0.0, 0.3)
    lon = np.array([205.0, 210.0]); lat = np.array([25.0, 30.0])
    lo, la, _, _, oob = run(f, lon, lat, 3600.0, 48)
    assert oob == 0
    assert np.array_equal(lo, lon)
    assert np.allclose(la, lat + 0.3 * 48 * 3600.0 / M, rtol=0, atol=1e-12)


def test_zonal_flow_along_parallel_exact():
    f = const_field(0.25, 0.0)
    lon = np.array([205.0]); lat = np.array([33.0])
    lo, la, _, _, _ = run(f, lon, lat, 3600.0, 24)
    assert la[0] == 33.0
    assert abs(lo[0] - (205.0 + 0.25 * 24 * 3600.0 / (M * np.cos(np.deg2rad(33.0))))) < 1e-12


def test_zonal_and_meridional_closed_form():
    # dlat/dt = v/M, dlon/dt = u / (M cos lat): lon(t) = lon0 + (u/v) * (gd(lat) - gd(lat0)) in degrees,
    # with gd(phi) = ln(sec phi + tan phi) in radians -> degrees
    u, v = 0.2, 0.3
    f = const_field(u, v)
    lon0, lat0 = 205.0, 25.0
    T = 96 * 3600.0
    lo, la, _, _, _ = run(f, np.array([lon0]), np.array([lat0]), 3600.0, 96)
    lat1 = lat0 + v * T / M
    gd = lambda d: np.log(1 / np.cos(np.deg2rad(d)) + np.tan(np.deg2rad(d)))
    lon1 = lon0 + (u / v) * np.rad2deg(gd(lat1) - gd(lat0))
    assert abs(la[0] - lat1) < 1e-12
    assert abs(lo[0] - lon1) < 1e-9  # RK4 truncation error at dt = 1 h is far below this


def test_steady_equals_constant_levels():
    w = configs.rps_10k_steady()
    f = w.velocity_field()
    f3 = VelocityField(f.lons, f.lats, np.repeat(f.u, 3, 0), np.repeat(f.v, 3, 0), np.arange(3) * 86400.0)
    lon, lat = w.particles()
    a = run(f, lon[:500], lat[:500], 3600.0, 24)
    b = run(f3, lon[:500], lat[:500], 3600.0, 24)
    # blending identical levels reproduces the level exactly, so the paths agree bit for bit
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_blend_weights():
    levels = np.random.default_rng(0).normal(size=(3, 4, 5, 2))
    out = oracle.field_blend_f64(levels, np.array([0, 1, 1]), np.array([0.0, 0.25, 1.0]))
    assert np.array_equal(out[0], levels[0])
    assert np.allclose(out[1], levels[1] + 0.25 * (levels[2] - levels[1]), atol=1e-15)
    assert np.allclose(out[2], levels[2], atol=1e-15)


def test_out_of_grid_is_counted_and_frozen():
    f = const_field(0.0, 1.0)
    lon = np.array([210.0, 210.0]); lat = np.array([30.0, 39.99])
    lo, la, _, _, oob = run(f, lon, lat, 3600.0, 2)
    assert oob == 2 and la[1] == 39.99 and la[0] > 30.0


@pytest.mark.parametrize("name,steps,stride", [("rps_10k_steady_24h", 24, 1), ("rps_490k_30day", 48, 245)])
def test_against_independent_restatement(name, steps, stride):
    w = configs.WORKLOADS[name]()
    f = w.velocity_field()
    lon, lat = w.particles()
    lon, lat = lon[::stride], lat[::stride]
    lo, la, rl, rt, oob = run(f, lon, lat, w.dt, steps)
    gv = advect_ref.GridVelocity(f.lons, f.lats, f.times, f.u, f.v)
    x, y, rec = advect_ref.rk4(gv, lon, lat, 0.0, w.dt, steps)
    assert oob == 0
    assert np.max(np.abs(lo - x)) < 1e-12 and np.max(np.abs(la - y)) < 1e-12
    assert np.max(np.abs(lo - lon)) > 0.05  # the field really moves particles
    assert np.array_equal(rl, rec[:, 0].astype(np.float32)) or np.max(np.abs(rl - rec[:, 0])) < 1e-4


def test_cos_polynomial_f32_against_numpy():
    lat = np.linspace(-90, 90, 20001)
    ours = np.array([oracle.cos_deg_f32(v) for v in lat], dtype=np.float64)
    # float32 rounding of the argument and of the result: a few units of 2^-24
    assert np.max(np.abs(ours - np.cos(np.deg2rad(lat)))) < 4e-7


def test_fp32_restatement_uniform_flows():
    # constant velocities: every stage sees the same float32 velocity, so the position moves by
    # 24 equal float32 increments accumulated in float64
    f = const_field(0.0, 0.3)
    lon = np.array([205.0]); lat = np.array([30.0])
    lev0, w = f.stage_schedule(0.0, 3600.0, 24)
    sf = f.interleaved()[:1].astype(np.float32)
    lo, la, _, _, oob = oracle.advect_rk4_f32(lon, lat, sf, grid_of(f), 24, 3600.0, steady=True)
    assert oob == 0 and lo[0] == 205.0
    vdeg = np.float32(np.float32(0.3) * np.float32(float.fromhex("0x1.2df73ap-17")))
    assert abs(la[0] - (30.0 + 24 * 3600.0 * float(vdeg))) < 1e-10


def test_fp32_restatement_close_to_fp64():
    # the FP32 path differs from the FP64 one by float32 interpolation and velocity rounding only:
    # over one day of the 1/12-degree 21-level field the positions agree to ~1e-8 degrees
    w = configs.rps_4m_60day()
    f = w.velocity_field()
    lon, lat = w.particles()
    sel = np.arange(0, w.n, 40009)[:100]
    lon, lat = lon[sel], lat[sel]
    steps = 48
    lev0, wt = f.stage_schedule(0.0, w.dt, steps)
    a = oracle.advect_rk4_f64(lon, lat, oracle.field_blend_f64(f.interleaved(), lev0, wt), grid_of(f), steps,
                              w.dt, record=False)
    b = oracle.advect_rk4_f32(lon, lat, oracle.field_blend_f32(f.interleaved().astype(np.float32), lev0, wt),
                              grid_of(f), steps, w.dt, record=False)
    d = np.hypot(a[0] - b[0], a[1] - b[1])
    assert a[4] == b[4] == 0
    assert d.max() < 5e-8, d.max()
    assert np.median(np.hypot(a[0] - lon, a[1] - lat)) > 0.05  # the particles do move
