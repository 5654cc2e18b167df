"""ORACLE (test infrastructure): independent RK4 restatement with library parts.

A second formulation of the advection step that shares no code with
lm_oracle.c: the velocity is interpolated by scipy.interpolate
RegularGridInterpolator(method="linear") over (time, lat, lon) -- linear in
time, bilinear in space, the semantics of parcels.Field(interp_method=
"linear") set up at particle_advecter.py:177-184 -- the unit conversion uses
numpy.cos, and the step is the generic explicit Runge-Kutta loop over the
classic RK4 Butcher tableau (parcels.AdvectionRK4, called at
particle_advecter.py:222-223).  Agreement with the C oracle to rounding level
pins the C oracle's indexing, weights, stage times and conversion.
"""
import numpy as np
from scipy.interpolate import RegularGridInterpolator

M_PER_DEG = 1852.0 * 60.0

RK4_A = [[], [0.5], [0.0, 0.5], [0.0, 0.0, 1.0]]
RK4_B = [1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0]
RK4_C = [0.0, 0.5, 0.5, 1.0]


class GridVelocity:
    """u, v [nt, ny, nx] (m/s) on lons [nx], lats [ny], times [nt] (s)."""

    def __init__(self, lons, lats, times, u, v):
        self.steady = len(times) == 1
        if self.steady:
            self.iu = RegularGridInterpolator((lats, lons), u[0], method="linear", bounds_error=True)
            self.iv = RegularGridInterpolator((lats, lons), v[0], method="linear", bounds_error=True)
        else:
            self.iu = RegularGridInterpolator((times, lats, lons), u, method="linear", bounds_error=True)
            self.iv = RegularGridInterpolator((times, lats, lons), v, method="linear", bounds_error=True)

    def deg_per_s(self, t, lon, lat):
        if self.steady:
            pts = np.stack([lat, lon], axis=-1)
        else:
            pts = np.stack([np.full_like(lon, t), lat, lon], axis=-1)
        u = self.iu(pts)
        v = self.iv(pts)
        return u / (M_PER_DEG * np.cos(np.deg2rad(lat))), v / M_PER_DEG


def rk4(field, lon, lat, t0, dt, nsteps):
    """Advance copies of lon/lat; returns (lon, lat, record[nsteps, 2, n])."""
    x = np.array(lon, dtype=np.float64, copy=True)
    y = np.array(lat, dtype=np.float64, copy=True)
    rec = np.zeros((nsteps, 2, x.size))
    t = float(t0)
    for s in range(nsteps):
        ks = []
        for st in range(4):
            xs = x + dt * sum(a * k[0] for a, k in zip(RK4_A[st], ks))
            ys = y + dt * sum(a * k[1] for a, k in zip(RK4_A[st], ks))
            ks.append(field.deg_per_s(t + RK4_C[st] * dt, xs, ys))
        x = x + dt * sum(b * k[0] for b, k in zip(RK4_B, ks))
        y = y + dt * sum(b * k[1] for b, k in zip(RK4_B, ks))
        t += dt
        rec[s, 0] = x
        rec[s, 1] = y
    return x, y, rec
