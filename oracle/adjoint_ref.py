"""ORACLE -- test infrastructure, not product code.

Differentiable CPU restatement of the RK4 advection for checking liblm's
hand-written adjoint: float64 torch operations, gradients by torch.autograd.

The forward is the same method as the C oracle's oracle_advect_rk4_f64
(lm_oracle.c), i.e. parcels.AdvectionRK4 with linear interpolation as
ParticleAdvecter.time_step sets it up (particle_advecter.py:177-187,
222-223):
  - bilinear interpolation in lon/lat inside cell (i, j) with i = trunc(fx)
    clamped to nx - 2 (j likewise), node-wise linear blending in time between
    level lev0[k] and lev0[k] + 1 with weight w[k] at stage time k;
  - u, v in m/s become deg/s through 1852*60 m per degree and cos(lat);
  - a step with any stage outside the grid keeps its position.
Arithmetic is written as multiply-then-add (torch has no fused form), so the
positions agree with the C oracle to rounding (about 1e-12 degrees after 24
steps), not bit for bit.

What pins it (tests/test_adjoint_oracle.py): its positions against the C
oracle; its gradients against central finite differences of its own forward
in float64; and closed forms of a uniform zonal flow, where
lon_T = lon0 + T U K / cos(lat0) gives dlon_T/dU summed over the nodes,
dlat_T/dV summed over the nodes and dlon_T/dlat0 exactly.
"""
import numpy as np
import torch

M_PER_DEG = 1852.0 * 60.0
K = 1.0 / M_PER_DEG
D2R = np.pi / 180.0


def _sample(fu, fv, lev, w, x, y, grid, nlev):
    """Velocity (deg/s) at (x, y) for one stage time; and an inside mask."""
    nx, ny, x0, y0, inv_dx, inv_dy = grid
    fx = (x - x0) * inv_dx
    fy = (y - y0) * inv_dy
    ok = (fx >= 0) & (fx <= nx - 1) & (fy >= 0) & (fy <= ny - 1)
    fxd = fx.detach().clamp(0, nx - 1)
    fyd = fy.detach().clamp(0, ny - 1)
    i = torch.clamp(fxd.to(torch.int64), max=nx - 2)
    j = torch.clamp(fyd.to(torch.int64), max=ny - 2)
    xi = fx - i.to(fx.dtype)
    eta = fy - j.to(fy.dtype)
    l1 = lev + 1 if nlev > 1 else lev

    def corner(f, jj, ii):
        a = f[lev, jj, ii]
        b = f[l1, jj, ii]
        return a + w * (b - a)

    def interp(f):
        c00, c01 = corner(f, j, i), corner(f, j, i + 1)
        c10, c11 = corner(f, j + 1, i), corner(f, j + 1, i + 1)
        bot = c00 + xi * (c01 - c00)
        top = c10 + xi * (c11 - c10)
        return bot + eta * (top - bot)

    u = interp(fu)
    v = interp(fv)
    return u * K / torch.cos(y * D2R), v * K, ok


def rk4(lon0, lat0, u, v, grid, lev0, w, dt, nsteps):
    """Positions after nsteps RK4 steps; differentiable in every tensor input.

    u, v: [nlev, ny, nx] (any float dtype; widened to float64); grid:
    (nx, ny, x0, y0, inv_dx, inv_dy); lev0, w: the 2*nsteps+1 stage levels and
    weights (numpy).  Returns (lon, lat, number of steps that left the grid).
    """
    fu, fv = u.to(torch.float64), v.to(torch.float64)
    nlev = fu.shape[0]
    x, y = lon0.to(torch.float64), lat0.to(torch.float64)
    hh, h, h6 = 0.5 * dt, dt, dt / 6.0
    oob = 0
    for s in range(nsteps):
        st = [(int(lev0[2 * s + q]), float(w[2 * s + q])) for q in range(3)]
        u1, v1, ok1 = _sample(fu, fv, st[0][0], st[0][1], x, y, grid, nlev)
        u2, v2, ok2 = _sample(fu, fv, st[1][0], st[1][1], x + u1 * hh, y + v1 * hh, grid, nlev)
        u3, v3, ok3 = _sample(fu, fv, st[1][0], st[1][1], x + u2 * hh, y + v2 * hh, grid, nlev)
        u4, v4, ok4 = _sample(fu, fv, st[2][0], st[2][1], x + u3 * h, y + v3 * h, grid, nlev)
        # later stages of a step that already left the grid are masked out with it
        ok = ok1 & ok2 & ok3 & ok4
        xn = x + (u1 + 2.0 * (u2 + u3) + u4) * h6
        yn = y + (v1 + 2.0 * (v2 + v3) + v4) * h6
        x = torch.where(ok, xn, x)
        y = torch.where(ok, yn, y)
        oob += int((~ok).sum())
    return x, y, oob
