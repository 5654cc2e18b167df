"""Differentiable RK4 advection through a learnable float32 velocity field.

``advect(lon0, lat0, u, v, plan)`` returns the positions after ``plan.nsteps``
RK4 steps, as a ``torch.autograd.Function``: the forward is liblm's
``lm_adjoint_forward`` (the FP64 RK4 of ``ParticleAdvecter.time_step``,
particle_advecter.py:177-187,222-223, on the float32 field widened to
double), and the backward is the hand-written reverse sweep
``lm_adjoint_backward``, which recomputes every segment of ``plan.ckpt``
steps from the positions the forward kept at its start.  Gradients flow to the
initial positions and to u and v.

The reference has no gradient path; this module adds one for fitting a
gridded field to observed trajectories, with ``great_circle_sq`` as the
distance the loss of such a fit is built on.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import LmGrid, check, ptr, require, stream_ptr
from .velocity_fields import M_PER_DEG

# Radius of the sphere on which one degree of arc is 1852*60 m, the
# convention the advection's m/s -> deg/s conversion uses.
SPHERE_RADIUS_M = M_PER_DEG * 180.0 / np.pi


class AdjointPlan:
    """Grid geometry, stage-time schedule and checkpoint interval of a run.

    field: a VelocityField giving the grid and the time levels (its own u, v
    are not used: the run reads the u, v tensors handed to ``advect``).
    The run covers nsteps RK4 steps of dt seconds from t0 (seconds on the
    field's time axis); a checkpoint is kept every ``ckpt`` steps.
    """

    def __init__(self, field, t0, dt, nsteps, ckpt=16, device=None):
        if nsteps < 0 or ckpt < 1:
            raise ValueError("nsteps must be >= 0 and ckpt >= 1")
        self.device = torch.device(device or "cuda")
        self.nx, self.ny, self.nlev = field.nx, field.ny, field.nt
        self.grid = LmGrid(field.nx, field.ny, field.x0, field.y0, 1.0 / field.dx, 1.0 / field.dy)
        lev0, w = field.stage_schedule(t0, dt, nsteps)
        self.lev0_host = np.ascontiguousarray(lev0, dtype=np.int32)
        self.lev0 = torch.from_numpy(self.lev0_host).to(self.device)
        self.w = torch.from_numpy(np.ascontiguousarray(w, dtype=np.float64)).to(self.device)
        self.dt, self.nsteps, self.ckpt = float(dt), int(nsteps), int(ckpt)
        self.nseg = (self.nsteps + self.ckpt - 1) // self.ckpt
        # steps of the forward runs that left the grid (device counter)
        self.oob = torch.zeros(1, dtype=torch.int32, device=self.device)

    def field_shape(self):
        return (self.nlev, self.ny, self.nx)

    def interleave(self, u, v):
        """(u, v) float32 [nlev, ny, nx] -> the device layout [nlev, ny, nx, 2]."""
        shape = self.field_shape()
        for name, t in (("u", u), ("v", v)):
            if tuple(t.shape) != shape:
                raise ValueError(f"{name} must be {shape} (got {tuple(t.shape)})")
            if t.dtype != torch.float32 or t.device.type != "cuda":
                raise TypeError(f"{name} must be a float32 GPU tensor")
        return torch.stack((u.detach(), v.detach()), dim=-1).contiguous()

    def lev0_host_ptr(self):
        return ctypes.c_void_p(self.lev0_host.ctypes.data)


# The forward blends the float64 stage fields of as many steps at once as
# this budget holds (all 481 stage times of 240 steps on an 81x81 grid take
# 50 MB).
FORWARD_STAGE_BUDGET = 256 << 20


def forward_scratch_bytes(plan):
    field_bytes = plan.nx * plan.ny * 16
    if plan.nlev == 1:
        return field_bytes
    return max(3, min(2 * plan.nsteps + 1, FORWARD_STAGE_BUDGET // field_bytes)) * field_bytes


def forward_positions(lon, lat, field, plan, checkpoints):
    """lm_adjoint_forward on lon/lat (float64, in place); field from
    ``plan.interleave``; checkpoints float64 [nseg, n, 2]."""
    n = lon.numel()
    require(lon, torch.float64, name="lon")
    require(lat, torch.float64, numel=n, name="lat")
    require(field, torch.float32, numel=plan.nlev * plan.ny * plan.nx * 2, name="field")
    require(checkpoints, torch.float64, min_numel=plan.nseg * n * 2, name="checkpoints")
    scratch = torch.empty(forward_scratch_bytes(plan), dtype=torch.uint8, device=lon.device)
    check(_lib.load().lm_adjoint_forward(n, ptr(lon), ptr(lat), ptr(field), plan.nlev, ctypes.byref(plan.grid),
                                         ptr(plan.lev0), ptr(plan.w), plan.lev0_host_ptr(), plan.nsteps, plan.dt,
                                         plan.ckpt, ptr(checkpoints), ptr(scratch), scratch.numel(), ptr(plan.oob),
                                         stream_ptr()))


def backward_positions(field, plan, checkpoints, glon, glat, gfield):
    """lm_adjoint_backward: glon/glat (float64) go from dL/d(final) to
    dL/d(initial) in place; dL/d(field) is added into gfield."""
    n = glon.numel()
    require(glon, torch.float64, name="glon")
    require(glat, torch.float64, numel=n, name="glat")
    require(field, torch.float32, numel=plan.nlev * plan.ny * plan.nx * 2, name="field")
    require(gfield, torch.float32, numel=field.numel(), name="gfield")
    require(checkpoints, torch.float64, min_numel=plan.nseg * n * 2, name="checkpoints")
    L = _lib.load()
    nbytes = L.lm_adjoint_scratch_bytes(n, plan.ckpt, ctypes.byref(plan.grid), plan.nlev)
    if nbytes < 0:
        raise ValueError("invalid particle count or checkpoint interval")
    scratch = torch.empty(max(int(nbytes), 16), dtype=torch.uint8, device=glon.device)
    check(L.lm_adjoint_backward(n, ptr(checkpoints), ptr(field), plan.nlev, ctypes.byref(plan.grid), ptr(plan.lev0),
                                ptr(plan.w), plan.lev0_host_ptr(), plan.nsteps, plan.dt, plan.ckpt, ptr(scratch),
                                scratch.numel(), ptr(glon), ptr(glat), ptr(gfield), stream_ptr()))


class Rk4Adjoint(torch.autograd.Function):
    """(lon0, lat0, u, v) -> (lon, lat) after plan.nsteps RK4 steps."""

    @staticmethod
    def forward(ctx, lon0, lat0, u, v, plan):
        field = plan.interleave(u, v)
        lon = lon0.detach().to(torch.float64).contiguous().clone()
        lat = lat0.detach().to(torch.float64).contiguous().clone()
        ck = torch.empty((plan.nseg, lon.numel(), 2), dtype=torch.float64, device=lon.device)
        forward_positions(lon, lat, field, plan, ck)
        ctx.save_for_backward(field, ck)
        ctx.plan = plan
        ctx.dtypes = (lon0.dtype, lat0.dtype)
        return lon, lat

    @staticmethod
    def backward(ctx, glon_out, glat_out):
        field, ck = ctx.saved_tensors
        plan = ctx.plan
        n = ck.shape[1]
        dev = field.device
        glon = (glon_out.to(torch.float64).contiguous().clone() if glon_out is not None
                else torch.zeros(n, dtype=torch.float64, device=dev))
        glat = (glat_out.to(torch.float64).contiguous().clone() if glat_out is not None
                else torch.zeros(n, dtype=torch.float64, device=dev))
        gfield = torch.zeros_like(field)
        backward_positions(field, plan, ck, glon, glat, gfield)
        need = ctx.needs_input_grad
        return (glon.to(ctx.dtypes[0]) if need[0] else None, glat.to(ctx.dtypes[1]) if need[1] else None,
                gfield[..., 0].contiguous() if need[2] else None, gfield[..., 1].contiguous() if need[3] else None,
                None)


def advect(lon0, lat0, u, v, plan):
    """Differentiable RK4 advection: see ``Rk4Adjoint``."""
    return Rk4Adjoint.apply(lon0, lat0, u, v, plan)


def great_circle_sq(lon, lat, lon_t, lat_t, radius=SPHERE_RADIUS_M):
    """Squared great-circle distance (haversine), in the units of radius^2."""
    d2r = np.pi / 180.0
    p1, p2 = lat * d2r, lat_t * d2r
    a = torch.sin(0.5 * (p2 - p1)) ** 2 + torch.cos(p1) * torch.cos(p2) * torch.sin(0.5 * (lon_t - lon) * d2r) ** 2
    # a is clamped away from 0, where the square root's derivative is infinite
    ang = 2.0 * torch.asin(torch.sqrt(a.clamp(min=1e-300, max=1.0)))
    return (radius * ang) ** 2


def trajectory_loss(lon, lat, lon_t, lat_t):
    """Mean squared great-circle distance to the targets, in km^2."""
    return great_circle_sq(lon, lat, lon_t, lat_t, radius=SPHERE_RADIUS_M / 1e3).mean()


def interact_final(lon, lat, species, radius, params, key, step, bbox, norm=2.0):
    """The rock-paper-scissors update on the positions after an advection,
    forward only (the discrete update has no gradient): the pair search on
    the float32-rounded positions, as the interaction pass reads the stored
    float32 record (interaction_simulator.py:88-99), then the pairs in
    sorted order through the rule of interactions.py:13-40 with Philox
    uniforms keyed by (key, step).  lon/lat: float64 GPU tensors (detached
    from any graph); species: int8 GPU tensor, updated in place.  Returns the
    pair count."""
    from .engine import InteractionEngine
    n = lon.numel()
    require(lat, torch.float64, numel=n, name="lat")
    require(species, torch.int8, numel=n, name="species")
    with torch.no_grad():
        rx = lon.detach().to(torch.float32).reshape(1, n).contiguous()
        ry = lat.detach().to(torch.float32).reshape(1, n).contiguous()
        eng = InteractionEngine(n, bbox, radius, norm, lon.device, steps_max=1)
        cnt = eng.run_chunk(rx, ry, species, params, key, step)
        return int(cnt[0].item())
