"""ORACLE -- test infrastructure, not product code.

CPU restatements of the lagrangian-microbes hot path used as the checker by
tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  The
product package (lagrangian_microbes_amd) never imports this package.

Contents and what pins each function (the checks live in tests/):

- ``advect_rk4_f64`` / ``field_blend_f64`` (C, lm_oracle.c): RK4 advection as
  parcels.AdvectionRK4 with linear interpolation (call sites
  particle_advecter.py:177-187,222-223).  Pinned by closed-form flows
  (uniform, meridional shear, zonal flow along a parallel), by the independent
  numpy/scipy restatement in :mod:`oracle.advect_ref` (scipy
  RegularGridInterpolator + generic RK4 tableau) and by the cosine polynomial
  against numpy.cos.  parcels itself (2.0.0beta2, environment.yml:80) is not
  installed, so agreement with parcels' own float32 arithmetic is parity
  unpinned.
- ``philox4x64_10`` / ``philox_uniforms`` (C) and :mod:`oracle.philox`
  (numpy): pinned against numpy.random.Philox and the Random123 known-answer
  vector.
- ``rps_sequential`` (C) and :func:`oracle.rps.rps_sequential_py` (Python):
  interactions.py:13-40 applied in sorted pair order
  (interaction_simulator.py:104-105).  Pinned by an independent cyclic-dominance
  formulation over every ordered species pair and both branches, p = 0 / p = 1
  limits and hand-worked order-dependent chains.
- ``pairs_bruteforce`` (C) and :mod:`oracle.pairs` (numpy): query_pairs
  semantics (interaction_simulator.py:93,98).  Pinned against
  scipy.spatial.cKDTree.query_pairs, the library the reference calls.
- ``advect_rk4_f32`` / ``field_blend_f32`` / ``cos_deg_f32`` (C): the FP32
  path's operations (float32 field, interpolation and velocities; float64
  positions), the GPU kernel's restatement.  Pinned by the cosine polynomial
  against numpy.cos in float32 rounding, closed-form flows, and its distance
  from ``advect_rk4_f64`` on the same inputs (float32 interpolation error
  only).
- ``distance_histogram`` / ``sin_deg`` (C): the pairwise haversine term and
  its binning against thresholds, as the device reduction computes them.
  Pinned by an independent numpy histogram of round(10 log10(max(1, d)))
  with numpy's sin/cos/arcsin (equal away from bin edges), the sine
  polynomial against numpy.sin, and the total n (n - 1) / 2.
- :mod:`oracle.adjoint_ref` (torch float64 on the CPU): the RK4 advection
  written with differentiable torch operations, whose autograd gradients the
  hand-written adjoint is checked against.  Pinned by its positions against
  ``advect_rk4_f64``, by central finite differences of its own forward, and by
  the closed forms of a uniform zonal flow.
"""
import ctypes
import os

import numpy as np

from . import build as _build

_lib = None


def lib():
    """Load (building first if needed) the oracle's C library."""
    global _lib
    if _lib is None:
        if not os.path.exists(_build.LIB):
            _build.build()
        L = ctypes.CDLL(_build.LIB)
        i64, i32, u64, dbl = ctypes.c_int64, ctypes.c_int32, ctypes.c_uint64, ctypes.c_double
        vp = ctypes.c_void_p
        L.oracle_philox4x64_10.argtypes = [vp, vp, vp]
        L.oracle_philox4x64_10.restype = None
        L.oracle_philox_uniforms.argtypes = [i64, u64, u64, u64, u64, u64, u64, ctypes.c_int, vp]
        L.oracle_philox_uniforms.restype = None
        L.oracle_cos_deg.argtypes = [dbl]
        L.oracle_cos_deg.restype = dbl
        L.oracle_field_blend_f64.argtypes = [vp, i64, vp, vp, i32, vp]
        L.oracle_field_blend_f64.restype = None
        L.oracle_advect_rk4_f64.argtypes = [i64, vp, vp, vp, i64, i32, i32, dbl, dbl, dbl, dbl, i32, dbl,
                                            vp, vp, i64, dbl, u64, u64, u64]
        L.oracle_advect_rk4_f64.restype = i64
        L.oracle_cos_deg_f32.argtypes = [ctypes.c_float]
        L.oracle_cos_deg_f32.restype = ctypes.c_float
        L.oracle_field_blend_f32.argtypes = [vp, i64, vp, vp, i32, vp]
        L.oracle_field_blend_f32.restype = None
        L.oracle_advect_rk4_f32.argtypes = [i64, vp, vp, vp, i64, i32, i32, dbl, dbl, dbl, dbl, i32, dbl,
                                            vp, vp, i64, dbl, u64, u64, u64]
        L.oracle_advect_rk4_f32.restype = i64
        L.oracle_sin_deg.argtypes = [dbl]
        L.oracle_sin_deg.restype = dbl
        L.oracle_distance_histogram.argtypes = [i64, vp, vp, i32, vp, vp]
        L.oracle_distance_histogram.restype = None
        L.oracle_rps_sequential.argtypes = [i64, vp, vp, vp, i64, dbl, dbl, dbl, u64, u64, u64]
        L.oracle_rps_sequential.restype = None
        L.oracle_pairs_bruteforce.argtypes = [i64, vp, vp, dbl, dbl, vp, vp, i64]
        L.oracle_pairs_bruteforce.restype = i64
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def philox4x64_10(counter, key):
    """Raw Philox4x64-10 block: 4 uint64 counter words, 2 key words -> 4 words."""
    c = np.ascontiguousarray(np.asarray(counter, dtype=np.uint64).reshape(4))
    k = np.ascontiguousarray(np.asarray(key, dtype=np.uint64).reshape(2))
    out = np.zeros(4, dtype=np.uint64)
    lib().oracle_philox4x64_10(_ptr(c), _ptr(k), _ptr(out))
    return out


def philox_uniforms(count, c0, c1, c2, c3, key0, key1, word=0):
    out = np.empty(int(count), dtype=np.float64)
    lib().oracle_philox_uniforms(int(count), int(c0), int(c1), int(c2), int(c3), int(key0), int(key1),
                                 int(word), _ptr(out))
    return out


def cos_deg(lat):
    return lib().oracle_cos_deg(float(lat))


def field_blend_f64(levels, lev0, w):
    """levels: float64 [nt, ny, nx, 2]; returns float64 [nstage, ny, nx, 2]."""
    levels = np.ascontiguousarray(levels, dtype=np.float64)
    lev0 = np.ascontiguousarray(lev0, dtype=np.int32)
    w = np.ascontiguousarray(w, dtype=np.float64)
    ln = int(np.prod(levels.shape[1:]))
    out = np.empty((len(lev0),) + levels.shape[1:], dtype=np.float64)
    lib().oracle_field_blend_f64(_ptr(levels), ln, _ptr(lev0), _ptr(w), len(lev0), _ptr(out))
    return out


def advect_rk4_f64(lon, lat, stage_fields, grid, nsteps, dt, steady=False, record=True,
                   walk_amp=0.0, key=(0, 0), step0=0):
    """RK4 on copies of lon/lat.

    stage_fields: float64 [2*nsteps+1, ny, nx, 2] (or [1, ny, nx, 2] with
    steady=True).  grid: (nx, ny, x0, y0, inv_dx, inv_dy).
    Returns (lon, lat, rec_lon [nsteps, n] or None, rec_lat, oob_count).
    """
    lon = np.array(lon, dtype=np.float64, copy=True)
    lat = np.array(lat, dtype=np.float64, copy=True)
    n = lon.size
    f = np.ascontiguousarray(stage_fields, dtype=np.float64)
    stride = 0 if steady else int(np.prod(f.shape[1:]))
    nx, ny, x0, y0, idx, idy = grid
    rl = np.zeros((nsteps, n), dtype=np.float32) if record else None
    rt = np.zeros((nsteps, n), dtype=np.float32) if record else None
    oob = lib().oracle_advect_rk4_f64(
        n, _ptr(lon), _ptr(lat), _ptr(f), stride, int(nx), int(ny), float(x0), float(y0), float(idx),
        float(idy), int(nsteps), float(dt), _ptr(rl) if record else None, _ptr(rt) if record else None,
        n, float(walk_amp), int(key[0]), int(key[1]), int(step0))
    return lon, lat, rl, rt, int(oob)


def cos_deg_f32(lat):
    return lib().oracle_cos_deg_f32(float(lat))


def field_blend_f32(levels, lev0, w):
    """levels: float32 [nt, ny, nx, 2]; returns float32 [nstage, ny, nx, 2]."""
    levels = np.ascontiguousarray(levels, dtype=np.float32)
    lev0 = np.ascontiguousarray(lev0, dtype=np.int32)
    w = np.ascontiguousarray(w, dtype=np.float64)
    ln = int(np.prod(levels.shape[1:]))
    out = np.empty((len(lev0),) + levels.shape[1:], dtype=np.float32)
    lib().oracle_field_blend_f32(_ptr(levels), ln, _ptr(lev0), _ptr(w), len(lev0), _ptr(out))
    return out


def advect_rk4_f32(lon, lat, stage_fields, grid, nsteps, dt, steady=False, record=True, walk_amp=0.0, key=(0, 0),
                   step0=0):
    """The FP32 path's RK4 on copies of lon/lat (float64); stage_fields float32
    [2*nsteps+1, ny, nx, 2] (or [1, ...] with steady=True).  Returns as
    advect_rk4_f64."""
    lon = np.array(lon, dtype=np.float64, copy=True)
    lat = np.array(lat, dtype=np.float64, copy=True)
    n = lon.size
    f = np.ascontiguousarray(stage_fields, dtype=np.float32)
    stride = 0 if steady else int(np.prod(f.shape[1:]))
    nx, ny, x0, y0, idx, idy = grid
    rl = np.zeros((nsteps, n), dtype=np.float32) if record else None
    rt = np.zeros((nsteps, n), dtype=np.float32) if record else None
    oob = lib().oracle_advect_rk4_f32(
        n, _ptr(lon), _ptr(lat), _ptr(f), stride, int(nx), int(ny), float(x0), float(y0), float(idx),
        float(idy), int(nsteps), float(dt), _ptr(rl) if record else None, _ptr(rt) if record else None,
        n, float(walk_amp), int(key[0]), int(key[1]), int(step0))
    return lon, lat, rl, rt, int(oob)


def sin_deg(x):
    return lib().oracle_sin_deg(float(x))


def distance_histogram(lon, lat, edges):
    """Counts of pairs i < j by the number of thresholds <= their haversine
    term (float32 positions).  O(n^2)."""
    lon = np.ascontiguousarray(lon, dtype=np.float32)
    lat = np.ascontiguousarray(lat, dtype=np.float32)
    edges = np.ascontiguousarray(edges, dtype=np.float64)
    hist = np.zeros(edges.size + 1, dtype=np.uint64)
    lib().oracle_distance_histogram(lon.size, _ptr(lon), _ptr(lat), edges.size + 1, _ptr(edges), _ptr(hist))
    return hist


def rps_sequential(species, pair_i, pair_j, pRS, pPR, pSP, key, step):
    """Sequential RPS over the sorted pair list; returns the updated copy."""
    sp = np.array(species, dtype=np.int8, copy=True)
    pi = np.ascontiguousarray(pair_i, dtype=np.int32)
    pj = np.ascontiguousarray(pair_j, dtype=np.int32)
    lib().oracle_rps_sequential(sp.size, _ptr(sp), _ptr(pi), _ptr(pj), pi.size, float(pRS), float(pPR),
                                float(pSP), int(key[0]), int(key[1]), int(step))
    return sp


def pairs_bruteforce(x, y, r, p=2.0):
    """All (i, j), i < j, within r of each other; float32 positions.  O(n^2)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    n = x.size
    cnt = lib().oracle_pairs_bruteforce(n, _ptr(x), _ptr(y), float(r), float(p), None, None, 0)
    pi = np.zeros(cnt, dtype=np.int32)
    pj = np.zeros(cnt, dtype=np.int32)
    lib().oracle_pairs_bruteforce(n, _ptr(x), _ptr(y), float(r), float(p), _ptr(pi), _ptr(pj), cnt)
    return pi, pj
