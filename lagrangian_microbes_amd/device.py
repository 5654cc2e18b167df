"""Device-side operators over torch tensors, one per liblm entry point.

These are thin: validate on the host, hand raw pointers and the current
stream to liblm, and return.  Nothing here synchronises with the host except
where a data-dependent size must come back (PairSearch.query_sync).
"""
import atexit
import collections
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import LmGrid, c_vp, check, ptr, require, stream_ptr

# Set once the interpreter starts to exit (atexit runs before module globals
# are cleared); workspace destructors skip freeing device memory from then on.
_EXITING = [False]
atexit.register(_EXITING.__setitem__, 0, True)


def lib():
    return _lib.load()


# ---------------------------------------------------------------------------
# velocity field on the device
# ---------------------------------------------------------------------------
class DeviceField:
    """A VelocityField's levels on the GPU as interleaved (u, v), float64 or
    (precision="f32") float32.

    ``stage_fields(t0_rel, dt, nsteps)`` returns the 2*nsteps+1 node-wise
    time-blended fields a run of RK4 steps samples (t0, t0+dt/2, ..., t0+n dt),
    or the single level of a steady field, with the stride to pass to the
    advection; ``advect(...)`` blends and advects with the kernel of the
    field's precision.  ``advect_steps(clock, ...)`` does the same for steps
    clock .. clock + nsteps - 1 of dt from the field's first level, from a
    schedule of every stage time of the field's span kept on the device
    (nothing uploaded per call).
    """

    def __init__(self, field, device=None, precision="f64"):
        if precision not in ("f64", "f32"):
            raise ValueError(f"precision must be 'f64' or 'f32' (got {precision!r})")
        self.field = field
        self.precision = precision
        self.dtype = torch.float64 if precision == "f64" else torch.float32
        self.device = torch.device(device or "cuda")
        inter = field.interleaved()  # [nt, ny, nx, 2] float64
        self.levels = torch.from_numpy(np.ascontiguousarray(inter)).to(self.device, self.dtype)
        self.level_len = int(np.prod(inter.shape[1:]))
        self.grid = LmGrid(field.nx, field.ny, field.x0, field.y0, 1.0 / field.dx, 1.0 / field.dy)
        self._stage_buf = None
        self._span = None  # (dt, lev0 host, lev0 device, w device): the stage schedule of the whole span

    @property
    def steady(self):
        return self.field.nt == 1

    def stage_fields(self, t0_rel, dt, nsteps):
        if self.steady:
            return self.levels[0], 0
        lev0, w = self.field.stage_schedule(t0_rel, dt, nsteps)
        ns = lev0.size
        need = ns * self.level_len
        if self._stage_buf is None or self._stage_buf.numel() < need:
            self._stage_buf = torch.empty(need, dtype=self.dtype, device=self.device)
        out = self._stage_buf[:need]
        # pinned, non-blocking uploads: a synchronous one would hold the host
        # until this stream drains, and with it a caller's launch loop
        lev0_d = torch.from_numpy(lev0.astype(np.int32)).pin_memory().to(self.device, non_blocking=True)
        w_d = torch.from_numpy(w.astype(np.float64)).pin_memory().to(self.device, non_blocking=True)
        blend = field_blend_f64 if self.precision == "f64" else field_blend_f32
        blend(self.levels, self.level_len, lev0_d, w_d, out, lev0_host=lev0)
        return out, self.level_len

    def advect(self, lon, lat, t0_rel, dt, nsteps, rec_lon=None, rec_lat=None, rec_stride=None, walk_amp=0.0,
               key=(0, 0), step0=0, oob=None):
        """nsteps RK4 steps from t0_rel (seconds on the field's time axis)."""
        stage, stride = self.stage_fields(t0_rel, dt, nsteps)
        run = advect_rk4_f64 if self.precision == "f64" else advect_rk4_f32
        return run(lon, lat, stage, stride, self.grid, nsteps, dt, rec_lon, rec_lat, rec_stride, walk_amp, key,
                   step0, oob)

    SPAN_STAGES_MAX = 1 << 22  # largest span schedule kept on the device (48 MiB with its host copy)

    def _span_schedule(self, dt):
        """Stage schedule of every half step of dt in the field's span (stage q
        at q dt/2 from the first level), on the host and the device, built once
        per dt; None when it would exceed SPAN_STAGES_MAX stages."""
        if self._span is None or self._span[0] != dt:
            if self._span is not None:  # launches on any stream may still read the old schedule
                torch.cuda.synchronize(self.device)
            span = float(self.field.times[-1] - self.field.times[0])
            nsteps = int(span // dt)
            if 2 * nsteps + 1 > self.SPAN_STAGES_MAX:
                self._span = (dt, None)
            else:
                lev0, w = self.field.stage_schedule(0.0, dt, nsteps)
                lev0 = np.ascontiguousarray(lev0, dtype=np.int32)
                self._span = (dt, (lev0, torch.from_numpy(lev0).to(self.device),
                                   torch.from_numpy(np.ascontiguousarray(w, dtype=np.float64)).to(self.device)))
        return self._span[1]

    def advect_steps(self, lon, lat, clock, dt, nsteps, rec_lon=None, rec_lat=None, rec_stride=None,
                     walk_amp=0.0, key=(0, 0), step0=0, oob=None):
        """advect() over steps clock .. clock + nsteps - 1 (t0 = clock * dt on
        the field's axis): the stage times are whole half steps of dt, so their
        schedule is a slice of the span's, already on the device.  The slice is
        taken when 2 dt is a whole number of seconds: then clock * dt + k dt/2
        and (2 clock + k) dt/2 are both exact, and the weights are bit for bit
        those advect() computes; other steps, and times past the span (which
        raise as stage_schedule does), go through advect()."""
        q0, q1 = 2 * int(clock), 2 * (int(clock) + int(nsteps)) + 1
        sched = None
        if not self.steady and dt > 0 and float(2.0 * dt).is_integer() and q1 * dt < 2.0 ** 50:
            sched = self._span_schedule(dt)
        if sched is None or q0 < 0 or q1 > sched[0].size:
            return self.advect(lon, lat, clock * dt, dt, nsteps, rec_lon, rec_lat, rec_stride, walk_amp, key,
                               step0, oob)
        lev0_h, lev0_d, w_d = sched
        need = (q1 - q0) * self.level_len
        if self._stage_buf is None or self._stage_buf.numel() < need:
            self._stage_buf = torch.empty(need, dtype=self.dtype, device=self.device)
        out = self._stage_buf[:need]
        blend = field_blend_f64 if self.precision == "f64" else field_blend_f32
        blend(self.levels, self.level_len, lev0_d[q0:q1], w_d[q0:q1], out, lev0_host=lev0_h[q0:q1])
        run = advect_rk4_f64 if self.precision == "f64" else advect_rk4_f32
        return run(lon, lat, out, self.level_len, self.grid, nsteps, dt, rec_lon, rec_lat, rec_stride, walk_amp,
                   key, step0, oob)


def field_blend_f64(levels, level_len, lev0, w, out, lev0_host=None):
    """Node-wise time blend (lm_field_blend_f64); lev0_host, if given, is the
    host copy of lev0 used for the range check (saves a device read-back)."""
    _field_blend(levels, level_len, lev0, w, out, lev0_host, torch.float64, lib().lm_field_blend_f64)


def field_blend_f32(levels, level_len, lev0, w, out, lev0_host=None):
    """The float32 blend of the FP32 path (lm_field_blend_f32)."""
    _field_blend(levels, level_len, lev0, w, out, lev0_host, torch.float32, lib().lm_field_blend_f32)


def _field_blend(levels, level_len, lev0, w, out, lev0_host, dtype, entry):
    require(levels, dtype, name="levels")
    require(lev0, torch.int32, name="lev0")
    require(w, torch.float64, numel=lev0.numel(), name="w")
    nstage = lev0.numel()
    require(out, dtype, min_numel=nstage * level_len, name="out")
    esize = out.element_size()
    nt = levels.numel() // level_len
    l0 = np.asarray(lev0_host) if lev0_host is not None else lev0.cpu().numpy()
    if l0.size != nstage:
        raise ValueError("lev0_host does not match lev0")
    if nstage and (l0.min() < 0 or l0.max() + 1 >= nt):
        raise ValueError("lev0 out of range of the field's time levels")
    for a in range(0, nstage, 65535):
        b = min(nstage, a + 65535)
        check(entry(ptr(levels), level_len, c_vp(lev0.data_ptr() + 4 * a), c_vp(w.data_ptr() + 8 * a), b - a,
                    c_vp(out.data_ptr() + esize * a * level_len), stream_ptr()))


def advect_rk4_f64(lon, lat, stage_fields, stage_stride, grid, nsteps, dt, rec_lon=None, rec_lat=None,
                   rec_stride=None, walk_amp=0.0, key=(0, 0), step0=0, oob=None):
    """RK4 steps on device lon/lat (float64, in place); see lm_advect_rk4_f64."""
    return _advect(lon, lat, stage_fields, stage_stride, grid, nsteps, dt, rec_lon, rec_lat, rec_stride, walk_amp,
                   key, step0, oob, torch.float64, lib().lm_advect_rk4_f64)


def advect_rk4_f32(lon, lat, stage_fields, stage_stride, grid, nsteps, dt, rec_lon=None, rec_lat=None,
                   rec_stride=None, walk_amp=0.0, key=(0, 0), step0=0, oob=None):
    """The FP32 path (lm_advect_rk4_f32): float32 stage fields, float64 positions."""
    return _advect(lon, lat, stage_fields, stage_stride, grid, nsteps, dt, rec_lon, rec_lat, rec_stride, walk_amp,
                   key, step0, oob, torch.float32, lib().lm_advect_rk4_f32)


def _advect(lon, lat, stage_fields, stage_stride, grid, nsteps, dt, rec_lon, rec_lat, rec_stride, walk_amp, key,
            step0, oob, dtype, entry):
    n = lon.numel()
    require(lon, torch.float64, name="lon")
    require(lat, torch.float64, numel=n, name="lat")
    require(stage_fields, dtype, name="stage_fields")
    flen = 2 * grid.nx * grid.ny
    if stage_stride == 0:
        if stage_fields.numel() < flen:
            raise ValueError("stage field smaller than the grid")
    elif stage_fields.numel() < (2 * nsteps) * stage_stride + flen:
        raise ValueError("not enough stage fields for nsteps")
    if oob is None:
        oob = torch.zeros(1, dtype=torch.int32, device=lon.device)
    require(oob, torch.int32, min_numel=1, name="oob")
    if rec_lon is not None:
        rec_stride = n if rec_stride is None else rec_stride
        require(rec_lon, torch.float32, name="rec_lon", min_numel=(nsteps - 1) * rec_stride + n)
        require(rec_lat, torch.float32, name="rec_lat", min_numel=(nsteps - 1) * rec_stride + n)
    else:
        rec_stride = 0
    check(entry(n, ptr(lon), ptr(lat), ptr(stage_fields), int(stage_stride), ctypes.byref(grid), int(nsteps),
                float(dt), ptr(rec_lon), ptr(rec_lat), int(rec_stride), float(walk_amp), int(key[0]), int(key[1]),
                int(step0), ptr(oob), stream_ptr()))
    return oob


def philox_uniforms(count, counter, key, word=0, device=None):
    """word `word` of Philox4x64-10((c0 + k, c1, c2, c3), key) as doubles in [0,1)."""
    out = torch.empty(int(count), dtype=torch.float64, device=device or "cuda")
    c0, c1, c2, c3 = (int(c) for c in counter)
    check(lib().lm_philox_uniforms(int(count), c0, c1, c2, c3, int(key[0]), int(key[1]), int(word),
                                   ptr(out), stream_ptr()))
    return out


def copy_bytes(src, dst, nbytes):
    check(lib().lm_copy_bytes(ptr(src), ptr(dst), int(nbytes), stream_ptr()))


def sort_pairs_u32(keys, vals, key_bits=32, seg_len=None):
    """Stable sort of each row of int32 ``keys`` [nseg, L] (as uint32) with
    ``vals``, in place; row g sorts its first ``seg_len[g]`` entries."""
    if not (keys.dtype == torch.int32 and vals.dtype == torch.int32 and keys.shape == vals.shape
            and keys.is_cuda and vals.is_cuda and keys.is_contiguous() and vals.is_contiguous()):
        raise ValueError("keys and vals must be contiguous int32 device tensors of one shape")
    k2 = keys if keys.dim() == 2 else keys.view(1, -1)
    v2 = vals if vals.dim() == 2 else vals.view(1, -1)
    nseg, L = k2.shape
    kt, vt = torch.empty_like(k2), torch.empty_like(v2)
    sl = None
    if seg_len is not None:
        sl = seg_len.to(device=k2.device, dtype=torch.int64).contiguous()
    check(lib().lm_sort_pairs_u32(nseg, ptr(sl) if sl is not None else None, L, L, L, int(key_bits),
                                  ptr(k2), ptr(v2), ptr(kt), ptr(vt), stream_ptr()))
    return keys, vals


# ---------------------------------------------------------------------------
# pair search
# ---------------------------------------------------------------------------
class PairSearch:
    """Fixed-radius pair search (replaces cKDTree(...).query_pairs).

    ``bbox`` = (lon_min, lon_max, lat_min, lat_max) where the particles are
    expected; particles outside are still handled exactly.
    """

    def __init__(self, n_max, bbox, max_radius, device=None, steps_max=1, cell_scale=1.0):
        self.n_max = int(n_max)
        self.steps_max = int(steps_max)
        self.max_radius = float(max_radius)
        self.cell_scale = float(cell_scale)
        self.bbox = tuple(float(b) for b in bbox)
        h = c_vp()
        check(lib().lm_pairs_ws_create_ex(self.n_max, self.steps_max, self.bbox[0], self.bbox[1], self.bbox[2],
                                          self.bbox[3], self.max_radius, self.cell_scale, ctypes.byref(h)))
        self._destroy = lib().lm_pairs_ws_destroy  # kept: module globals may be gone when __del__ runs at exit
        self._exiting = _EXITING
        self._ws = h
        self.device = torch.device(device or "cuda")
        self.pair_i = None
        self.pair_j = None
        self.npairs = torch.zeros(1, dtype=torch.int64, device=self.device)

    def __del__(self):
        ws, destroy = getattr(self, "_ws", None), getattr(self, "_destroy", None)
        if ws and destroy is not None and not self._exiting[0]:  # at exit the process frees everything
            destroy(ws)
            self._ws = None

    @property
    def workspace_bytes(self):
        return int(lib().lm_pairs_ws_bytes(self._ws))

    @property
    def buckets(self):
        """Number of cell buckets of the cell list."""
        return int(lib().lm_pairs_ws_buckets(self._ws))

    def query(self, lon, lat, radius, p, pair_i, pair_j, npairs):
        """Asynchronous: pairs into caller buffers (capacity = their length)."""
        n = lon.numel()
        require(lon, torch.float32, name="lon")
        require(lat, torch.float32, numel=n, name="lat")
        require(pair_i, torch.int32, name="pair_i")
        require(pair_j, torch.int32, numel=pair_i.numel(), name="pair_j")
        require(npairs, torch.int64, min_numel=1, name="npairs")
        if n > self.n_max:
            raise ValueError(f"{n} particles exceed the workspace's {self.n_max}")
        check(lib().lm_query_pairs(self._ws, n, ptr(lon), ptr(lat), float(radius), float(p), ptr(pair_i),
                                   ptr(pair_j), pair_i.numel(), ptr(npairs), stream_ptr()))

    def query_batch(self, lon, lat, radius, p, pair_i, pair_j, npairs, runs=None):
        """Asynchronous, one snapshot per row: lon/lat [S, n] float32, pairs
        [S, cap] int32 (capacity cap per step), npairs [S] int64.  With
        ``runs`` (EnsembleRuns) only pairs within a run are returned."""
        if lon.dim() != 2:
            raise ValueError("lon/lat must be [steps, n]")
        S, n = lon.shape
        require(lon, torch.float32, name="lon")
        require(lat, torch.float32, numel=S * n, name="lat")
        if pair_i.dim() != 2 or pair_i.shape[0] < S:
            raise ValueError("pair buffers must be [steps, capacity]")
        require(pair_i, torch.int32, name="pair_i")
        require(pair_j, torch.int32, numel=pair_i.numel(), name="pair_j")
        require(npairs, torch.int64, min_numel=S, name="npairs")
        if n > self.n_max or S > self.steps_max:
            raise ValueError(f"{S} x {n} exceeds the workspace's {self.steps_max} x {self.n_max}")
        if runs is not None:
            runs.check(n)
            check(lib().lm_query_pairs_runs(self._ws, S, n, ptr(lon), ptr(lat), n, ptr(runs.run), float(radius),
                                            float(p), ptr(pair_i), ptr(pair_j), pair_i.shape[1], ptr(npairs),
                                            stream_ptr()))
            return
        check(lib().lm_query_pairs_batch(self._ws, S, n, ptr(lon), ptr(lat), n, float(radius), float(p),
                                         ptr(pair_i), ptr(pair_j), pair_i.shape[1], ptr(npairs), stream_ptr()))

    def ensure_capacity(self, cap):
        if self.pair_i is None or self.pair_i.numel() < cap:
            cap = max(int(cap), 1024)
            self.pair_i = torch.empty(cap, dtype=torch.int32, device=self.device)
            self.pair_j = torch.empty(cap, dtype=torch.int32, device=self.device)

    def query_sync(self, lon, lat, radius, p=2.0):
        """Pairs as exact-length device tensors (reads the count back)."""
        n = lon.numel()
        self.ensure_capacity(max(2 * n, 1024))
        for _ in range(2):
            self.query(lon, lat, radius, p, self.pair_i, self.pair_j, self.npairs)
            P = int(self.npairs.item())
            if P <= self.pair_i.numel():
                return self.pair_i[:P], self.pair_j[:P]
            self.ensure_capacity(P + P // 4 + 1024)
        raise RuntimeError("pair buffer growth did not converge")


# ---------------------------------------------------------------------------
# ensembles
# ---------------------------------------------------------------------------
RUN_PARAMS = np.dtype([("key0", "<u8"), ("key1", "<u8"), ("pRS", "<f8"), ("pPR", "<f8"), ("pSP", "<f8")])


class EnsembleRuns:
    """Independent runs sharing one index range: run r holds particles
    [start[r], start[r+1]) with its own pair probabilities and Philox key.

    sizes: particles per run; params: one {"pRS", "pPR", "pSP"} per run;
    keys: one (key0, key1) per run.  Device arrays: ``run`` (int16 per
    particle, read as uint16), ``run_start`` (int64, nruns + 1) and
    ``params`` (lm_run_params per run).
    """

    def __init__(self, sizes, params, keys, device=None):
        sizes = [int(v) for v in sizes]
        if not sizes or min(sizes) < 0 or len(sizes) > 32767:
            raise ValueError("1..32767 runs of non-negative size")
        if len(params) != len(sizes) or len(keys) != len(sizes):
            raise ValueError("one params dict and one key per run")
        self.device = torch.device(device or "cuda")
        self.sizes = sizes
        self.nruns = len(sizes)
        self.start = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        self.n = int(self.start[-1])
        self.run = torch.from_numpy(np.repeat(np.arange(self.nruns, dtype=np.int16), sizes)).to(self.device)
        self.run_start = torch.from_numpy(self.start).to(self.device)
        table = np.zeros(self.nruns, dtype=RUN_PARAMS)
        for r, (p, k) in enumerate(zip(params, keys)):
            table[r] = (int(k[0]), int(k[1]), float(p["pRS"]), float(p["pPR"]), float(p["pSP"]))
        self.params = torch.from_numpy(table.view(np.uint8).copy()).to(self.device)

    def check(self, n):
        if n != self.n:
            raise ValueError(f"{n} particles, but the runs hold {self.n}")

    def slice(self, r):
        return slice(int(self.start[r]), int(self.start[r + 1]))


# ---------------------------------------------------------------------------
# rock-paper-scissors
# ---------------------------------------------------------------------------
class RPSUpdater:
    """Rock-paper-scissors over a sorted device pair list (replaces the
    per-pair Python loop over rock_paper_scissors_interaction)."""

    def __init__(self, n_max, device=None, steps_max=1):
        self.n_max = int(n_max)
        self.steps_max = int(steps_max)
        h = c_vp()
        check(lib().lm_rps_ws_create(self.n_max, self.steps_max, ctypes.byref(h)))
        self._destroy = lib().lm_rps_ws_destroy  # kept: module globals may be gone when __del__ runs at exit
        self._exiting = _EXITING
        self._ws = h
        self.device = torch.device(device or "cuda")
        self._stats = torch.zeros(8, dtype=torch.int64, device=self.device)

    def __del__(self):
        ws, destroy = getattr(self, "_ws", None), getattr(self, "_destroy", None)
        if ws and destroy is not None and not self._exiting[0]:  # at exit the process frees everything
            destroy(ws)
            self._ws = None

    @property
    def workspace_bytes(self):
        return int(lib().lm_rps_ws_bytes(self._ws))

    def update(self, species, pair_i, pair_j, npairs, pRS, pPR, pSP, key, step):
        n = species.numel()
        require(species, torch.int8, name="species")
        require(pair_i, torch.int32, name="pair_i")
        require(pair_j, torch.int32, numel=pair_i.numel(), name="pair_j")
        require(npairs, torch.int64, min_numel=1, name="npairs")
        if n > self.n_max:
            raise ValueError(f"{n} particles exceed the workspace's {self.n_max}")
        check(lib().lm_rps_update(self._ws, n, ptr(species), ptr(pair_i), ptr(pair_j), ptr(npairs),
                                  pair_i.numel(), float(pRS), float(pPR), float(pSP), int(key[0]),
                                  int(key[1]), int(step), stream_ptr()))

    def prepare(self, n, pair_i, pair_j, npairs, pRS, pPR, pSP, key, step0, runs=None):
        """Group the pairs of S steps ([S, cap] buffers, counts npairs[S]).
        With ``runs`` (EnsembleRuns) each run uses its own key and p."""
        if pair_i.dim() != 2:
            raise ValueError("pair buffers must be [steps, capacity]")
        S, cap = pair_i.shape
        require(pair_i, torch.int32, name="pair_i")
        require(pair_j, torch.int32, numel=pair_i.numel(), name="pair_j")
        require(npairs, torch.int64, min_numel=S, name="npairs")
        if n > self.n_max or S > self.steps_max:
            raise ValueError(f"{S} x {n} exceeds the workspace's {self.steps_max} x {self.n_max}")
        if runs is not None:
            runs.check(n)
            check(lib().lm_rps_prepare_runs(self._ws, S, int(n), ptr(pair_i), ptr(pair_j), cap, ptr(npairs),
                                            ptr(runs.run), runs.nruns, ptr(runs.run_start), ptr(runs.params),
                                            int(step0), stream_ptr()))
            return
        check(lib().lm_rps_prepare(self._ws, S, int(n), ptr(pair_i), ptr(pair_j), cap, ptr(npairs), float(pRS),
                                   float(pPR), float(pSP), int(key[0]), int(key[1]), int(step0), stream_ptr()))

    def set_play_blocks(self, wide, narrow):
        """Launch shape of the replay (see lm_rps_ws_set_play_blocks)."""
        check(lib().lm_rps_ws_set_play_blocks(self._ws, int(wide), int(narrow)))

    def reserve(self, pairs_per_step):
        """Size the per-pair arrays for pairs_per_step pairs per step now, so
        that a prepare at that capacity does not stop the stream to grow them."""
        check(lib().lm_rps_ws_reserve(self._ws, int(pairs_per_step), stream_ptr()))

    def play(self, step, species_in, species_out):
        """Step `step` of the last prepare: species_out = species_in after its pairs."""
        n = species_in.numel()
        require(species_in, torch.int8, name="species_in")
        require(species_out, torch.int8, numel=n, name="species_out")
        check(lib().lm_rps_play(self._ws, int(step), n, ptr(species_in), ptr(species_out), stream_ptr()))

    def component_stats(self, step=0):
        """Component sizes of step ``step`` of the last prepare (a ComponentStats)."""
        check(lib().lm_rps_component_stats(self._ws, int(step), ptr(self._stats), stream_ptr()))
        return ComponentStats(*(int(v) for v in self._stats.cpu().tolist()))


# ---------------------------------------------------------------------------
# per-kernel timing
# ---------------------------------------------------------------------------
ComponentStats = collections.namedtuple(
    "ComponentStats", "components max_pairs max_particles pairs wide_components wide_pairs wide_particles touched")


class KernelTimer:
    """Context manager: every liblm launch inside is timed with events on its
    own dispatch; ``summary()`` returns {kernel: (launches, total_ms)}."""

    active = False  # launches carry timing events: no graph capture meanwhile

    def __enter__(self):
        check(lib().lm_timing_enable(1))
        KernelTimer.active = True
        return self

    def __exit__(self, *exc):
        self._summary = self._read()
        check(lib().lm_timing_enable(0))
        KernelTimer.active = False
        return False

    @staticmethod
    def _read():
        nk = 64
        counts = np.zeros(nk, dtype=np.int64)
        ms = np.zeros(nk, dtype=np.float64)
        got = lib().lm_timing_summary(counts.ctypes.data_as(c_vp), ms.ctypes.data_as(c_vp), nk)
        if got < 0:
            check(got)
        out = {}
        for k in range(min(got, nk)):
            if counts[k]:
                out[lib().lm_timing_kernel_name(k).decode()] = (int(counts[k]), float(ms[k]))
        return out

    def summary(self):
        return getattr(self, "_summary", None) or self._read()
