"""Fused advection + interactions on the GPU (no intermediate files).

The reference runs two passes joined through particle_data.nc: advection
(ParticleAdvecter.time_step, particle_advecter.py:123-260) then interactions
(InteractionSimulator.time_step, interaction_simulator.py:54-122), where step i
of the interactions uses the float32 positions recorded after advection step
i.  ``FusedRunner`` does both per output chunk on the device: one RK4 launch
advances every particle through the chunk and records float32 positions,
then each step's pair search and rock-paper-scissors update read those
records.  The species and pair lists are therefore identical to the two-pass
result on the same inputs.  ``MicrobeSimulator`` wraps it with the outputs
of both passes (particle_data.nc and microbe_data.nc).
"""
import logging
import os
from time import time

import numpy as np
import torch

from . import device as dev
from .engine import InteractionEngine, new_key, walk_amplitude
from .interactions import rock_paper_scissors_interaction
from .netcdf_io import ParticleDataset, copy_variables
from .records import DEFAULT_RING_BYTES, DeviceRecords, RecordStream, device_record_budget, write_device_records
from .utils import pretty_time
from .velocity_fields import VelocityField

logger = logging.getLogger(__name__)


# Smallest chunk the tapered end of a pipelined run uses.
TAPER_MIN = 8


def chunk_schedule(total, chunk, taper_min=TAPER_MIN):
    """Chunk lengths of a pipelined run of `total` steps: full chunks, then
    the last (up to) two chunks' worth of steps in halving chunks down to
    taper_min.  Each chunk's replay overlaps the next chunk's preparation, so
    only the last replay is exposed; a short last chunk keeps it short."""
    sizes, left = [], int(total)
    while left > 2 * chunk:
        sizes.append(chunk)
        left -= chunk
    while left > 0:
        k = min(chunk, left if left <= taper_min else max(taper_min, (left + 1) // 2))
        sizes.append(k)
        left -= k
    return sizes


class FusedRunner:
    """Device state and buffers for a fused run.

    lon0/lat0: float64 initial positions; species0: int8; field: VelocityField;
    params: {"pRS", "pPR", "pSP"}; key: Philox key of the pair uniforms;
    walk_key / Kh (deg^2/s): the random walk (off when Kh == 0);
    precision: "f64" (the FP64 path) or "f32" (float32 field, interpolation
    and velocities; float64 positions and pair test); runs: an
    EnsembleRuns splitting the particles into independent runs (then each
    run's own p and key apply and params/key are unused); cell_scale: the
    pair search's cell side over the radius; play_blocks: (wide, narrow)
    workgroups of the replay; graphs: replay each chunk's per-step launches
    from a captured HIP graph; pair_capacity: initial pair buffer length per
    step (default 2 n; buffers grow when a step finds more pairs).
    """

    PIPELINE_SETS = 3  # chunk buffer sets of run(): replay, preparation, advection

    def __init__(self, lon0, lat0, species0, field, radius, params, key, dt_s, norm=2.0, chunk=100, Kh=0.0,
                 walk_key=(0, 0x6B68), device=None, bbox=None, precision="f64", runs=None, cell_scale=1.0,
                 play_blocks=None, graphs=False, pair_capacity=None):
        self.device = torch.device(device or "cuda")
        self.n = int(np.asarray(lon0).size)
        self.lon0 = np.asarray(lon0, dtype=np.float64)
        self.lat0 = np.asarray(lat0, dtype=np.float64)
        self.species0 = np.asarray(species0, dtype=np.int8)
        self.params = dict(params)
        self.key = tuple(key)
        self.dt_s = float(dt_s)
        self.chunk = int(chunk)
        self.walk_amp = walk_amplitude(Kh, self.dt_s)
        self.walk_key = tuple(walk_key)
        self.dfield = dev.DeviceField(field, self.device, precision)
        self.engine = InteractionEngine(self.n, bbox or field.bounds, radius, norm, self.device,
                                        steps_max=self.chunk, nsets=self.PIPELINE_SETS, runs=runs,
                                        cell_scale=cell_scale, initial_capacity=pair_capacity,
                                        play_blocks=play_blocks, graphs=graphs)
        self.rec_x = torch.empty((self.chunk, self.n), dtype=torch.float32, device=self.device)
        self.rec_y = torch.empty((self.chunk, self.n), dtype=torch.float32, device=self.device)
        self.rec_s = torch.empty((self.chunk, self.n), dtype=torch.int8, device=self.device)
        self.npairs = torch.zeros(self.chunk, dtype=torch.int64, device=self.device)
        self.oob = torch.zeros(1, dtype=torch.int32, device=self.device)
        self.reset()

    def reset(self):
        """Back to the initial state, step 0 and the field's first time level."""
        self.lon = torch.from_numpy(self.lon0.copy()).to(self.device)
        self.lat = torch.from_numpy(self.lat0.copy()).to(self.device)
        self.species = torch.from_numpy(self.species0.copy()).to(self.device)
        self.step = 0          # global step index: keys the uniforms
        self.clock_steps = 0   # steps since the field's first time level
        self.oob.zero_()

    def advance(self, nsteps):
        """Run up to self.chunk steps; returns the number of steps done.

        After the call rec_x/rec_y/rec_s[:nsteps] hold the step records and
        npairs[:nsteps] the pair counts.  Nothing is synchronised except the
        chunk-end check of the pair counts inside the interaction engine.
        """
        k = min(int(nsteps), self.chunk)
        if k <= 0:
            return 0
        self.dfield.advect_steps(self.lon, self.lat, self.clock_steps, self.dt_s, k, self.rec_x, self.rec_y,
                                 self.n, self.walk_amp, self.walk_key, self.step, self.oob)
        self.engine.run_chunk(self.rec_x[:k], self.rec_y[:k], self.species, self.params, self.key, self.step,
                              rec_species=self.rec_s[:k], npairs=self.npairs[:k])
        self.step += k
        self.clock_steps += k
        return k

    def run(self, nsteps, on_chunk=None, records=None):
        """Run nsteps steps with consecutive chunks overlapped on the device.

        Three streams, three sets of chunk buffers (set c % 3 for chunk c):
        the advection of chunk c + 2 (FP64 arithmetic) runs on its own stream
        beside the pair search and grouping of chunk c + 1 (memory-latency
        bound) on a side stream, while chunk c replays the species on the
        caller's stream.  The advection of a chunk waits for the replay that
        last used its set; the grouping waits for its advection.  After each
        chunk, ``on_chunk(k, rec_x, rec_y, rec_s, npairs)`` receives that
        chunk's records (device views, valid until the callback returns; if
        it returns a torch.cuda.Event, e.g. after copies on another stream
        that read them, until that event).  With ``records`` (rec_x, rec_y,
        rec_s: device float32, float32, int8 arrays [>= nsteps, n]) step t of
        the call writes its records to row t of those instead of the chunk
        buffers, so a run whose records stay on the device copies nothing.
        The state and step counters end as after ``advance`` calls over the
        same steps; returns the number of steps done.
        """
        total = max(0, int(nsteps))
        if total == 0:
            return 0
        self._ensure_pipeline()
        eng = self.engine
        nsets = self.PIPELINE_SETS
        main = torch.cuda.current_stream(self.device)
        side, adv = self._side, self._adv
        sizes = chunk_schedule(total, self.chunk)
        if max(sizes) > self.chunk:  # the chunk buffers hold self.chunk steps
            raise RuntimeError(f"chunk schedule exceeds the buffers: {sizes}")
        self.last_schedule = list(sizes)  # chunk lengths of this call, one batched launch each
        offs = [sum(sizes[:c]) for c in range(len(sizes))]
        if records is not None:
            for t, dt in zip(records, (torch.float32, torch.float32, torch.int8)):
                on_device = t.device.type == self.device.type and self.device.index in (None, t.device.index)
                if (t.dtype != dt or t.dim() != 2 or t.shape[0] < total or t.shape[1] != self.n
                        or not t.is_contiguous() or not on_device):
                    raise ValueError(f"records must be contiguous {dt} [>= {total}, {self.n}] on {self.device}")

        def recs(c):  # chunk c's records: rows of the caller's arrays, or its buffer set
            k = sizes[c]
            if records is not None:
                return tuple(t[offs[c]:offs[c] + k] for t in records)
            return tuple(t[:k] for t in self._rec_sets[c % nsets])

        step0 = [self.step + o for o in offs]
        clock0 = [self.clock_steps + o for o in offs]
        advected = [None] * len(sizes)   # event after chunk c's advection
        prepared = [None] * len(sizes)   # event after chunk c's preparation
        played = [None] * nsets          # event after the replay that last used set b
        released = [None] * nsets        # event the callback returned for set b's last chunk
        # The events are the runner's own, recorded again every chunk: a set's
        # event is re-recorded only after every wait on its previous record
        # has been issued (chunk c + 3 reuses chunk c's set)
        ev_adv, ev_prep, ev_play, ev_count, ev_join = self._events
        ev_join[0].record(main)          # the state as the caller left it
        side.wait_event(ev_join[0])
        adv.wait_event(ev_join[0])

        def advect(c):
            b, k = c % nsets, sizes[c]
            with torch.cuda.stream(adv):
                if played[b] is not None:
                    adv.wait_event(played[b])
                if released[b] is not None:
                    adv.wait_event(released[b])
                rx, ry, _ = recs(c)
                self.dfield.advect_steps(self.lon, self.lat, clock0[c], self.dt_s, k, rx, ry, self.n,
                                         self.walk_amp, self.walk_key, step0[c], self.oob)
                ev = ev_adv[b]
                ev.record(adv)
                advected[c] = ev

        def group(c):
            b = c % nsets
            rx, ry, _ = recs(c)
            with torch.cuda.stream(side):
                side.wait_event(advected[c])
                eng.prepare_set(b, rx, ry, self.params, self.key, step0[c])
                ev = ev_prep[b]
                ev.record(side)
                prepared[c] = ev

        # chunk 0's preparation goes out before chunk 1's advection: it waits
        # for nothing else
        advect(0)
        group(0)
        if len(sizes) > 1:
            advect(1)
        backup, worst_host = self._backup, self._worst_host
        for c, k in enumerate(sizes):
            b = c % nsets
            if c + 1 < len(sizes):
                group(c + 1)
            if c + 2 < len(sizes):
                advect(c + 2)
            main.wait_event(prepared[c])
            rx, ry, rs = recs(c)
            # The replay goes out before the host has seen the chunk's pair
            # counts, so it starts as soon as the preparation ends; the counts
            # come back (a pinned copy) while it runs.  A step that overflowed
            # the pair buffers voids the replay: the chunk is then grouped
            # again with larger buffers (positions are kept; so is the next
            # prepared chunk, whose buffers are replaced) and replayed from its
            # starting species.
            backup.copy_(self.species)
            worst_host.copy_(eng.npairs_sets[b][:k].max(), non_blocking=True)
            counted = ev_count
            counted.record(main)
            eng.play_set(b, k, self.species, rs)
            counted.synchronize()
            worst = int(worst_host.item())
            while worst > eng.capacity:
                torch.cuda.synchronize(self.device)
                eng.fits(worst)
                group(c)
                if c + 1 < len(sizes):
                    group(c + 1)
                main.wait_event(prepared[c])
                self.species.copy_(backup)
                eng.play_set(b, k, self.species, rs)
                worst = eng.worst_count(b, k)
            ret = None
            if on_chunk is not None:
                ret = on_chunk(k, rx, ry, rs, eng.npairs_sets[b][:k])
            ev = ev_play[b]  # after the callback's own work on this stream
            ev.record(main)
            played[b] = ev
            released[b] = ret if isinstance(ret, torch.cuda.Event) else None
            self.step += k
            self.clock_steps += k
        ev_join[1].record(side)
        ev_join[2].record(adv)
        main.wait_event(ev_join[1])
        main.wait_event(ev_join[2])
        return total

    def _ensure_pipeline(self):
        if getattr(self, "_side", None) is None:
            extra = [tuple(torch.empty_like(t) for t in (self.rec_x, self.rec_y, self.rec_s))
                     for _ in range(self.PIPELINE_SETS - 1)]
            self._rec_sets = [(self.rec_x, self.rec_y, self.rec_s)] + extra
            self._side = torch.cuda.Stream(self.device)
            self._adv = torch.cuda.Stream(self.device)
            self._backup = torch.empty_like(self.species)  # a chunk's starting species
            self._worst_host = torch.zeros((), dtype=torch.int64, pin_memory=True)
            sets = range(self.PIPELINE_SETS)
            self._events = ([torch.cuda.Event() for _ in sets], [torch.cuda.Event() for _ in sets],
                            [torch.cuda.Event() for _ in sets], torch.cuda.Event(),
                            [torch.cuda.Event() for _ in range(3)])

    def check_domain(self):
        n_oob = int(self.oob.item())
        if n_oob:
            raise RuntimeError(f"{n_oob} particle-step(s) left the velocity field's domain "
                               f"{self.dfield.field.bounds}")


class MicrobeSimulator:
    """Advection and pair interactions of one population, fused on the GPU.

    Arguments follow ParticleAdvecter and InteractionSimulator; the
    interaction must be the rock-paper-scissors plugin tuple.  ``time_step``
    writes particle_data.nc and microbe_data.nc with the layout of the two
    reference passes (write_outputs=False skips the files and the host
    copies of the records) and leaves the final species in
    microbe_properties["species"].  A run whose records fit the device
    budget (``device_record_bytes``, see records.device_record_budget) keeps
    them in HBM and writes the files after the run (``DeviceRecords``,
    ``write_device_records``); a larger one moves them to the host chunk by
    chunk through a ``RecordStream`` (pinned ring of ``ring_bytes``, copy
    stream, writer thread) while the run goes on.  ``timings`` holds the
    device run's wall time and the output stages'.
    """

    def __init__(self, particle_lons, particle_lats, velocity_field, pair_interaction, interaction_radius,
                 interaction_norm=2, output_dir=".", output_chunk_iters=100, Kh=0, seed=None, philox_key=None,
                 device=None, ring_bytes=DEFAULT_RING_BYTES, device_record_bytes=None):
        assert isinstance(velocity_field, VelocityField), "velocity_field must be a VelocityField"
        assert particle_lons.size == particle_lats.size
        fn, params, props = pair_interaction
        if fn is not rock_paper_scissors_interaction:
            raise ValueError("the fused GPU path runs the rock-paper-scissors interaction only; use "
                             "ParticleAdvecter + InteractionSimulator for other callbacks")
        assert output_chunk_iters >= 1
        self.velocity_field = velocity_field
        self.microbe_properties = props
        self.params = params
        self.output_dir = os.path.abspath(output_dir)
        os.makedirs(self.output_dir, exist_ok=True)
        self.radius = float(interaction_radius)
        self.norm = float(interaction_norm)
        self.chunk = int(output_chunk_iters)
        self.Kh = Kh / 1e10
        self.particle_lons = np.asarray(particle_lons, dtype=np.float64)
        self.particle_lats = np.asarray(particle_lats, dtype=np.float64)
        self.philox_key = tuple(philox_key) if philox_key is not None else new_key()
        self.walk_key = (new_key(seed)[0], 0x6B68)
        self.device = torch.device(device or "cuda")
        self.ring_bytes = int(ring_bytes)
        self.device_record_bytes = device_record_bytes
        self.iteration = 0
        self.timings = {}

    def time_step(self, start_time, end_time, dt, write_outputs=True):
        dt_s = dt.total_seconds()
        iters = max(0, (end_time - start_time) // dt)
        runner = FusedRunner(self.particle_lons, self.particle_lats, self.microbe_properties["species"],
                             self.velocity_field, self.radius, self.params, self.philox_key, dt_s, self.norm,
                             self.chunk, self.Kh, self.walk_key, self.device)
        runner.step = self.iteration  # uniforms keyed by the global step; the clock restarts at 0
        times = [start_time + k * dt for k in range(iters)]
        paths = (os.path.join(self.output_dir, "microbe_data.nc"), os.path.join(self.output_dir, "particle_data.nc"))
        resident = (write_outputs and iters > 0 and DeviceRecords.nbytes(runner.n, iters)
                    <= device_record_budget(self.device, self.device_record_bytes))
        if resident:
            self.timings = self._run_resident(runner, iters, times, paths)
        else:
            self.timings = self._run_streamed(runner, iters, times, paths, write_outputs)
        t = self.timings
        logger.info("Fused advection + interactions, {:d} steps: {:s} (outputs: {:s} more)".format(
            iters, pretty_time(t["run_s"]), pretty_time(t["output_s"])))
        self.microbe_properties["species"][:] = runner.species.cpu().numpy()
        self.final_lons = runner.lon.cpu().numpy()
        self.final_lats = runner.lat.cpu().numpy()
        # a later call continues from the float64 end state
        self.particle_lons, self.particle_lats = self.final_lons, self.final_lats
        self.iteration += iters

    def _run_resident(self, runner, iters, times, paths):
        """The run with its records in HBM, then the files written from there."""
        n = runner.n
        tic = time()
        recs = DeviceRecords(n, iters, self.device)
        done = [0]

        def keep(k, rx, ry, rs, cnt):  # the records are already in place; the pair counts join them
            recs.npairs[done[0]:done[0] + k].copy_(cnt)
            done[0] += k

        t_setup = time() - tic
        tic = time()
        runner.run(iters, on_chunk=keep, records=recs.arrays)
        torch.cuda.synchronize(self.device)
        t_run = time() - tic
        tic = time()
        self.pairs_per_step = recs.npairs.cpu().numpy()
        with ParticleDataset(paths[0], n, times, {"longitude": np.float32, "latitude": np.float32,
                                                  "species": np.int8}, threads=1) as microbe, \
                ParticleDataset(paths[1], n, times, {"longitude": np.float32, "latitude": np.float32},
                                threads=1) as particle:
            wc, ww = write_device_records({"longitude": recs.longitude, "latitude": recs.latitude,
                                           "species": recs.species},
                                          [(microbe, ("longitude", "latitude", "species")),
                                           (particle, ("longitude", "latitude"))])
        t_out = time() - tic
        del recs
        runner.check_domain()
        return {"records": "device", "setup_s": t_setup, "run_s": t_run, "output_s": t_out, "copy_wait_s": wc,
                "write_wait_s": ww, "device_record_bytes": DeviceRecords.nbytes(n, iters)}

    def _run_streamed(self, runner, iters, times, paths, write_outputs):
        """The run with each chunk's records moved to the host as it goes
        (only the pair counts without outputs)."""
        n = runner.n
        microbe_nc, particle_nc = paths
        self.pairs_per_step = np.zeros(iters, dtype=np.int64)
        tic = time()
        ds = None
        fields = {"npairs": (torch.int64, ())}
        if write_outputs:
            ds = ParticleDataset(microbe_nc, n, times, {"longitude": np.float32, "latitude": np.float32,
                                                        "species": np.int8})
            fields.update({"longitude": (torch.float32, (n,)), "latitude": (torch.float32, (n,)),
                           "species": (torch.int8, (n,))})

        def sink(t0, k, arrays):
            self.pairs_per_step[t0:t0 + k] = arrays["npairs"]
            if ds is not None:
                ds.write_block(t0, {name: arrays[name] for name in ("longitude", "latitude", "species")})

        stream = RecordStream(fields, self.chunk, sink, self.device, chunks=len(chunk_schedule(iters, self.chunk)),
                              ring_bytes=self.ring_bytes)
        t_setup = time() - tic
        done = [0]

        def keep(k, rx, ry, rs, cnt):
            recs = {"npairs": cnt}
            if ds is not None:
                recs.update({"longitude": rx, "latitude": ry, "species": rs})
            ev = stream.submit(done[0], k, recs)
            done[0] += k
            return ev

        tic = time()
        try:
            runner.run(iters, on_chunk=keep)
            torch.cuda.synchronize(self.device)
            t_run = time() - tic
        finally:
            try:
                stream.finish()  # the writer drains what was queued, also after an error
            finally:
                if ds is not None:
                    ds.close()
        t_drain = time() - tic - t_run
        runner.check_domain()
        tic = time()
        if write_outputs:
            # particle_data.nc holds the same longitude and latitude columns
            ParticleDataset(particle_nc, n, times, {"longitude": np.float32, "latitude": np.float32}).close()
            copy_variables(microbe_nc, particle_nc, ["longitude", "latitude"])
        t_copy = time() - tic
        return {"records": "stream" if write_outputs else "none", "setup_s": t_setup, "run_s": t_run,
                "output_s": t_drain + t_copy, "writer_drain_s": t_drain, "particle_file_s": t_copy,
                "writer_s": stream.write_s, "submit_wait_s": stream.wait_s, "ring_bytes": stream.ring_bytes}
