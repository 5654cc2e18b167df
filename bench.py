"""Throughput of the fused hot path: particle-steps/s (advect + pairs + RPS).

    python bench.py [--steps K] [--warmup W] [--workload NAME] [--no-cpu] [--no-fp32] [--no-ensemble]
                    [--no-adjoint] [--window-only]

One step = one RK4 step of every particle, the fixed-radius pair search on
the recorded float32 positions, and the rock-paper-scissors update over the
sorted pairs, with positions and species recorded on the device.  The default
workload is 490,000 microbes on a 700x700 lattice in a seeded 30-day
synthetic field (720 hourly steps).  The timed region runs K steps with
inputs already resident in HBM: all 720 from the initial state by default; a
shorter window starts at mid-run (K = 20: steps 360-379), since the first
steps of the run have almost no pairs.  Prints one JSON line.

``roofline`` is the kernel with the most time among those that move
algorithmic bytes (advection, pair count, pair emission, RPS replay): its
bytes over the window (from the chunk lengths actually launched and the
window's pair counts) over its launch time measured with events on its own
dispatches; ``roofline.path`` is the whole step against the same per-unit
figures.  ``--window-only`` runs just the window once (the run a counter or
trace profile of the window is taken from) and prints its launch counts.

Further objects of the line:
``fp32_path``: the FP32 path's workload (4M particles, 2,880 half-hour steps,
1/12-degree field) in FP32 and FP64 from identical inputs, with both rates
and the distance between their positions.  ``ensemble``: 64 independent runs
(14.7M particles) batched with the autotuner's launch shape against the same
runs one at a time.  ``adjoint``: the differentiable advection of 1,000,000
particles over 240 hourly steps through 20 float32 field levels
(lagrangian_microbes_amd.adjoint): the forward with checkpoints every 16
steps, and loss.backward() of the mean squared great-circle distance to
targets from a second seeded field (recompute plus reverse sweep), then the
RPS update on the final positions; ``adjoint_forward_ratio`` is backward
time over forward time.
"""
import argparse
import concurrent.futures
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s (advect+pairs+RPS) at 490k; adjoint/forward time ratio"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=720)
    ap.add_argument("--warmup", type=int, default=24)
    ap.add_argument("--workload", default="rps_490k_30day")
    ap.add_argument("--chunk", type=int, default=100)
    ap.add_argument("--cell-scale", type=float, default=1.0,
                    help="pair-search cell side in radii (changes speed, never the pairs)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint forward/backward timing")
    ap.add_argument("--adjoint-reps", type=int, default=3)
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 path at 4M particles")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 64-run ensemble")
    ap.add_argument("--ensemble-tune", choices=("short", "full"), default="short",
                    help="autotuner grid of the ensemble leg (short: 8 launch shapes, full: 27)")
    ap.add_argument("--window-only", action="store_true",
                    help="run only the timed window once with per-launch events and print its launch counts "
                         "(for counter and trace profiles of the window)")
    ap.add_argument("--outputs", action="store_true",
                    help="also time MicrobeSimulator at 490k with and without the output files (writes ~6 GB)")
    ap.add_argument("--fp32-steps", type=int, default=2880)
    ap.add_argument("--cpu-steps", type=int, default=48)
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the adjoint leg and print its object (for trace and counter profiles)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel on BYTES bytes 10 times (byte-counter calibration)")
    return ap.parse_args()


def copy_calibration(nbytes):
    """Plain copy of a known size, to scale hardware byte counters."""
    import torch
    from lagrangian_microbes_amd import device as dev
    nbytes = (nbytes // 16) * 16
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(10):
        dev.copy_bytes(src, dst, nbytes)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": nbytes, "launches": 10,
                      "read_bytes_per_launch": nbytes, "write_bytes_per_launch": nbytes}), flush=True)


def window_start(total, steps):
    """First step of the timed window.  A window as long as the workload
    starts at step 0; a shorter one is placed at mid-run and ends inside the
    run: at step 0 the lattice spacing exceeds the radius, so the first steps
    have almost no pairs and no RPS work, unlike the rest of the run."""
    if steps >= total:
        return 0
    return min(total // 2, total - steps)


def position(runner, start):
    """Back to the initial state, then (untimed, pipelined as the timed run)
    to step `start`."""
    import torch
    runner.reset()
    if start:
        runner.run(start)
    torch.cuda.synchronize()


def run_steps(runner, k, total, on_chunk=None):
    """Run k steps from the runner's state (chunks pipelined), restarting the
    workload at its end.  Returns the chunk lengths launched (one batched
    launch of the advection and of each preparation kernel per chunk)."""
    done, sizes = 0, []
    while done < k:
        if runner.step >= total:
            runner.reset()
        cb = None if on_chunk is None else on_chunk()
        done += runner.run(min(k - done, total - runner.step), on_chunk=cb)
        sizes += runner.last_schedule
    return sizes


def window_bytes(kernel, n, K, L, sum_pairs, F):
    """Algorithmic bytes all launches of `kernel` move over a window of K steps
    launched as L chunks, from the per-unit figures of the path: advection
    reads and writes the float64 position once per launch (32 B per particle),
    writes the float32 record (8 B per particle and step) and reads two field
    levels per step (F bytes); the pair search reads each position once
    (16 B per particle and step) and writes the pair list (8 B per pair); the
    RPS update reads the pair list (8 B per pair) and reads and writes the
    species (2 B per particle and step).  Kernels that only organise the work
    (sorts, cell tables, grouping, scans) have none: their HBM traffic is
    overhead, reported from the byte counters instead."""
    return {
        "advect_rk4_f64": 32 * n * L + (8 * n + F) * K,
        "pair_count": 16 * n * K,
        "pair_emit": 8 * sum_pairs,
        "rps_play": 8 * sum_pairs + 2 * n * K,
    }.get(kernel)


def step_statistics(runner, start, steps, total):
    """Per-step pair counts and component sizes over the window, from a pass
    run like the timed one (chunks pipelined)."""
    import torch
    position(runner, start)
    rows = []

    def per_call():
        chunks = [0]

        def collect(k, rx, ry, rs, cnt):
            counts = cnt.cpu().numpy()
            rps = runner.engine.rps_sets[chunks[0] % runner.PIPELINE_SETS]
            for s in range(k):
                st = rps.component_stats(s)
                rows.append((counts[s], st.touched, st.wide_pairs, st.wide_particles))
            chunks[0] += 1
        return collect

    run_steps(runner, steps, total, on_chunk=per_call)
    torch.cuda.synchronize()
    a = np.array(rows, dtype=np.float64)
    return {"P": a[:, 0].mean(), "sum_P": int(a[:, 0].sum()), "U": a[:, 1].mean(), "Pw": a[:, 2].mean(),
            "Uw": a[:, 3].mean()}


def traffic_file(workload, start, steps):
    return os.path.join(ROOT, "profiles", f"traffic_{workload}_steps{start}-{start + steps}.json")


def measured_traffic(workload, start, steps):
    """HBM bytes per launch of each kernel over this window, from the
    committed counter summary of the same window (FETCH_SIZE + WRITE_SIZE
    scaled by a plain copy), if one exists."""
    path = traffic_file(workload, start, steps)
    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        t = json.load(fh)
    w = t.get("window", {})
    if (w.get("workload"), w.get("start"), w.get("steps")) != (workload, start, steps):
        return {}
    return {k: int(v["bytes_per_launch"]) for k, v in t.get("kernels", {}).items()}


def cpu_baseline(w, runner, steps):
    """The oracle's CPU path on a bounded sample: `steps` steps from the
    device state at mid-run.  As the reference does, the advection runs over
    contiguous particle tiles in parallel (its joblib tiles; here the C RK4 on
    one thread per host core, up to 16), the pair search (scipy cKDTree) and
    the per-pair loop on one thread."""
    import torch
    import oracle
    from oracle.pairs import sorted_kdtree_pairs
    from oracle.rps import rps_sequential_py

    f = w.velocity_field()
    lon = runner.lon.cpu().numpy()
    lat = runner.lat.cpu().numpy()
    sp = runner.species.cpu().numpy()
    step0 = runner.step
    lev0, wts = f.stage_schedule(runner.clock_steps * w.dt, w.dt, steps)
    stage = oracle.field_blend_f64(f.interleaved(), lev0, wts)
    grid = (f.nx, f.ny, f.x0, f.y0, 1.0 / f.dx, 1.0 / f.dy)
    assert getattr(runner, "walk_amp", 0.0) == 0.0  # walk counters are per call: tiles only without a walk
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    bounds = np.linspace(0, w.n, cores + 1).astype(np.int64)

    def advect_tile(k):  # the C call releases the GIL
        a, b = bounds[k], bounds[k + 1]
        return oracle.advect_rk4_f64(lon[a:b], lat[a:b], stage, grid, steps, w.dt)

    t0 = time.perf_counter()
    with concurrent.futures.ThreadPoolExecutor(cores) as ex:
        parts = list(ex.map(advect_tile, range(cores)))
    t_adv = time.perf_counter() - t0
    rl = np.concatenate([q[2] for q in parts], axis=1)
    rt = np.concatenate([q[3] for q in parts], axis=1)
    t_pairs = t_rps = 0.0
    npairs = 0
    cpu_species, cpu_pairs = [], []
    for s in range(steps):
        a = time.perf_counter()
        pi, pj = sorted_kdtree_pairs(rl[s], rt[s], w.radius, w.norm)
        b = time.perf_counter()
        sp = rps_sequential_py(sp, pi, pj, w.p, w.p, w.p, w.key, step0 + s)
        c = time.perf_counter()
        t_pairs += b - a
        t_rps += c - b
        npairs += len(pi)
        cpu_species.append(sp)
        cpu_pairs.append(len(pi))
    total = t_adv + t_pairs + t_rps
    # the GPU over the same sample, through the timed (pipelined) path, for a
    # parity check of the sample: every step's pair count and species, and the
    # last step's recorded positions
    got = {"species": [], "pairs": []}

    def keep(k, rx, ry, rs, cnt):
        got["species"] += [rs[s].cpu().numpy() for s in range(k)]
        got["pairs"] += cnt.cpu().tolist()
        got["last"] = (rx[k - 1].cpu().numpy(), ry[k - 1].cpu().numpy())

    runner.run(steps, on_chunk=keep)
    torch.cuda.synchronize()
    parity = bool(got["pairs"] == cpu_pairs
                  and all(np.array_equal(g, c) for g, c in zip(got["species"], cpu_species))
                  and np.array_equal(got["last"][0], rl[steps - 1])
                  and np.array_equal(got["last"][1], rt[steps - 1]))
    return {
        "value": w.n * steps / total,
        "unit": "particle-steps/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"{steps} steps x {w.n} particles from step {step0} (mid-run, {npairs // steps} pairs/step): "
                   f"oracle C RK4 over {cores} particle tiles on {cores} threads {t_adv:.2f}s + scipy "
                   f"cKDTree.query_pairs {t_pairs:.2f}s + per-pair Python RPS loop {t_rps:.2f}s (one thread each, "
                   f"as in the reference), {platform.processor() or 'cpu'}"),
        "parity_with_gpu": parity,
    }


def fp32_leg(steps):
    """The FP32 path on its workload (4M particles, 1/12-degree 21-level field,
    half-hour steps): the fused run in FP32 and in FP64 from identical inputs,
    timed, with the distance between their positions after 1 day, 10 days and
    the whole run."""
    import torch
    from lagrangian_microbes_amd import configs
    from lagrangian_microbes_amd.simulator import FusedRunner

    w = configs.rps_4m_60day()
    f = w.velocity_field()
    lon, lat = w.particles()
    marks = [m for m in (48, 480) if m < steps] + [steps]
    out, snaps = {}, {}
    for prec in ("f32", "f64"):
        r = FusedRunner(lon, lat, w.species(), f, w.radius, w.params, w.key, w.dt, chunk=48, precision=prec)
        r.run(48)
        torch.cuda.synchronize()
        r.reset()
        elapsed, done, snaps[prec] = 0.0, 0, []
        for m in marks:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r.run(m - done)
            torch.cuda.synchronize()
            elapsed += time.perf_counter() - t0
            done = m
            snaps[prec].append((r.lon.cpu().numpy(), r.lat.cpu().numpy()))
        out[prec] = {"ms_per_step": round(1e3 * elapsed / steps, 4),
                     "particle_steps_per_s": w.n * steps / elapsed,
                     "particle_steps_outside_field": int(r.oob.item())}
        log(f"[bench] {w.name} {prec}: {out[prec]['ms_per_step']} ms/step")
        del r
        torch.cuda.empty_cache()
    err = {}
    for m, a, b in zip(marks, snaps["f32"], snaps["f64"]):
        d = np.hypot(a[0] - b[0], a[1] - b[1])
        err[f"after_{m}_steps"] = {"median_deg": float(np.median(d)), "p99_deg": float(np.percentile(d, 99)),
                                   "max_deg": float(d.max())}
    return {"workload": w.name, "particles": w.n, "steps": steps, "dt_s": w.dt, "field": w.describe()["field"],
            "f32": out["f32"], "f64": out["f64"], "f32_vs_f64_position_error": err}


def ensemble_leg(tune="short"):
    """The 64-run ensemble (14.7M particles, uneven runs, 240 hourly steps,
    FP32): the launch shape picked by the autotuner, the batched run timed
    from the initial state (best of two), then every run alone through the
    plain FP32 path (one FusedRunner per run, setup outside the timing, best
    of two), for the speedup."""
    import torch
    from lagrangian_microbes_amd import autotune, configs, device as dev
    from lagrangian_microbes_amd.simulator import FusedRunner

    e = configs.ensemble_64runs()
    field = e.velocity_field()
    lon, lat = e.particles()
    species = e.species()
    w0 = e.runs[0]

    def make(cell_scale, chunk, play_blocks):
        runs = dev.EnsembleRuns([w.n for w in e.runs], [w.params for w in e.runs], [w.key for w in e.runs])
        return FusedRunner(lon, lat, species, field, w0.radius, {"pRS": 0, "pPR": 0, "pSP": 0}, (0, 0), w0.dt,
                           chunk=chunk, precision="f32", runs=runs, cell_scale=cell_scale, play_blocks=play_blocks,
                           graphs=True)

    cands = None if tune == "full" else autotune.candidates((1.0, 1.5), (8, 24), ((2048, 2048), (4096, 2048)))
    log(f"[bench] ensemble {e.n} particles in {len(e.runs)} runs: autotuning")
    def timed(make_runner, cand, steps):
        t = autotune.time_runner(make_runner, cand, steps)
        log(f"[bench] ensemble autotune {cand}: {1e3 * t:.3f} ms/step")
        return t

    best, table = autotune.tune(make, steps=24, cands=cands, timer=timed)
    log(f"[bench] ensemble autotuner picked {best}")
    r = make(**best)
    r.run(FusedRunner.PIPELINE_SETS * best["chunk"])  # warm-up: every chunk buffer set's replay graph captured
    torch.cuda.synchronize()
    batched = float("inf")
    for _ in range(2):  # best of two, as for the runs alone below
        r.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r.run(e.steps)
        torch.cuda.synchronize()
        batched = min(batched, time.perf_counter() - t0)
    oob = int(r.oob.item())
    del r
    torch.cuda.empty_cache()
    per_run = []
    for w in e.runs:
        lo, la = w.particles()
        rr = FusedRunner(lo, la, w.species(), field, w.radius, w.params, w.key, w.dt, precision="f32")
        rr.run(FusedRunner.PIPELINE_SETS * rr.chunk)  # warm-up: every chunk buffer set
        torch.cuda.synchronize()
        t_run = float("inf")
        for _ in range(2):  # best of two: a run alone now and then stalls for ~1 s outside its kernels
            rr.reset()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rr.run(w.steps)
            torch.cuda.synchronize()
            t_run = min(t_run, time.perf_counter() - t0)
        per_run.append(t_run)
        if len(per_run) % 16 == 0:
            log(f"[bench] ensemble: {len(per_run)} of {len(e.runs)} runs timed alone")
        del rr
    torch.cuda.empty_cache()
    seq = float(sum(per_run))
    ms = np.array(per_run) * 1e3 / e.steps
    log(f"[bench] ensemble: batched {batched:.3f}s, runs alone {seq:.3f}s, speedup {seq / batched:.2f}")
    return {**e.describe(), "autotuned": {**best, "play_blocks": list(best["play_blocks"])},
            "tuning_ms_per_step": [[c["cell_scale"], c["chunk"], list(c["play_blocks"]), round(t * 1e3, 4)]
                                   for c, t in table],
            "ms_per_step": round(1e3 * batched / e.steps, 4), "particle_steps_per_s": e.n * e.steps / batched,
            "particle_steps_outside_field": oob, "runs_alone_s": round(seq, 4),
            "runs_alone_particle_steps_per_s": e.n * e.steps / seq,
            "runs_alone_ms_per_step": {"min": round(float(ms.min()), 4), "median": round(float(np.median(ms)), 4),
                                       "max": round(float(ms.max()), 4)},
            "speedup_vs_runs_alone": round(seq / batched, 3)}


def outputs_leg(steps):
    """MicrobeSimulator on the 490k workload without output files, with them
    (particle_data.nc and microbe_data.nc in a temporary directory; the
    records stay in HBM during the run and are written after it) and with
    them streamed through the pinned host ring during the run instead: the
    device run's time each way, the output stages' times and the file
    sizes.  The files are deleted afterwards."""
    import shutil
    import tempfile
    from datetime import datetime, timedelta
    from lagrangian_microbes_amd import configs
    from lagrangian_microbes_amd.interactions import rock_paper_scissors
    from lagrangian_microbes_amd.simulator import MicrobeSimulator

    w = configs.rps_490k_30day()
    lon, lat = w.particles()
    f = w.velocity_field()
    start, dt = datetime(2017, 1, 1), timedelta(seconds=w.dt)
    out = {}
    tmp = tempfile.mkdtemp(prefix="lm_outputs_")
    try:
        free = shutil.disk_usage(tmp).free
        need = 2 * w.n * steps * 9
        if free < 2 * need:
            return {"skipped": f"{free / 2**30:.1f} GiB free, {need / 2**30:.1f} GiB of files"}
        for mode in ("warm-up", "off", "on", "on-streamed"):
            rps = rock_paper_scissors(w.n, w.p, w.p, w.p, seed=w.species_seed)
            kw = {"device_record_bytes": 0} if mode == "on-streamed" else {}
            sim = MicrobeSimulator(lon, lat, f, rps, w.radius, output_dir=tmp, output_chunk_iters=100,
                                   philox_key=w.key, **kw)
            t0 = time.perf_counter()
            sim.time_step(start, start + steps * dt, dt, write_outputs=mode.startswith("on"))
            wall = time.perf_counter() - t0
            if mode != "warm-up":
                out[mode] = {**{k: (round(v, 4) if isinstance(v, float) else v) for k, v in sim.timings.items()},
                             "call_s": round(wall, 3)}
                log(f"[bench] outputs {mode}: run {sim.timings['run_s']:.3f}s, call {wall:.3f}s")
            if mode == "on":
                sizes = {name: os.path.getsize(os.path.join(tmp, name))
                         for name in ("microbe_data.nc", "particle_data.nc")}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return {"workload": w.name, "particles": w.n, "steps": steps, "file_bytes": sizes,
            "run_s_without_outputs": out["off"]["run_s"], "run_s_with_outputs": out["on"]["run_s"],
            "run_slowdown": round(out["on"]["run_s"] / out["off"]["run_s"] - 1.0, 4),
            "run_slowdown_streamed": round(out["on-streamed"]["run_s"] / out["off"]["run_s"] - 1.0, 4),
            "timings": out}


def adjoint_leg(reps):
    """The adjoint workload (1M particles, 240 steps, 20 float32 levels): times the
    differentiable forward and loss.backward() with events on the stream they run on,
    and the liblm kernels of each pass with their own dispatch events."""
    import torch
    from lagrangian_microbes_amd import adjoint, configs, device as dev

    w = configs.adjoint_1m_10day()
    f, tf = w.velocity_field(), w.target_velocity_field()
    lon, lat = w.particles()
    lon_d, lat_d = torch.tensor(lon, device="cuda"), torch.tensor(lat, device="cuda")
    plan = adjoint.AdjointPlan(f, 0.0, w.dt, w.steps, ckpt=16)
    tplan = adjoint.AdjointPlan(tf, 0.0, w.dt, w.steps, ckpt=16)
    tu = torch.tensor(tf.u, dtype=torch.float32, device="cuda")
    tv = torch.tensor(tf.v, dtype=torch.float32, device="cuda")
    with torch.no_grad():
        TX, TY = adjoint.advect(lon_d, lat_d, tu, tv, tplan)
    u = torch.tensor(f.u, dtype=torch.float32, device="cuda", requires_grad=True)
    v = torch.tensor(f.v, dtype=torch.float32, device="cuda", requires_grad=True)
    x0 = lon_d.clone().requires_grad_(True)
    y0 = lat_d.clone().requires_grad_(True)

    def one(timed):
        u.grad = v.grad = x0.grad = y0.grad = None
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        e[0].record()
        X, Y = adjoint.advect(x0, y0, u, v, plan)
        e[1].record()
        L = adjoint.trajectory_loss(X, Y, TX, TY)
        L.backward()
        e[2].record()
        torch.cuda.synchronize()
        return e[0].elapsed_time(e[1]), e[1].elapsed_time(e[2]), float(L.detach())

    one(False)
    runs = [one(True) for _ in range(reps)]
    fwd = min(r[0] for r in runs)
    bwd = min(r[1] for r in runs)
    with dev.KernelTimer() as timer:
        one(True)
    kt = {k: {"launches": c, "total_ms": round(ms, 3)} for k, (c, ms) in timer.summary().items()}
    log(f"[bench] adjoint: forward {fwd:.2f} ms, backward {bwd:.2f} ms, ratio {bwd / fwd:.2f}")
    # the RPS update, forward only, on the final positions of the last pass
    with torch.no_grad():
        X, Y = adjoint.advect(x0, y0, u, v, plan)
    sp = torch.tensor(w.species(), device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    npairs = adjoint.interact_final(X, Y, sp, w.radius, w.params, w.key, w.steps, f.bounds)
    torch.cuda.synchronize()
    t_rps = time.perf_counter() - t0
    counts = np.bincount(sp.cpu().numpy().astype(np.int64), minlength=4)[1:4]
    return {"workload": w.name, "particles": w.n, "steps": w.steps, "checkpoint_every": plan.ckpt,
            "levels": f.nt, "grid": [f.ny, f.nx], "loss_km2": runs[-1][2], "forward_ms": round(fwd, 3),
            "backward_ms": round(bwd, 3), "ratio": round(bwd / fwd, 3),
            "particle_steps_per_s_forward": w.n * w.steps / (fwd * 1e-3), "kernels": kt,
            "final_interaction": {"pairs": npairs, "species_counts": [int(c) for c in counts],
                                  "ms": round(1e3 * t_rps, 3)}}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the path is built for one GPU, "
              "its sharding across GPUs is not", file=sys.stderr)
        sys.exit(2)

    import torch
    from lagrangian_microbes_amd import _lib, configs, device as dev
    from lagrangian_microbes_amd.simulator import FusedRunner

    _lib.load()
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.adjoint_only:
        print(json.dumps(adjoint_leg(args.adjoint_reps)), flush=True)
        return
    w = configs.WORKLOADS[args.workload]()
    lon, lat = w.particles()
    log(f"[bench] workload {w.name}: {w.n} particles, {w.steps} steps available; building runner")
    runner = FusedRunner(lon, lat, w.species(), w.velocity_field(), w.radius, w.params, w.key, w.dt, w.norm,
                         chunk=args.chunk, cell_scale=args.cell_scale)
    start = window_start(w.steps, args.steps)

    # warm-up (compiles nothing, but touches every kernel and sizes buffers)
    run_steps(runner, args.warmup, w.steps)
    torch.cuda.synchronize()
    position(runner, start)
    log(f"[bench] warm-up {args.warmup} steps done; timing {args.steps} steps from step {start}")

    if args.window_only:
        # the timed window once, with per-launch events, and nothing after it:
        # the run a counter or trace profile of the window is taken from
        with dev.KernelTimer() as timer:
            sizes = run_steps(runner, args.steps, w.steps)
            torch.cuda.synchronize()
        kt = timer.summary()
        print(json.dumps({"workload": w.name, "start": start, "steps": args.steps, "chunks": sizes,
                          "launches": {k: c for k, (c, ms) in kt.items()},
                          "avg_us": {k: round(1e3 * ms / c, 2) for k, (c, ms) in kt.items()}}), flush=True)
        return

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sizes = run_steps(runner, args.steps, w.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    runner.check_domain()
    log(f"[bench] timed region {elapsed:.3f}s ({1e3 * elapsed / args.steps:.3f} ms/step), chunks {sizes}")

    # the same window again with events on each launch's own dispatch
    position(runner, start)
    with dev.KernelTimer() as timer:
        run_steps(runner, args.steps, w.steps)
        torch.cuda.synchronize()
    kt = timer.summary()
    kernels = {k: {"launches": c, "avg_us": 1e3 * ms / c, "total_ms": ms} for k, (c, ms) in kt.items()}
    stats = step_statistics(runner, start, args.steps, w.steps)
    P, K, L = stats["P"], args.steps, len(sizes)
    F = 2 * 2 * runner.dfield.level_len * 8  # two levels of interleaved (u, v) float64 per step
    traffic = measured_traffic(w.name, start, K)
    for k, v in kernels.items():
        tb = window_bytes(k, w.n, K, L, stats["sum_P"], F)
        v["bytes_per_launch"] = None if tb is None else tb / v["launches"]
        v["gbs"] = None if tb is None else tb / (v["total_ms"] * 1e-3) / 1e9
        t = traffic.get(k)
        v["hbm_gbs"] = None if t is None else t / (v["avg_us"] * 1e-6) / 1e9
    # the roofline kernel: the one with the most time among those that carry
    # algorithmic bytes (the organising kernels have none)
    with_bytes = [k for k, v in kernels.items() if v["gbs"] is not None]
    rk = max(with_bytes, key=lambda k: kernels[k]["total_ms"])
    rv = kernels[rk]
    roofline = {"bound": "hbm", "achieved": round(rv["gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(rv["gbs"] / HBM_PEAK_GBS, 5), "traffic": traffic.get(rk), "kernel": rk,
                "avg_us": round(rv["avg_us"], 2), "algorithmic_bytes_per_launch": int(rv["bytes_per_launch"]),
                "launches": rv["launches"]}
    # the whole step against the same per-unit figures (records included)
    path_bytes = (59 * w.n + F) * K + 16 * stats["sum_P"]
    path_gbs = path_bytes / elapsed / 1e9
    roofline["path"] = {"algorithmic_bytes_per_step": int(path_bytes / K), "achieved": round(path_gbs, 1),
                        "frac": round(path_gbs / HBM_PEAK_GBS, 5)}

    value = w.n * K / elapsed
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": w.name, "particles": w.n, "lattice": "700x700", "dt_s": w.dt,
                   "radius_deg": w.radius, "p": w.p, "field": w.describe()["field"],
                   "window": {"first_step": start, "steps": K, "chunks": sizes},
                   "mean_pairs_per_step": round(P, 1), "mean_particles_with_pairs": round(stats["U"], 1),
                   "chunk": args.chunk, "cell_scale": args.cell_scale},
        "roofline": roofline,
        "kernels": {k: {"launches": v["launches"], "avg_us": round(v["avg_us"], 2),
                        "algorithmic_gbs": None if v["gbs"] is None else round(v["gbs"], 1),
                        "hbm_gbs": None if v["hbm_gbs"] is None else round(v["hbm_gbs"], 1)}
                    for k, v in kernels.items()},
        "adjoint_forward_ratio": None,
    }
    log(f"[bench] roofline kernel {rk}: {rv['gbs']:.1f} GB/s algorithmic; path {path_gbs:.1f} GB/s")

    if not args.no_cpu:
        position(runner, w.steps // 2)
        log(f"[bench] CPU baseline: {args.cpu_steps} steps from step {runner.step}")
        result["cpu_baseline"] = cpu_baseline(w, runner, args.cpu_steps)
    else:
        result["cpu_baseline"] = None
    del runner
    torch.cuda.empty_cache()
    if not args.no_fp32:
        result["fp32_path"] = fp32_leg(args.fp32_steps)
    if not args.no_ensemble:
        result["ensemble"] = ensemble_leg(args.ensemble_tune)
    if args.outputs:
        result["outputs"] = outputs_leg(w.steps)
    if not args.no_adjoint:
        result["adjoint"] = adjoint_leg(args.adjoint_reps)
        result["adjoint_forward_ratio"] = result["adjoint"]["ratio"]
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
