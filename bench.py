"""Throughput of the fused hot path: particle-steps/s (advect + pairs + RPS).

    python bench.py [--steps K] [--warmup W] [--workload NAME] [--no-cpu]

One step = one RK4 step of every particle, the fixed-radius pair search on
the recorded float32 positions, and the rock-paper-scissors update over the
sorted pairs, with positions and species recorded on the device.  The default
workload is 490,000 microbes on a 700x700 lattice in a seeded 30-day
synthetic field (720 hourly steps); the timed region runs K steps from the
initial state with inputs already resident in HBM.  Prints one JSON line.

The line's ``adjoint`` object times the differentiable advection of 1,000,000
particles over 240 hourly steps through 20 float32 field levels
(lagrangian_microbes_amd.adjoint): the forward with checkpoints every 16
steps, and loss.backward() of the mean squared great-circle distance to
targets from a second seeded field (recompute plus reverse sweep);
``adjoint_forward_ratio`` is backward time over forward time.

Opt-in legs (their kernels share names with the 490k run's, so they stay out
of the default run that the kernel-trace summary profiles): ``--fp32`` adds
the FP32 path's workload (4M particles, 2,880 half-hour steps, 1/12-degree
field) in FP32 and FP64 from identical inputs, with both rates and the
distance between their positions; ``--ensemble`` adds 64 independent runs
(14.7M particles) batched with the autotuner's launch shape against the same
runs one at a time.
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
    ap.add_argument("--fp32", action="store_true", help="also run the FP32 path at 4M particles (2,880 steps)")
    ap.add_argument("--ensemble", action="store_true", help="also run the 64-run ensemble")
    ap.add_argument("--fp32-steps", type=int, default=2880)
    ap.add_argument("--cpu-steps", type=int, default=48)
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


def run_steps(runner, k, total):
    """Run k steps from the runner's state (chunks pipelined), restarting the
    workload at its end."""
    done = 0
    while done < k:
        if runner.step >= total:
            runner.reset()
        done += runner.run(min(k - done, total - runner.step))
    return done


def algorithmic_bytes(kernel, m):
    """Algorithmic bytes one launch of `kernel` moves, from the per-unit figures
    of the path: advection reads and writes the float64 position (32 B per
    particle), writes the float32 record (8 B per particle and step) and reads
    two field levels per step (F); the pair search reads each position once
    (16 B per particle) and writes the pair list (8 B per pair); the RPS update
    reads the pair list (8 B per pair) and reads and writes the species (2 B
    per particle).  Kernels that only exist to organise the work (sorts, cell
    tables, grouping, scans) have none: their HBM traffic is overhead, reported
    from the byte counters instead.

    Per step: n particles, P pairs.  Batched kernels cover `chunk` steps per
    launch; rps_play runs once per step.
    """
    n, P, chunk, F = (m[k] for k in ("n", "P", "chunk", "F"))
    per_launch = {
        "advect_rk4_f64": 32 * n + (8 * n + F) * chunk,
        "pair_count": 16 * n * chunk,
        "pair_emit": 8 * P * chunk,
        "rps_play": 8 * P + 2 * n,
    }
    return per_launch.get(kernel)


def step_statistics(runner, steps, total, chunk):
    """Mean per-step pair count and component sizes over `steps` steps, from a
    pass run like the timed one (chunks pipelined)."""
    runner.reset()
    rows = []
    chunks = [0]

    def collect(k, rx, ry, rs, cnt):
        counts = cnt.cpu().numpy()
        rps = runner.engine.rps_sets[chunks[0] % runner.PIPELINE_SETS]
        for s in range(k):
            st = rps.component_stats(s)
            rows.append((counts[s], st.touched, st.wide_pairs, st.wide_particles))
        chunks[0] += 1

    runner.run(min(steps, total), on_chunk=collect)
    P, U, Pw, Uw = np.array(rows, dtype=np.float64).mean(axis=0)
    return {"P": P, "U": U, "Pw": Pw, "Uw": Uw}


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary of
    this workload (FETCH_SIZE + WRITE_SIZE, scaled by a plain copy), if any."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        t = json.load(fh)
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else int(k["bytes_per_launch"])


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
    for s in range(steps):
        a = time.perf_counter()
        pi, pj = sorted_kdtree_pairs(rl[s], rt[s], w.radius, w.norm)
        b = time.perf_counter()
        sp = rps_sequential_py(sp, pi, pj, w.p, w.p, w.p, w.key, step0 + s)
        c = time.perf_counter()
        t_pairs += b - a
        t_rps += c - b
        npairs += len(pi)
    total = t_adv + t_pairs + t_rps
    # GPU on the same sample, for a parity check of the sample
    runner.advance(steps)
    torch.cuda.synchronize()
    parity = bool(np.array_equal(runner.species.cpu().numpy(), sp)
                  and np.array_equal(runner.rec_x[steps - 1].cpu().numpy(), rl[steps - 1]))
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


def ensemble_leg():
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

    log(f"[bench] ensemble {e.n} particles in {len(e.runs)} runs: autotuning")
    best, table = autotune.tune(make, steps=24)
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
    return {"workload": w.name, "particles": w.n, "steps": w.steps, "checkpoint_every": plan.ckpt,
            "levels": f.nt, "grid": [f.ny, f.nx], "loss_km2": runs[-1][2], "forward_ms": round(fwd, 3),
            "backward_ms": round(bwd, 3), "ratio": round(bwd / fwd, 3),
            "particle_steps_per_s_forward": w.n * w.steps / (fwd * 1e-3), "kernels": kt}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the path's multi-GPU sharding "
              "is described in DESIGN.md and not built", file=sys.stderr)
        sys.exit(2)

    import torch
    from lagrangian_microbes_amd import _lib, configs, device as dev
    from lagrangian_microbes_amd.simulator import FusedRunner

    _lib.load()
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    w = configs.WORKLOADS[args.workload]()
    lon, lat = w.particles()
    log(f"[bench] workload {w.name}: {w.n} particles, {w.steps} steps available; building runner")
    runner = FusedRunner(lon, lat, w.species(), w.velocity_field(), w.radius, w.params, w.key, w.dt, w.norm,
                         chunk=args.chunk, cell_scale=args.cell_scale)

    # warm-up (compiles nothing, but touches every kernel and sizes buffers)
    run_steps(runner, args.warmup, w.steps)
    torch.cuda.synchronize()
    runner.reset()
    log(f"[bench] warm-up {args.warmup} steps done; timing {args.steps} steps")

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run_steps(runner, args.steps, w.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    runner.check_domain()
    log(f"[bench] timed region {elapsed:.3f}s ({1e3 * elapsed / args.steps:.3f} ms/step)")

    # same steps again with per-kernel events on each launch's own dispatch
    runner.reset()
    torch.cuda.synchronize()
    with dev.KernelTimer() as timer:
        run_steps(runner, args.steps, w.steps)
        torch.cuda.synchronize()
    kt = timer.summary()
    kernels = {k: {"launches": c, "avg_us": 1e3 * ms / c, "total_ms": ms} for k, (c, ms) in kt.items()}
    dominant = max(kernels, key=lambda k: kernels[k]["total_ms"])
    stats = step_statistics(runner, args.steps, w.steps, args.chunk)
    P = stats["P"]
    # two levels of u and v per step (float64)
    m = dict(stats, n=w.n, chunk=args.chunk, F=2 * 8 * runner.dfield.level_len)
    ab = algorithmic_bytes(dominant, m)
    avg_s = kernels[dominant]["avg_us"] * 1e-6
    roofline = None
    if ab is not None:
        achieved = ab / avg_s / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(w.name, dominant),
                    "kernel": dominant, "avg_us": round(kernels[dominant]["avg_us"], 2),
                    "algorithmic_bytes": int(ab)}
    for k, v in kernels.items():
        b, t = algorithmic_bytes(k, m), measured_traffic(w.name, k)
        v["gbs"] = None if b is None else round(b / (v["avg_us"] * 1e-6) / 1e9, 1)
        v["hbm_gbs"] = None if t is None else round(t / (v["avg_us"] * 1e-6) / 1e9, 1)
    # the whole step against the same per-unit figures (records included)
    path_bytes = 50 * w.n + 16 * P + m["F"] + 9 * w.n

    value = w.n * args.steps / elapsed
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": w.name, "particles": w.n, "lattice": "700x700", "dt_s": w.dt,
                   "radius_deg": w.radius, "p": w.p, "field": w.describe()["field"],
                   "mean_pairs_per_step": round(P, 1), "mean_particles_with_pairs": round(stats["U"], 1),
                   "chunk": args.chunk, "cell_scale": args.cell_scale},
        "roofline": roofline,
        "kernels": {k: {"launches": v["launches"], "avg_us": round(v["avg_us"], 2), "algorithmic_gbs": v["gbs"],
                        "hbm_gbs": v["hbm_gbs"]} for k, v in kernels.items()},
        "path": {"algorithmic_bytes_per_step": int(path_bytes),
                 "achieved_gbs": round(path_bytes / (elapsed / args.steps) / 1e9, 1)},
        "adjoint_forward_ratio": None,
    }

    if not args.no_cpu:
        runner.reset()
        run_steps(runner, w.steps // 2, w.steps)
        torch.cuda.synchronize()
        log(f"[bench] CPU baseline: {args.cpu_steps} steps from step {runner.step}")
        result["cpu_baseline"] = cpu_baseline(w, runner, args.cpu_steps)
    else:
        result["cpu_baseline"] = None
    del runner
    torch.cuda.empty_cache()
    if args.fp32:
        result["fp32_path"] = fp32_leg(args.fp32_steps)
    if args.ensemble:
        result["ensemble"] = ensemble_leg()
    if not args.no_adjoint:
        result["adjoint"] = adjoint_leg(args.adjoint_reps)
        result["adjoint_forward_ratio"] = result["adjoint"]["ratio"]
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
