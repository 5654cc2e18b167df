"""Throughput of the fused hot path: particle-steps/s (advect + pairs + RPS).

    python bench.py [--steps K] [--warmup W] [--workload NAME] [--no-cpu]

One step = one RK4 step of every particle, the fixed-radius pair search on
the recorded float32 positions, and the rock-paper-scissors update over the
sorted pairs, with positions and species recorded on the device.  The default
workload is 490,000 microbes on a 700x700 lattice in a seeded 30-day
synthetic field (720 hourly steps); the timed region runs K steps from the
initial state with inputs already resident in HBM.  Prints one JSON line.
"""
import argparse
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
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-steps", type=int, default=8)
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
    """Run k steps from the runner's state, restarting the workload at its end."""
    done = 0
    while done < k:
        if runner.step >= total:
            runner.reset()
        done += runner.advance(min(k - done, total - runner.step))
    return done


def algorithmic_bytes(kernel, n, pairs, chunk, T, field_len):
    """Compulsory bytes one launch of `kernel` must move.

    Per step: n particles, P pairs, T cell buckets, U = min(n, 2P) particles
    with a pair.  Batched kernels cover `chunk` steps per launch; the replay
    kernels (rps_play*) run once per step.  The advection launch reads and
    writes the float64 state once and writes float32 records every step.
    """
    P, U = pairs, min(n, 2 * pairs)
    per_step = {
        "bin_count": 8 * n + 8 * min(n, T),                     # positions, bucket counters
        "bin_scatter": 8 * n + 8 * min(n, T) + 16 * n,          # positions, starts, records
        "pair_count": 16 * n + 4 * n + 4 * P,                   # records, counts, neighbour rows
        "pair_emit": 8 * n + 4 * P + 8 * P,                     # counts+offsets, rows, pair list
        "rps_init": 8 * n,                                      # parent, slot
        "rps_union": 8 * P + P + 8 * U,                         # pairs, outcome bits, parents
        "rps_claim": 8 * P + 24 * U,                            # pairs, rows, slots, labels, counts
        "rps_group": 4 * P + 20 * U,                            # heads, ids, offsets
        "rps_order": 20 * U,                                    # heads, rows, row offsets
        "rps_fill": 8 * P + P + 12 * U + 8 * P,                 # pairs, bits, lookups, records
        "rps_deps": 16 * P,                                     # records in, masks out
    }
    per_launch = {
        "advect_rk4_f64": 32 * n + 8 * n * chunk + 8 * field_len * (2 * chunk + 1),
        "field_blend_f64": 24 * field_len * (2 * chunk + 1),
        # untouched particles copied (slot check + species), components: ids,
        # species in and out, pair records
        "rps_play": 4 * n + 2 * (n - U) + 6 * U + 8 * P,
        "rps_play_wide": 6 * U + 16 * P,
    }
    if kernel in per_launch:
        return per_launch[kernel]
    if kernel in per_step:
        return per_step[kernel] * chunk
    return None


def cpu_baseline(w, runner, steps):
    """The oracle's CPU path on a bounded sample: `steps` steps from the
    device state at mid-run, C RK4 + scipy cKDTree + the Python pair loop."""
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
    t0 = time.perf_counter()
    lo, la, rl, rt, _ = oracle.advect_rk4_f64(lon, lat, stage, grid, steps, w.dt)
    t_adv = time.perf_counter() - t0
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
        "cores": 1,
        "kind": "port",
        "sample": (f"{steps} steps x {w.n} particles from step {step0} (mid-run, {npairs // steps} pairs/step): "
                   f"oracle C RK4 {t_adv:.2f}s + scipy cKDTree.query_pairs {t_pairs:.2f}s + "
                   f"per-pair Python RPS loop {t_rps:.2f}s, single thread, {platform.processor() or 'cpu'}"),
        "parity_with_gpu": parity,
    }


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
                         chunk=args.chunk)

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
    pairs_per_step = float(runner.npairs[:min(args.chunk, args.steps)].double().mean().item())
    log(f"[bench] timed region {elapsed:.3f}s ({1e3 * elapsed / args.steps:.3f} ms/step)")

    # same steps again with per-kernel events (dispatch-bracketed) for the roofline
    runner.reset()
    torch.cuda.synchronize()
    with dev.KernelTimer() as timer:
        run_steps(runner, args.steps, w.steps)
        torch.cuda.synchronize()
    kt = timer.summary()
    all_pairs = []
    runner.reset()
    kernels = {k: {"launches": c, "avg_us": 1e3 * ms / c, "total_ms": ms} for k, (c, ms) in kt.items()}
    dominant = max(kernels, key=lambda k: kernels[k]["total_ms"])
    T = runner.engine.search.workspace_bytes  # noqa: F841 (reported below)
    # average pair count over the whole run, from a pass that keeps every step's count
    counts = []
    runner.reset()
    while runner.step < min(args.steps, w.steps):
        k = runner.advance(min(args.chunk, w.steps - runner.step))
        counts.append(runner.npairs[:k].cpu().numpy().copy())
    P = float(np.concatenate(counts).mean()) if counts else 0.0
    field_len = runner.dfield.level_len
    grid_T = 1 << max(10, int(np.ceil(np.log2(max(w.n, 1)))))
    ab = algorithmic_bytes(dominant, w.n, P, args.chunk, grid_T, field_len)
    avg_s = kernels[dominant]["avg_us"] * 1e-6
    roofline = None
    if ab is not None:
        achieved = ab / avg_s / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dominant,
                    "avg_us": round(kernels[dominant]["avg_us"], 2), "algorithmic_bytes": int(ab)}

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
                   "mean_pairs_per_step": round(P, 1), "chunk": args.chunk},
        "roofline": roofline,
        "kernels": {k: {"launches": v["launches"], "avg_us": round(v["avg_us"], 2)} for k, v in kernels.items()},
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
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
