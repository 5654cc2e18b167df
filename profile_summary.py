"""Short text summaries of rocprofv3 CSV output for liblm kernels.

    python profile_summary.py trace DIR [--title T] [--command C] [--window WINDOW_JSON]
        [--passes KERNEL=N1,N2,...[:NAME1,NAME2,...]]
    python profile_summary.py traffic FETCH_DIR WRITE_DIR \
        COPY_FETCH_DIR COPY_WRITE_DIR COPY_BYTES [--json OUT] [--command C] [--window WINDOW_JSON]

``trace`` reads the ``*_kernel_stats.csv`` of a ``rocprofv3 --kernel-trace
--stats --output-format csv`` run and prints a markdown table of the liblm
kernels (calls, average and total duration).  ``--passes`` splits one
kernel's dispatches, in launch order (``*_kernel_trace.csv``), into
consecutive groups of N1, N2, ... launches and prints each group's average:
the passes of a benchmark run over one workload.

    python profile_summary.py legs BENCH_JSON OUT_DIR [--command C]
    python profile_summary.py units PMC_DIR [PMC_DIR ...] [--title T] [--command C] [--window WINDOW_JSON]

``units`` reads ``--pmc`` runs of SQ_WAVES, SQ_INSTS_VALU, TA_TA_BUSY_sum,
TA_ADDR_STALLED_BY_TC_CYCLES_sum and GRBM_GUI_ACTIVE (any passes) and prints
per kernel the VALU issue share and how busy the address unit was.

``legs`` renders the ``fp32_path`` and ``ensemble`` objects of a ``bench.py
--fp32 --ensemble`` JSON line as ``bench_<workload>.md`` files in OUT_DIR.

``traffic`` reads four ``rocprofv3 --pmc FETCH_SIZE`` / ``--pmc WRITE_SIZE``
runs (``*_counter_collection.csv``): the workload and a plain copy of
COPY_BYTES bytes each way (``bench.py --copy-calibration COPY_BYTES``).  Each
counter (KiB per dispatch) is scaled by COPY_BYTES over what it reads for the
copy, which turns it into HBM bytes per dispatch.  ``--json`` writes the
per-kernel bytes per launch that bench.py reports as ``roofline.traffic``.

``--window WINDOW_JSON`` (both ``trace`` and ``traffic``) keeps only the
timed window's dispatches: the JSON line of ``bench.py --window-only`` names
the number of launches of each kernel in the window, and since nothing runs
after the window in that mode, they are the last that many dispatches of the
kernel.  Averages are then over the window, comparable with the live
per-launch timing of the same window in ``bench.py``'s JSON line.
"""
import argparse
import collections
import csv
import glob
import json
import os
import sys

# rocprofv3 kernel name -> the name liblm's own launch timer uses
TIMER_NAMES = {
    "advect_rk4_f64_kernel": "advect_rk4_f64", "field_blend_f64_kernel": "field_blend_f64",
    "bin_keys_kernel": "bin_keys", "bin_table_kernel": "bin_table", "pair_count_kernel": "pair_count",
    "pair_emit_kernel": "pair_emit", "radix_hist_kernel": "sort_hist", "radix_scatter_kernel": "sort_scatter",
    "scan_reduce_kernel": "scan_reduce", "scan_apply_kernel": "scan_apply", "rps_init_kernel": "rps_init",
    "rps_tile_union_kernel": "rps_tile",
    "rps_union_kernel": "rps_union", "rps_keys_kernel": "rps_keys", "rps_group_blk_kernel": "rps_group",
    "rps_comps_kernel": "rps_comps", "rps_fill_kernel": "rps_fill", "rps_deps_kernel": "rps_deps",
    "rps_levels_kernel": "rps_levels", "rps_play_kernel": "rps_play", "copy16_kernel": "copy16",
    "philox_kernel": "philox", "seg_copy_kernel": "copy16",
    "advect_rk4_f32_kernel": "advect_rk4_f32", "field_blend_f32_kernel": "field_blend_f32",
    "adj_forward_kernel": "adj_forward", "adj_recompute_kernel": "adj_recompute",
    "adj_backward_kernel": "adj_backward", "adj_blend_kernel": "adj_blend", "adj_slab_kernel": "adj_slab",
    "adj_zero_kernel": "adj_slab", "species_count_kernel": "species_count", "dist_hist_kernel": "dist_hist",
}


def short_name(full):
    """'void lm::radix_scatter_kernel<10>(...)' -> 'radix_scatter_kernel'."""
    name = full.split("(")[0].replace("void ", "").strip()
    if not name.startswith("lm::"):
        return None
    return name[4:].split("<")[0]


def one_file(d, pattern):
    """The newest file matching pattern under d (a directory may hold several runs)."""
    files = glob.glob(os.path.join(d, "**", pattern), recursive=True)
    if not files:
        sys.exit(f"no {pattern} under {d}")
    return max(files, key=os.path.getmtime)


def window_launches(path):
    """{rocprof short kernel name: launches in the timed window} from the JSON
    line of ``bench.py --window-only``."""
    if not path:
        return None
    with open(path) as fh:
        line = [ln for ln in fh if ln.lstrip().startswith("{")][-1]
    w = json.loads(line)
    by_timer = w["launches"]
    out = {}
    for k, tn in TIMER_NAMES.items():
        if tn in by_timer:
            out[k] = by_timer[tn]
    return out


def trace(args):
    acc = {}
    win = window_launches(args.window)
    if win is None:
        for r in csv.DictReader(open(one_file(args.dir, "*_kernel_stats.csv"))):
            k = short_name(r["Name"])
            if k is None:
                continue
            tot, calls = acc.get(k, (0.0, 0))
            acc[k] = (tot + float(r["TotalDurationNs"]), calls + int(r["Calls"]))
    else:
        per = collections.defaultdict(list)
        for r in csv.DictReader(open(one_file(args.dir, "*_kernel_trace.csv"))):
            k = short_name(r["Kernel_Name"])
            if k is not None:
                per[k].append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]) - int(r["Start_Timestamp"])))
        for k, rows in per.items():
            # kernels sharing a timer name (copy16, adj_slab) split the window's launches
            # between them; only kernels the window names are kept
            n = win.get(k)
            if not n:
                continue
            rows.sort()
            last = rows[-n:]
            acc[k] = (float(sum(d for _, d in last)), len(last))
    rows = [(tot, k, calls, tot / calls) for k, (tot, calls) in acc.items()]
    total = sum(r[0] for r in rows)
    out = [f"# {args.title}", ""]
    if args.command:
        out += ["Command:", "", "```", args.command, "```", ""]
    out += ["| kernel | timer name | calls | avg (us) | total (ms) | share |", "|---|---|---:|---:|---:|---:|"]
    for tot, k, calls, avg in sorted(rows, reverse=True):
        out.append(f"| {k} | {TIMER_NAMES.get(k, '')} | {calls} | {avg / 1e3:.2f} | {tot / 1e6:.2f} | "
                   f"{100 * tot / total:.1f}% |")
    if args.passes:
        kern, spec = args.passes.split("=", 1)
        counts, _, names = spec.partition(":")
        counts = [int(c) for c in counts.split(",")]
        names = names.split(",") if names else [f"pass {k}" for k in range(len(counts))]
        rows_k = [r for r in csv.DictReader(open(one_file(args.dir, "*_kernel_trace.csv")))
                  if short_name(r["Kernel_Name"]) == kern]
        rows_k.sort(key=lambda r: int(r["Start_Timestamp"]))
        dur = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in rows_k]
        out += ["", f"{kern} by pass ({len(dur)} dispatches in launch order):", "",
                "| pass | dispatches | avg (us) |", "|---|---:|---:|"]
        i = 0
        for c, nm in zip(counts, names):
            seg = dur[i:i + c]
            i += c
            if seg:
                out.append(f"| {nm} | {len(seg)} | {sum(seg) / len(seg):.2f} |")
    print("\n".join(out))


def counters(d, name, win=None):
    """Mean of counter `name` per dispatch, per kernel (over the last win[k]
    dispatches of kernel k when a window is given)."""
    acc = collections.defaultdict(dict)
    for r in csv.DictReader(open(one_file(d, "*_counter_collection.csv"))):
        if r["Counter_Name"] != name:
            continue
        k = short_name(r["Kernel_Name"])
        if k is not None:
            did = int(r["Dispatch_Id"])
            acc[k][did] = acc[k].get(did, 0.0) + float(r["Counter_Value"])
    out, calls = {}, {}
    for k, per in acc.items():
        vals = [per[i] for i in sorted(per)]
        if win is not None and k != "copy16_kernel":
            n = win.get(k)
            if not n:
                continue
            vals = vals[-n:]
        out[k] = sum(vals) / len(vals)
        calls[k] = len(vals)
    return out, calls


def traffic(args):
    win = window_launches(args.window)
    fetch, calls = counters(args.fetch_dir, "FETCH_SIZE", win)
    write, _ = counters(args.write_dir, "WRITE_SIZE", win)
    cf, _ = counters(args.copy_fetch_dir, "FETCH_SIZE")
    cw, _ = counters(args.copy_write_dir, "WRITE_SIZE")
    nb = args.copy_bytes
    sf = nb / (1024.0 * cf["copy16_kernel"])
    sw = nb / (1024.0 * cw["copy16_kernel"])
    out = ["# HBM traffic per dispatch (FETCH_SIZE + WRITE_SIZE)", ""]
    if args.command:
        out += ["Commands:", "", "```", args.command, "```", ""]
    out += [f"Copy calibration: a plain copy of {nb} bytes each way reads FETCH_SIZE "
            f"{cf['copy16_kernel']:.0f} KiB and WRITE_SIZE {cw['copy16_kernel']:.0f} KiB per dispatch, so "
            f"bytes = FETCH_SIZE x 1024 x {sf:.4f} and WRITE_SIZE x 1024 x {sw:.4f}.", "",
            "| kernel | timer name | dispatches | FETCH_SIZE (KiB) | WRITE_SIZE (KiB) | HBM bytes / dispatch |",
            "|---|---|---:|---:|---:|---:|"]
    res = {}
    for k in sorted(set(fetch) | set(write)):
        if k == "copy16_kernel":
            continue
        f, w = fetch.get(k, 0.0), write.get(k, 0.0)
        b = f * 1024 * sf + w * 1024 * sw
        tn = TIMER_NAMES.get(k, k)
        res[tn] = {"dispatches": calls.get(k, 0), "fetch_kib": round(f, 1), "write_kib": round(w, 1),
                   "bytes_per_launch": int(round(b))}
        out.append(f"| {k} | {tn} | {calls.get(k, 0)} | {f:.1f} | {w:.1f} | {b:,.0f} |")
    print("\n".join(out))
    if args.json:
        with open(args.json, "w") as fh:
            doc = {"copy_bytes": nb, "fetch_scale": sf, "write_scale": sw, "kernels": res}
            if args.window:
                with open(args.window) as wf:
                    w = json.loads([ln for ln in wf if ln.lstrip().startswith("{")][-1])
                doc["window"] = {"workload": w["workload"], "start": w["start"], "steps": w["steps"]}
            json.dump(doc, fh, indent=1)


# MI355X: 256 CUs (one address unit each) in 8 XCDs, 4 SIMDs per CU; a wave64
# VALU instruction issues once per 4 cycles per SIMD; clock 2.4 GHz
CUS, XCDS, SIMDS, CLOCK_HZ = 256, 8, 1024, 2.4e9


def units(args):
    """Per kernel over the window: VALU instructions per wavefront, the share
    of the VALU issue rate used, and the share of time the vector memory
    path's address unit (TA) was busy or stalled by the cache, from one or
    more ``--pmc`` passes (durations from each pass's own dispatches)."""
    win = window_launches(args.window)
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    dur = collections.defaultdict(dict)
    for d in args.dirs:
        by_kernel = collections.defaultdict(lambda: collections.defaultdict(dict))
        for r in csv.DictReader(open(one_file(d, "*_counter_collection.csv"))):
            k = short_name(r["Kernel_Name"])
            if k is None:
                continue
            did = int(r["Dispatch_Id"])
            by_kernel[k][did][r["Counter_Name"]] = (by_kernel[k][did].get(r["Counter_Name"], 0.0)
                                                     + float(r["Counter_Value"]))
            dur[(d, k)][did] = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-9
        for k, disp in by_kernel.items():
            ids = sorted(disp)
            if win is not None:
                n = win.get(k)
                if not n:
                    continue
                ids = ids[-n:]
            for did in ids:
                for name, v in disp[did].items():
                    per[k][name] += v / len(ids)
            per[k]["_dispatches"] = len(ids)
            per[k][f"_dur:{d}"] = sum(dur[(d, k)][i] for i in ids) / len(ids)
    out = [f"# {args.title}", ""]
    if args.command:
        out += ["Commands:", "", "```", args.command, "```", ""]
    out += [f"Per dispatch of the window. VALU share = SQ_INSTS_VALU x 4 cycles / (duration x {SIMDS} SIMDs x "
            f"{CLOCK_HZ / 1e9:.1f} GHz); TA busy = (TA_TA_BUSY_sum / {CUS} address units) / (GRBM_GUI_ACTIVE / "
            f"{XCDS} XCDs), TA stalled likewise from TA_ADDR_STALLED_BY_TC_CYCLES_sum.", "",
            "| kernel | timer name | dispatches | avg us | VALU / wave | VALU share | TA busy | TA stalled by cache |",
            "|---|---|---:|---:|---:|---:|---:|---:|"]
    rows = []
    for k, c in per.items():
        durs = [v for name, v in c.items() if name.startswith("_dur:")]
        t = sum(durs) / len(durs) if durs else 0.0
        valu = c.get("SQ_INSTS_VALU")
        waves = c.get("SQ_WAVES")
        grbm = c.get("GRBM_GUI_ACTIVE")
        ta = c.get("TA_TA_BUSY_sum")
        st = c.get("TA_ADDR_STALLED_BY_TC_CYCLES_sum")
        f = lambda x: "" if x is None else f"{x:.2f}"
        rows.append((t, f"| {k} | {TIMER_NAMES.get(k, '')} | {int(c['_dispatches'])} | {t * 1e6:.1f} | "
                        f"{'' if not (valu and waves) else f'{valu / waves:.0f}'} | "
                        f"{f(None if valu is None or not t else valu * 4 / (t * SIMDS * CLOCK_HZ))} | "
                        f"{f(None if not (ta and grbm) else (ta / CUS) / (grbm / XCDS))} | "
                        f"{f(None if not (st and grbm) else (st / CUS) / (grbm / XCDS))} |"))
    out += [r for _, r in sorted(rows, reverse=True)]
    print("\n".join(out))


def _fmt_rate(x):
    return f"{x:.2e}".replace("e+0", "e+")


def legs(args):
    d = json.load(open(args.bench_json))
    cmd = ["Command:", "", "```", args.command, "```", ""] if args.command else []
    f = d.get("fp32_path")
    if f:
        out = [f"# FP32 path: {f['workload']} ({f['particles']:,} particles, {f['steps']:,} steps of "
               f"{f['dt_s']:g} s; field: {f['field']})", ""] + cmd
        out += ["Fused run (advection, pair search, RPS, records) from identical inputs in both precisions, "
                "chunks pipelined:", "", "| precision | ms/step | particle-steps/s | steps outside the field |",
                "|---|---:|---:|---:|"]
        for prec in ("f32", "f64"):
            r = f[prec]
            out.append(f"| {prec} | {r['ms_per_step']:.4f} | {_fmt_rate(r['particle_steps_per_s'])} | "
                       f"{r['particle_steps_outside_field']} |")
        out += ["", "Distance between the FP32 and FP64 positions (degrees):", "",
                "| after | median | 99th percentile | max |", "|---|---:|---:|---:|"]
        for k, v in f["f32_vs_f64_position_error"].items():
            out.append(f"| {k.replace('after_', '').replace('_', ' ')} | {v['median_deg']:.3g} | "
                       f"{v['p99_deg']:.3g} | {v['max_deg']:.3g} |")
        out += ["", "JSON object:", "", "```", json.dumps(f, indent=1), "```"]
        with open(os.path.join(args.out_dir, f"bench_{f['workload']}.md"), "w") as fh:
            fh.write("\n".join(out) + "\n")
    e = d.get("ensemble")
    if e:
        out = [f"# Ensemble: {e['workload']} ({e['runs']} runs, {e['particles']:,} particles, "
               f"{e['steps']} steps; field: {e['field']}; FP32)", ""] + cmd
        a = e["autotuned"]
        out += [f"Autotuner choice: cell side {a['cell_scale']} radius, chunk {a['chunk']} steps, "
                f"replay workgroups {a['play_blocks']}.", "",
                f"| | seconds for {e['steps']} steps | particle-steps/s |", "|---|---:|---:|",
                f"| batched ensemble (best of two) | {e['ms_per_step'] * e['steps'] / 1e3:.3f} | "
                f"{_fmt_rate(e['particle_steps_per_s'])} |",
                f"| the {e['runs']} runs alone, one after another (best of two each) | {e['runs_alone_s']:.3f} | "
                f"{_fmt_rate(e['runs_alone_particle_steps_per_s'])} |", "",
                f"Speedup {e['speedup_vs_runs_alone']:.2f}x. A run alone takes "
                f"{e['runs_alone_ms_per_step']['min']}-{e['runs_alone_ms_per_step']['max']} ms per step "
                f"(median {e['runs_alone_ms_per_step']['median']}).", "",
                "Autotuner table (cell scale, chunk, replay workgroups, ms/step over 24 steps), fastest first:", "",
                "| cell scale | chunk | replay (wide, narrow) | ms/step |", "|---:|---:|---|---:|"]
        for c, k, pb, t in sorted(e["tuning_ms_per_step"], key=lambda r: r[3]):
            out.append(f"| {c} | {k} | ({pb[0]}, {pb[1]}) | {t} |")
        slim = {k: v for k, v in e.items() if k != "tuning_ms_per_step"}
        out += ["", "JSON object (without the tuning table):", "", "```", json.dumps(slim, indent=1), "```"]
        with open(os.path.join(args.out_dir, f"bench_{e['workload']}.md"), "w") as fh:
            fh.write("\n".join(out) + "\n")


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("dir")
    t.add_argument("--title", default="Kernel trace")
    t.add_argument("--command", default="")
    t.add_argument("--passes", default="")
    t.add_argument("--window", default="")
    c = sub.add_parser("traffic")
    for a in ("fetch_dir", "write_dir", "copy_fetch_dir", "copy_write_dir"):
        c.add_argument(a)
    c.add_argument("copy_bytes", type=int)
    c.add_argument("--json")
    c.add_argument("--command", default="")
    c.add_argument("--window", default="")
    u = sub.add_parser("units")
    u.add_argument("dirs", nargs="+")
    u.add_argument("--title", default="Compute-unit pipes per kernel")
    u.add_argument("--command", default="")
    u.add_argument("--window", default="")
    g = sub.add_parser("legs")
    g.add_argument("bench_json")
    g.add_argument("out_dir")
    g.add_argument("--command", default="")
    args = ap.parse_args()
    {"trace": trace, "traffic": traffic, "units": units, "legs": legs}[args.cmd](args)


if __name__ == "__main__":
    main()
