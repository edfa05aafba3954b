"""Summaries of rocprofv3 output for the profiles/ folder.

`stats`: the kernel-stats CSV of a `rocprofv3 --kernel-trace --stats -f csv` run as a
Markdown table, with the command that made it.

`valu`: executed VALU instructions per kernel launch from a counter pass over SQ_INSTS_VALU,
SQ_INSTS_VALU_INT32/INT64, SQ_ACTIVE_INST_VALU, SQ_BUSY_CYCLES and GRBM_GUI_ACTIVE: the
integer instruction mix, the issue rate and the VALU-busy fraction of each kernel.

`traffic`: HBM traffic per kernel launch from two counter passes (FETCH_SIZE, WRITE_SIZE).
Each pass is a `rocprofv3 --pmc <counter> -f csv` run of `bench.py --calibrate-copy-mb M`,
which copies M MiB device to device before the timed work.  The counters' readings for that
plain copy (M MiB read, M MiB written) give one scale factor per counter; every kernel's mean
reading per dispatch is multiplied by it.  Writes a JSON summary (read by bench.py for
`roofline.traffic`) and a Markdown table.

    python3 -m fk20_amd.profsum traffic --fetch F.csv --write W.csv --copy-mb 256 \
        --config N=4096,rows=512,window=12 --json out.json --md out.md --command "..."
    python3 -m fk20_amd.profsum stats --csv run_kernel_stats.csv --title "..." \
        --command "..." --md out.md [--note "..."]
"""
import argparse
import collections
import csv
import json

COPY_KERNEL = "__amd_rocclr_copyBuffer"


def _kernel(name):
    name = name.split("(")[0]
    return name[5:] if name.startswith("void ") else name


def _grid(r):
    """Work-items of a dispatch, from a counter-collection row (Grid_Size) or a kernel-trace
    row (Grid_Size_X/Y/Z)."""
    if "Grid_Size" in r:
        return int(r["Grid_Size"])
    return int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])


def _workload(rows):
    """The dispatches of the benchmarked workload: per kernel, those with at least a quarter of
    the kernel's largest grid.  This drops the setup's small launches of the same kernels (the
    16-row G1 FFT of fk20_setup2xext_fft) while keeping every launch of a pass (the twiddle
    layers of one pass differ in size by at most 2x)."""
    big = collections.defaultdict(int)
    for r in rows:
        big[_kernel(r["Kernel_Name"])] = max(big[_kernel(r["Kernel_Name"])], _grid(r))
    return [r for r in rows if 4 * _grid(r) >= big[_kernel(r["Kernel_Name"])]]


def read_pass(path, counter):
    """{kernel: [KiB per dispatch]} for one counter pass (workload dispatches only)."""
    out = collections.defaultdict(list)
    with open(path) as f:
        rows = [r for r in csv.DictReader(f) if r["Counter_Name"] == counter]
    for r in _workload(rows):
        out[_kernel(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return out


def summarize(fetch, write, copy_mb):
    copy_bytes = copy_mb << 20
    # the calibration copy is the largest copy dispatch of the run
    f_copy, w_copy = max(fetch[COPY_KERNEL]), max(write[COPY_KERNEL])
    f_scale = copy_bytes / (f_copy * 1024)
    w_scale = copy_bytes / (w_copy * 1024)
    kernels = {}
    for k in sorted(set(fetch) | set(write)):
        if k.startswith("__amd_rocclr"):
            continue
        fv, wv = fetch.get(k, []), write.get(k, [])
        kernels[k] = {
            "dispatches": len(fv),
            "fetch_bytes": round(sum(fv) / max(1, len(fv)) * 1024 * f_scale),
            "write_bytes": round(sum(wv) / max(1, len(wv)) * 1024 * w_scale),
        }
    return {"fetch_scale": round(f_scale, 4), "write_scale": round(w_scale, 4),
            "copy_fetch_kib": f_copy, "copy_write_kib": w_copy, "copy_mb": copy_mb,
            "kernels": kernels}


VALU_COUNTERS = ("SQ_INSTS_VALU", "SQ_INSTS_VALU_INT32", "SQ_INSTS_VALU_INT64", "SQ_ACTIVE_INST_VALU",
                 "SQ_INSTS_SALU", "SQ_WAVES", "SQ_BUSY_CYCLES", "GRBM_GUI_ACTIVE")
SIMDS = 1024  # 256 CUs x 4 SIMDs
XCDS = 8      # GRBM_GUI_ACTIVE is summed over the 8 XCDs


def valu_summary(counter_csv, trace_csv):
    """Per kernel: mean counters per dispatch, mean duration (from a kernel trace of the same
    workload), lane-instruction rate and VALU-busy fraction."""
    cnt = collections.defaultdict(lambda: collections.defaultdict(list))
    with open(counter_csv) as f:
        for r in _workload(list(csv.DictReader(f))):
            cnt[_kernel(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    dur = collections.defaultdict(list)
    with open(trace_csv) as f:
        for r in _workload(list(csv.DictReader(f))):
            dur[_kernel(r["Kernel_Name"])].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-9)
    out = {}
    for k, c in cnt.items():
        if k.startswith("__amd_rocclr") or "SQ_INSTS_VALU" not in c:
            continue
        m = {n: sum(v) / len(v) for n, v in c.items()}
        if not m["SQ_INSTS_VALU"]:  # a launch too small to register (a fill of a few words)
            continue
        t = sum(dur[k]) / len(dur[k]) if dur.get(k) else None
        cycles = m.get("GRBM_GUI_ACTIVE", 0) / XCDS
        e = {"dispatches": len(c["SQ_INSTS_VALU"]),
             "valu_wave_instr": m["SQ_INSTS_VALU"],
             "int64_share": m.get("SQ_INSTS_VALU_INT64", 0) / m["SQ_INSTS_VALU"],
             "int32_share": m.get("SQ_INSTS_VALU_INT32", 0) / m["SQ_INSTS_VALU"],
             "salu_instr": m.get("SQ_INSTS_SALU", 0), "waves": m.get("SQ_WAVES", 0),
             "gpu_cycles": cycles,
             # SQ_ACTIVE_INST_VALU counts quad-cycles: one per wave64 instruction on a SIMD16
             "valu_busy": 4 * m.get("SQ_ACTIVE_INST_VALU", 0) / (SIMDS * cycles) if cycles else None}
        if t:
            e["avg_s"] = t
            e["clock_ghz"] = cycles / t / 1e9 if cycles else None
            e["lane_instr_per_s"] = m["SQ_INSTS_VALU"] * 64 / t
        out[k] = e
    return out


def stats_md(csv_path, title, command, note=""):
    rows = list(csv.DictReader(open(csv_path)))
    lines = ["# " + title, "", "Command (from the repository root on one MI355X):", "", "```", command.strip(), "```", ""]
    if note:
        lines += [note.strip(), ""]
    lines += ["| kernel | calls | avg (ms) | min (ms) | max (ms) | total (ms) | % |", "|---|---:|---:|---:|---:|---:|---:|"]
    for r in rows:
        lines.append("| `%s` | %s | %.3f | %.3f | %.3f | %.1f | %.2f |" % (
            _kernel(r["Name"]), r["Calls"], float(r["AverageNs"]) / 1e6, float(r["MinNs"]) / 1e6,
            float(r["MaxNs"]) / 1e6, float(r["TotalDurationNs"]) / 1e6, float(r["Percentage"])))
    return "\n".join(lines) + "\n"


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--copy-mb", type=int, required=True)
    t.add_argument("--config", required=True, help="workload, e.g. N=4096,rows=512,window=12")
    t.add_argument("--json", required=True)
    t.add_argument("--md", required=True)
    t.add_argument("--command", default="", help="the rocprofv3 commands, for the record")
    v = sub.add_parser("valu")
    v.add_argument("--counters", required=True)
    v.add_argument("--trace", required=True, help="kernel_trace.csv of a run of the same workload")
    v.add_argument("--config", required=True)
    v.add_argument("--json", required=True)
    v.add_argument("--md", required=True)
    v.add_argument("--command", default="")
    k = sub.add_parser("stats")
    k.add_argument("--csv", required=True)
    k.add_argument("--title", required=True)
    k.add_argument("--command", required=True)
    k.add_argument("--note", default="")
    k.add_argument("--md", required=True)
    args = ap.parse_args()
    if args.cmd == "valu":
        kern = valu_summary(args.counters, args.trace)
        cfg = {a: int(b) for a, b in (x.split("=") for x in args.config.split(","))}
        with open(args.json, "w") as f:
            json.dump({"config": cfg, "command": args.command, "kernels": kern}, f, indent=1, sort_keys=True)
        lines = ["# Executed VALU instructions per launch (rocprofv3 SQ counters)", "",
                 "Workload: %s.  Counters are summed over all SEs; GRBM_GUI_ACTIVE over the 8 XCDs "
                 "(cycles = value / 8); SQ_ACTIVE_INST_VALU is in quad-cycles, so VALU busy = "
                 "4 x SQ_ACTIVE_INST_VALU / (1024 SIMDs x cycles).  Durations come from the kernel "
                 "trace of the same workload.  Per kernel, only dispatches with at least a quarter of "
                 "its largest grid are counted (the setup's 16-row launches of the transform kernels "
                 "are left out)." % ", ".join("%s=%d" % kv for kv in sorted(cfg.items())), "",
                 "| kernel | dispatches | VALU wave-instr / launch | INT64 share | INT32 share | avg ms | "
                 "T lane-instr/s | clock GHz | VALU busy |", "|---|---:|---:|---:|---:|---:|---:|---:|---:|"]
        for name, e in sorted(kern.items(), key=lambda kv: -kv[1]["valu_wave_instr"] * kv[1]["dispatches"]):
            lines.append("| `%s` | %d | %.4g | %.3f | %.3f | %s | %s | %s | %s |" % (
                name, e["dispatches"], e["valu_wave_instr"], e["int64_share"], e["int32_share"],
                "%.3f" % (e["avg_s"] * 1e3) if e.get("avg_s") else "-",
                "%.2f" % (e["lane_instr_per_s"] / 1e12) if e.get("lane_instr_per_s") else "-",
                "%.2f" % e["clock_ghz"] if e.get("clock_ghz") else "-",
                "%.3f" % e["valu_busy"] if e.get("valu_busy") is not None else "-"))
        lines += ["", "Commands:", "", "```", args.command.strip(), "```", ""]
        with open(args.md, "w") as f:
            f.write("\n".join(lines))
        return
    if args.cmd == "stats":
        with open(args.md, "w") as f:
            f.write(stats_md(args.csv, args.title, args.command, args.note))
        return
    s = summarize(read_pass(args.fetch, "FETCH_SIZE"), read_pass(args.write, "WRITE_SIZE"), args.copy_mb)
    s["config"] = {k: int(v) for k, v in (x.split("=") for x in args.config.split(","))}
    s["command"] = args.command
    with open(args.json, "w") as f:
        json.dump(s, f, indent=1, sort_keys=True)
    lines = ["# HBM traffic per launch (rocprofv3 FETCH_SIZE / WRITE_SIZE, calibrated)", "",
             "Workload: %s." % ", ".join("%s=%d" % kv for kv in sorted(s["config"].items())), "",
             "Calibration: a %d MiB device-to-device copy read FETCH_SIZE = %.1f KiB and WRITE_SIZE = "
             "%.1f KiB, so FETCH_SIZE is scaled by %.4f and WRITE_SIZE by %.4f.  Per kernel, only "
             "dispatches with at least a quarter of its largest grid are counted (the setup's 16-row "
             "launches of the transform kernels are left out)." %
             (args.copy_mb, s["copy_fetch_kib"], s["copy_write_kib"], s["fetch_scale"], s["write_scale"]), "",
             "| kernel | dispatches | fetch MB / launch | write MB / launch |", "|---|---:|---:|---:|"]
    for name, v in s["kernels"].items():
        lines.append("| `%s` | %d | %.2f | %.2f |" % (name, v["dispatches"], v["fetch_bytes"] / 1e6, v["write_bytes"] / 1e6))
    lines += ["", "Commands:", "", "```", args.command.strip(), "```", ""]
    with open(args.md, "w") as f:
        f.write("\n".join(lines))


if __name__ == "__main__":
    main()
