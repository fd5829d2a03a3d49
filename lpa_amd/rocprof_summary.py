"""Text summaries of rocprofv3 output for the benchmark's headline phase.

The benchmark's profiled command (bench.py --headline-only ...) builds the
graph, lays it out (with --no-learned the degree schedule; otherwise the
judged learned schedule), the last k_relayout dispatch, and then runs
warmup + steps iterations.
Everything dispatched after that last relayout belongs to those iterations,
except the copy kernel of the calibration (bench.py --copy-calibration:
k_calibration_copy over 256 MiB), whose FETCH_SIZE / WRITE_SIZE readings
give the factor between the counters and bytes actually moved.

Usage:
  python -m lpa_amd.rocprof_summary --trace DIR [--fetch DIR --write DIR]
      --iters N --steps S [--pmc-iters N --pmc-steps S] --command "..."
      --workload W --K K --out PREFIX
writes PREFIX_kernels.txt and PREFIX_traffic.json.
"""
import argparse
import collections
import csv
import json
import os
import re

COPY_BYTES = 256 * 1024 * 1024


def family(name):
    """kernel family (bin kind) of a dispatch, from its demangled name"""
    m = re.search(r"lpa::(k_\w+)(<[^(]*>)?", name)
    if not m:
        return "other"
    k, targs = m.group(1), (m.group(2) or "")
    args = [a.strip() for a in targs.strip("<>").split(",")] if targs else []
    if k == "k_lane":
        return "lane"
    if k == "k_group":
        return "block" if args and args[0] == "256" else "wave"
    if k in ("k_heavy_decode", "k_heavy_clear", "k_heavy_finish", "k_heavy_local"):
        return "cms" if args and args[0] == "1" else "global"
    if k == "k_heavy_chunks":
        return "global" if args and args[0] == "0" else "cms"
    if k == "k_calibration_copy":
        return "copy"
    if k in ("k_fill_u64",):
        return "step"
    if k.startswith("k_decode_csr"):
        return "cglp-decode"
    return "other"


def _rows(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _headline(rows, iters, steps):
    """step-kernel rows of the last `steps` of the `iters` iterations that
    follow the last k_relayout dispatch (every iteration dispatches the same
    kernel sequence), plus the calibration copies"""
    rows = sorted(rows, key=lambda r: int(r["Dispatch_Id"]))
    last = max((i for i, r in enumerate(rows) if "k_relayout" in r["Kernel_Name"]), default=-1)
    rows = rows[last + 1:]
    step = [r for r in rows if family(r["Kernel_Name"]) not in ("copy", "other")]
    if len(step) % iters:
        raise SystemExit(f"{len(step)} step dispatches do not split into {iters} iterations")
    per = len(step) // iters
    keep = step[(iters - steps) * per:]
    return keep + [r for r in rows if family(r["Kernel_Name"]) == "copy"]


def trace_summary(trace_dir, iters, steps):
    rows = _headline(_rows(os.path.join(trace_dir, "run_kernel_trace.csv")), iters, steps)
    per = collections.defaultdict(lambda: [0, 0.0])
    for r in rows:
        d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3   # us
        k = r["Kernel_Name"]
        per[k][0] += 1
        per[k][1] += d
    fam = collections.defaultdict(float)
    for k, (c, t) in per.items():
        if family(k) != "copy":
            fam[family(k)] += t / steps / 1e3   # ms per iteration
    return per, fam


def counter_summary(dirname, counter, iters, steps):
    rows = [r for r in _rows(os.path.join(dirname, "run_counter_collection.csv")) if r["Counter_Name"] == counter]
    rows = _headline(rows, iters, steps)
    fam = collections.defaultdict(float)
    copy = []
    for r in rows:
        f = family(r["Kernel_Name"])
        v = float(r["Counter_Value"]) * 1024.0   # KiB -> bytes
        if f == "copy":
            copy.append(v)
        else:
            fam[f] += v / steps
    return fam, copy


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--iters", type=int, required=True, help="warmup + timed iterations of the trace run")
    ap.add_argument("--steps", type=int, required=True, help="timed iterations of the trace run (summarised)")
    ap.add_argument("--pmc-iters", type=int, help="warmup + timed iterations of the counter runs")
    ap.add_argument("--pmc-steps", type=int, help="timed iterations of the counter runs")
    ap.add_argument("--command", required=True, help="profiled command lines, ';'-separated")
    ap.add_argument("--workload", required=True)
    ap.add_argument("--K", type=int, required=True)
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    per, fam = trace_summary(a.trace, a.iters, a.steps)
    lines = [f"rocprofv3 kernel trace, headline phase of {a.workload} K={a.K}: the {a.steps} timed iterations "
             f"(of {a.iters}) after the last layout build", "commands:"]
    lines += [f"  {c.strip()}" for c in a.command.split(";") if c.strip()]
    lines.append("")
    lines.append(f"{'kernel':78s} {'calls':>6s} {'avg_us':>9s} {'us/iter':>9s} family")
    for k, (c, t) in sorted(per.items(), key=lambda x: -x[1][1]):
        lines.append(f"{k[:78]:78s} {c:6d} {t / c:9.1f} {t / a.steps:9.1f} {family(k)}")
    lines.append("")
    lines.append("per family (ms per iteration): " +
                 ", ".join(f"{f} {v:.3f}" for f, v in sorted(fam.items(), key=lambda x: -x[1])))
    out = {"workload": a.workload, "K": a.K, "steps": a.steps, "ms_per_iter": dict(fam)}
    if a.fetch and a.write:
        pi, ps = a.pmc_iters or a.iters, a.pmc_steps or a.steps
        fetch, cf = counter_summary(a.fetch, "FETCH_SIZE", pi, ps)
        write, cw = counter_summary(a.write, "WRITE_SIZE", pi, ps)
        ff = COPY_BYTES / (sum(cf) / len(cf)) if cf else 1.0
        fw = COPY_BYTES / (sum(cw) / len(cw)) if cw else 1.0
        lines.append("")
        lines.append(f"copy calibration (256 MiB k_calibration_copy): FETCH_SIZE factor {ff:.3f}, WRITE_SIZE factor {fw:.3f}"
                     f" ({len(cf)} / {len(cw)} copies)")
        lines.append(f"{'family':10s} {'fetch MB/iter':>14s} {'write MB/iter':>14s} {'HBM GB/s':>10s}")
        traffic = {}
        for f in sorted(set(fetch) | set(write)):
            fb, wb = fetch.get(f, 0.0) * ff, write.get(f, 0.0) * fw
            ms = fam.get(f, 0.0)
            gbs = (fb + wb) / (ms / 1e3) / 1e9 if ms else 0.0
            traffic[f] = {"fetch_bytes": fb, "write_bytes": wb, "ms": ms}
            lines.append(f"{f:10s} {fb / 1e6:14.1f} {wb / 1e6:14.1f} {gbs:10.1f}")
        out["traffic_per_iter"] = traffic
        out["copy_factor"] = {"fetch": ff, "write": fw}
    out["commands"] = [c.strip() for c in a.command.split(";") if c.strip()]
    with open(a.out + "_kernels.txt", "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(a.out + "_traffic.json", "w") as f:
        json.dump(out, f, indent=1)
    print("\n".join(lines))


if __name__ == "__main__":
    main()
