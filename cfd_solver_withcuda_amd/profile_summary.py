"""Turns rocprofv3 CSV output into the short text summaries kept under profiles/.

    python -m cfd_solver_withcuda_amd.profile_summary \\
        --title "..." --command "rocprofv3 ..." --stats <dir>/<name>_kernel_stats.csv \\
        [--fetch <fetch counter_collection.csv> --write <write counter_collection.csv> \\
         --copy-fetch <csv> --copy-write <csv> --copy-bytes 536870912 \\
         --kernel flux_tile_kernel --algorithmic-bytes N --config "tile=128 ..."] > profiles/x.md

The FETCH_SIZE / WRITE_SIZE counters are reported in KiB per dispatch. They are scaled by the
ratio of the bytes a plain device copy of known size moves to what the same counter reads for
that copy (`--copy-*`), so the traffic line is in HBM bytes per launch.
"""
from __future__ import annotations

import argparse
import csv
import sys


def read_stats(path):
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            rows.append(dict(name=r["Name"], calls=int(r["Calls"]), total_ns=float(r["TotalDurationNs"]),
                             avg_ns=float(r["AverageNs"]), pct=float(r["Percentage"]),
                             min_ns=float(r["MinNs"]), max_ns=float(r["MaxNs"])))
    return rows


def counter_per_dispatch(path, kernel, counter, ordinals=None):
    """Average counter value (KiB) over the dispatches whose kernel name contains `kernel`;
    `ordinals` (start, stop) keeps only those dispatches of that kernel, counted from 0 in
    dispatch order."""
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == counter and kernel in r["Kernel_Name"]:
                rows.append((int(r["Dispatch_Id"]), float(r["Counter_Value"])))
    rows.sort()
    vals = [v for _, v in rows]
    if ordinals:
        vals = vals[ordinals[0]:ordinals[1]]
    if not vals:
        raise SystemExit(f"no {counter} rows for {kernel} in {path}")
    return sum(vals) / len(vals), len(vals)


def durations(path, kernel):
    """Durations (us) of the dispatches whose kernel name contains `kernel`, in dispatch order,
    from a rocprofv3 kernel-trace CSV."""
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            if kernel in r["Kernel_Name"]:
                rows.append((int(r["Dispatch_Id"]), (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3))
    rows.sort()
    return [d for _, d in rows]


def short(name):
    return name.replace("HIP_vector_type<double, 4u>", "double4").replace("cfd::", "")


def main(argv=None):
    p = argparse.ArgumentParser()
    p.add_argument("--title", required=True)
    p.add_argument("--command", action="append", required=True)
    p.add_argument("--stats", required=True)
    p.add_argument("--fetch")
    p.add_argument("--write")
    p.add_argument("--copy-fetch")
    p.add_argument("--copy-write")
    p.add_argument("--copy-bytes", type=float, default=512 * 1024 * 1024)
    p.add_argument("--copy-kernel", default="copy")
    p.add_argument("--kernel", default="flux_tile_kernel")
    p.add_argument("--algorithmic-bytes", type=float)
    p.add_argument("--ordinals", default="",
                   help="start:stop of the kernel's dispatches to average (default: all)")
    p.add_argument("--config", default="")
    p.add_argument("--trace", help="kernel-trace CSV of the --stats run: durations of single dispatches")
    p.add_argument("--trace-ordinals", default="",
                   help="start:stop of the kernel's dispatches in the trace that bench.py times alone")
    p.add_argument("--stages", default="",
                   help="start:stop:stride: per-stage average duration over those dispatches")
    p.add_argument("--extra-kernel", action="append", default=[],
                   help="NAME=ALGORITHMIC_BYTES: a further kernel's traffic (all its dispatches)")
    p.add_argument("--note", action="append", default=[])
    a = p.parse_args(argv)

    out = [f"# {a.title}", "", "Commands (run from the repository root):", ""]
    out += [f"    {c}" for c in a.command]
    out += ["", "## Kernel time (rocprofv3 --kernel-trace --stats)", "",
            "| kernel | calls | avg us | min us | max us | % time |", "|---|---|---|---|---|---|"]
    for r in read_stats(a.stats):
        out.append(f"| `{short(r['name'])}` | {r['calls']} | {r['avg_ns'] / 1e3:.1f} | "
                   f"{r['min_ns'] / 1e3:.1f} | {r['max_ns'] / 1e3:.1f} | {r['pct']:.2f} |")
    if a.trace and a.trace_ordinals:
        d = durations(a.trace, a.kernel)
        t0, t1 = (int(x) for x in a.trace_ordinals.split(":"))
        sel = d[t0:t1]
        out += ["", f"## Launches of `{a.kernel}` timed alone (rocprofv3 kernel trace, dispatches {t0}-{t1 - 1})", "",
                f"- {len(sel)} launches, average {sum(sel) / len(sel):.1f} us "
                f"({', '.join(f'{x:.1f}' for x in sel)})"]
        if a.stages:
            s0, s1, k = (int(x) for x in a.stages.split(":"))
            means = [sum(d[s0 + j:s1:k]) / len(d[s0 + j:s1:k]) for j in range(k)]
            out.append(f"- average per stage over dispatches {s0}-{s1 - 1}: "
                       + ", ".join(f"stage {j} {m:.1f} us" for j, m in enumerate(means)))
    if a.fetch and a.write and a.copy_fetch and a.copy_write:
        ords = tuple(int(x) for x in a.ordinals.split(":")) if a.ordinals else None
        f_kib, nf = counter_per_dispatch(a.fetch, a.kernel, "FETCH_SIZE", ords)
        w_kib, nw = counter_per_dispatch(a.write, a.kernel, "WRITE_SIZE", ords)
        cf_kib, _ = counter_per_dispatch(a.copy_fetch, a.copy_kernel, "FETCH_SIZE")
        cw_kib, _ = counter_per_dispatch(a.copy_write, a.copy_kernel, "WRITE_SIZE")
        kf = a.copy_bytes / (cf_kib * 1024)
        kw = a.copy_bytes / (cw_kib * 1024)
        fetch_b = f_kib * 1024 * kf
        write_b = w_kib * 1024 * kw
        traffic = fetch_b + write_b
        out += ["", f"## HBM traffic of `{a.kernel}` (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE)", "",
                f"- copy calibration: a device copy of {a.copy_bytes:.0f} B reads FETCH_SIZE "
                f"{cf_kib:.0f} KiB (factor {kf:.3f}) and WRITE_SIZE {cw_kib:.0f} KiB (factor {kw:.3f})",
                f"- FETCH_SIZE {f_kib:.0f} KiB/dispatch over {nf} dispatches -> {fetch_b / 1e6:.1f} MB read",
                f"- WRITE_SIZE {w_kib:.0f} KiB/dispatch over {nw} dispatches -> {write_b / 1e6:.1f} MB written"]
        if a.algorithmic_bytes:
            out.append(f"- algorithmic bytes per launch {a.algorithmic_bytes / 1e6:.1f} MB; "
                       f"measured / algorithmic = {traffic / a.algorithmic_bytes:.3f}")
        out += ["", f"traffic_bytes_per_launch: {traffic:.0f}", f"traffic_config: {a.config}"]
        for spec in a.extra_kernel:
            name, _, alg = spec.rpartition("=")
            f2, n2 = counter_per_dispatch(a.fetch, name, "FETCH_SIZE")
            w2, _ = counter_per_dispatch(a.write, name, "WRITE_SIZE")
            t2 = f2 * 1024 * kf + w2 * 1024 * kw
            out += ["", f"## HBM traffic of `{name}`", "",
                    f"- FETCH_SIZE {f2:.0f} KiB, WRITE_SIZE {w2:.0f} KiB per dispatch ({n2} dispatches) -> "
                    f"{f2 * 1024 * kf / 1e6:.1f} MB read, {w2 * 1024 * kw / 1e6:.1f} MB written",
                    f"- algorithmic bytes per launch {float(alg) / 1e6:.1f} MB; measured / algorithmic = "
                    f"{t2 / float(alg):.3f}"]
    if a.note:
        out += [""] + [f"- {n}" for n in a.note]
    sys.stdout.write("\n".join(out) + "\n")


if __name__ == "__main__":
    main()
