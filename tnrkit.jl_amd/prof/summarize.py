"""Turn rocprofv3 CSV output into the short summaries kept under profiles/.

  python tnrkit.jl_amd/prof/summarize.py stats KERNEL_STATS_CSV --command CMD \
      [--trace KERNEL_TRACE_CSV --kernel NAME --phases name:count,... [--also K1,K2]] \
      > profiles/X.md
  python tnrkit.jl_amd/prof/summarize.py traffic FETCH_CSV WRITE_CSV CAL_FETCH_CSV CAL_WRITE_CSV \
      --command CMD > profiles/traffic_X.json

traffic: FETCH_SIZE / WRITE_SIZE are KiB per dispatch.  Each is scaled by the
ratio of the known byte count of a plain device copy (bench.py --workload
copy: 256 MiB read, 256 MiB written) to what the same counter reads for it.
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import re

COPY_KIB = 256 * 1024
# the FP64 GEMM launches: LDS-DMA tiles and the 32-deep register-staged tile
FP64_GEMM = ("k_gemm_dl", "k_gemm_f64")


def short(name):
    m = re.match(r"(?:void )?tnr::(k_\w+)(<[^>]*>)?", name)
    return (m.group(1) + (m.group(2) or "")) if m else name.split("(")[0][:48]


def stats(path, command):
    rows = list(csv.DictReader(open(path)))
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    out = ["# rocprofv3 kernel statistics", "", "Command:", "", "```", command, "```", "",
           "| kernel | calls | total ms | avg us | min us | max us | % |",
           "|---|---|---|---|---|---|---|"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        out.append("| %s | %s | %.3f | %.1f | %.1f | %.1f | %.2f |" % (
            short(r["Name"]), r["Calls"], float(r["TotalDurationNs"]) / 1e6,
            float(r["AverageNs"]) / 1e3, float(r["MinNs"]) / 1e3, float(r["MaxNs"]) / 1e3,
            100.0 * float(r["TotalDurationNs"]) / total))
    out.append("")
    out.append("Total kernel time: %.1f ms" % (total / 1e6))
    return "\n".join(out)


def _union_ms(spans):
    """Length of the union of [start, end) intervals (ns in, ms out)."""
    busy, a, b = 0, 0, -1
    for s, e in sorted(spans):
        if s > b:
            if b > a:
                busy += b - a
            a, b = s, e
        elif e > b:
            b = e
    if b > a:
        busy += b - a
    return busy / 1e6


def phases(trace, kernel, spec):
    """Per consecutive group of the dispatches of `kernel` (a name or a
    comma-separated list of names) in start order (e.g. warmup, timed steps,
    a last pass), from the trace: the average duration, the summed durations
    and the union of the dispatch intervals (dispatches on different streams
    overlap; the union is the time the kernels were on the device)."""
    names = set(kernel.split(","))
    rows = [r for r in csv.DictReader(open(trace)) if short(r["Kernel_Name"]).split("<")[0] in names]
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    spans = [(int(r["Start_Timestamp"]), int(r["End_Timestamp"])) for r in rows]
    dur = [(e - s) / 1e3 for s, e in spans]
    out = ["", "Per phase (`%s`, dispatches in start order, from the kernel trace):" % kernel, "",
           "| phase | dispatches | avg us | summed ms | union ms |", "|---|---|---|---|---|"]
    i = 0
    for item in spec.split(","):
        name, cnt = item.split(":")
        seg = dur[i:i + int(cnt)]
        un = _union_ms(spans[i:i + int(cnt)])
        i += int(cnt)
        out.append("| %s | %d | %.1f | %.1f | %.1f |" % (name, len(seg), sum(seg) / max(1, len(seg)),
                                                        sum(seg) / 1e3, un))
    if i != len(dur):
        out.append("")
        out.append("(%d dispatches in the trace, %d assigned to phases)" % (len(dur), i))
    return "\n".join(out)


def phases_also(trace, kernel, spec, others):
    """Average launch duration of each kernel in `others` (comma-separated)
    per phase, the phases bounded by the start of the first dispatch of
    `kernel` in each phase of `spec`."""
    names = set(kernel.split(","))
    allrows = sorted(csv.DictReader(open(trace)), key=lambda r: int(r["Start_Timestamp"]))
    prim = [int(r["Start_Timestamp"]) for r in allrows
            if short(r["Kernel_Name"]).split("<")[0] in names]
    bounds, i = [], 0
    labels = []
    for item in spec.split(","):
        name, cnt = item.split(":")
        bounds.append(prim[i])
        labels.append(name)
        i += int(cnt)
    out = ["", "Other kernels per phase (phases start at the first `%s` dispatch of each):" % kernel,
           "", "| kernel | " + " | ".join("%s: launches, avg us" % x for x in labels) + " |",
           "|---|" + "---|" * len(labels)]
    for k in others.split(","):
        cells = []
        rows = [r for r in allrows if short(r["Kernel_Name"]).split("<")[0] == k]
        for p in range(len(bounds)):
            lo = bounds[p]
            hi = bounds[p + 1] if p + 1 < len(bounds) else None
            d = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in rows
                 if int(r["Start_Timestamp"]) >= lo and (hi is None or int(r["Start_Timestamp"]) < hi)]
            cells.append("%d, %.1f" % (len(d), sum(d) / len(d)) if d else "0, -")
        out.append("| %s | %s |" % (k, " | ".join(cells)))
    return "\n".join(out)


def _per_kernel(path):
    d = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        d[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return d


def _copy_factor(path):
    v = [x for k, xs in _per_kernel(path).items() if "copyBuffer" in k for x in xs]
    return COPY_KIB / (sum(v) / len(v))


def traffic(fetch, write, cal_fetch, cal_write, command):
    sf, sw = _copy_factor(cal_fetch), _copy_factor(cal_write)
    # the two passes are separate runs: average each counter over its own
    # dispatches (a tile-shape autotuner may split GEMM launches differently)
    per = collections.defaultdict(lambda: [0, 0.0, 0, 0.0])
    for col, data in ((0, _per_kernel(fetch)), (2, _per_kernel(write))):
        for name, xs in data.items():
            keys = {short(name), short(name).split("<")[0]}
            if short(name).split("<")[0] in FP64_GEMM:
                keys.add("fp64_gemm")
            for key in keys:
                per[key][col] += len(xs)
                per[key][col + 1] += sum(xs)
    kernels = {}
    for key, (nf, f, nw, w) in per.items():
        if key.startswith("__amd") or nf == 0 or nw == 0:
            continue
        fb, wb = f / nf * sf * 1024, w / nw * sw * 1024
        kernels[key] = {"dispatches": nf, "fetch_bytes_per_launch": fb,
                        "write_bytes_per_launch": wb, "bytes_per_launch": fb + wb}
    return json.dumps({"command": command, "fetch_scale": sf, "write_scale": sw,
                       "calibration": "256 MiB device copy (bench.py --workload copy)",
                       "kernels": kernels}, indent=1, sort_keys=True)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="mode", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--command", required=True)
    s.add_argument("--trace")
    s.add_argument("--kernel")
    s.add_argument("--phases")
    s.add_argument("--also", help="comma-separated kernels timed per phase as well")
    t = sub.add_parser("traffic")
    for x in ("fetch", "write", "cal_fetch", "cal_write"):
        t.add_argument(x)
    t.add_argument("--command", required=True)
    a = ap.parse_args()
    if a.mode == "stats":
        print(stats(a.csv, a.command))
        if a.trace:
            print(phases(a.trace, a.kernel, a.phases))
            if a.also:
                print(phases_also(a.trace, a.kernel, a.phases, a.also))
    else:
        print(traffic(a.fetch, a.write, a.cal_fetch, a.cal_write, a.command))


if __name__ == "__main__":
    main()
