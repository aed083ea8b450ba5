"""Summaries of rocprofv3 output (kernel-trace statistics and per-dispatch
PMC counters) as short text tables.

  python -m examinimd_amd.profiling stats <kernel_stats.csv> [command...]
  python -m examinimd_amd.profiling pmc <counter_collection.csv> <counter> <kernel-substring>
      [--calib <copy-kernel-substring> <copy-bytes>] [command...]
"""
import csv
import json
import sys
from collections import defaultdict


def kernel_stats(path, top=25):
    rows = list(csv.DictReader(open(path)))
    out = []
    for r in rows:
        name = r.get("Name") or r.get("KernelName") or r.get("name")
        calls = int(r.get("Calls") or r.get("calls") or 0)
        tot = float(r.get("TotalDurationNs") or r.get("total_duration") or 0)
        avg = float(r.get("AverageNs") or r.get("average") or (tot / max(calls, 1)))
        pct = float(r.get("Percentage") or r.get("percentage") or 0)
        out.append((name, calls, tot, avg, pct))
    out.sort(key=lambda t: -t[2])
    return out[:top]


def format_stats(stats, command):
    lines = ["```", "command: " + command, "```", "",
             "| kernel | calls | total ms | avg us | % |", "|---|---:|---:|---:|---:|"]
    for name, calls, tot, avg, pct in stats:
        short = name if len(name) < 90 else name[:87] + "..."
        lines.append(f"| `{short}` | {calls} | {tot / 1e6:.3f} | {avg / 1e3:.2f} | {pct:.1f} |")
    return "\n".join(lines) + "\n"


def pmc_per_kernel(path, counter):
    """mean counter value per dispatch, grouped by kernel name"""
    acc = defaultdict(list)
    for r in csv.DictReader(open(path)):
        cname = r.get("Counter_Name") or r.get("counter_name")
        if cname != counter:
            continue
        kname = r.get("Kernel_Name") or r.get("kernel_name")
        acc[kname].append(float(r.get("Counter_Value") or r.get("counter_value")))
    return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}


def main(argv):
    if argv[0] == "stats":
        print(format_stats(kernel_stats(argv[1]), " ".join(argv[2:])))
    elif argv[0] == "pmc":
        path, counter, ksub = argv[1], argv[2], argv[3]
        rest = argv[4:]
        calib = None
        if rest[:1] == ["--calib"]:
            calib = (rest[1], float(rest[2]))
            rest = rest[3:]
        per = pmc_per_kernel(path, counter)
        res = {"counter": counter, "command": " ".join(rest)}
        for k, (mean, n) in per.items():
            if ksub in k:
                res["kernel"] = k
                res["mean_per_dispatch"] = mean
                res["dispatches"] = n
            if calib and calib[0] in k:
                res["calib_kernel"] = k
                res["calib_counter"] = mean
                res["calib_bytes"] = calib[1]
        print(json.dumps(res, indent=1))


if __name__ == "__main__":
    main(sys.argv[1:])
