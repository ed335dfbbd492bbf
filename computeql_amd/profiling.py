"""Summaries of rocprofv3 output for the kernels of this library.

    python -m computeql_amd.profiling stats <kernel_stats.csv> [--command CMD]
        markdown table of per-kernel dispatch counts and durations from a
        `rocprofv3 --kernel-trace --stats --output-format csv` run.

    python -m computeql_amd.profiling traffic --fetch <counter_collection.csv>
        --write <counter_collection.csv> --rows N [--command CMD] [--json OUT]
        HBM bytes per dispatch from two `rocprofv3 --pmc FETCH_SIZE` /
        `--pmc WRITE_SIZE` runs of `bench.py --calibrate`.  Each counter is
        scaled by what it reads for a kernel of known traffic in the same
        run: FETCH_SIZE by the calibration read of one 4-byte column
        (k_bandwidth_read, rows x 4 bytes), WRITE_SIZE by the table
        generator (k_gen_events, four 4-byte columns written, rows x 16
        bytes).
"""
from __future__ import annotations

import argparse
import csv
import json
import sys
from collections import defaultdict

# kernel-name prefixes bench.py reports (rocprofv3 names carry the k_ prefix
# and the template-instance suffix)
KERNELS = ["aggregate_global", "scan_simple", "scan_filter", "aggregate_dense", "qk_scan", "join_part_hist",
           "join_part_scatter", "join_part_probe", "join_part_build", "join_probe", "join_build", "bandwidth_read",
           "gen_events", "gen_fact_fk", "fill_u32"]


def _key(name: str) -> str | None:
    n = name.split("(")[0].strip()
    if n.startswith("k_"):
        n = n[2:]
    best = None
    for k in KERNELS:
        # per-query kernels compiled at run time are named qk_main
        if n == k or n.startswith(k + "_") or (k == "qk_scan" and n.startswith("qk_")):
            if best is None or len(k) > len(best):
                best = k
    return best


def stats_table(path: str, command: str = "") -> str:
    rows = list(csv.DictReader(open(path)))
    out = []
    if command:
        out.append(f"Command: `{command}`\n")
    out.append("| kernel | calls | avg us | min us | max us | total ms | % |")
    out.append("|---|---:|---:|---:|---:|---:|---:|")
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        out.append(f"| {r['Name'].split('(')[0]} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                   f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | "
                   f"{float(r['TotalDurationNs']) / 1e6:.2f} | {float(r['Percentage']):.1f} |")
    return "\n".join(out) + "\n"


def counter_means(path: str, counter: str) -> dict:
    """mean counter value per dispatch, by kernel key"""
    acc = defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        k = _key(r["Kernel_Name"])
        if k:
            acc[k].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def traffic(fetch_csv: str, write_csv: str, rows: int) -> dict:
    fetch = counter_means(fetch_csv, "FETCH_SIZE")
    write = counter_means(write_csv, "WRITE_SIZE")
    if "bandwidth_read" not in fetch or "gen_events" not in write:
        raise SystemExit("calibration kernels missing: run bench.py --calibrate under the profiler")
    # KiB read per known byte, KiB written per known byte
    fetch_scale = (rows * 4) / (fetch["bandwidth_read"] * 1024.0)
    write_scale = (rows * 16) / (write["gen_events"] * 1024.0)
    kernels = {}
    for k in sorted(set(fetch) | set(write)):
        f = fetch.get(k, 0.0) * 1024.0 * fetch_scale
        w = write.get(k, 0.0) * 1024.0 * write_scale
        kernels[k] = {"fetch_kib": round(fetch.get(k, 0.0), 1), "write_kib": round(write.get(k, 0.0), 1),
                      "bytes": int(f + w)}
    return {"rows": rows, "fetch_scale": round(fetch_scale, 5), "write_scale": round(write_scale, 5),
            "kernels": kernels}


def traffic_table(rec: dict, command: str = "") -> str:
    out = []
    if command:
        out.append(f"Command: `{command}`\n")
    out.append(f"rows = {rec['rows']}; FETCH_SIZE scale {rec['fetch_scale']}, WRITE_SIZE scale {rec['write_scale']} "
               "(known bytes / counter bytes of the calibration kernels)\n")
    out.append("| kernel | FETCH_SIZE KiB/dispatch | WRITE_SIZE KiB/dispatch | HBM bytes/dispatch (scaled) | bytes/row |")
    out.append("|---|---:|---:|---:|---:|")
    for k, v in rec["kernels"].items():
        out.append(f"| {k} | {v['fetch_kib']:.0f} | {v['write_kib']:.0f} | {v['bytes']} | "
                   f"{v['bytes'] / rec['rows']:.2f} |")
    return "\n".join(out) + "\n"


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--command", default="")
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--rows", type=int, required=True)
    t.add_argument("--command", default="")
    t.add_argument("--json", default="")
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        sys.stdout.write(stats_table(a.csv, a.command))
    else:
        rec = traffic(a.fetch, a.write, a.rows)
        if a.command:
            rec["command"] = a.command
        if a.json:
            with open(a.json, "w") as f:
                json.dump(rec, f, indent=1)
        sys.stdout.write(traffic_table(rec, a.command))
    return 0


if __name__ == "__main__":
    sys.exit(main())
