"""Summaries of rocprofv3 output for the kernels of this library.

    python -m computeql_amd.profiling stats <kernel_stats.csv> [--command CMD]
        markdown table of per-kernel dispatch counts and durations from a
        `rocprofv3 --kernel-trace --stats --output-format csv` run.

    python -m computeql_amd.profiling traffic --read <counter_collection.csv>
        --write <counter_collection.csv> --rows N [--command CMD] [--json OUT]
        HBM bytes per dispatch from two rocprofv3 --pmc runs: the TCC's read
        requests to memory by size (TCC_EA0_RDREQ_32B/_64B/_128B_sum, with
        TCC_EA0_RDREQ_sum, whose count they must add up to) and its write
        requests (TCC_EA0_WRREQ_sum, 32 bytes each unless counted in
        TCC_EA0_WRREQ_64B_sum).  bytes = 32 n32 + 64 n64 + 128 n128 read +
        32 (w - w64) + 64 w64 written.  No scaling: a plain read of a 4 GB
        column (k_bandwidth_read) and the table generator (k_gen_events,
        16 GB written) come out at their known sizes, which the summary
        shows as the check.  (The derived FETCH_SIZE of this ROCm takes
        gfx950's 128-byte reads from TCC_BUBBLE, which does not count
        them, so it undercounts streaming kernels by 2-4x.)
"""
from __future__ import annotations

import argparse
import csv
import json
import sys
from collections import defaultdict

# kernel-name prefixes bench.py reports (rocprofv3 names carry the k_ prefix
# and the template-instance suffix)
KERNELS = ["aggregate_global", "aggregate_global_fold", "scan_simple", "scan_filter", "aggregate_dense",
           "aggregate_hash", "hash_compact", "qk_scan", "join_part_hist", "join_part_offsets", "join_part_scatter",
           "join_part_probe", "join_part_build", "join_probe", "join_build", "bandwidth_read", "bandwidth_copy",
           "gen_events", "gen_fact_fk", "fill_u32", "zonemap_build"]


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


def counter_values(path: str) -> dict:
    """every counter's value per dispatch, in dispatch order, by kernel key"""
    acc = defaultdict(lambda: defaultdict(dict))
    for r in csv.DictReader(open(path)):
        k = _key(r["Kernel_Name"])
        if k:
            acc[k][r["Counter_Name"]][int(r["Dispatch_Id"])] = float(r["Counter_Value"])
    return {k: {c: [v[i] for i in sorted(v)] for c, v in d.items()} for k, d in acc.items()}


READ_COUNTERS = ("TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_128B_sum", "TCC_EA0_RDREQ_sum")
WRITE_COUNTERS = ("TCC_EA0_WRREQ_sum", "TCC_EA0_WRREQ_64B_sum")


def traffic(read_csv: str, write_csv: str, rows: int) -> dict:
    rd = counter_values(read_csv)
    wr = counter_values(write_csv)
    kernels = {}
    mean = lambda v: sum(v) / len(v) if v else 0.0
    for k in sorted(set(rd) | set(wr)):
        r, w = rd.get(k, {}), wr.get(k, {})
        n32, n64, n128, n = (r.get(c, []) for c in READ_COUNTERS)
        wn, w64 = (w.get(c, []) for c in WRITE_COUNTERS)
        reads = [32 * a + 64 * b + 128 * c for a, b, c in zip(n32, n64, n128)]
        writes = [32 * (a - b) + 64 * b for a, b in zip(wn, w64)]
        per = [int(x + y) for x, y in zip(reads, writes)] if len(reads) == len(writes) else []
        kernels[k] = {"read_bytes": int(mean(reads)), "write_bytes": int(mean(writes)),
                      "bytes": int(mean(reads) + mean(writes)), "read_requests": int(mean(n)),
                      "sized_requests": int(mean(n32) + mean(n64) + mean(n128)),
                      "dispatch_bytes": per}
    return {"rows": rows, "method": "TCC read requests by size + write requests, no scaling", "kernels": kernels}


def traffic_table(rec: dict, command: str = "") -> str:
    out = []
    if command:
        out.append(f"Command: `{command}`\n")
    out.append(f"rows = {rec['rows']}; {rec['method']}. Read bytes = 32/64/128 x the TCC's read requests of that "
               "size; write bytes = 32 x 32-byte + 64 x 64-byte write requests.\n")
    out.append("| kernel | read requests (sized / all) | read GB | write GB | HBM GB/dispatch | bytes/row |")
    out.append("|---|---:|---:|---:|---:|---:|")
    for k, v in rec["kernels"].items():
        out.append(f"| {k} | {v['sized_requests']} / {v['read_requests']} | {v['read_bytes'] / 1e9:.3f} | "
                   f"{v['write_bytes'] / 1e9:.3f} | {v['bytes'] / 1e9:.3f} | {v['bytes'] / rec['rows']:.2f} |")
    return "\n".join(out) + "\n"


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--command", default="")
    t = sub.add_parser("traffic")
    t.add_argument("--read", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--rows", type=int, required=True)
    t.add_argument("--command", default="")
    t.add_argument("--json", default="")
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        sys.stdout.write(stats_table(a.csv, a.command))
    else:
        rec = traffic(a.read, a.write, a.rows)
        if a.command:
            rec["command"] = a.command
        if a.json:
            with open(a.json, "w") as f:
                json.dump(rec, f, indent=1)
        sys.stdout.write(traffic_table(rec, a.command))
    return 0


if __name__ == "__main__":
    sys.exit(main())
