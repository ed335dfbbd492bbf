#!/usr/bin/env python3
"""Warm GB/s and % of HBM peak of the scan/filter -> group-by/aggregate ->
join path on one MI355X, over a 1B-row 4 x u32 table (plus a u32 foreign
key) and two dimension tables, all resident in HBM.

A step runs the query suite once (algorithmic bytes = rows x 4 B x u32
columns the query reads in full):
  q1 SUM/MIN/MAX(amount)                                    amount              4 B/row
  q2 COUNT(*) WHERE amount < 200                            amount              4 B/row
  q3 region, COUNT(*), SUM(amount) GROUP BY region          region, amount      8 B/row
  q4 amount, COUNT/SUM/MIN/MAX(amount) GROUP BY amount      amount              4 B/row
  q5 COUNT(*) WHERE (region = r AND amount < a) OR ts > t   region, amount, ts 12 B/row
     (literals drawn per step from a seeded RNG)
  q6 JOIN dim A (1M unique keys spread over u32) GROUP BY segment, SUM(amount)
                                                            fk, amount          8 B/row
  q7 JOIN dim B (50 rows) WHERE attr < 3, COUNT(*)          region              4 B/row
  q8 point lookup id = k (zone maps leave one 8192-row zone): latency only
value = algorithmic bytes of q1..q7 / wall time of the timed steps (GB/s),
inputs resident in HBM before the timed region starts.

The same JSON line also carries, measured after the timed steps:
  encoded  the same steps with id, region and amount also kept narrow
           (sequence, 1-byte, 2-byte codes): effective GB/s of the same
           algorithmic bytes, never a fraction of HBM peak
  streamed q1-q3 over a host-backed copy of region and amount (pinned host
           memory): cold (every 16M-row chunk copied over PCIe, copies
           overlapping the previous chunk's kernels; PCIe-inclusive GB/s)
           and warm (chunks held by the resident chunk cache)
  batched  a 64-query mixed batch over the 10M-row 30-column event table,
           one query per launch vs one shared-scan batch (QPS, speedup)
  served   the request loop under Poisson arrivals at rates rising by half
           until p99 passes 50 ms, plain (one query per launch) and batched
           (with identical queries of a batch folded and without): the
           highest rate whose p99 latency stays within 50 ms, and the ratio
           of the two; on 2048 event-table queries with literals drawn per
           query, and (served.mix) on those plus configs 1-3's eight shapes
           (aggregates, filters, point lookups, both joins) over each of
           4 x u32 fact tables of 1M, 100M and 1B rows
  roofline the query with the longest kernel time per step (the partitioned
           join q6): its algorithmic bytes / the summed hipEvent time of its
           kernels, and the HBM bytes they move (rocprofv3 TCC counters,
           profiles/traffic.json); config.per_query carries the same for
           every query
  roofline.ceilings  what this device sustains in the same run: a plain
           4 GB read and a 4 GB copy (read + write bytes); the dominant
           query's fraction of the matching one is frac_of_measured_ceiling
  extra    shapes timed beside the suite: COUNT/SUM of region and
           COUNT/SUM/MIN/MAX of ts grouped by amount (10k groups), the
           point lookup pruned vs in scan form, and the first compile of a
           generated scan kernel
  c1_10m   q1-q3 at 10M rows, with the last-level cache flushed between
           runs and warm (those tables fit the 256 MB cache)
  cpu_baseline  the oracle (C restatement of the CPU operator path) on a
           bounded sample of the q1-q7 tables

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--rows N] [--no-served] [--no-encoded]
                    [--no-streamed] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_GBS = 8000.0
METRIC = "warm GB/s & % HBM peak at 1B-row u32 agg/filter/join; served QPS at p99 latency"
SEED = 0x1234567890ABCDEF  # bench_common.h:59
DIM_A_ROWS = 1_000_000
QUERY_BYTES_PER_ROW = {"q1": 4, "q2": 4, "q3": 8, "q4": 4, "q5": 12, "q6": 8, "q7": 4}
# the kernels each query launches over the fact table, with the bytes per
# fact row each one must move (reads + writes; the partitioned join's
# scatter reads the key and amount and writes both back in partition order)
QUERY_KERNELS = {
    "q1": [("aggregate_global", 4), ("aggregate_global_fold", 0)],
    "q2": [("scan_simple", 4)],
    "q3": [("aggregate_dense", 8)],
    "q4": [("aggregate_dense", 4)],
    "q5": [("qk_scan*", 12)],
    "q6": [("join_part_hist*", 4), ("join_part_offsets", 0), ("join_part_scatter*", 16), ("join_part_probe*", 8),
           ("join_probe", 8)],
    "q7": [("aggregate_dense", 4)],
    "q8": [("scan_simple", 0)],
}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def dim_b_attr():
    import numpy as np

    return (((np.arange(50, dtype=np.uint64) * 0x9E3779B1) >> 7) % 10).astype(np.uint32)


def build_tables(cq, rows: int, seed: int):
    import numpy as np

    t = cq.Table(rows)
    for name in ("id", "region", "amount", "ts", "fk"):
        t.add_column(name, cq.U32)
    t.gen_events(seed, id=0, region=1, amount=2, ts=3)
    dima = cq.Table(DIM_A_ROWS)
    dima.add_column("key", cq.U32)
    dima.add_column("segment", cq.U32)
    dima.gen_dim(seed ^ 0xA, 100, spread_keys=True, key=0, attr=1)
    t.gen_fact_fk(4, seed, dima, 0)
    dimb = cq.Table(50)
    dimb.add_column("key", cq.U32, np.arange(50, dtype=np.uint32))
    dimb.add_column("attr", cq.U32, dim_b_attr())
    return t, dima, dimb


def q5_literals(step: int, seed: int):
    import numpy as np

    r = np.random.default_rng([seed & 0xFFFFFFFF, step])
    return int(r.integers(0, 50)), int(r.integers(100, 10000)), int(r.integers(2**31, 2**32 - 2**28))


def point_key(step: int, rows: int) -> int:
    return (step * 2654435761) % rows


def run_suite(cq, lib, t, dima, dimb, step: int, seed: int, rows: int):
    from computeql_amd.ir import col

    res = {}

    def tag(q):
        lib.gpu_hip_profile_tag(q.encode())

    tag("q1"); res["q1"] = t.aggregate([cq.SUM, cq.MIN, cq.MAX], [2, 2, 2])
    tag("q2"); res["q2"] = t.scan(col(2) < 200, count_only=True)
    tag("q3"); res["q3"] = t.aggregate([cq.COUNT, cq.SUM], [None, 2], group=1)
    tag("q4"); res["q4"] = t.aggregate([cq.COUNT, cq.SUM, cq.MIN, cq.MAX], [None, 2, 2, 2], group=2)
    r, a, ts = q5_literals(step, seed)
    tag("q5"); res["q5"] = t.scan(((col(1) == r) & (col(2) < a)) | (col(3) > ts), count_only=True)
    tag("q6"); res["q6"] = cq.aggregate_join(t, 4, dima, 0, [cq.SUM], [2], dim_group=1)
    tag("q7"); res["q7"] = cq.aggregate_join(t, 1, dimb, 0, [cq.COUNT], [None], dim_where=col(1) < 3)
    tag("q8"); res["q8"] = t.scan(col(0) == point_key(step, rows))
    tag("")
    return res


def trace_of(r):
    return r.trace if hasattr(r, "trace") else r[2]


def check_invariants(res, rows: int, step: int) -> list:
    """Size-independent checks at full size: group counts sum to N, group
    sums to the global sum, join A covers every fact row once, q2 equals the
    first 200 groups of q4, the point lookup finds its one row."""
    errs = []
    total = res["q1"].column(0)[0]
    if int(res["q3"].column(0).sum()) != rows or float(res["q3"].column(1).sum()) != total:
        errs.append("q3 totals")
    if int(res["q4"].column(0).sum()) != rows or float(res["q4"].column(1).sum()) != total:
        errs.append("q4 totals")
    if res["q4"].groups != 10000 or res["q4"].column(2)[123] != 123.0 or res["q1"].column(2)[0] != 9999.0:
        errs.append("q4 groups")
    if float(res["q6"].column(0).sum()) != total or res["q6"].groups != 100:
        errs.append("q6 sum")
    if res["q2"][0] != int(res["q4"].column(0)[:200].sum()):
        errs.append("q2 vs q4")
    if res["q7"].groups != 1 or int(res["q7"].cells[0, 0]) != int(res["q3"].column(0)[dim_b_attr() < 3].sum()):
        errs.append("q7 vs q3")
    n8, ids8, _ = res["q8"]
    if n8 != 1 or int(ids8[0]) != point_key(step, rows):
        errs.append("q8")
    return errs


def timed_cpu_baseline(sample_rows: int, seed: int, workers: int, min_seconds: float):
    """The oracle (C restatement of the reference CPU operator path) over a
    bounded sample of the same tables: the q1-q7 suite repeated on all host
    cores until min_seconds have passed, then once on one core.  Returns
    (GB/s of algorithmic bytes, repetitions, per-query seconds, 1-core GB/s)."""
    import numpy as np

    from computeql_amd.ir import col
    from oracle import oracle as orc

    id_, region, amount, ts = orc.gen_events(sample_rows, seed=seed, nthreads=workers)
    dk, seg = orc.gen_dim(DIM_A_ROWS, 100, seed=seed ^ 0xA)
    fk = orc.gen_fact_fk(sample_rows, dk, seed=seed, nthreads=workers)
    bk = np.arange(50, dtype=np.uint32)
    battr = dim_b_attr()
    cols = [id_, region, amount, ts]
    r, a, tsl = q5_literals(0, seed)
    where5 = ((col(1) == r) & (col(2) < a)) | (col(3) > tsl)

    def suite(nt, secs):
        def timeit(name, fn):
            t0 = time.perf_counter()
            fn()
            secs[name] = secs.get(name, 0.0) + time.perf_counter() - t0

        timeit("q1", lambda: orc.aggregate_global([amount], workers=nt))
        timeit("q2", lambda: orc.scan_chunked_count(cols, [(2, 1, 200.0)], workers=nt))
        timeit("q3", lambda: orc.dense_group_u32(region, amount, 0, 50, workers=nt))
        timeit("q4", lambda: orc.dense_group_u32(amount, amount, 0, 10000, workers=nt))
        timeit("q5", lambda: orc.scan(cols, where5, nthreads=nt, count_only=True))
        timeit("q6", lambda: orc.join_aggregate(fk, [amount], [1], dk, dim_group=seg, nthreads=nt))
        timeit("q7", lambda: orc.join_aggregate(region, [None], [0], bk, dim_pass=(battr < 3), nthreads=nt))

    bytes_per_suite = sum(QUERY_BYTES_PER_ROW.values()) * sample_rows
    secs = {}
    reps = 0
    while reps < 1000 and (reps == 0 or sum(secs.values()) < min_seconds):
        suite(workers, secs)
        reps += 1
    one = {}
    suite(1, one)
    return (bytes_per_suite * reps / sum(secs.values()) / 1e9, reps, secs,
            bytes_per_suite / sum(one.values()) / 1e9)


TRAFFIC_SOURCE = ("profiles/traffic.json: the TCC's HBM read requests by size and write requests, per dispatch, of "
                  "`bench.py --steps 1 --warmup 0 --calibrate` under two rocprofv3 --pmc runs on this 1B-row "
                  "workload (not this run; counters cannot be read from inside the process)")
# dispatch order of one suite run: which dispatch of a shared kernel belongs to which query
TRAFFIC_DISPATCH = {("q3", "aggregate_dense"): 0, ("q4", "aggregate_dense"): 1, ("q7", "aggregate_dense"): 2,
                    ("q2", "scan_simple"): 0}


def _traffic_record(rows: int):
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    return rec if rec.get("rows") == rows else None


def measured_traffic(query: str, kernels, rows: int):
    """HBM bytes one run of `query` moves through `kernels` (profile names,
    a trailing * a prefix), from profiles/traffic.json when it covers these
    kernels at this row count; else None."""
    rec = _traffic_record(rows)
    if not rec:
        return None
    total = 0
    for k in kernels:
        base = k.split(":", 1)[-1].rstrip("*")
        e = rec["kernels"].get(base)
        if not e:
            return None
        # the suite's dispatch of the kernel: its first in the profiled run
        # (the extra shapes, the 10M-row runs and the calibration follow)
        i = TRAFFIC_DISPATCH.get((query, base), 0)
        if i >= len(e.get("dispatch_bytes", [])):
            return None
        total += e["dispatch_bytes"][i]
    return total


EVENT_ROWS = 10_000_000
P99_LIMIT_MS = 50.0


def served_sweep(Server, name, kw, start, drive, seconds_per_rate, dedupe=None):
    """offered rates rising by half from `start` until p99 passes the limit
    or the achieved rate falls 10% below the offered one (at most 16
    rates), then two rates bisecting (geometrically) between the last one
    that held and the first that did not; the highest rate that held.
    dedupe: QE_BATCH_DEDUPE for the sweep (identical queries of a batch run
    once), None = as set."""
    import computeql_amd as cq

    old = cq.get_setting("QE_BATCH_DEDUPE", 1)
    if dedupe is not None:
        cq.set_setting("QE_BATCH_DEDUPE", int(dedupe))
    points, best, crossed = [], None, None
    try:
        with Server(**kw) as srv:
            drive(srv, start, 0.3, 1)

            def point(r):
                st = drive(srv, r, seconds_per_rate, int(r))
                points.append({"offered": round(r), "qps": round(st["qps"], 1), "p50_ms": round(st["p50_ms"], 3),
                               "p99_ms": round(st["p99_ms"], 3), "mean_batch": round(st["mean_batch"], 1),
                               "joins": st["joins"], "scans": st["scans"]})
                log(f"{name} offered {r:.0f}: {st['qps']:.0f} qps p99 {st['p99_ms']:.2f} ms")
                if st["failed"]:
                    raise RuntimeError(f"{name}: {st['failed']} requests failed")
                return st["p99_ms"] <= P99_LIMIT_MS and st["qps"] >= 0.9 * r, st

            r = float(start)
            for _ in range(16):
                held, st = point(r)
                if not held:
                    crossed = "p99" if st["p99_ms"] > P99_LIMIT_MS else "rate"
                    lo, hi = (best["offered"] if best else None), r
                    for _ in range(2 if lo else 0):
                        mid = (lo * hi) ** 0.5
                        held, _ = point(mid)
                        if held:
                            best, lo = points[-1], mid
                        else:
                            hi = mid
                    break
                best = points[-1]
                r *= 1.5
    finally:
        cq.set_setting("QE_BATCH_DEDUPE", old)
    return {"config": {k: v for k, v in kw.items() if k != "hbm_budget"},
            "dedupe": old if dedupe is None else int(dedupe), "rates": points,
            "max_qps_at_p99": best["qps"] if best else 0.0, "p99_ms_there": best["p99_ms"] if best else None,
            "stopped_by": crossed or "last rate"}


def best_sweep(Server, name, candidates, start, drive, seconds_per_rate, dedupe=None):
    """served_sweep over each candidate configuration; the one with the
    highest QPS at the p99 limit, and every candidate's figure.  Both the
    plain and the batched side are chosen this way, so neither is compared
    at a worker count or batch size that holds it back (the plain side
    gains from more workers, whose host work then overlaps)."""
    runs = [served_sweep(Server, f"{name} {c['workers']}w/{c['max_batch']}", c, start, drive, seconds_per_rate,
                         dedupe) for c in candidates]
    best = dict(max(runs, key=lambda x: x["max_qps_at_p99"]))
    best["candidates"] = [{"config": x["config"], "max_qps_at_p99": x["max_qps_at_p99"]} for x in runs]
    return best


def mix_requests(cq, t4, dima, dimb, variants: int, seed: int):
    """configs 1-3's query shapes over a 4 x u32 fact table (id, region,
    amount, ts, fk) with literals drawn per request: SUM/MIN/MAX(amount);
    COUNT(*) WHERE amount < a; COUNT, SUM(amount) GROUP BY region;
    COUNT/SUM/MIN/MAX(amount) GROUP BY amount; COUNT(*) under the
    three-predicate AND/OR filter; the point lookup id = k (row ids); join
    A GROUP BY segment, SUM(amount); join B WHERE attr < b, COUNT(*)"""
    import numpy as np

    from computeql_amd.ir import col
    from computeql_amd.serve import JoinSpec, ScanSpec

    r = np.random.default_rng(seed)
    out = []
    for _ in range(variants):
        reg, amt, ts = int(r.integers(0, 50)), int(r.integers(100, 10000)), int(r.integers(2**31, 2**32 - 2**28))
        out += [cq.AggSpec([cq.SUM, cq.MIN, cq.MAX], [2, 2, 2]),
                cq.AggSpec([cq.COUNT], [None], col(2) < int(r.integers(1, 10000))),
                cq.AggSpec([cq.COUNT, cq.SUM], [None, 2], None, 1),
                cq.AggSpec([cq.COUNT, cq.SUM, cq.MIN, cq.MAX], [None, 2, 2, 2], None, 2, cap=10_000),
                cq.AggSpec([cq.COUNT], [None], ((col(1) == reg) & (col(2) < amt)) | (col(3) > ts)),
                ScanSpec(col(0) == int(r.integers(1, t4.rows + 1)), cap=16),
                JoinSpec(dima, 4, 0, [cq.SUM], [2], dim_group=1, cap=128),
                JoinSpec(dimb, 1, 0, [cq.COUNT], [None], dim_where=col(1) < int(r.integers(1, 10)))]
    return [(t4, sp) for sp in out]


def batched_and_served(cq, lib, seconds_per_rate: float, served_too: bool = True, fact=None, dima=None,
                       dimb=None):
    """The batched and served forms on the 10M-row 30-column event table
    (synthetic, seeded): one 64-query mixed batch run plain (one query per
    launch) and as a shared scan, then the request loop under Poisson load
    at rising rates, plain (max_batch 1) and batched, each up to the last
    rate whose p99 stays within 50 ms and whose achieved rate is within 10%
    of the offered one."""
    from computeql_amd import workloads as wl
    from computeql_amd.serve import Server

    log(f"batched/served: building the {EVENT_ROWS}-row 30-column event table")
    ev = wl.upload(wl.event_table(EVENT_ROWS, seed=7), wl.event_nulls(EVENT_ROWS, seed=7))  # 2% NULL timestamps
    # GROUP BY user_id (Zipf over 1M users) on the hash path: the
    # reference's "group by user (1M groups)" (README.md:46, 56 ms)
    cq.profile(True)
    wall, kms, res = _timed(lib, lambda: ev.aggregate([cq.COUNT, cq.SUM], [None, wl.M0], group=wl.USER))
    by_user = {"groups": res.groups, "path": cq.AGG_PATHS.get(res.trace["path"]), "wall_ms": round(wall, 3),
               "kernel_ms": round(kms, 3)}
    log(f"group by user: {res.groups} groups, {kms:.2f} ms kernel, {wall:.2f} ms wall")
    specs = wl.mixed_queries(64, seed=11)
    # the plain and batched sides, and the served sweeps below, are timed
    # without per-dispatch hipEvents: those cost each launch of the
    # one-query-per-launch side far more than the batched side's few
    cq.profile(False)
    for _ in range(2):
        for sp in specs:
            ev.aggregate(sp.funcs, sp.args, where=sp.where, group=sp.group)
        cq.aggregate_batch(ev, specs)
    reps = 10
    t0 = time.perf_counter()
    for _ in range(reps):
        for sp in specs:
            ev.aggregate(sp.funcs, sp.args, where=sp.where, group=sp.group)
    plain_s = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        _, tr = cq.aggregate_batch(ev, specs, with_trace=True)
    batch_s = (time.perf_counter() - t0) / reps
    # the shared scan's device time from a separate profiled pass.  Batches
    # run as HIP graphs (QE_BATCH_GRAPH): their device time is the graph
    # launch (shared_scan + shared_reduce), timed as "graph:..."
    cq.profile(True)
    lib.gpu_hip_profile_reset()
    for _ in range(reps):
        cq.aggregate_batch(ev, specs)
    lib.gpu_hip_wait()
    cq.profile(False)
    scan_ns, _ = profile_stats(lib, "shared_scan*")
    timed = "shared_scan kernels"
    if not scan_ns:
        scan_ns, _ = profile_stats(lib, "graph:*")
        timed = "graph launches (shared_scan + shared_reduce)"
    batched = {"workload": f"{EVENT_ROWS}-row 30-column event table (2% NULLs in t0-t4), 64 mixed filter/GROUP BY aggregates",
               "plain_ms_per_batch": round(plain_s * 1e3, 3), "batched_ms_per_batch": round(batch_s * 1e3, 3),
               "plain_qps": round(64 / plain_s, 1), "batched_qps": round(64 / batch_s, 1),
               "speedup": round(plain_s / batch_s, 2), "launches": tr["launches"],
               "union_bytes_per_batch": tr["bytes_read"],
               "shared_scan_gbs": round(tr["bytes_read"] / (scan_ns / reps) if scan_ns else 0.0, 1),
               "shared_scan_timed": timed, "group_by_user_hash": by_user}
    if not served_too:
        ev.close()
        return batched, None
    # served: request k runs pool[k % len(pool)]; the pool holds 2048
    # queries of the mixed shapes with literals drawn per query, so a batch
    # (at most a few hundred consecutive arrivals) holds distinct requests
    # except where two draws coincide.  Batched sweeps run with identical
    # queries folded (QE_BATCH_DEDUPE 1, the default) and without.
    pool = wl.mixed_queries(2048, seed=12)
    served = {"workload": f"{EVENT_ROWS}-row 30-column event table, Poisson arrivals over 2048 mixed queries "
                          "with literals drawn per query",
              "p99_limit_ms": P99_LIMIT_MS, "seconds_per_rate": seconds_per_rate}

    def drive_ev(srv, r, secs, seed):
        return srv.drive(ev, pool, rate_qps=r, seconds=secs, seed=seed)

    # each side at its best configuration (best_sweep); the batched side
    # without folding runs at the configuration chosen with it
    served["plain"] = best_sweep(Server, "served plain", [dict(workers=w, max_batch=1) for w in (4, 8)], 8000,
                                 drive_ev, seconds_per_rate)
    served["batched"] = best_sweep(Server, "served batched", [dict(workers=3, max_batch=256),
                                                              dict(workers=3, max_batch=512),
                                                              dict(workers=4, max_batch=512)], 18000, drive_ev,
                                   seconds_per_rate, dedupe=1)
    batched_kw = served["batched"]["config"]
    served["batched_no_folding"] = served_sweep(Server, "served batched, no folding", batched_kw, 18000, drive_ev,
                                                seconds_per_rate, dedupe=0)
    p = served["plain"]["max_qps_at_p99"]
    served["speedup_at_p99"] = round(served["batched"]["max_qps_at_p99"] / p, 2) if p else None
    served["speedup_at_p99_no_folding"] = round(served["batched_no_folding"]["max_qps_at_p99"] / p, 2) if p else None
    # the served mix (config 5): the event table's queries and configs 1-3's
    # eight shapes (aggregates, filters, point lookups, both joins) over 4 x
    # u32 fact tables of 1M, 100M and 1B rows, literals drawn per request
    mix = None
    if fact is not None:
        import numpy as np

        small, mid = cq.Table(1_000_000), cq.Table(100_000_000)
        for t4 in (small, mid):
            for name in ("id", "region", "amount", "ts", "fk"):
                t4.add_column(name, cq.U32)
            t4.gen_events(SEED, id=0, region=1, amount=2, ts=3)
            t4.gen_fact_fk(4, SEED, dima, 0)
        pairs = [(ev, sp) for sp in wl.mixed_queries(1024, seed=13)]
        for k, t4 in enumerate((small, mid, fact)):
            pairs += mix_requests(cq, t4, dima, dimb, 64, seed=100 + k)
        order = np.random.default_rng(5).permutation(len(pairs))
        pairs = [pairs[i] for i in order]
        shapes = ["sum_min_max", "count_amount_lt", "by_region", "by_amount", "and_or_count", "point_lookup",
                  "join_a_by_segment", "join_b_count"]

        def drive_mix(srv, r, secs, seed):
            return srv.drive_mix(pairs, rate_qps=r, seconds=secs, seed=seed)

        mix = {"workload": f"{len(pairs)} requests cycled in a seeded order: 1024 of the event table's mixed queries "
                           f"and 64 draws of each of eight shapes ({', '.join(shapes)}) on each of 4 x u32 fact tables "
                           "of 1M, 100M and 1B rows (join A: 1M-key dim, join B: 50-row dim); literals drawn per "
                           "request in the shapes that have one (count_amount_lt, and_or_count, point_lookup, "
                           "join_b_count; the other four have none); Poisson arrivals",
               "p99_limit_ms": P99_LIMIT_MS, "seconds_per_rate": seconds_per_rate}
        budget = 96 << 30
        mix["plain"] = best_sweep(Server, "served mix plain",
                                  [dict(workers=w, max_batch=1, hbm_budget=budget) for w in (2, 6)], 500, drive_mix,
                                  seconds_per_rate)
        mix["batched"] = best_sweep(Server, "served mix batched",
                                    [dict(workers=w, max_batch=256, hbm_budget=budget) for w in (3, 6)], 1000,
                                    drive_mix, seconds_per_rate, dedupe=1)
        mix_batched_kw = dict(mix["batched"]["config"], hbm_budget=budget)
        mix["batched_no_folding"] = served_sweep(Server, "served mix batched, no folding", mix_batched_kw, 1000,
                                                 drive_mix, seconds_per_rate, dedupe=0)
        p = mix["plain"]["max_qps_at_p99"]
        mix["speedup_at_p99"] = round(mix["batched"]["max_qps_at_p99"] / p, 2) if p else None
        mix["speedup_at_p99_no_folding"] = round(mix["batched_no_folding"]["max_qps_at_p99"] / p, 2) if p else None
        small.close()
        mid.close()
    served["mix"] = mix
    ev.close()
    return batched, served


def encoded_run(cq, lib, t, dima, dimb, seed: int, rows: int, steps: int):
    """The same q1-q8 steps after the fact table's id, region and amount
    columns are also kept narrow (qe_table_encode_column: id a sequence,
    region 1-byte and amount 2-byte codes; ts and fk span the u32 range and
    stay full width).  Reported as effective GB/s: the full-width
    algorithmic bytes over the time, never as a fraction of HBM peak."""
    encs = {name: t.encode(i) for i, name in enumerate(("id", "region", "amount", "ts", "fk"))}
    lib.gpu_hip_wait()
    errs = []
    res = run_suite(cq, lib, t, dima, dimb, 0, seed, rows)
    errs += check_invariants(res, rows, 0)
    wall = {q: 0.0 for q in QUERY_KERNELS}
    t0 = time.perf_counter()
    for s in range(steps):
        res = run_suite(cq, lib, t, dima, dimb, 1 + s, seed, rows)
        for q in wall:
            wall[q] += trace_of(res[q])["wall_ns"] / 1e9
    lib.gpu_hip_wait()
    elapsed = time.perf_counter() - t0
    errs += check_invariants(res, rows, steps)
    total = sum(QUERY_BYTES_PER_ROW.values()) * rows * steps
    log(f"encoded: {1e3 * elapsed / steps:.2f} ms/step")
    return {"encodings": encs, "effective_gbs": round(total / elapsed / 1e9, 1),
            "ms_per_step": round(1e3 * elapsed / steps, 3),
            "per_query_wall_ms": {q: round(1e3 * v / steps, 3) for q, v in wall.items()},
            "invariants": "ok" if not errs else sorted(set(errs)),
            "note": "narrow codes read by scan_simple, the global aggregate and the dense GROUP BY (q1-q4, q7, "
                    "q8); q5's generated kernel and q6's join read full width"}


def streamed_run(cq, lib, t, rows: int, repeats: int = 3):
    """The cold path: region and amount copied out of the fact table into a
    host-backed table (pinned host memory), then q1-q3 over it cold (resident
    chunk cache emptied, every chunk copied over PCIe while the previous
    chunk's kernels run) and warm (every chunk in the cache).  Each result
    must equal the same query over the resident table.  Cold GB/s include the
    PCIe copies and are never a fraction of HBM peak."""
    import numpy as np

    t0 = time.perf_counter()
    region, amount = t.read(1), t.read(2)
    h = cq.Table(rows)
    h.add_host_column("region", cq.U32, region)
    h.add_host_column("amount", cq.U32, amount)
    del region, amount
    lib.gpu_hip_wait()
    build_s = time.perf_counter() - t0
    from computeql_amd.ir import col

    queries = {
        "q1": (lambda x, reg, amt: x.aggregate([cq.SUM, cq.MIN, cq.MAX], [amt] * 3), 4),
        "q2": (lambda x, reg, amt: x.scan(col(amt) < 200, count_only=True), 4),
        "q3": (lambda x, reg, amt: x.aggregate([cq.COUNT, cq.SUM], [None, amt], group=reg), 8),
    }

    def same(a, b):
        if isinstance(a, tuple):
            return a[0] == b[0]
        oa, ob = np.argsort(a.keys, kind="stable"), np.argsort(b.keys, kind="stable")
        return np.array_equal(a.keys[oa], b.keys[ob]) and np.array_equal(a.cells[oa], b.cells[ob])

    out = {"rows": rows, "chunk_rows": int(cq.get_setting("QE_SCAN_STREAM_CHUNK_ROWS", 16777216)),
           "build_s": round(build_s, 2)}
    errs = []
    for q, (fn, bpr) in queries.items():
        ref = fn(t, 1, 2)
        t1 = time.perf_counter()
        fn(t, 1, 2)
        resident_ms = 1e3 * (time.perf_counter() - t1)
        cq.resident_clear()
        t1 = time.perf_counter()
        cold = fn(h, 0, 1)
        cold_ms = 1e3 * (time.perf_counter() - t1)
        st = cq.stream_stats()
        warm_ms = []
        for _ in range(repeats):
            t1 = time.perf_counter()
            warm = fn(h, 0, 1)
            warm_ms.append(1e3 * (time.perf_counter() - t1))
        wst = cq.stream_stats()
        if not (same(cold, ref) and same(warm, ref)):
            errs.append(f"streamed {q} differs from resident")
        nbytes = bpr * rows
        w = min(warm_ms)
        out[q] = {"cold_ms": round(cold_ms, 2), "cold_gbs_pcie_inclusive": round(nbytes / cold_ms / 1e6, 1),
                  "bytes_copied": st["bytes_copied"], "copy_wait_ms": round(st["copy_wait_ns"] / 1e6, 2),
                  "warm_ms": round(w, 3), "warm_gbs": round(nbytes / w / 1e6, 1),
                  "warm_bytes_copied": wst["bytes_copied"], "resident_ms": round(resident_ms, 3)}
        log(f"streamed {q}: cold {cold_ms:.1f} ms, warm {w:.2f} ms, resident {resident_ms:.2f} ms")
    h.close()
    cq.resident_clear()
    out["invariants"] = "ok" if not errs else errs
    return out


def _timed(lib, fn, reps: int = 3, flush=None):
    """median wall ms and summed hipEvent kernel ms of fn over reps runs
    (after one warm run); flush() runs before each timed run"""
    fn()
    walls, kns, out = [], [], None
    for _ in range(reps):
        if flush:
            flush()
        lib.gpu_hip_wait()
        lib.gpu_hip_profile_reset()
        t0 = time.perf_counter()
        out = fn()
        lib.gpu_hip_wait()
        walls.append(1e3 * (time.perf_counter() - t0))
        ns, _ = profile_stats(lib, "*")
        kns.append(ns / 1e6)
    walls.sort()
    kns.sort()
    return walls[len(walls) // 2], kns[len(kns) // 2], out


def extra_measurements(cq, lib, t, rows: int) -> dict:
    """Rows of the 1B-row configuration outside the timed suite:
    - GROUP BY amount (10k groups) over non-key arguments: the reference's
      published COUNT(*), SUM(region) (bench_select_scale.c:1153-1156) and
      COUNT/SUM/MIN/MAX(ts), 8 B/row each;
    - the point lookup's scan form (zone maps off: all 4 GB of id read);
    - the first hiprtc compile of a new WHERE shape (disk cache off) and
      the same shape again."""
    from computeql_amd.ir import col

    out = {}
    for name, funcs, args in (("count_sum_region_by_amount", [cq.COUNT, cq.SUM], [None, 1]),
                              ("count_sum_min_max_ts_by_amount", [cq.COUNT, cq.SUM, cq.MIN, cq.MAX], [None, 3, 3, 3])):
        wall, kms, res = _timed(lib, lambda: t.aggregate(funcs, args, group=2))
        nbytes = 8 * rows
        out[name] = {"groups": res.groups, "path": cq.AGG_PATHS.get(res.trace["path"]), "wall_ms": round(wall, 3),
                     "kernel_ms": round(kms, 3), "kernel_gbs": round(nbytes / kms / 1e6, 1),
                     "frac": round(nbytes / kms / 1e6 / PEAK_GBS, 4)}
        log(f"{name}: {out[name]['path']} {kms:.2f} ms kernel, {wall:.2f} ms wall")
    k = point_key(7, rows)
    cq.set_setting("QE_ZONEMAP", 0)
    try:
        wall, kms, res = _timed(lib, lambda: t.scan(col(0) == k))
    finally:
        cq.set_setting("QE_ZONEMAP", 1)
    pruned_wall, _, _ = _timed(lib, lambda: t.scan(col(0) == k))
    out["point_lookup"] = {"pruned_wall_ms": round(pruned_wall, 3), "scan_form_wall_ms": round(wall, 3),
                           "scan_form_kernel_ms": round(kms, 3), "scan_form_gbs": round(4 * rows / kms / 1e6, 1),
                           "found": int(res[0])}
    cq.set_setting("QE_KERNEL_DISK_CACHE", 0)
    try:
        # a shape no other query of this run has: the first call compiles
        w = ((col(1) == 11) | (col(2) == 4242) | (col(1) == 13)) & ((col(3) < 12345) | (col(2) > 9990))
        _, _, tr1 = t.scan(w, count_only=True)
        t0 = time.perf_counter()
        _, _, tr2 = t.scan(w, count_only=True)
        second = 1e3 * (time.perf_counter() - t0)
    finally:
        cq.set_setting("QE_KERNEL_DISK_CACHE", 1)
    out["hiprtc_first_compile"] = {"compile_ms": round(tr1["compile_ns"] / 1e6, 1),
                                   "first_wall_ms": round(tr1["wall_ns"] / 1e6, 1),
                                   "second_wall_ms": round(second, 3), "second_compile_ms": tr2["compile_ns"] / 1e6}
    log(f"hiprtc first compile {tr1['compile_ns'] / 1e6:.1f} ms")
    return out


def c1_measurements(cq, lib, big, seed: int) -> dict:
    """Config 1: the 10M-row table (40 MB a column), q1-q3.  40 MB sits in
    the 256 MB last-level cache, so each query is timed both with that cache
    flushed first (a 4 GB column of the big table read in between) and
    warm; only the flushed figure is an HBM rate."""
    from computeql_amd.ir import col

    n = 10_000_000
    t = cq.Table(n)
    for name in ("id", "region", "amount", "ts"):
        t.add_column(name, cq.U32)
    t.gen_events(seed, id=0, region=1, amount=2, ts=3)
    flush = lambda: cq.bandwidth_read(big, 3)
    qs = {"q1": (lambda: t.aggregate([cq.SUM, cq.MIN, cq.MAX], [2, 2, 2]), 4),
          "q2": (lambda: t.scan(col(2) < 200, count_only=True), 4),
          "q3": (lambda: t.aggregate([cq.COUNT, cq.SUM], [None, 2], group=1), 8)}
    out = {"rows": n}
    for q, (fn, bpr) in qs.items():
        cw, ck, _ = _timed(lib, fn, reps=5, flush=flush)
        ww, wk, _ = _timed(lib, fn, reps=5)
        nbytes = bpr * n
        out[q] = {"llc_flushed": {"wall_us": round(1e3 * cw, 1), "kernel_us": round(1e3 * ck, 1),
                                  "kernel_gbs": round(nbytes / ck / 1e6, 1)},
                  "llc_warm_cache_bound": {"wall_us": round(1e3 * ww, 1), "kernel_us": round(1e3 * wk, 1),
                                           "kernel_gbs": round(nbytes / wk / 1e6, 1)}}
    t.close()
    log(f"c1: q1 {out['q1']['llc_flushed']['kernel_us']} us flushed, {out['q1']['llc_warm_cache_bound']['kernel_us']} us warm")
    return out


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def device_ceilings(cq, lib, t, reps: int = 3) -> dict:
    """What this device sustains for the two traffic shapes of the suite,
    measured in the same run with hipEvents: a plain read of one 4 GB column
    (bandwidth_read) and a copy of it into a scratch buffer (bandwidth_copy,
    counted as read + write bytes), each the best of three grid sizes."""
    cus = cq.compute_units()
    out = {}
    for name, fn, factor in (("read", cq.bandwidth_read, 4.0), ("copy", cq.bandwidth_copy, 8.0)):
        best = None
        for blocks in (2 * cus, 4 * cus, 8 * cus):
            fn(t, 2, blocks)
            lib.gpu_hip_profile_reset()
            for _ in range(reps):
                fn(t, 2, blocks)
            ns, d = profile_stats(lib, f"bandwidth_{name}")
            if d:
                gbs = factor * t.rows / (ns / d)
                best = gbs if best is None or gbs > best else best
        out[f"{name}_gbs"] = round(best, 1) if best else None
    out["note"] = ("read: one 4 GB column read; copy: that column copied to a scratch buffer, read + write bytes; "
                   "best of 2, 4 and 8 blocks per CU")
    return out


def profile_stats(lib, name: str):
    d = ctypes.c_uint64(0)
    ns = lib.gpu_hip_profile_query(name.encode(), ctypes.byref(d))
    return int(ns), int(d.value)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--rows", type=int, default=1_000_000_000)
    ap.add_argument("--cpu-sample-rows", type=int, default=64_000_000)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-served", action="store_true", help="skip the batched and served measurements")
    ap.add_argument("--served-seconds", type=float, default=1.0)
    ap.add_argument("--no-encoded", action="store_true", help="skip the narrow-encoded rerun of the steps")
    ap.add_argument("--no-streamed", action="store_true", help="skip the host-backed (cold/warm) streamed runs")
    ap.add_argument("--only", choices=["batched", "served"],
                    help="run the 64-query batched comparison (batched), or it and the served sweeps over "
                         "tables of --rows (served), alone and print their JSON (not a bench line)")
    ap.add_argument("--calibrate", action="store_true",
                    help="after the timed steps, read one column once (the known-size read that scales "
                         "rocprofv3 FETCH_SIZE readings)")
    ap.add_argument("--set", action="append", default=[], metavar="KEY=VALUE",
                    help="set a library setting (qe_settings_set_u64) before anything runs, e.g. "
                         "QE_BATCH_DEDUPE=0; repeatable; the JSON line's config lists them")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {a.gpus}); this build has no multi-GPU path, "
              "run it with --gpus 1.", file=sys.stderr)
        return 2

    import computeql_amd as cq

    lib = cq.load()
    settings = {}
    for kv in a.set:
        k, _, v = kv.partition("=")
        settings[k.strip()] = int(v, 0)
        cq.set_setting(k.strip(), settings[k.strip()])
    if a.only == "batched":
        cq.profile(True)
        batched, _ = batched_and_served(cq, lib, a.served_seconds, served_too=False)
        print(json.dumps({"batched": batched}), flush=True)
        return 0
    seed, rows = SEED, a.rows
    log(f"building tables: {rows} fact rows x 5 u32 columns in HBM")
    t0 = time.perf_counter()
    t, dima, dimb = build_tables(cq, rows, seed)
    lib.gpu_hip_wait()
    log(f"tables ready in {time.perf_counter() - t0:.1f}s")
    if a.only == "served":
        batched, served = batched_and_served(cq, lib, a.served_seconds, fact=t, dima=dima, dimb=dimb)
        print(json.dumps({"batched": batched, "served": served}), flush=True)
        for x in (t, dima, dimb):
            x.close()
        return 0

    cq.profile(True)
    errs = []
    for w in range(a.warmup):
        res = run_suite(cq, lib, t, dima, dimb, w, seed, rows)
        errs += check_invariants(res, rows, w)
        log(f"warmup {w} done")
    lib.gpu_hip_wait()
    lib.gpu_hip_profile_reset()
    wall = {q: 0.0 for q in QUERY_KERNELS}
    t_start = time.perf_counter()
    for s in range(a.steps):
        ts0 = time.perf_counter()
        res = run_suite(cq, lib, t, dima, dimb, a.warmup + s, seed, rows)
        for q in wall:
            wall[q] += trace_of(res[q])["wall_ns"] / 1e9
        log(f"step {s}: {1e3 * (time.perf_counter() - ts0):.2f} ms")
    lib.gpu_hip_wait()
    elapsed = time.perf_counter() - t_start
    errs += check_invariants(res, rows, a.warmup + a.steps - 1)

    total_bytes = sum(QUERY_BYTES_PER_ROW.values()) * rows * a.steps
    value = total_bytes / elapsed / 1e9
    # per-query kernel times from HIP events recorded on the launch stream
    kernels = {}
    kernel_ns = {q: 0 for q in QUERY_KERNELS}
    kernel_names = {q: [] for q in QUERY_KERNELS}
    for q, ks in QUERY_KERNELS.items():
        for k, bpr in ks:
            ns, d = profile_stats(lib, f"{q}:{k}")
            if not d:
                continue
            kernel_ns[q] += ns
            kernel_names[q].append(k)
            kernels[f"{q}:{k}"] = {"dispatches": d, "avg_us": round(ns / d / 1e3, 2), "bytes_per_row": bpr}
    ceilings = device_ceilings(cq, lib, t)
    log(f"ceilings: read {ceilings['read_gbs']} GB/s, copy {ceilings['copy_gbs']} GB/s")
    # each query's roofline on the algorithmic bytes of SURVEY 8(d): rows x
    # 4 B per u32 column it reads in full, over its kernels' time per run
    per_query = {}
    for q, bpr in QUERY_BYTES_PER_ROW.items():
        kms = kernel_ns[q] / a.steps / 1e6
        nbytes = bpr * rows
        gbs = nbytes / kms / 1e6 if kms else None
        traffic = measured_traffic(q, kernel_names[q], rows)
        per_query[q] = {"algorithmic_bytes": nbytes, "kernel_ms": round(kms, 3),
                        "wall_ms": round(1e3 * wall[q] / a.steps, 3), "kernel_gbs": round(gbs, 1) if gbs else None,
                        "frac": round(gbs / PEAK_GBS, 4) if gbs else None,
                        "frac_of_read_ceiling": round(gbs / ceilings["read_gbs"], 4) if gbs and ceilings["read_gbs"] else None,
                        "wall_gbs": round(nbytes / (wall[q] / a.steps) / 1e9, 1) if wall[q] else None,
                        "traffic_bytes": traffic,
                        "traffic_over_algorithmic": round(traffic / nbytes, 2) if traffic else None,
                        "kernels": kernel_names[q]}
    # the dominant unit: the query with the most device time, its kernels
    # together (their bytes are the query's algorithmic bytes)
    dom_q = max(per_query, key=lambda q: per_query[q]["kernel_ms"])
    dom = per_query[dom_q]
    chain = " + ".join(k.rstrip("*") for k in kernel_names[dom_q])
    if dom_q == "q6":
        chain = chain.replace("join_part_probe", res["q6"].trace["kernel"])
    roofline = {"bound": "hbm", "achieved": dom["kernel_gbs"], "peak": PEAK_GBS, "unit": "GB/s", "frac": dom["frac"],
                "traffic": dom["traffic_bytes"], "traffic_source": TRAFFIC_SOURCE,
                "kernel": f"{dom_q}: {chain}", "bytes_per_launch": dom["algorithmic_bytes"],
                "avg_us": round(1e3 * dom["kernel_ms"], 1),
                "note": "achieved = the query's algorithmic bytes (rows x 4 B x u32 columns read in full) / the summed "
                        "hipEvent time of its kernels per run; traffic = HBM bytes those kernels move per run",
                "ceilings": ceilings,
                "frac_of_measured_ceiling": round(dom["kernel_gbs"] / ceilings["read_gbs"], 4)
                if ceilings["read_gbs"] else None}
    if a.calibrate:
        # a plain read of one known 4-byte column: its counter reading scales
        # the FETCH_SIZE of the kernels in a --pmc run of this command
        log("calibration read")
        cq.bandwidth_read(t, 2)
        lib.gpu_hip_wait()

    extra = extra_measurements(cq, lib, t, rows)
    c1 = c1_measurements(cq, lib, t, seed)

    encoded = None
    if not a.no_encoded:
        encoded = encoded_run(cq, lib, t, dima, dimb, seed, rows, a.steps)
        errs += [] if encoded["invariants"] == "ok" else [f"encoded {e}" for e in encoded["invariants"]]

    streamed = None
    if not a.no_streamed:
        streamed = streamed_run(cq, lib, t, rows)
        errs += [] if streamed["invariants"] == "ok" else streamed["invariants"]

    batched = served = None
    if not a.no_served:
        batched, served = batched_and_served(cq, lib, a.served_seconds, fact=t, dima=dima, dimb=dimb)

    cpu = None
    if not a.no_cpu_baseline:
        try:
            workers = len(os.sched_getaffinity(0))
        except Exception:
            workers = os.cpu_count() or 1
        log(f"cpu baseline: {a.cpu_sample_rows} rows, {workers} threads, at least {a.cpu_seconds} s")
        gbs, reps, secs, gbs1 = timed_cpu_baseline(a.cpu_sample_rows, seed, workers, a.cpu_seconds)
        cpu = {"value": round(gbs, 3), "unit": "GB/s", "cores": workers, "kind": "port", "cpu_model": cpu_model(),
               "sample": f"q1-q7 run {reps} times over a {a.cpu_sample_rows}-row sample of the same seeded tables "
                         f"(the oracle's C restatement of the CPU operator path, {workers} threads, "
                         f"{sum(secs.values()):.1f} s)",
               "value_1_core": round(gbs1, 3),
               "per_query_s": {k: round(v / reps, 4) for k, v in secs.items()}}

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * elapsed / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic: splitmix64-seeded tables generated in HBM",
        "config": {"workload": f"{rows}-row 4xu32 fact table + u32 fk, dim A 1M rows, dim B 50 rows; q1-q7 per step",
                   "rows": rows, "pct_hbm_peak": round(100 * value / PEAK_GBS, 2),
                   "per_query_wall_ms": {q: round(1e3 * v / a.steps, 3) for q, v in wall.items()},
                   "kernels": kernels, "per_query": per_query,
                   "invariants": "ok" if not errs else sorted(set(errs)),
                   "settings": settings},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
        "c1_10m": c1,
        "encoded": encoded,
        "streamed": streamed,
        "batched": batched,
        "served": served,
    }
    print(json.dumps(out), flush=True)
    t.close()
    dima.close()
    dimb.close()
    return 0 if not errs else 1


if __name__ == "__main__":
    sys.exit(main())
