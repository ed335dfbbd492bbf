"""The served request loop (qe_server_* of include/cqlhip.h) from Python.

    srv = Server(workers=2, max_batch=64)
    ticket = srv.submit(table, AggSpec([COUNT], [None], col(1) < 5))
    result = ticket.result()          # AggResult, as Table.aggregate returns
    srv.submit(fact, JoinSpec(dim, 4, 0, [SUM], [2], dim_group=1)).result()
    srv.submit(table, ScanSpec(col(0) == 42)).result()   # (count, ids)
    stats = srv.drive(table, specs, rate_qps=20000, seconds=2.0)
    srv.close()

Aggregates batch into shared scans; joins on one fact table batch through
qe_aggregate_join_batch (identical joins once, a fact side shared); row-id
scans (qe_scan_filter) run one per launch beside them.  drive() and
drive_mix() run the native Poisson load driver (arrivals generated in C, so
the offered rate is not bounded by Python) and return the server's stats
over the run: completed, batches, joins, scans, p50/p99/max latency (ms)
and achieved QPS.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import ir
from .engine import AggResult, AggSpec, QueryError, Table, _err, group_list, shape_keys
from .native import (REQUEST_AGGREGATE, REQUEST_JOIN, REQUEST_SCAN, QE_BatchQuery, QE_JoinBatchTrace, QE_JoinQuery,
                     QE_ScanQuery, QE_ServeRequest, QE_ServerConfig, QE_ServerStats, load)


@dataclass
class JoinSpec:
    """A served join: what aggregate_join takes, the fact table being the
    one the request is submitted on."""
    dim: Table
    fact_key: int
    dim_key: int
    funcs: list
    args: list
    dim_where: object = None
    dim_group: object = None
    cap: int | None = None


@dataclass
class ScanSpec:
    """A served scan: what Table.scan takes (row ids, or the count alone)."""
    where: object
    count_only: bool = False
    cap: int = 1024


def _query(spec: AggSpec, keep: list, cap: int | None = None):
    funcs = list(spec.funcs)
    e = len(funcs)
    fa = (C.c_uint32 * e)(*funcs)
    aa = (C.c_int32 * e)(*[(-1 if a is None else int(a)) for a in spec.args])
    nodes = ir.to_nodes(spec.where)
    gl = group_list(spec.group)
    if cap is None:
        cap = spec.cap if spec.cap is not None else (1 if not gl else 1 << 17)
    q = QE_BatchQuery()
    q.where = nodes.ctypes.data if nodes is not None else None
    q.where_count = len(nodes) if nodes is not None else 0
    q.group_count = len(gl)
    for j, c in enumerate(gl):
        q.group_cols[j] = c
    q.funcs = C.addressof(fa)
    q.arg_cols = C.addressof(aa)
    q.expr_count = e
    q.cap = cap
    keep += [fa, aa, nodes]
    return q, funcs, cap


def _join_query(spec: JoinSpec, keep: list, cap: int | None = None):
    funcs = list(spec.funcs)
    e = len(funcs)
    fa = (C.c_uint32 * e)(*funcs)
    aa = (C.c_int32 * e)(*[(-1 if a is None else int(a)) for a in spec.args])
    nodes = ir.to_nodes(spec.dim_where)
    gl = group_list(spec.dim_group)
    if cap is None:
        cap = spec.cap if spec.cap is not None else (1 if not gl else 1 << 16)
    q = QE_JoinQuery()
    q.dim = spec.dim.handle
    q.fact_key_col = spec.fact_key
    q.dim_key_col = spec.dim_key
    q.dim_where = nodes.ctypes.data if nodes is not None else None
    q.dim_where_count = len(nodes) if nodes is not None else 0
    q.group_count = len(gl)
    for j, c in enumerate(gl):
        q.group_cols[j] = c
    q.funcs = C.addressof(fa)
    q.arg_cols = C.addressof(aa)
    q.expr_count = e
    q.cap = cap
    keep += [fa, aa, nodes, spec.dim]
    return q, funcs, cap


def aggregate_join_batch(fact: Table, specs, with_trace: bool = False):
    """qe_aggregate_join_batch: many joins over one fact table, identical
    ones run once and a fact side (partitioned rows, per-key totals)
    shared by the joins that need the same.  A list of AggResult (and the
    batch trace: queries, runs, fact_passes, wall_ns)."""
    lib = load()
    keep: list = []
    arr = (QE_JoinQuery * max(len(specs), 1))()
    outs = []
    for i, sp in enumerate(specs):
        q, funcs, cap = _join_query(sp, keep)
        keys = np.empty(max(cap, 1) * max(len(group_list(sp.dim_group)), 1), dtype=np.float64)
        cells = np.empty((max(cap, 1), len(funcs)), dtype=np.uint64)
        count = C.c_uint64(0)
        q.out_group_count = C.addressof(count)
        q.out_keys = keys.ctypes.data
        q.out_cells = cells.ctypes.data
        arr[i] = q
        outs.append((funcs, keys, cells, count, sp))
    tr = QE_JoinBatchTrace()
    if not lib.qe_aggregate_join_batch(fact.handle, arr, len(specs), C.byref(tr)):
        raise _err("aggregate_join_batch")
    res = []
    for i, (funcs, keys, cells, count, sp) in enumerate(outs):
        if not arr[i].ok:
            raise QueryError(f"aggregate_join_batch: join {i} failed: {lib.qe_last_error().decode()}")
        n = int(count.value)
        if n > len(cells):
            raise QueryError(f"aggregate_join_batch: join {i} has {n} groups, more than its cap {len(cells)}")
        res.append(AggResult(funcs, shape_keys(sp.dim_group, keys, n), cells[:n].copy(), {}))
    return (res, tr.as_dict()) if with_trace else res


def _scan_query(spec: ScanSpec, keep: list):
    nodes = ir.to_nodes(spec.where)
    q = QE_ScanQuery()
    q.where = nodes.ctypes.data if nodes is not None else None
    q.where_count = len(nodes) if nodes is not None else 0
    q.count_only = 1 if spec.count_only else 0
    q.cap = 0 if spec.count_only else max(int(spec.cap), 1)
    keep.append(nodes)
    return q


class ScanTicket:
    def __init__(self, server: "Server", ticket: int, q, ids, count, keep):
        self._server, self._ticket, self._q, self._ids, self._count, self._keep = server, ticket, q, ids, count, keep
        self._done = False

    def result(self, timeout: float = 60.0):
        """(count, ids or None): the row ids are the first min(count, cap)
        matches in ascending order"""
        if not self._done:
            if not self._server._lib.qe_server_wait(self._server.handle, self._ticket, int(timeout * 1e9)):
                raise TimeoutError("request not finished")
            self._done = True
        if not self._q.ok:
            raise QueryError("served scan failed")
        n = int(self._count.value)
        if self._ids is None:
            return n, None
        return n, self._ids[: min(n, len(self._ids))].copy()


class Ticket:
    def __init__(self, server: "Server", ticket: int, q, funcs, keys, cells, count, keep, group=None):
        self._server, self._ticket, self._q = server, ticket, q
        self._funcs, self._keys, self._cells, self._count, self._keep = funcs, keys, cells, count, keep
        self._group = group
        self._done = False

    def result(self, timeout: float = 60.0) -> AggResult:
        if not self._done:
            if not self._server._lib.qe_server_wait(self._server.handle, self._ticket, int(timeout * 1e9)):
                raise TimeoutError("request not finished")
            self._done = True
        if not self._q.ok:
            raise QueryError("served request failed")
        n = int(self._count.value)
        if n > len(self._cells):
            raise QueryError(f"{n} groups exceed the request's capacity {len(self._cells)}")
        return AggResult(self._funcs, shape_keys(self._group, self._keys, n), self._cells[:n].copy(),
                         {"fused": int(getattr(self._q, "fused", 0))})


class Server:
    def __init__(self, workers: int = 2, max_batch: int = 64, hbm_budget: int = 1 << 30):
        self._lib = load()
        cfg = QE_ServerConfig(workers, max_batch, hbm_budget)
        self.handle = self._lib.qe_server_start(C.byref(cfg))
        if not self.handle:
            raise _err("qe_server_start")

    def submit(self, table: Table, spec):
        """an AggSpec or JoinSpec (-> Ticket, whose result() is an
        AggResult) or a ScanSpec (-> ScanTicket)"""
        keep: list = []
        if isinstance(spec, ScanSpec):
            q = _scan_query(spec, keep)
            ids = None if spec.count_only else np.empty(q.cap, dtype=np.uint64)
            count = C.c_uint64(0)
            q.out_ids = ids.ctypes.data if ids is not None else None
            q.out_count = C.addressof(count)
            t = self._lib.qe_server_submit_scan(self.handle, table.handle, C.byref(q))
            if not t:
                raise _err("qe_server_submit_scan")
            return ScanTicket(self, t, q, ids, count, keep)
        if isinstance(spec, JoinSpec):
            q, funcs, cap = _join_query(spec, keep)
            keys = np.empty(max(cap, 1) * max(len(group_list(spec.dim_group)), 1), dtype=np.float64)
            cells = np.empty((max(cap, 1), len(funcs)), dtype=np.uint64)
            count = C.c_uint64(0)
            q.out_group_count = C.addressof(count)
            q.out_keys = keys.ctypes.data
            q.out_cells = cells.ctypes.data
            t = self._lib.qe_server_submit_join(self.handle, table.handle, C.byref(q))
            if not t:
                raise _err("qe_server_submit_join")
            return Ticket(self, t, q, funcs, keys, cells, count, keep, spec.dim_group)
        q, funcs, cap = _query(spec, keep)
        keys = np.empty(max(cap, 1) * max(len(group_list(spec.group)), 1), dtype=np.float64)
        cells = np.empty((max(cap, 1), len(funcs)), dtype=np.uint64)
        count = C.c_uint64(0)
        q.out_group_count = C.addressof(count)
        q.out_keys = keys.ctypes.data
        q.out_cells = cells.ctypes.data
        t = self._lib.qe_server_submit(self.handle, table.handle, C.byref(q))
        if not t:
            raise _err("qe_server_submit")
        return Ticket(self, t, q, funcs, keys, cells, count, keep, spec.group)

    def stats(self) -> dict:
        st = QE_ServerStats()
        self._lib.qe_server_stats(self.handle, C.byref(st))
        return _stats_dict(st)

    def reset_stats(self) -> None:
        self._lib.qe_server_stats_reset(self.handle)

    def drive(self, table: Table, specs, rate_qps: float, seconds: float, seed: int = 1, cap: int = 1024) -> dict:
        keep: list = []
        pool = (QE_BatchQuery * len(specs))()
        for i, sp in enumerate(specs):
            q, _, _ = _query(sp, keep, cap=cap)
            pool[i] = q
        st = QE_ServerStats()
        ok = self._lib.qe_server_drive(self.handle, table.handle, pool, len(specs), float(rate_qps),
                                       int(seconds * 1e9), int(seed), C.byref(st))
        if not ok:
            raise _err("qe_server_drive")
        d = _stats_dict(st)
        d["offered_qps"] = rate_qps
        return d

    def drive_mix(self, pairs, rate_qps: float, seconds: float, seed: int = 1, cap: int = 1024) -> dict:
        """drive() over a mix: pairs is a list of (table, spec), each spec
        an AggSpec, JoinSpec or ScanSpec; request k runs pairs[k %
        len(pairs)]."""
        keep: list = []
        n = len(pairs)
        pool = (QE_ServeRequest * n)()
        for i, (t, sp) in enumerate(pairs):
            if isinstance(sp, ScanSpec):
                q, kind = _scan_query(sp, keep), REQUEST_SCAN
            elif isinstance(sp, JoinSpec):
                q, kind = _join_query(sp, keep, cap=sp.cap if sp.cap is not None else cap)[0], REQUEST_JOIN
            else:
                q, kind = _query(sp, keep, cap=cap)[0], REQUEST_AGGREGATE
            keep.append(q)
            pool[i].kind = kind
            pool[i].table = t.handle
            pool[i].query = C.addressof(q)
        st = QE_ServerStats()
        ok = self._lib.qe_server_drive_requests(self.handle, pool, n, float(rate_qps), int(seconds * 1e9), int(seed),
                                                C.byref(st))
        if not ok:
            raise _err("qe_server_drive_requests")
        d = _stats_dict(st)
        d["offered_qps"] = rate_qps
        return d

    def close(self) -> None:
        if self.handle:
            self._lib.qe_server_stop(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _stats_dict(st: QE_ServerStats) -> dict:
    d = st.as_dict()
    secs = d["elapsed_ns"] / 1e9
    d["qps"] = d["completed"] / secs if secs > 0 else 0.0
    for k in ("p50_ns", "p99_ns", "max_ns", "mean_ns"):
        d[k.replace("_ns", "_ms")] = d.pop(k) / 1e6
    d["mean_batch"] = d["batched_queries"] / d["batches"] if d["batches"] else 0.0
    return d
