"""Host-side mirror of ComputeQL's operator interface for the hot path.

Thin Python over the C ABI (include/cqlhip.h): device-resident tables, the
scan/filter operator, global and grouped aggregates and the join+aggregate
pushdown.  Every call goes to the HIP library; an error in the library
raises `QueryError` with its message (the reference's GPU operators return
0 to their caller in the same situations).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import ir
from .native import (MAX_GROUP_COLS, QE_BatchConfig, QE_BatchQuery, QE_BatchTrace, QE_ResidentStats, QE_StreamStats,
                     QE_Trace, load)

def _np_format(t):
    if issubclass(t, C.Array):
        return (_np_format(t._type_), (t._length_,))
    return "<u8" if t in (C.c_void_p, C.c_uint64) else ("<i4" if t is C.c_int32 else "<u4")


# QE_BatchQuery as a numpy record: a batch's records are filled column-wise
_BQ_DTYPE = np.dtype({
    "names": [f for f, _ in QE_BatchQuery._fields_],
    "formats": [_np_format(t) for _, t in QE_BatchQuery._fields_],
    "offsets": [getattr(QE_BatchQuery, f).offset for f, _ in QE_BatchQuery._fields_],
    "itemsize": C.sizeof(QE_BatchQuery),
})


def group_list(group) -> list:
    """GROUP BY columns as a list: None -> [], a column -> [it], a list or
    tuple of up to 4 columns as given."""
    if group is None:
        return []
    if isinstance(group, (list, tuple)):
        if len(group) > MAX_GROUP_COLS:
            raise ValueError(f"at most {MAX_GROUP_COLS} GROUP BY columns")
        return [int(g) for g in group]
    return [int(group)]


def shape_keys(group, keys: np.ndarray, n: int) -> np.ndarray:
    """A result's keys as the caller asked for them: (n, K) for a list of
    columns, (n,) for one column, n zeros without a GROUP BY.  `keys` holds
    at least n * K values, K key values per group."""
    k = len(group_list(group))
    if k == 0:
        return np.zeros(n, dtype=np.float64)
    flat = np.asarray(keys, dtype=np.float64).reshape(-1)[: n * k]
    return flat.reshape(n, k).copy() if isinstance(group, (list, tuple)) else flat.copy()

U32, U64, F32, F64, I32, I64 = 1, 2, 3, 4, 7, 8
COUNT, SUM, AVG, MIN, MAX = 0, 1, 2, 3, 4
FUNC_NAMES = {COUNT: "COUNT", SUM: "SUM", AVG: "AVG", MIN: "MIN", MAX: "MAX"}
_NP = {U32: np.uint32, U64: np.uint64, F32: np.float32, F64: np.float64, I32: np.int32, I64: np.int64}
SCAN_PATHS = {0: "none", 1: "generated_simple", 2: "generated", 3: "simple", 4: "interpreter", 5: "zone_empty",
              6: "constant"}
AGG_PATHS = {10: "global", 11: "fused_global", 12: "dense", 13: "dense_key_fold", 14: "fused_dense", 20: "join_dense",
             21: "join_probe_group", 22: "join_probe_row", 23: "join_part_group", 24: "join_part_row",
             30: "hash"}


class QueryError(RuntimeError):
    pass


def _err(what: str) -> QueryError:
    lib = load()
    return QueryError(f"{what}: {lib.qe_last_error().decode(errors='replace')}")


def set_setting(key: str, value: int) -> None:
    load().qe_settings_set_u64(key.encode(), int(value))


def get_setting(key: str, default: int = 0) -> int:
    return int(load().qe_settings_u64(key.encode(), int(default)))


def reset_settings() -> None:
    load().qe_settings_reset()


def profile(enabled: bool = True) -> None:
    load().gpu_hip_profile_enable(1 if enabled else 0)


def compute_units() -> int:
    return int(load().gpu_hip_compute_units())


@dataclass
class AggResult:
    funcs: list
    keys: np.ndarray   # (groups,) for one key (zeros without one), (groups, K) for a list of K keys
    cells: np.ndarray  # uint64 [groups, exprs]: COUNT as U64, others F64 bits
    trace: dict = field(default_factory=dict)

    @property
    def groups(self) -> int:
        return len(self.cells)

    def column(self, e: int) -> np.ndarray:
        c = self.cells[:, e]
        return c.copy() if self.funcs[e] == COUNT else c.view(np.float64).copy()

    def rows(self) -> list:
        out = []
        for g in range(self.groups):
            row = [float(k) for k in np.atleast_1d(self.keys[g])]
            for e, f in enumerate(self.funcs):
                v = self.cells[g, e]
                row.append(int(v) if f == COUNT else float(np.array([v], dtype=np.uint64).view(np.float64)[0]))
            out.append(tuple(row))
        return out


class Table:
    """A table whose columns live whole in HBM (CQL_Table, cql.h:206-322)."""

    def __init__(self, rows: int):
        self._lib = load()
        self.handle = self._lib.qe_table_alloc(int(rows))
        if not self.handle:
            raise QueryError("qe_table_alloc failed")
        self.rows = int(rows)
        self.names: list[str] = []
        self.types: list[int] = []

    def close(self) -> None:
        if self.handle:
            self._lib.qe_table_release(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_column(self, name: str, ctype: int, data: np.ndarray | None = None, nulls: np.ndarray | None = None) -> int:
        """nulls: a byte per row, nonzero = NULL (cql.h:283).  WHEREs that are
        an AND of comparisons and IS [NOT] NULL tests evaluate such a column
        (a NULL row fails its comparison); other uses of it (a key, an
        argument, OR/IN shapes, joins, gathers) are refused."""
        ptr = None
        if data is not None:
            data = np.ascontiguousarray(data, dtype=_NP[ctype])
            if data.shape != (self.rows,):
                raise ValueError(f"column {name}: expected {self.rows} rows, got {data.shape}")
            ptr = data.ctypes.data
        nptr = None
        if nulls is not None:
            nulls = np.ascontiguousarray(nulls, dtype=np.uint8)
            if nulls.shape != (self.rows,):
                raise ValueError(f"column {name}: expected {self.rows} NULL flags, got {nulls.shape}")
            nptr = nulls.ctypes.data
        idx = self._lib.qe_table_add_column(self.handle, name.encode(), ctype, ptr, nptr)
        if idx < 0:
            raise _err("add_column")
        self.names.append(name)
        self.types.append(ctype)
        return int(idx)

    def add_host_column(self, name: str, ctype: int, data: np.ndarray) -> int:
        """qe_table_add_host_column: the rows stay in pinned host memory and
        scans/aggregates stream them through the resident chunk cache.  A
        table holds either host-backed or device columns."""
        data = np.ascontiguousarray(data, dtype=_NP[ctype])
        if data.shape != (self.rows,):
            raise ValueError(f"column {name}: expected {self.rows} rows, got {data.shape}")
        idx = self._lib.qe_table_add_host_column(self.handle, name.encode(), ctype, data.ctypes.data)
        if idx < 0:
            raise _err("add_host_column")
        self.names.append(name)
        self.types.append(ctype)
        return int(idx)

    def col(self, name: str) -> ir.Column:
        return ir.col(self.names.index(name))

    def index(self, name: str) -> int:
        return self.names.index(name)

    def read(self, column: int | str, row0: int = 0, rows: int | None = None) -> np.ndarray:
        c = self.index(column) if isinstance(column, str) else column
        n = self.rows - row0 if rows is None else rows
        out = np.empty(n, dtype=_NP[self.types[c]])
        if not self._lib.qe_table_column_read(self.handle, c, row0, n, out.ctypes.data):
            raise _err("column_read")
        return out

    def encode(self, column, allow_sequence: bool = True) -> int:
        """qe_table_encode_column: keep the column also as narrow codes (1 u8,
        2 u16) or as a sequence (3); 0 when none applies."""
        r = self._lib.qe_table_encode_column(self.handle, self._col_index(column), 1 if allow_sequence else 0)
        if r < 0:
            raise _err("encode")
        return int(r)

    def build_zonemap(self, column: int) -> None:
        if not self._lib.qe_table_build_zonemap(self.handle, column):
            raise _err("build_zonemap")

    # ---- generators (data produced in HBM) ----
    def gen_events(self, seed: int, id=-1, region=-1, amount=-1, ts=-1) -> None:
        if not self._lib.qe_gen_events(self.handle, seed & (2**64 - 1), id, region, amount, ts):
            raise _err("gen_events")

    def gen_scale_events(self, id=-1, region=-1, amount=-1, key=-1) -> None:
        if not self._lib.qe_gen_scale_events(self.handle, id, region, amount, key):
            raise _err("gen_scale_events")

    def gen_dim(self, seed: int, attr_range: int, spread_keys: bool = True, key: int = 0, attr: int = -1) -> None:
        if not self._lib.qe_gen_dim(self.handle, seed & (2**64 - 1), attr_range, 1 if spread_keys else 0, key, attr):
            raise _err("gen_dim")

    def gen_fact_fk(self, fk_col: int, seed: int, dim: "Table", dim_key_col: int) -> None:
        if not self._lib.qe_gen_fact_fk(self.handle, fk_col, seed & (2**64 - 1), dim.handle, dim_key_col):
            raise _err("gen_fact_fk")

    # ---- operators ----
    def scan(self, where, count_only: bool = False, cap: int | None = None):
        """qe_scan_filter_selected: returns (count, ids or None, trace)."""
        nodes = ir.to_nodes(where)
        nptr = nodes.ctypes.data if nodes is not None else None
        ncount = len(nodes) if nodes is not None else 0
        cnt = C.c_uint64(0)
        tr = QE_Trace()
        if count_only:
            ok = self._lib.qe_scan_filter(self.handle, nptr, ncount, 1, None, 0, C.byref(cnt), C.byref(tr))
            if not ok:
                raise _err("scan")
            return int(cnt.value), None, tr.as_dict()
        cap = self.rows if cap is None else cap
        ids = np.empty(max(cap, 1), dtype=np.uint64)
        ok = self._lib.qe_scan_filter(self.handle, nptr, ncount, 0, ids.ctypes.data, cap, C.byref(cnt), C.byref(tr))
        if not ok:
            raise _err("scan")
        n = int(cnt.value)
        return n, ids[: min(n, cap)], tr.as_dict()

    def count(self, where) -> int:
        return self.scan(where, count_only=True)[0]

    def _col_index(self, c):
        return self.index(c) if isinstance(c, str) else int(c)

    def gather(self, columns, rows, stride: int = 1, offset: int = 0) -> list:
        """qe_gather: the values of `columns` at rows[i * stride + offset]
        (0 where the id is all ones), one array per column."""
        cols = [self._col_index(c) for c in columns]
        rows = np.ascontiguousarray(rows, dtype=np.uint64).reshape(-1)
        n = (len(rows) - offset + stride - 1) // stride if len(rows) > offset else 0
        outs = [np.empty(max(n, 1), dtype=_NP[self.types[c]]) for c in cols]
        ca = (C.c_uint32 * max(len(cols), 1))(*cols)
        pa = (C.c_void_p * max(len(cols), 1))(*[o.ctypes.data for o in outs])
        if not self._lib.qe_gather(self.handle, ca, len(cols), rows.ctypes.data, n, stride, offset, pa):
            raise _err("gather")
        return [o[:n] for o in outs]

    def select(self, columns, where=None, cap: int | None = None) -> list:
        """qe_scan_materialize: SELECT columns WHERE ..., in row order."""
        cols = [self._col_index(c) for c in columns]
        nodes = ir.to_nodes(where)
        cap = self.rows if cap is None else cap
        outs = [np.empty(max(cap, 1), dtype=_NP[self.types[c]]) for c in cols]
        ca = (C.c_uint32 * max(len(cols), 1))(*cols)
        pa = (C.c_void_p * max(len(cols), 1))(*[o.ctypes.data for o in outs])
        cnt = C.c_uint64(0)
        ok = self._lib.qe_scan_materialize(self.handle, nodes.ctypes.data if nodes is not None else None,
                                           len(nodes) if nodes is not None else 0, ca, len(cols), cap, C.byref(cnt),
                                           pa)
        if not ok:
            raise _err("select")
        n = min(int(cnt.value), cap)
        return [o[:n] for o in outs]

    def aggregate(self, funcs, args, where=None, group=None, cap: int | None = None) -> AggResult:
        """qe_aggregate_impl for the hot-path shapes (args[e] < 0 or None =
        COUNT(*)); group: None, a column, or a list of up to 4 columns."""
        funcs = list(funcs)
        e = len(funcs)
        fa = (C.c_uint32 * e)(*funcs)
        aa = (C.c_int32 * e)(*[(-1 if a is None else int(a)) for a in args])
        gl = group_list(group)
        ga = (C.c_int32 * max(len(gl), 1))(*gl)
        nodes = ir.to_nodes(where)
        nptr = nodes.ctypes.data if nodes is not None else None
        ncount = len(nodes) if nodes is not None else 0
        if cap is None:
            cap = 1 if not gl else 1 << 17
        keys = np.empty(max(cap, 1) * max(len(gl), 1), dtype=np.float64)
        cells = np.empty((max(cap, 1), e), dtype=np.uint64)
        g = C.c_uint64(0)
        tr = QE_Trace()
        ok = self._lib.qe_aggregate(self.handle, nptr, ncount, ga, len(gl), fa, aa, e, cap, C.byref(g),
                                    keys.ctypes.data, cells.ctypes.data, C.byref(tr))
        if not ok:
            raise _err("aggregate")
        n = int(g.value)
        if n > cap:
            return self.aggregate(funcs, args, where, group, cap=n)
        return AggResult(funcs, shape_keys(group, keys, n), cells[:n].copy(), tr.as_dict())


@dataclass
class AggSpec:
    """One query of a batch: what Table.aggregate takes."""
    funcs: list
    args: list
    where: object = None
    group: object = None  # None, a column, or a list of up to 4 columns
    cap: int | None = None

    def prepared(self):
        """(funcs, u32 function codes, i32 argument columns, IR nodes, and
        their addresses: (nodes, node count, codes, columns)) as the C ABI
        takes them, built once per spec (a spec is not changed after
        it is built), so a query submitted again converts nothing."""
        p = self.__dict__.get("_prepared")
        if p is None:
            funcs = list(self.funcs)
            fa = np.asarray(funcs, dtype=np.uint32)
            aa = np.asarray([-1 if a is None else int(a) for a in self.args], dtype=np.int32)
            nodes = ir.to_nodes(self.where)
            ptrs = (nodes.ctypes.data if nodes is not None else 0, len(nodes) if nodes is not None else 0,
                    fa.ctypes.data, aa.ctypes.data)
            p = (funcs, fa, aa, nodes, ptrs)
            self._prepared = p
        return p


def _batch_records(specs):
    """The QE_BatchQuery array of a batch as one numpy record array (the C
    struct's layout) and the outputs: every query's keys and cells are
    views of one arena.  A grouped query gets room for 4096 groups unless
    its spec says otherwise (the shared scan fuses at most 1024; a larger
    result is run again with room for all of it)."""
    n = len(specs)
    recs = np.zeros(max(n, 1), dtype=_BQ_DTYPE)
    counts = np.zeros(max(n, 1), dtype=np.uint64)
    preps = [sp.prepared() for sp in specs]
    groups = [group_list(sp.group) for sp in specs]
    caps = [sp.cap if sp.cap is not None else (1 if not gl else 1 << 12) for sp, gl in zip(specs, groups)]
    caps = [max(c, 1) for c in caps]
    widths = [len(p[0]) for p in preps]
    kw = [max(len(gl), 1) for gl in groups]
    sizes = [c * (k + e) for c, k, e in zip(caps, kw, widths)]
    arena = np.empty(max(sum(sizes), 1), dtype=np.uint64)
    base = arena.ctypes.data
    offs = np.zeros(n, dtype=np.uint64)
    if n > 1:
        offs[1:] = np.cumsum(sizes[:-1], dtype=np.uint64)
    outs = []
    for (funcs, _, _, _, _), cap, k, e, o in zip(preps, caps, kw, widths, offs.tolist()):
        outs.append((funcs, arena[o:o + cap * k].view(np.float64), arena[o + cap * k:o + cap * (k + e)].reshape(cap, e),
                     cap))
    if n:
        recs["where"][:n] = [p[4][0] for p in preps]
        recs["where_count"][:n] = [p[4][1] for p in preps]
        recs["group_count"][:n] = [len(gl) for gl in groups]
        recs["group_cols"][:n] = [gl + [-1] * (MAX_GROUP_COLS - len(gl)) for gl in groups]
        recs["funcs"][:n] = [p[4][2] for p in preps]
        recs["arg_cols"][:n] = [p[4][3] for p in preps]
        recs["expr_count"][:n] = widths
        recs["cap"][:n] = caps
        recs["out_group_count"][:n] = counts.ctypes.data + 8 * np.arange(n, dtype=np.uint64)
        o = base + 8 * offs
        recs["out_keys"][:n] = o
        recs["out_cells"][:n] = o + 8 * np.asarray(caps, dtype=np.uint64) * np.asarray(kw, dtype=np.uint64)
    return recs, counts, outs


def aggregate_batch(table: Table, specs, with_trace: bool = False):
    """qe_aggregate_batch: many aggregate queries over one table in shared
    scans.  Returns one AggResult per spec (trace holds 'fused'), and the
    batch trace when with_trace is set."""
    lib = load()
    specs = [s if isinstance(s, AggSpec) else AggSpec(**s) for s in specs]
    n = len(specs)
    recs, counts, outs = _batch_records(specs)
    tr = QE_BatchTrace()
    ok = lib.qe_aggregate_batch(table.handle, recs.ctypes.data, n, C.byref(tr))
    results = []
    oks, fused = recs["ok"], recs["fused"]
    for i, (funcs, keys, cells, cap) in enumerate(outs):
        if not oks[i]:
            raise _err(f"aggregate_batch query {i}")
        ng = int(counts[i])
        if ng > cap:
            sp = specs[i]
            r = table.aggregate(sp.funcs, sp.args, sp.where, sp.group, cap=ng)
            r.trace["fused"] = 0
            results.append(r)
            continue
        results.append(AggResult(funcs, shape_keys(specs[i].group, keys, ng), cells[:ng], {"fused": int(fused[i])}))
    if not ok:
        raise _err("aggregate_batch")
    return (results, tr.as_dict()) if with_trace else results


def autotune_batch(table: Table, specs) -> dict:
    """qe_aggregate_batch_tune: times the shared-scan configurations on this
    batch and records the fastest for its workload; returns it with its
    time in ms (used by aggregate_batch when QE_BATCH_AUTOTUNE is 1)."""
    lib = load()
    specs = [s if isinstance(s, AggSpec) else AggSpec(**s) for s in specs]
    qs = (QE_BatchQuery * max(len(specs), 1))()
    keep = []
    for i, sp in enumerate(specs):
        funcs = list(sp.funcs)
        e = len(funcs)
        fa = (C.c_uint32 * e)(*funcs)
        aa = (C.c_int32 * e)(*[(-1 if a is None else int(a)) for a in sp.args])
        nodes = ir.to_nodes(sp.where)
        gl = group_list(sp.group)
        cap = sp.cap if sp.cap is not None else (1 if not gl else 1 << 17)
        keys = np.empty(max(cap, 1) * max(len(gl), 1), dtype=np.float64)
        cells = np.empty((max(cap, 1), e), dtype=np.uint64)
        g = C.c_uint64(0)
        keep += [fa, aa, nodes, keys, cells, g]
        q = qs[i]
        q.where = nodes.ctypes.data if nodes is not None else None
        q.where_count = len(nodes) if nodes is not None else 0
        q.group_count = len(gl)
        for j, c in enumerate(gl):
            q.group_cols[j] = c
        q.funcs, q.arg_cols, q.expr_count, q.cap = C.addressof(fa), C.addressof(aa), e, cap
        q.out_group_count, q.out_keys, q.out_cells = C.addressof(g), keys.ctypes.data, cells.ctypes.data
    best = QE_BatchConfig()
    ms = C.c_double(0.0)
    if not lib.qe_aggregate_batch_tune(table.handle, qs, len(specs), C.byref(best), C.byref(ms)):
        raise _err("autotune_batch")
    out = best.as_dict()
    out["ms"] = ms.value
    return out


def aggregate_join(fact: Table, fact_key: int, dim: Table, dim_key: int, funcs, fact_args, dim_where=None,
                   dim_group=None, cap: int = 1 << 16) -> AggResult:
    """qe_aggregate_join_pushdown: aggregates of fact rows grouped by dim
    columns (dim_group: None, a column, or a list of up to 4 columns)."""
    lib = load()
    funcs = list(funcs)
    e = len(funcs)
    fa = (C.c_uint32 * e)(*funcs)
    aa = (C.c_int32 * e)(*[(-1 if a is None else int(a)) for a in fact_args])
    gl = group_list(dim_group)
    ga = (C.c_int32 * max(len(gl), 1))(*gl)
    nodes = ir.to_nodes(dim_where)
    nptr = nodes.ctypes.data if nodes is not None else None
    ncount = len(nodes) if nodes is not None else 0
    keys = np.empty(max(cap, 1) * max(len(gl), 1), dtype=np.float64)
    cells = np.empty((max(cap, 1), e), dtype=np.uint64)
    g = C.c_uint64(0)
    tr = QE_Trace()
    ok = lib.qe_aggregate_join(fact.handle, fact_key, dim.handle, dim_key, nptr, ncount, ga, len(gl), fa, aa, e, cap,
                               C.byref(g), keys.ctypes.data, cells.ctypes.data, C.byref(tr))
    if not ok:
        raise _err("aggregate_join")
    n = int(g.value)
    if n > cap:
        return aggregate_join(fact, fact_key, dim, dim_key, funcs, fact_args, dim_where, dim_group, cap=n)
    return AggResult(funcs, shape_keys(dim_group, keys, n), cells[:n].copy(), tr.as_dict())


JOIN_TYPES = {"inner": 0, "left": 1, "semi": 2, "anti": 3}
BUILD_NULL = (1 << 64) - 1


def hash_join(probe: Table, probe_key: int, build: Table, build_key: int, join_type: str = "inner",
              probe_rows=None, build_rows=None, cap: int | None = None, with_trace: bool = False):
    """qe_hash_join_impl: the equi-join probe.key = build.key.

    inner / left: an (n, 2) u64 array of (probe position, build row) pairs,
    a left join's unmatched position paired with BUILD_NULL; semi / anti:
    the probe positions kept, ascending.  Sorted by probe position, then
    build position.  probe_rows / build_rows (u64 arrays) restrict the
    sides; probe positions index probe_rows when it is given."""
    lib = load()
    jt = JOIN_TYPES[join_type]
    keep = []

    def rows_arg(rows):
        if rows is None:
            return None, 0
        arr = np.ascontiguousarray(rows, dtype=np.uint64)
        keep.append(arr)
        return arr.ctypes.data, len(arr)

    pr, pc = rows_arg(probe_rows)
    br, bc = rows_arg(build_rows)
    rec = 1 if jt >= 2 else 2
    cap = (pc if probe_rows is not None else probe.rows) if cap is None else cap
    out = np.empty(max(cap, 1) * rec, dtype=np.uint64)
    n = C.c_uint64(0)
    tr = QE_Trace()
    ok = lib.qe_hash_join(probe.handle, probe_key, pr, pc, build.handle, build_key, br, bc, jt, out.ctypes.data, cap,
                          C.byref(n), C.byref(tr))
    if not ok:
        raise _err("hash_join")
    total = int(n.value)
    if total > cap:
        return hash_join(probe, probe_key, build, build_key, join_type, probe_rows, build_rows, cap=total,
                         with_trace=with_trace)
    res = out[: total * rec].copy()
    if rec == 2:
        res = res.reshape(total, 2)
    return (res, tr.as_dict()) if with_trace else res


def bandwidth_read(table: Table, column: int, blocks: int = 0) -> dict:
    lib = load()
    buf = lib.qe_table_column_buffer(table.handle, column)
    sink = C.c_uint64(0)
    tr = QE_Trace()
    if not lib.qe_bandwidth_read(buf, table.rows * 4, blocks, C.byref(sink), C.byref(tr)):
        raise _err("bandwidth_read")
    return tr.as_dict()


def bandwidth_copy(table: Table, column: int, blocks: int = 0) -> dict:
    """qe_bandwidth_copy of one 4-byte column into a scratch buffer of its
    size (the buffer is freed after the copy)."""
    lib = load()
    src = lib.qe_table_column_buffer(table.handle, column)
    nbytes = table.rows * 4
    dst = lib.gpu_hip_buffer_alloc(nbytes + 64, (1 << 2) | (1 << 4), None)
    if not dst:
        raise _err("bandwidth_copy")
    try:
        tr = QE_Trace()
        if not lib.qe_bandwidth_copy(src, dst, nbytes, blocks, C.byref(tr)):
            raise _err("bandwidth_copy")
        return tr.as_dict()
    finally:
        lib.gpu_hip_buffer_release(dst)


def resident_stats() -> dict:
    """Counters of the resident chunk cache (qe_resident_stats)."""
    st = QE_ResidentStats()
    load().qe_resident_stats(C.byref(st))
    return st.as_dict()


def resident_clear() -> None:
    """Free every cached chunk no query holds (a cold start)."""
    load().qe_resident_clear()


def stream_stats() -> dict:
    """What the calling thread's last streamed scan or aggregate did."""
    st = QE_StreamStats()
    load().qe_stream_last_stats(C.byref(st))
    return st.as_dict()
