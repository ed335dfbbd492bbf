"""ctypes binding of the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
SEED = 0x1234567890ABCDEF  # bench_common.h:59

U32, U64, F32, F64, I32, I64 = 1, 2, 3, 4, 7, 8
_NP = {U32: np.uint32, U64: np.uint64, F32: np.float32, F64: np.float64, I32: np.int32, I64: np.int64}
_TYPE_OF = {np.dtype(v): k for k, v in _NP.items()}

_lib = None


class ExactnessError(ArithmeticError):
    """A SUM over a whole-number column reached 2^53: the CPU path fails
    the query (qe_aggregate_check_sum_exact, query_exec.c:7028-7049)."""


class OrcColumn(C.Structure):
    _fields_ = [("type", C.c_uint32), ("pad", C.c_uint32), ("data", C.c_void_p), ("null_flags", C.c_void_p)]


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    P, u32, u64, f64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_double, C.c_int
    sig = {
        "orc_splitmix64_at": (u64, [u64, u64]),
        "orc_gen_events": (None, [u64, u64, u64, P, P, P, P, i32]),
        "orc_gen_fact_fk": (None, [u64, u64, u64, P, u64, P, i32]),
        "orc_gen_dim": (None, [u64, u64, u32, i32, P, P]),
        "orc_gen_scale_events": (None, [u64, u64, P, P, P, P]),
        "orc_ss_sum_amount": (f64, [u64]),
        "orc_ss_count_amount_above": (u64, [u64, u64]),
        "orc_ss_count_region": (u64, [u64, u64]),
        "orc_ss_region_amount_sum": (f64, [u64, u64]),
        "orc_ss_amount_rows": (u64, [u64, u64]),
        "orc_ss_weight_sum": (f64, [u64]),
        "orc_ss_key_of_row": (u32, [u64]),
        "orc_row_condition_eval": (i32, [P, u32, P, u64]),
        "orc_cpu_scan_filter": (u64, [P, u32, P, u64, P, u64, i32]),
        "orc_scan_int_interval": (i32, [u32, f64, f64, f64, P, P]),
        "orc_aggregate_global_chunked": (None, [P, u32, u64, i32, P, P, P]),
        "orc_aggregate_cpu_reduce": (u64, [u64, P, P, u32, P, P, u32, P, P]),
        "orc_group_aggregate": (u64, [P, u64, P, u32, P, u32, P, P, u32, P, P, u64]),
        "orc_join_aggregate": (u64, [u64, P, P, P, u32, u64, P, P, P, u32, P, P, u64, i32]),
        "orc_check_sum_exact": (i32, [P, P, u32, P, u64]),
        "orc_scan_chunked_count": (u64, [P, P, P, P, u32, u64, i32]),
        "orc_hash_join": (u64, [P, P, u64, P, P, u64, u32, P, u64]),
        "orc_dense_group_u32": (None, [P, P, u64, u32, u32, i32, P, P, P, P]),
        "orc_hw_threads": (i32, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def threads() -> int:
    return int(load().orc_hw_threads())


# ---------------- generators ----------------
def splitmix64_at(seed: int, k: int) -> int:
    return int(load().orc_splitmix64_at(seed, k))


def gen_events(n: int, seed: int = SEED, row0: int = 0, nthreads: int | None = None):
    """The seeded 4 x u32 table: id, region, amount, ts."""
    cols = [np.empty(n, dtype=np.uint32) for _ in range(4)]
    load().orc_gen_events(seed, row0, n, *[c.ctypes.data for c in cols], nthreads or threads())
    return cols


def gen_dim(n: int, attr_range: int, seed: int, spread: bool = True):
    keys = np.empty(n, dtype=np.uint32)
    attr = np.empty(n, dtype=np.uint32)
    load().orc_gen_dim(seed, n, attr_range, 1 if spread else 0, keys.ctypes.data, attr.ctypes.data)
    return keys, attr


def gen_fact_fk(n: int, dim_keys: np.ndarray, seed: int = SEED, row0: int = 0, nthreads: int | None = None):
    dim_keys = np.ascontiguousarray(dim_keys, dtype=np.uint32)
    fk = np.empty(n, dtype=np.uint32)
    load().orc_gen_fact_fk(seed, row0, n, dim_keys.ctypes.data, len(dim_keys), fk.ctypes.data, nthreads or threads())
    return fk


def gen_scale_events(n: int, row0: int = 0):
    cols = [np.empty(n, dtype=np.uint32) for _ in range(4)]
    load().orc_gen_scale_events(row0, n, *[c.ctypes.data for c in cols])
    return cols


# ---------------- operator semantics ----------------
def _columns(cols, nulls=None):
    """orc_column[] over the arrays; nulls: {column index: byte-per-row
    flags, nonzero = NULL} (cql.h:283)"""
    arr = (OrcColumn * len(cols))()
    keep = []
    for i, c in enumerate(cols):
        c = np.ascontiguousarray(c)
        keep.append(c)
        arr[i].type = _TYPE_OF[c.dtype]
        arr[i].data = c.ctypes.data
        arr[i].null_flags = None
        if nulls and nulls.get(i) is not None:
            f = np.ascontiguousarray(nulls[i], dtype=np.uint8)
            keep.append(f)
            arr[i].null_flags = f.ctypes.data
    return arr, keep


IR_DTYPE = np.dtype([("type", "<u4"), ("op", "<u4"), ("column", "<u4"), ("child_count", "<u4"), ("value", "<f8")])


def _nodes(where):
    """The WHERE tree as orc_ir_node[]: an object with .nodes() or an array."""
    if where is None:
        return None, 0
    n = where.nodes() if hasattr(where, "nodes") else where
    n = np.ascontiguousarray(n).view(IR_DTYPE) if np.asarray(n).dtype != IR_DTYPE else np.ascontiguousarray(n)
    assert n.dtype.itemsize == 24
    return n, len(n)


def scan(cols, where, nthreads: int | None = None, count_only: bool = False, nulls=None):
    """qe_cpu_scan_filter: ascending row ids (or the count)."""
    arr, keep = _columns(cols, nulls)
    nodes, nn = _nodes(where)
    rows = len(cols[0])
    lib = load()
    nptr = nodes.ctypes.data if nodes is not None else None
    if count_only:
        return int(lib.orc_cpu_scan_filter(nptr, nn, arr, rows, None, 0, nthreads or threads()))
    out = np.empty(max(rows, 1), dtype=np.uint64)
    n = lib.orc_cpu_scan_filter(nptr, nn, arr, rows, out.ctypes.data, rows, nthreads or threads())
    del keep
    return out[:n]


def aggregate_global(cols, workers: int | None = None):
    """qe_aggregate_global_chunked: (sum, min, max) per column as f64."""
    arr, keep = _columns(cols)
    k = len(cols)
    s, mn, mx = (np.empty(k), np.empty(k), np.empty(k))
    load().orc_aggregate_global_chunked(arr, k, len(cols[0]), workers or threads(), s.ctypes.data, mn.ctypes.data,
                                        mx.ctypes.data)
    del keep
    return s, mn, mx


def _check_exact(funcs, arg_types, res, groups, check):
    if not check or not groups:
        return
    fa = np.ascontiguousarray(funcs, dtype=np.uint32)
    ta = np.ascontiguousarray(arg_types, dtype=np.uint32)
    r = np.ascontiguousarray(res[:groups], dtype=np.float64)
    if not load().orc_check_sum_exact(fa.ctypes.data, ta.ctypes.data, len(fa), r.ctypes.data, groups):
        raise ExactnessError("a SUM over a whole-number column reached 2^53")


def group_aggregate(cols, key_col, funcs, arg_cols, where=None, max_groups=None, nulls=None, check_exact=True):
    """qe_aggregate_cpu_reduce over the (filtered) rows: (keys, results[g, e]),
    groups in first-appearance order.  key_col: None = one group, a column
    (keys shape (g,)), or a list of up to 4 columns (keys shape (g, K));
    nulls as for scan (WHERE columns only: NULL arguments are not
    aggregated here).  Raises ExactnessError when a SUM over a whole-number
    column reaches 2^53 (check_exact False returns the rounded value)."""
    arr, keep = _columns(cols, nulls)
    nodes, nn = _nodes(where)
    rows = len(cols[0])
    e = len(funcs)
    kl = [] if key_col is None else (list(key_col) if isinstance(key_col, (list, tuple)) else [key_col])
    ka = np.array(kl or [0], dtype=np.int32)
    fa = np.array(funcs, dtype=np.uint32)
    aa = np.array([-1 if a is None else a for a in arg_cols], dtype=np.int32)
    cap = max_groups or max(rows, 1)
    keys = np.empty(cap * max(len(kl), 1))
    res = np.empty((cap, e))
    g = load().orc_group_aggregate(arr, rows, nodes.ctypes.data if nodes is not None else None, nn, ka.ctypes.data,
                                   len(kl), fa.ctypes.data, aa.ctypes.data, e, keys.ctypes.data, res.ctypes.data, cap)
    types = [0 if a is None else _TYPE_OF[np.asarray(cols[a]).dtype] for a in arg_cols]
    del keep
    _check_exact(funcs, types, res, min(g, cap), check_exact)
    if isinstance(key_col, (list, tuple)):
        return keys[: g * len(kl)].reshape(g, len(kl)).copy(), res[:g].copy()
    return (keys[:g].copy() if kl else np.zeros(g)), res[:g].copy()


def join_aggregate(fact_key, fact_args, funcs, dim_key, dim_pass=None, dim_group=None, nthreads=None,
                   max_groups=1 << 20, check_exact=True):
    """Join + aggregate semantics: (group values, results[g, e]) in first
    fact-appearance order.  dim_group: None (one group), one u32 array
    (keys shape (g,)) or a list of up to 4 (keys shape (g, K)).  Raises
    ExactnessError when a SUM reaches 2^53, as group_aggregate."""
    fk = np.ascontiguousarray(fact_key, dtype=np.uint32)
    e = len(funcs)
    args = [None if a is None else np.ascontiguousarray(a, dtype=np.uint32) for a in fact_args]
    argp = (C.c_void_p * e)(*[None if a is None else a.ctypes.data for a in args])
    fa = np.array(funcs, dtype=np.uint32)
    dk = np.ascontiguousarray(dim_key, dtype=np.uint32)
    dp = None if dim_pass is None else np.ascontiguousarray(dim_pass, dtype=np.uint8)
    many = isinstance(dim_group, (list, tuple))
    gl = [] if dim_group is None else (list(dim_group) if many else [dim_group])
    gl = [np.ascontiguousarray(x, dtype=np.uint32) for x in gl]
    gp = (C.c_void_p * max(len(gl), 1))(*[x.ctypes.data for x in gl])
    k = max(len(gl), 1)
    keys = np.empty(max_groups * k)
    res = np.empty((max_groups, e))
    g = load().orc_join_aggregate(len(fk), fk.ctypes.data, argp, fa.ctypes.data, e, len(dk), dk.ctypes.data,
                                  None if dp is None else dp.ctypes.data, gp, len(gl), keys.ctypes.data,
                                  res.ctypes.data, max_groups, nthreads or threads())
    _check_exact(funcs, [0 if a is None else U32 for a in fact_args], res, min(g, max_groups), check_exact)
    if many:
        return keys[: g * len(gl)].reshape(g, len(gl)).copy(), res[:g].copy()
    return (keys[:g].copy() if gl else np.zeros(g)), res[:g].copy()


JOIN_TYPES = {"inner": 0, "left": 1, "semi": 2, "anti": 3}


def hash_join(probe_key, build_key, join_type="inner", probe_rows=None, build_rows=None):
    """Row-pair equi-join: (n, 2) u64 (probe position, build row) pairs for
    inner/left, u64 probe positions for semi/anti (orc_hash_join)."""
    jt = JOIN_TYPES[join_type]
    cols, keep = _columns([probe_key, build_key])
    pr = None if probe_rows is None else np.ascontiguousarray(probe_rows, dtype=np.uint64)
    br = None if build_rows is None else np.ascontiguousarray(build_rows, dtype=np.uint64)
    npr = len(pr) if pr is not None else len(probe_key)
    nbr = len(br) if br is not None else len(build_key)
    lib = load()
    args = (C.byref(cols[0]), None if pr is None else pr.ctypes.data, npr, C.byref(cols[1]),
            None if br is None else br.ctypes.data, nbr, jt)
    n = lib.orc_hash_join(*args, None, 0)
    rec = 1 if jt >= 2 else 2
    out = np.empty(max(n, 1) * rec, dtype=np.uint64)
    lib.orc_hash_join(*args, out.ctypes.data, n)
    out = out[: n * rec]
    return out.reshape(n, 2) if rec == 2 else out


def scan_chunked_count(cols, leaves, workers=None):
    """Numeric-AND count; leaves = [(col, op 0..4, target)]."""
    arr, keep = _columns(cols)
    lc = np.array([l[0] for l in leaves], dtype=np.uint32)
    lo = np.array([l[1] for l in leaves], dtype=np.uint32)
    lt = np.array([l[2] for l in leaves], dtype=np.float64)
    n = load().orc_scan_chunked_count(arr, lc.ctypes.data, lo.ctypes.data, lt.ctypes.data, len(leaves), len(cols[0]),
                                      workers or threads())
    del keep
    return int(n)


def dense_group_u32(key, arg, key_min, rng, workers=None):
    key = np.ascontiguousarray(key, dtype=np.uint32)
    arg = None if arg is None else np.ascontiguousarray(arg, dtype=np.uint32)
    cnt = np.empty(rng, dtype=np.uint64)
    s = np.empty(rng, dtype=np.uint64)
    mn = np.empty(rng, dtype=np.uint32)
    mx = np.empty(rng, dtype=np.uint32)
    load().orc_dense_group_u32(key.ctypes.data, None if arg is None else arg.ctypes.data, len(key), key_min, rng,
                               workers or threads(), cnt.ctypes.data, s.ctypes.data, mn.ctypes.data, mx.ctypes.data)
    return cnt, s, mn, mx


def scan_int_interval(op, target, code_offset, max_code):
    lo, span = C.c_uint32(0), C.c_uint32(0)
    ok = load().orc_scan_int_interval(op, target, code_offset, max_code, C.byref(lo), C.byref(span))
    return (int(lo.value), int(span.value)) if ok else None


# closed forms, bench_select_scale.c:98-168
def ss_sum_amount(rows):
    return float(load().orc_ss_sum_amount(rows))


def ss_count_amount_above(rows, threshold):
    return int(load().orc_ss_count_amount_above(rows, threshold))


def ss_count_region(rows, region):
    return int(load().orc_ss_count_region(rows, region))


def ss_region_amount_sum(rows, region):
    return float(load().orc_ss_region_amount_sum(rows, region))


def ss_amount_rows(rows, amount):
    return int(load().orc_ss_amount_rows(rows, amount))


def ss_weight_sum(count):
    return float(load().orc_ss_weight_sum(count))


def ss_key_of_row(row):
    return int(load().orc_ss_key_of_row(row))
