"""Any AND / OR WHERE of at most 8 leaves (comparisons, !=, IN / NOT IN
lists, IS [NOT] NULL) and 16 conjunctions runs fused in the aggregate
kernels, the library scan and the shared-scan batches: the leaf kernels
evaluate each leaf (a NULL row failing it, != and NOT IN as complements,
a small integer IN list as one 64-value set) and match the rows against
the WHERE's conjunctions.  Every result is checked against the oracle's
row-by-row evaluation (qe_row_condition_eval semantics: a leaf over a
NULL is false, AND / OR two-valued), over columns with NULLs, I32 and
F64 columns, and narrow-encoded columns.  Counts, integer sums, min/max,
keys and row ids bit-exact; f64 sums within 1e-12 relative."""
import numpy as np
import pytest

import computeql_amd as cq
from computeql_amd.ir import col

pytestmark = pytest.mark.gpu

ROWS = 3_000_017
RTOL = 1e-12


@pytest.fixture(autouse=True)
def _settings():
    cq.reset_settings()
    yield
    cq.reset_settings()


@pytest.fixture(scope="module")
def mixed():
    """0 k u32 [0,50)  1 s i32 [-40,40)  2 v u32 [0,10000)  3 f f64
    4 n u32 [0,1000) with 5% NULLs  5 m i32 [-500,500) with 3% NULLs
    6 e u32 [0,50) (narrow-encoded)  7 x u32 uniform"""
    r = np.random.default_rng(77)
    cols = [r.integers(0, 50, ROWS).astype(np.uint32), r.integers(-40, 40, ROWS).astype(np.int32),
            r.integers(0, 10_000, ROWS).astype(np.uint32), r.lognormal(2.0, 1.0, ROWS),
            r.integers(0, 1000, ROWS).astype(np.uint32), r.integers(-500, 500, ROWS).astype(np.int32),
            r.integers(0, 50, ROWS).astype(np.uint32), r.integers(0, 1 << 32, ROWS, dtype=np.uint64).astype(np.uint32)]
    nulls = {4: (r.random(ROWS) < 0.05).astype(np.uint8), 5: (r.random(ROWS) < 0.03).astype(np.uint8)}
    types = [cq.U32, cq.I32, cq.U32, cq.F64, cq.U32, cq.I32, cq.U32, cq.U32]
    t = cq.Table(ROWS)
    for i, (ty, c) in enumerate(zip(types, cols)):
        t.add_column(f"c{i}", ty, c, nulls=nulls.get(i))
    assert t.encode(6) == 1
    yield t, cols, nulls
    t.close()


WHERES = {
    "or_of_ands": ((col(0) == 7) & (col(2) < 5000)) | (col(1) > 30),
    "in_small": col(0).isin([3, 5, 8, 13, 21, 34]),
    "in_i32_negative": col(1).isin([-40, -7, 0, 7, 39, 100, -1.5]),
    "in_wide_int": col(2).isin([1, 5000, 9999]),
    "in_f64": col(3).isin([1.0, 2.5]) | (col(3) < 1.5),
    "not_in": col(0).isin([1, 2, 3], negate=True) & (col(2) < 300),
    "ne": (col(0) != 4) & ((col(1) != -3) | (col(2) >= 9000)),
    "nested": ((col(0) < 10) | (col(2) > 9000)) & ((col(1) < 0) | (col(0) > 40)),
    "null_or": (col(4) < 100) | (col(5) > 400),
    "null_in": col(4).isin([1, 2, 3, 500]) | col(5).is_null(),
    "null_not_in": col(5).isin([-1, 0, 1], negate=True) & (col(0) < 25),
    "null_ne": (col(4) != 10) | (col(0) == 3),
    "is_not_null_or": col(4).is_not_null() & ((col(5) < -450) | (col(0) == 0)),
    "encoded": col(6).isin([0, 10, 20, 30, 40, 49]) | ((col(6) != 7) & (col(0) == 1)),
    "empty_in": col(0).isin([]) | (col(2) < 10),
    "empty_not_in": col(4).isin([], negate=True) & (col(0) < 3),
    "never": col(0).isin([]),
    "dup_leaves": ((col(0) < 5) & (col(2) < 100)) | ((col(0) < 5) & (col(1) > 35)) | (col(0) < 5),
}


def _vals(res):
    return np.stack([res.column(e).astype(np.float64) for e in range(len(res.funcs))], axis=1) if res.groups else \
        np.zeros((0, len(res.funcs)))


def _same(res, rk, rv, funcs, args, float_args, what):
    keys = res.keys if res.keys.ndim == 1 else res.keys
    if keys.ndim == 1:
        og, orf = np.argsort(keys, kind="stable"), np.argsort(rk, kind="stable")
    else:
        og, orf = np.lexsort(keys.T[::-1]), np.lexsort(rk.T[::-1])
    assert res.groups == len(rk), what
    assert np.array_equal(keys[og], rk[orf]), what
    gv = _vals(res)
    for e in range(len(funcs)):
        g, r = gv[og, e], rv[orf, e]
        if args[e] in float_args and funcs[e] in (cq.SUM, cq.AVG):
            assert np.allclose(g, r, rtol=RTOL, atol=0.0), (what, e)
        else:
            assert np.array_equal(g, r), (what, e)


@pytest.mark.parametrize("name", sorted(WHERES))
def test_global_fused(mixed, orc, name):
    t, cols, nulls = mixed
    w = WHERES[name]
    funcs, args = [cq.COUNT, cq.SUM, cq.MIN, cq.MAX, cq.SUM], [None, 2, 1, 1, 3]
    res = t.aggregate(funcs, args, where=w)
    assert cq.AGG_PATHS[res.trace["path"]] == "fused_global", res.trace
    rk, rv = orc.group_aggregate(cols, None, funcs, args, where=w, nulls=nulls)
    _same(res, rk, rv, funcs, args, {3}, name)


@pytest.mark.parametrize("name", ["or_of_ands", "in_i32_negative", "null_or", "null_not_in", "encoded", "never"])
def test_dense_and_narrow_fused(mixed, orc, name):
    t, cols, nulls = mixed
    w = WHERES[name]
    # dense by k (50 groups), narrow by v (10000 groups, one u32 argument)
    for group, funcs, args in ((0, [cq.COUNT, cq.SUM, cq.MAX], [None, 2, 1]),
                               (2, [cq.COUNT, cq.SUM, cq.MIN, cq.MAX], [None, 7, 7, 7]),
                               ([0, 6], [cq.COUNT, cq.SUM], [None, 2])):
        res = t.aggregate(funcs, args, where=w, group=group)
        assert cq.AGG_PATHS[res.trace["path"]] in ("fused_dense", "dense"), res.trace
        rk, rv = orc.group_aggregate(cols, group, funcs, args, where=w, nulls=nulls)
        _same(res, rk, rv, funcs, args, set(), f"{name} by {group}")


@pytest.mark.parametrize("name", ["nested", "null_in", "in_f64", "ne"])
def test_hash_fused(mixed, orc, name):
    t, cols, nulls = mixed
    w = WHERES[name]
    funcs, args = [cq.COUNT, cq.SUM, cq.MIN], [None, 3, 2]
    res = t.aggregate(funcs, args, where=w, group=7, cap=ROWS)  # ~ one group a row
    assert cq.AGG_PATHS[res.trace["path"]] == "hash" and res.trace["rows_scanned"] == ROWS
    rk, rv = orc.group_aggregate(cols, 7, funcs, args, where=w, nulls=nulls)
    _same(res, rk, rv, funcs, args, {3}, name)


@pytest.mark.parametrize("name", ["or_of_ands", "null_or", "null_in", "null_not_in", "encoded", "empty_not_in"])
def test_scan_ids(mixed, orc, name):
    t, cols, nulls = mixed
    w = WHERES[name]
    n, ids, tr = t.scan(w)
    ref = orc.scan(cols, w, nulls=nulls)
    assert n == len(ref) and np.array_equal(ids, ref), name
    assert t.count(w) == len(ref)


@pytest.mark.parametrize("compact", [0, 1])
def test_batch_fuses_every_form(mixed, orc, compact):
    t, cols, nulls = mixed
    cq.set_setting("QE_BATCH_COMPACT", compact)
    specs = []
    for i, name in enumerate(sorted(WHERES)):
        specs.append(cq.AggSpec([cq.COUNT, cq.SUM, cq.MIN], [None, 2, 1], WHERES[name], 0 if i % 2 else None))
        specs.append(cq.AggSpec([cq.COUNT, cq.SUM, cq.MAX], [None, 3, 3], WHERES[name], 1 if i % 3 else None))
    results, tr = cq.aggregate_batch(t, specs, with_trace=True)
    assert tr["fused"] == len(specs) and tr["fallback"] == 0, tr
    for sp, res in zip(specs, results):
        rk, rv = orc.group_aggregate(cols, sp.group, sp.funcs, sp.args, where=sp.where, nulls=nulls)
        _same(res, rk, rv, sp.funcs, sp.args, {3}, f"batch {sp.where}")


def test_join_dim_where_or_over_nulls(orc):
    """the dimension's WHERE: an OR over a nullable attribute"""
    r = np.random.default_rng(3)
    n, dn = 2_000_003, 5000
    dk = r.permutation(1 << 20)[:dn].astype(np.uint32)
    attr = r.integers(0, 100, dn).astype(np.uint32)
    an = (r.random(dn) < 0.1).astype(np.uint8)
    seg = r.integers(0, 20, dn).astype(np.uint32)
    fk = dk[r.integers(0, dn, n)]
    amt = r.integers(0, 10_000, n).astype(np.uint32)
    f, d = cq.Table(n), cq.Table(dn)
    try:
        f.add_column("fk", cq.U32, fk)
        f.add_column("amount", cq.U32, amt)
        d.add_column("key", cq.U32, dk)
        d.add_column("attr", cq.U32, attr, nulls=an)
        d.add_column("seg", cq.U32, seg)
        w = (col(1) < 10) | col(1).isin([50, 51, 52]) | (col(2) == 7)
        res = cq.aggregate_join(f, 0, d, 0, [cq.COUNT, cq.SUM], [None, 1], dim_where=w, dim_group=2)
        pas = np.zeros(dn, dtype=np.uint8)
        pas[orc.scan([dk, attr, seg], w, nulls={1: an}).astype(np.int64)] = 1
        rk, rv = orc.join_aggregate(fk, [None, amt], [cq.COUNT, cq.SUM], dk, dim_pass=pas, dim_group=seg)
        o = np.argsort(rk)
        assert np.array_equal(res.keys, rk[o])
        assert np.array_equal(res.column(0), rv[o, 0].astype(np.uint64)) and np.array_equal(res.column(1), rv[o, 1])
    finally:
        f.close()
        d.close()


def test_limits(mixed, orc):
    """nine leaves: over columns without NULLs the WHERE takes the scan and
    its device row ids; over a nullable column it is refused"""
    t, cols, nulls = mixed
    w = col(0) == 0
    for v in range(1, 9):
        w = w | (col(2) == v * 1000 + 1)
    funcs, args = [cq.COUNT, cq.SUM], [None, 2]
    res = t.aggregate(funcs, args, where=w)
    assert cq.AGG_PATHS[res.trace["path"]] == "hash"
    rk, rv = orc.group_aggregate(cols, None, funcs, args, where=w)
    _same(res, rk, rv, funcs, args, set(), "nine leaves")
    wn = w | (col(4) == 5)
    with pytest.raises(cq.QueryError):
        t.aggregate(funcs, args, where=wn)
