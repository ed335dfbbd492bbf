"""Dense GROUP BY over arguments wider than 32-bit integers (f32, f64,
i64, u64, mixed with i32): the wide LDS tables (aggregate_dense_wide:
f64 sums for the wider types, exact integer sums for 32-bit integers,
min/max as order-preserving keys of the f64 value).  Against the oracle:
counts, integer sums, min/max and keys bit-exact, f64 sums within 1e-12
relative (atomic summation order); and against the hash path
(QE_AGG_DENSE_WIDE 0) on the same query."""
import numpy as np
import pytest

import computeql_amd as cq
from computeql_amd import workloads as wl
from computeql_amd.ir import col

pytestmark = pytest.mark.gpu

ROWS = 2_000_003
RTOL = 1e-12


@pytest.fixture(autouse=True)
def _settings():
    cq.reset_settings()
    yield
    cq.reset_settings()


@pytest.fixture(scope="module")
def wide():
    """0 k u32 [0,40)  1 j i32 [-3,3)  2 f f64 lognormal (negatives too)
    3 g f32  4 n i64 [-2^40, 2^40)  5 u u64 [0, 2^40)  6 s i32"""
    r = np.random.default_rng(17)
    f = r.lognormal(3.0, 1.0, ROWS) * np.where(r.random(ROWS) < 0.2, -1.0, 1.0)
    cols = [r.integers(0, 40, ROWS).astype(np.uint32), r.integers(-3, 3, ROWS).astype(np.int32), f,
            r.standard_normal(ROWS).astype(np.float32),
            r.integers(-(1 << 40), 1 << 40, ROWS, dtype=np.int64), r.integers(0, 1 << 40, ROWS, dtype=np.uint64),
            r.integers(-(1 << 31), 1 << 31, ROWS).astype(np.int32)]
    types = [cq.U32, cq.I32, cq.F64, cq.F32, cq.I64, cq.U64, cq.I32]
    t = cq.Table(ROWS)
    for i, (ty, c) in enumerate(zip(types, cols)):
        t.add_column(f"c{i}", ty, c)
    yield t, cols
    t.close()


def _vals(res):
    return np.stack([res.column(e).astype(np.float64) for e in range(len(res.funcs))], axis=1) if res.groups else \
        np.zeros((0, len(res.funcs)))


def _check(res, rk, rv, funcs, args, float_args, what):
    if res.keys.ndim == 1:
        og, orf = np.argsort(res.keys, kind="stable"), np.argsort(rk, kind="stable")
    else:
        og, orf = np.lexsort(res.keys.T[::-1]), np.lexsort(rk.T[::-1])
    assert res.groups == len(rk), what
    assert np.array_equal(res.keys[og], rk[orf]), what
    gv = _vals(res)
    for e in range(len(funcs)):
        g, r = gv[og, e], rv[orf, e]
        if args[e] in float_args and funcs[e] in (cq.SUM, cq.AVG):
            assert np.allclose(g, r, rtol=RTOL, atol=0.0), (what, e)
        else:
            assert np.array_equal(g, r), (what, e)


CASES = [
    (0, [cq.COUNT, cq.SUM, cq.MIN, cq.MAX, cq.AVG], [None, 2, 2, 2, 2], None),
    (1, [cq.SUM, cq.MIN, cq.MAX], [3, 3, 3], col(0) < 20),
    ([0, 1], [cq.COUNT, cq.SUM, cq.MAX, cq.SUM], [None, 4, 4, 6], None),
    (0, [cq.SUM, cq.MIN, cq.MAX, cq.AVG], [5, 5, 5, 5], (col(1) == 2) | (col(2) > 50.0)),
    (1, [cq.SUM, cq.SUM, cq.MIN, cq.MAX], [2, 6, 6, 3], col(6) > 0),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_wide_dense_against_oracle_and_hash(wide, orc, case):
    t, cols = wide
    group, funcs, args, where = CASES[case]
    float_args = {2, 3, 4, 5}  # summed as f64
    res = t.aggregate(funcs, args, where=where, group=group)
    assert cq.AGG_PATHS[res.trace["path"]] in ("dense", "fused_dense") and res.trace["kernel"] == "aggregate_dense_wide"
    rk, rv = orc.group_aggregate(cols, group, funcs, args, where=where)
    _check(res, rk, rv, funcs, args, float_args, f"case {case} vs oracle")
    cq.set_setting("QE_AGG_DENSE_WIDE", 0)
    h = t.aggregate(funcs, args, where=where, group=group, cap=4096)
    assert cq.AGG_PATHS[h.trace["path"]] == "hash"
    _check(res, h.keys, _vals(h), funcs, args, float_args, f"case {case} vs hash")


def test_event_table_mix_queries_run_dense(orc):
    """the batched mix's f64-by-status and i64-by-region shapes, plain"""
    n = 1_000_003
    cols = wl.event_table(n, seed=7)
    t = wl.upload(cols)
    try:
        for group, funcs, args in ((wl.STATUS, [cq.COUNT, cq.SUM, cq.AVG], [None, wl.M0 + 3, wl.M0 + 3]),
                                   (wl.REGION, [cq.SUM, cq.MAX], [wl.T0 + 6, wl.T0 + 6])):
            res = t.aggregate(funcs, args, group=group)
            assert res.trace["kernel"] == "aggregate_dense_wide"
            rk, rv = orc.group_aggregate(cols, group, funcs, args)
            _check(res, rk, rv, funcs, args, set(args) - {None}, f"event by {group}")
    finally:
        t.close()
