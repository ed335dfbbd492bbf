"""qe_aggregate_noalloc (the entry point that allocates nothing) against
qe_aggregate, through the C ABI: the same cells bit for bit, on two calls
in a row (the second over a smaller table, so the shared scratch holds the
first call's partials when it starts), no device allocation made, and the
shapes beyond its scratch refused with an error."""
import ctypes as C

import numpy as np
import pytest

import computeql_amd as cq
from computeql_amd import ir
from computeql_amd.ir import col
from computeql_amd.native import GPU_MemoryStats, QE_Trace

pytestmark = pytest.mark.gpu


def _noalloc(t, funcs, args, where=None):
    lib = cq.load()
    e = len(funcs)
    fa = (C.c_uint32 * e)(*funcs)
    aa = (C.c_int32 * e)(*[-1 if a is None else a for a in args])
    nodes = ir.to_nodes(where)
    cells = np.zeros(e, dtype=np.uint64)
    tr = QE_Trace()
    ok = lib.qe_aggregate_noalloc(t.handle, nodes.ctypes.data if nodes is not None else None,
                                  len(nodes) if nodes is not None else 0, fa, aa, e, cells.ctypes.data, C.byref(tr))
    if not ok:
        raise cq.QueryError(lib.qe_last_error().decode())
    return cells


def _alloc_calls():
    st = cq.load().gpu_hip_memory_stats()
    return int(st.alloc_calls)


QUERIES = [
    ([cq.SUM, cq.MIN, cq.MAX], [2, 2, 2], None),
    ([cq.COUNT], [None], col(2) < 200),
    ([cq.COUNT, cq.SUM, cq.AVG, cq.MAX], [None, 2, 3, 3], (col(1) <= 25) & (col(2) >= 100)),
    ([cq.SUM, cq.MIN, cq.MAX, cq.SUM, cq.AVG], [4, 4, 5, 5, 6], col(3) > 2**31),
    ([cq.COUNT, cq.SUM], [None, 3], col(1) == 7.5),  # a literal no integer equals
    ([cq.COUNT, cq.SUM, cq.MIN], [None, 2, 4], ((col(1) == 3) & (col(2) < 5000)) | col(1).isin([7, 8, 40])),
]


def _table(orc, n):
    t = cq.Table(n)
    for name in ("id", "region", "amount", "ts"):
        t.add_column(name, cq.U32)
    t.gen_events(orc.SEED, id=0, region=1, amount=2, ts=3)
    r = np.random.default_rng(n)
    t.add_column("i", cq.I32, r.integers(-(2**31), 2**31, n).astype(np.int32))
    t.add_column("f", cq.F64, r.lognormal(3, 1, n))
    t.add_column("g", cq.F32, r.random(n).astype(np.float32))
    return t


def test_noalloc_equals_aggregate_on_two_calls(orc):
    big, small = _table(orc, 3_000_017), _table(orc, 10_007)
    try:
        for funcs, args, where in QUERIES:
            want = [t.aggregate(funcs, args, where=where).cells[0] for t in (big, small)]
            before = _alloc_calls()
            got_big = _noalloc(big, funcs, args, where)
            got_small = _noalloc(small, funcs, args, where)
            assert _alloc_calls() == before, "the no-allocation entry point allocated device memory"
            assert np.array_equal(got_big, want[0]), (funcs, where)
            assert np.array_equal(got_small, want[1]), (funcs, where)
        # empty table: one row of zeros, as qe_aggregate
        e = cq.Table(0)
        e.add_column("a", cq.U32)
        try:
            assert np.array_equal(_noalloc(e, [cq.COUNT, cq.SUM], [None, 0]),
                                  e.aggregate([cq.COUNT, cq.SUM], [None, 0]).cells[0])
        finally:
            e.close()
    finally:
        big.close()
        small.close()


def test_noalloc_refuses_what_its_scratch_does_not_hold(orc):
    t = _table(orc, 100_003)
    try:
        nine = col(2) < 9000
        for i in range(8):
            nine = nine & (col(2) != i)
        with pytest.raises(cq.QueryError, match="at most 8"):
            _noalloc(t, [cq.COUNT], [None], nine)
        # eight leaves in 18 conjunctions
        wide = (((col(1) == 1) | (col(1) == 2) | (col(1) == 30)) & ((col(2) < 5) | (col(2) > 9000) | (col(3) < 7))
                & ((col(0) < 100) | (col(0) > 200)))
        with pytest.raises(cq.QueryError, match="conjunctions"):
            _noalloc(t, [cq.COUNT], [None], wide)
        with pytest.raises(cq.QueryError, match="4 distinct"):
            _noalloc(t, [cq.SUM] * 5, [0, 1, 2, 3, 4])
        # the refusals left the scratch usable
        assert np.array_equal(_noalloc(t, [cq.COUNT], [None], col(1) < 10),
                              t.aggregate([cq.COUNT], [None], where=col(1) < 10).cells[0])
    finally:
        t.close()
