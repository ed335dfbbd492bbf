"""The served loop on the GPU: requests submitted together (continuous
batching into shared scans) give each query the result the plain operator
gives it alone; counts, integer sums, min/max and keys bit-exact, f64 sums
within 1e-12 relative.  Groups compared sorted by key."""
import numpy as np
import pytest

import computeql_amd as cq
from computeql_amd import workloads as wl
from computeql_amd.ir import col
from computeql_amd.serve import Server

pytestmark = pytest.mark.gpu

ROWS = 1_000_003
RTOL = 1e-12


@pytest.fixture(scope="module")
def events():
    cols = wl.event_table(ROWS, seed=7)
    t = wl.upload(cols)
    yield t, cols
    t.close()


@pytest.fixture(scope="module")
def small():
    t = cq.Table(200_003)
    for name in ("id", "region", "amount", "ts"):
        t.add_column(name, cq.U32)
    t.gen_events(0x1234567890ABCDEF, id=0, region=1, amount=2, ts=3)
    yield t
    t.close()


def _float_expr(spec, e):
    a = spec.args[e]
    return a is not None and wl.EVENT_COLUMNS[a][1] in (cq.F64, cq.I64) and spec.funcs[e] in (cq.SUM, cq.AVG)


def _same(spec, got, ref, what, float_cols=True):
    og, orf = np.argsort(got.keys, kind="stable"), np.argsort(ref.keys, kind="stable")
    assert got.groups == ref.groups, what
    assert np.array_equal(got.keys[og], ref.keys[orf]), what
    for e in range(len(spec.funcs)):
        g, r = got.column(e)[og], ref.column(e)[orf]
        if float_cols and _float_expr(spec, e):
            assert np.allclose(g, r, rtol=RTOL, atol=0.0), (what, e)
        else:
            assert np.array_equal(g, r), (what, e)


def test_mix_equals_alone(events):
    t, _ = events
    specs = wl.mixed_queries(64, seed=21)
    with Server(workers=2, max_batch=64) as srv:
        tickets = [srv.submit(t, sp) for sp in specs]
        mix = [tk.result() for tk in tickets]
        st = srv.stats()
    assert st["completed"] == 64 and st["failed"] == 0
    assert st["mean_batch"] > 1.0  # requests submitted together were batched
    with Server(workers=1, max_batch=1) as alone_srv:
        for i, sp in enumerate(specs):
            alone = alone_srv.submit(t, sp).result()
            plain = t.aggregate(sp.funcs, sp.args, where=sp.where, group=sp.group)
            _same(sp, mix[i], alone, f"q{i} mix vs alone")
            _same(sp, alone, plain, f"q{i} alone vs plain")


def test_two_tables_interleaved(events, small):
    t, _ = events
    specs_a = wl.mixed_queries(16, seed=3)
    specs_b = [cq.AggSpec([cq.COUNT, cq.SUM, cq.MAX], [None, 2, 2], col(1) < r, 1) for r in range(1, 17)]
    with Server(workers=2, max_batch=8) as srv:
        tk = []
        for a, b in zip(specs_a, specs_b):
            tk.append((t, a, srv.submit(t, a)))
            tk.append((small, b, srv.submit(small, b)))
        for table, sp, ticket in tk:
            got = ticket.result()
            ref = table.aggregate(sp.funcs, sp.args, where=sp.where, group=sp.group)
            _same(sp, got, ref, "interleaved", float_cols=table is t)


def test_memory_budget_admits_one_at_a_time(small):
    specs = [cq.AggSpec([cq.COUNT, cq.SUM], [None, 2], col(2) < 100 * (i + 1), 1) for i in range(12)]
    with Server(workers=2, max_batch=64, hbm_budget=1) as srv:
        got = [srv.submit(small, sp) for sp in specs]
        res = [g.result() for g in got]
        st = srv.stats()
    assert st["batches"] == 12 and st["completed"] == 12
    for sp, r in zip(specs, res):
        ref = small.aggregate(sp.funcs, sp.args, where=sp.where, group=sp.group)
        _same(sp, r, ref, "budget", float_cols=False)


def test_drive_poisson_load(small):
    specs = [cq.AggSpec([cq.COUNT, cq.SUM], [None, 2], col(2) < 100 * (i + 1), 1 if i % 2 else None)
             for i in range(32)]
    with Server(workers=2, max_batch=64) as srv:
        st = srv.drive(small, specs, rate_qps=3000, seconds=1.0, seed=5)
    assert st["failed"] == 0 and st["completed"] == st["submitted"] > 1000
    assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]


def test_drive_mix_over_two_tables(events, small):
    t, _ = events
    pairs = [(t, sp) for sp in wl.mixed_queries(8, seed=5)]
    pairs += [(small, cq.AggSpec([cq.COUNT, cq.SUM], [None, 2], col(2) < 100 * (i + 1), 1)) for i in range(4)]
    with Server(workers=2, max_batch=64) as srv:
        st = srv.drive_mix(pairs, rate_qps=2000, seconds=0.5, seed=3)
    assert st["failed"] == 0 and st["completed"] == st["submitted"] > 500
    assert 0 < st["p50_ms"] <= st["p99_ms"]
