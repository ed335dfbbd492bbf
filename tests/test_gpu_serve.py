"""The served loop on the GPU: requests submitted together (continuous
batching into shared scans, joins and scans beside the batches) give each
query the result the oracle gives it and the one the plain operator gives
it alone; counts, integer sums, min/max, keys and row ids bit-exact, f64
sums within 1e-12 relative.  Groups compared sorted by key.  At 1B rows
(the reference's closed-form table) every served answer is checked against
its closed form."""
import numpy as np
import pytest

import computeql_amd as cq
from computeql_amd import workloads as wl
from computeql_amd.ir import col
from computeql_amd.serve import JoinSpec, ScanSpec, Server

pytestmark = pytest.mark.gpu

ROWS = 1_000_003
RTOL = 1e-12


@pytest.fixture(scope="module")
def events():
    cols = wl.event_table(ROWS, seed=7)
    nulls = wl.event_nulls(ROWS, seed=7)
    t = wl.upload(cols, nulls)
    yield t, cols, nulls
    t.close()


@pytest.fixture(scope="module")
def small(orc):
    """4 x u32 (id, region, amount, ts) + fk into a 20000-key dim (key,
    segment uniform over 100) and a 50-row dim (region key, attr)"""
    n = 2_000_003
    t = cq.Table(n)
    for name in ("id", "region", "amount", "ts", "fk"):
        t.add_column(name, cq.U32)
    t.gen_events(orc.SEED, id=0, region=1, amount=2, ts=3)
    dim = cq.Table(20_000)
    dim.add_column("key", cq.U32)
    dim.add_column("segment", cq.U32)
    dim.gen_dim(orc.SEED ^ 0xA, 100, spread_keys=True, key=0, attr=1)
    t.gen_fact_fk(4, orc.SEED, dim, 0)
    dimb = cq.Table(50)
    dimb.add_column("region", cq.U32, np.arange(50, dtype=np.uint32))
    dimb.add_column("attr", cq.U32, (np.arange(50, dtype=np.uint32) * 7) % 10)
    cols = list(orc.gen_events(n, seed=orc.SEED))
    dk, seg = orc.gen_dim(20_000, 100, seed=orc.SEED ^ 0xA)
    cols.append(orc.gen_fact_fk(n, dk, seed=orc.SEED))
    yield t, dim, dimb, cols, (dk, seg)
    for x in (t, dim, dimb):
        x.close()


def _float_expr(spec, e):
    a = spec.args[e]
    return a is not None and wl.EVENT_COLUMNS[a][1] in (cq.F64, cq.I64) and spec.funcs[e] in (cq.SUM, cq.AVG)


def _vals(res):
    return np.stack([res.column(e).astype(np.float64) for e in range(len(res.funcs))], axis=1) if res.groups else \
        np.zeros((0, len(res.funcs)))


def _same_oracle(spec, got, rk, rv, what, float_cols=True):
    og, orf = np.argsort(got.keys, kind="stable"), np.argsort(rk, kind="stable")
    assert got.groups == len(rk), what
    assert np.array_equal(got.keys[og], rk[orf]), what
    gv = _vals(got)
    for e in range(len(spec.funcs)):
        g, r = gv[og, e], rv[orf, e]
        if float_cols and _float_expr(spec, e):
            assert np.allclose(g, r, rtol=RTOL, atol=0.0), (what, e)
        else:
            assert np.array_equal(g, r), (what, e)


def _same(spec, got, ref, what, float_cols=True):
    _same_oracle(spec, got, ref.keys, _vals(ref), what, float_cols)


def test_mix_equals_alone_and_oracle(events, orc):
    t, cols, nulls = events
    specs = wl.mixed_queries(64, seed=21)
    with Server(workers=2, max_batch=64) as srv:
        tickets = [srv.submit(t, sp) for sp in specs]
        mix = [tk.result() for tk in tickets]
        st = srv.stats()
    assert st["completed"] == 64 and st["failed"] == 0
    assert st["mean_batch"] > 1.0  # requests submitted together were batched
    with Server(workers=1, max_batch=1) as alone_srv:
        for i, sp in enumerate(specs):
            rk, rv = orc.group_aggregate(cols, sp.group, sp.funcs, sp.args, where=sp.where, nulls=nulls)
            _same_oracle(sp, mix[i], rk, rv, f"q{i} mix vs oracle")
            alone = alone_srv.submit(t, sp).result()
            _same(sp, mix[i], alone, f"q{i} mix vs alone")
            plain = t.aggregate(sp.funcs, sp.args, where=sp.where, group=sp.group)
            _same(sp, alone, plain, f"q{i} alone vs plain")


def test_two_tables_interleaved(events, small, orc):
    t, cols, nulls = events
    s, _, _, scols, _ = small
    specs_a = wl.mixed_queries(16, seed=3)
    specs_b = [cq.AggSpec([cq.COUNT, cq.SUM, cq.MAX], [None, 2, 2], col(1) < r, 1) for r in range(1, 17)]
    with Server(workers=2, max_batch=8) as srv:
        tk = []
        for a, b in zip(specs_a, specs_b):
            tk.append((t, cols, nulls, a, srv.submit(t, a)))
            tk.append((s, scols, None, b, srv.submit(s, b)))
        for table, tcols, tnulls, sp, ticket in tk:
            got = ticket.result()
            rk, rv = orc.group_aggregate(tcols, sp.group, sp.funcs, sp.args, where=sp.where, nulls=tnulls)
            _same_oracle(sp, got, rk, rv, "interleaved vs oracle", float_cols=table is t)
            ref = table.aggregate(sp.funcs, sp.args, where=sp.where, group=sp.group)
            _same(sp, got, ref, "interleaved vs plain", float_cols=table is t)


def test_joins_scans_and_aggregates_served_together(small, orc):
    """Config-3 joins (GROUP BY segment over a 20000-key dim; the filtered
    count over the 50-row dim), point lookups and a row-id scan, submitted
    with aggregates on the same fact table: each equals the oracle."""
    t, dim, dimb, cols, (dk, seg) = small
    bcols = [np.arange(50, dtype=np.uint32), (np.arange(50, dtype=np.uint32) * 7) % 10]
    reqs = []
    for i in range(6):
        reqs.append(("join_a", JoinSpec(dim, 4, 0, [cq.COUNT, cq.SUM, cq.MIN, cq.MAX], [None, 2, 2, 2],
                                        dim_where=col(1) < 20 + 10 * i, dim_group=1), 20 + 10 * i))
        reqs.append(("join_b", JoinSpec(dimb, 1, 0, [cq.COUNT], [None], dim_where=col(1) < 1 + i), 1 + i))
        reqs.append(("lookup", ScanSpec(col(0) == 1 + 333_331 * i), None))
        reqs.append(("scan", ScanSpec((col(2) < 3) & (col(1) == i), cap=1 << 16), None))
        reqs.append(("count", ScanSpec((col(1) == 7 + i) | (col(3) > 4_000_000_000 - i), count_only=True), None))
        reqs.append(("agg", cq.AggSpec([cq.COUNT, cq.SUM], [None, 2], col(2) < 500 * (i + 1), 1), None))
    with Server(workers=3, max_batch=16, hbm_budget=8 << 30) as srv:
        tickets = [(kind, sp, lit, srv.submit(t, sp)) for kind, sp, lit in reqs]
        results = [(kind, sp, lit, tk.result()) for kind, sp, lit, tk in tickets]
        st = srv.stats()
    assert st["failed"] == 0 and st["completed"] == len(reqs)
    assert st["joins"] == 12 and st["scans"] == 18
    for kind, sp, lit, res in results:
        if kind == "join_a":
            rk, rv = orc.join_aggregate(cols[4], [None, cols[2], cols[2], cols[2]], sp.funcs, dk,
                                        dim_pass=(seg < lit).astype(np.uint8), dim_group=seg)
            _same_oracle(sp, res, rk, rv, f"join A seg < {lit}", float_cols=False)
        elif kind == "join_b":
            rk, rv = orc.join_aggregate(cols[1], [None], sp.funcs, bcols[0], dim_pass=(bcols[1] < lit).astype(np.uint8))
            assert res.groups == 1 and int(res.cells[0, 0]) == int(rv[0, 0])
        elif kind == "agg":
            rk, rv = orc.group_aggregate(cols, sp.group, sp.funcs, sp.args, where=sp.where)
            _same_oracle(sp, res, rk, rv, "agg", float_cols=False)
        elif kind == "count":
            cnt, ids = res
            assert ids is None and cnt == orc.scan(cols, sp.where, count_only=True)
        else:
            cnt, ids = res
            ref = orc.scan(cols, sp.where)
            assert cnt == len(ref) and np.array_equal(ids, ref[: sp.cap])
            if kind == "lookup":
                assert cnt == 1


def test_memory_budget_admits_one_at_a_time(small, orc):
    s, _, _, cols, _ = small
    specs = [cq.AggSpec([cq.COUNT, cq.SUM], [None, 2], col(2) < 100 * (i + 1), 1) for i in range(12)]
    with Server(workers=2, max_batch=64, hbm_budget=1) as srv:
        got = [srv.submit(s, sp) for sp in specs]
        res = [g.result() for g in got]
        st = srv.stats()
    assert st["batches"] == 12 and st["completed"] == 12
    for sp, r in zip(specs, res):
        rk, rv = orc.group_aggregate(cols, sp.group, sp.funcs, sp.args, where=sp.where)
        _same_oracle(sp, r, rk, rv, "budget", float_cols=False)


def test_drive_poisson_load(small):
    s = small[0]
    specs = [cq.AggSpec([cq.COUNT, cq.SUM], [None, 2], col(2) < 100 * (i + 1), 1 if i % 2 else None)
             for i in range(32)]
    with Server(workers=2, max_batch=64) as srv:
        st = srv.drive(s, specs, rate_qps=3000, seconds=1.0, seed=5)
    assert st["failed"] == 0 and st["completed"] == st["submitted"] > 1000
    assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]


def test_drive_mix_with_joins_and_scans(events, small):
    t = events[0]
    s, dim, dimb = small[:3]
    pairs = [(t, sp) for sp in wl.mixed_queries(8, seed=5)]
    pairs += [(s, cq.AggSpec([cq.COUNT, cq.SUM], [None, 2], col(2) < 100 * (i + 1), 1)) for i in range(4)]
    pairs += [(s, JoinSpec(dim, 4, 0, [cq.SUM], [2], dim_group=1)), (s, JoinSpec(dimb, 1, 0, [cq.COUNT], [None],
                                                                                 dim_where=col(1) < 3)),
              (s, ScanSpec(col(0) == 77)), (s, ScanSpec(col(2) < 2, count_only=True))]
    with Server(workers=3, max_batch=64, hbm_budget=8 << 30) as srv:
        st = srv.drive_mix(pairs, rate_qps=2000, seconds=0.5, seed=3)
    assert st["failed"] == 0 and st["completed"] == st["submitted"] > 500
    assert st["joins"] > 0 and st["scans"] > 0
    assert 0 < st["p50_ms"] <= st["p99_ms"]


# ---------------------------------------------------------------------------
# served at 1B rows: the reference's closed-form table (bench_select_scale.c:39-54)
# ---------------------------------------------------------------------------
BIG = 1_000_000_000


@pytest.fixture(scope="module")
def scale_1b():
    t = cq.Table(BIG)
    for name in ("id", "region_id", "amount", "needle_key"):
        t.add_column(name, cq.U32)
    t.gen_scale_events(id=0, region=1, amount=2, key=3)
    regions = cq.Table(50)
    regions.add_column("region_id", cq.U32, np.arange(50, dtype=np.uint32))
    regions.add_column("zone", cq.U32, np.arange(50, dtype=np.uint32) // 10)
    yield t, regions
    regions.close()
    t.close()


def test_served_1b_closed_forms(scale_1b, orc):
    """The closed-form queries of bench_select_scale.c:1123-1169 submitted
    together to the loop: the aggregates batch, the joins and lookups run
    beside them, and every answer is its closed form."""
    t, regions = scale_1b
    reqs = [
        ("global", cq.AggSpec([cq.SUM, cq.MIN, cq.MAX], [2, 2, 2])),
        ("above", ScanSpec(col(2) > 9990, count_only=True)),
        ("region7", ScanSpec(col(1) == 7, count_only=True)),
        ("by_region", cq.AggSpec([cq.COUNT, cq.SUM], [None, 2], None, 1)),
        ("by_amount", cq.AggSpec([cq.COUNT, cq.SUM], [None, 1], None, 2, cap=10_000)),
        ("zone2", JoinSpec(regions, 1, 0, [cq.COUNT, cq.SUM], [None, 2], dim_where=col(1) == 2)),
    ]
    rows = (0, 123_456_789, BIG - 1)
    reqs += [(("lookup", r), ScanSpec(col(3) == orc.ss_key_of_row(r))) for r in rows]
    with Server(workers=3, max_batch=16, hbm_budget=64 << 30) as srv:
        got = {k: srv.submit(t, sp).result(timeout=120) for k, sp in reqs}
        st = srv.stats()
    assert st["failed"] == 0 and st["joins"] == 1 and st["scans"] == 5
    g = got["global"]
    assert (g.column(0)[0], g.column(1)[0], g.column(2)[0]) == (orc.ss_sum_amount(BIG), 0.0, 9999.0)
    assert got["above"][0] == orc.ss_count_amount_above(BIG, 9990)
    assert got["region7"][0] == orc.ss_count_region(BIG, 7)
    g = got["by_region"]
    o = np.argsort(g.keys)
    assert np.array_equal(g.keys[o], np.arange(50))
    for r in range(50):
        assert g.column(0)[o][r] == orc.ss_count_region(BIG, r)
        assert g.column(1)[o][r] == orc.ss_region_amount_sum(BIG, r)
    g = got["by_amount"]
    o = np.argsort(g.keys)
    a = np.arange(10_000, dtype=np.float64)
    n = np.array([orc.ss_amount_rows(BIG, int(x)) for x in a], dtype=np.float64)
    assert np.array_equal(g.keys[o], a)
    assert np.array_equal(g.column(0)[o], n.astype(np.uint64)) and np.array_equal(g.column(1)[o], n * (a % 50))
    j = got["zone2"]
    assert int(j.cells[0, 0]) == sum(orc.ss_count_region(BIG, r) for r in range(20, 30))
    assert j.column(1)[0] == sum(orc.ss_region_amount_sum(BIG, r) for r in range(20, 30))
    for r in rows:
        cnt, ids = got[("lookup", r)]
        assert cnt == 1 and int(ids[0]) == r
