"""GPU parity of the batched shared scan: every query of a 64-query mixed
batch over the 30-column event table equals the plain operator's result
for it alone and the CPU oracle's.  Counts, integer sums, min/max and
keys bit-exact; f64 sums and averages within 1e-12 relative (a batch adds a
query's values per thread, wave and block, the plain kernels in their own
order).  Groups are compared sorted by key."""
import functools

import numpy as np
import pytest

import computeql_amd as cq
from computeql_amd import workloads as wl
from computeql_amd.ir import col

pytestmark = pytest.mark.gpu

ROWS = 1_000_003
RTOL = 1e-12


@pytest.fixture(autouse=True)
def _settings():
    cq.reset_settings()
    yield
    cq.reset_settings()


@pytest.fixture(scope="module")
def events():
    cols = wl.event_table(ROWS, seed=7)
    t = wl.upload(cols)
    yield t, cols
    t.close()


FLOAT_ARGS = {cq.F64, cq.I64}


def _is_float(spec, e):
    a = spec.args[e]
    return a is not None and wl.EVENT_COLUMNS[a][1] in FLOAT_ARGS and spec.funcs[e] in (cq.SUM, cq.AVG)


def _same(spec, got_keys, got_vals, ref_keys, ref_vals, what):
    og, orf = np.argsort(got_keys, kind="stable"), np.argsort(ref_keys, kind="stable")
    assert len(got_keys) == len(ref_keys), what
    assert np.array_equal(got_keys[og], ref_keys[orf]), what
    for e in range(len(spec.funcs)):
        g, r = got_vals[og, e], ref_vals[orf, e]
        if _is_float(spec, e):
            assert np.allclose(g, r, rtol=RTOL, atol=0.0), (what, e)
        else:
            assert np.array_equal(g, r), (what, e)


def _vals(res):
    return np.stack([res.column(e).astype(np.float64) for e in range(len(res.funcs))], axis=1) if res.groups else \
        np.zeros((0, len(res.funcs)))


def test_mixed_batch_matches_plain_and_oracle(events, orc):
    t, cols = events
    specs = wl.mixed_queries(64, seed=11)
    results, tr = cq.aggregate_batch(t, specs, with_trace=True)
    assert tr["fused"] == 64 and tr["fallback"] == 0 and tr["launches"] >= 1
    for i, (sp, res) in enumerate(zip(specs, results)):
        assert res.trace["fused"] == 1
        plain = t.aggregate(sp.funcs, sp.args, where=sp.where, group=sp.group)
        _same(sp, res.keys, _vals(res), plain.keys, _vals(plain), f"plain q{i}")
        if i % 4 == 0:
            rk, rv = orc.group_aggregate(cols, sp.group, sp.funcs, sp.args, where=sp.where)
            _same(sp, res.keys, _vals(res), rk, rv, f"oracle q{i}")


@pytest.mark.parametrize("items,bpc,lds,wg,dbl", [(1, 1, 81920, 0, 2), (8, 4, 81920, 0, 2), (2, 2, 24576, 0, 2),
                                                  (4, 2, 0, 256, 1), (2, 1, 0, 512, 0), (1, 2, 0, 1024, 0)])
def test_batch_configs_agree(events, items, bpc, lds, wg, dbl):
    t, _ = events
    specs = wl.mixed_queries(24, seed=5)
    base = cq.aggregate_batch(t, specs)
    cq.set_setting("QE_BATCH_ITEMS", items)
    cq.set_setting("QE_BATCH_BLOCKS_PER_CU", bpc)
    cq.set_setting("QE_BATCH_DOUBLE_BUFFER", dbl)
    if lds:
        cq.set_setting("QE_BATCH_LDS_BYTES", lds)
    if wg:
        cq.set_setting("QE_BATCH_WG", wg)
    got, tr = cq.aggregate_batch(t, specs, with_trace=True)
    if lds == 24576:
        assert tr["launches"] > 1
    for i, (sp, a, b) in enumerate(zip(specs, base, got)):
        _same(sp, b.keys, _vals(b), a.keys, _vals(a), f"q{i}")


@pytest.mark.parametrize("regs", [0, 1])
def test_register_and_lds_staging_agree(events, orc, regs):
    """the 1024-thread shared scan fills its stage from registers (the next
    tile's loads in flight while a tile is evaluated, QE_BATCH_STAGE_REGS 1)
    or by direct copies into LDS (0); both match the oracle"""
    t, cols = events
    specs = wl.mixed_queries(32, seed=21)
    cq.set_setting("QE_BATCH_WG", 1024)
    cq.set_setting("QE_BATCH_COST", 0)
    cq.set_setting("QE_BATCH_STAGE_REGS", regs)
    got, tr = cq.aggregate_batch(t, specs, with_trace=True)
    assert tr["fused"] == 32
    for i, (sp, res) in enumerate(zip(specs, got)):
        rk, rv = orc.group_aggregate(cols, sp.group, sp.funcs, sp.args, where=sp.where)
        _same(sp, res.keys, _vals(res), rk, rv, f"oracle q{i}")


def test_alone_equals_in_the_mix(events):
    t, _ = events
    specs = wl.mixed_queries(16, seed=9)
    mix = cq.aggregate_batch(t, specs)
    for i, sp in enumerate(specs):
        alone = cq.aggregate_batch(t, [sp])[0]
        _same(sp, alone.keys, _vals(alone), mix[i].keys, _vals(mix[i]), f"q{i}")


def test_unfusable_queries_fall_back(events, orc):
    t, cols = events
    specs = [
        cq.AggSpec([cq.COUNT, cq.SUM], [None, wl.M0], col(wl.STATUS) == 2, wl.USER),  # 1M-range key: hash path
        # nine leaves: more than the leaf kernels take
        cq.AggSpec([cq.COUNT], [None], functools.reduce(lambda a, b: a | b, [col(wl.REGION) == r for r in range(9)])),
        cq.AggSpec([cq.COUNT, cq.SUM], [None, wl.U0], col(wl.REGION) < 10, wl.STATUS),
    ]
    res, tr = cq.aggregate_batch(t, specs, with_trace=True)
    assert tr["fallback"] == 2 and tr["fused"] == 1
    assert [r.trace["fused"] for r in res] == [0, 0, 1]
    for i, (sp, r) in enumerate(zip(specs, res)):
        rk, rv = orc.group_aggregate(cols, sp.group, sp.funcs, sp.args, where=sp.where)
        _same(sp, r.keys, _vals(r), rk, rv, f"q{i}")


def test_empty_table_batch():
    t = cq.Table(0)
    t.add_column("a", cq.U32)
    t.add_column("b", cq.F64)
    specs = [cq.AggSpec([cq.COUNT, cq.SUM, cq.MIN], [None, 1, 1]), cq.AggSpec([cq.COUNT], [None], col(0) < 5)]
    res = cq.aggregate_batch(t, specs)
    assert res[0].groups == 1 and int(res[0].cells[0, 0]) == 0 and res[0].column(1)[0] == 0.0
    assert res[1].groups == 1 and int(res[1].cells[0, 0]) == 0
    t.close()


def test_autotuner_records_and_applies(events, tmp_path, monkeypatch):
    t, _ = events
    monkeypatch.setenv("CQLHIP_TUNE_FILE", str(tmp_path / "tune.txt"))
    specs = wl.mixed_queries(16, seed=13)
    best = cq.autotune_batch(t, specs)
    assert best["wg"] in (256, 512, 1024) and best["ms"] > 0
    lines = (tmp_path / "tune.txt").read_text().split("\n")
    assert any(line.split()[3:5] == [str(best["wg"]), str(best["items"])] for line in lines if line)
    base = cq.aggregate_batch(t, specs)
    cq.set_setting("QE_BATCH_AUTOTUNE", 1)
    got, tr = cq.aggregate_batch(t, specs, with_trace=True)
    assert tr["tuned_wg"] == best["wg"] and tr["tuned_buffers"] == best["buffers"]
    for i, (sp, a, b) in enumerate(zip(specs, base, got)):
        _same(sp, b.keys, _vals(b), a.keys, _vals(a), f"q{i}")


def _graph_specs(r):
    """A batch whose shape is fixed and whose literals change with r."""
    return [
        cq.AggSpec([cq.COUNT, cq.SUM], [None, wl.U0], (col(wl.REGION) < 5 + r) & (col(wl.STATUS) == r % 5), wl.STATUS),
        cq.AggSpec([cq.COUNT, cq.MIN, cq.MAX], [None, wl.M0, wl.M0], col(wl.U0 + 1) < 1000 * r + 7),
        cq.AggSpec([cq.SUM, cq.AVG], [wl.M0 + 1, wl.U0 + 2], col(wl.REGION) == 7 + r, wl.REGION),
        cq.AggSpec([cq.COUNT], [None], (col(wl.STATUS) >= r % 3) & (col(wl.U0 + 3) > 123 * r)),
    ]


def test_graph_replays_match_eager(events):
    """The batch runs as a HIP graph: the first round captures it, the later
    rounds (new literals, same shape) replay it, with eager batches and
    plain queries run in between; every replay equals the eager result
    bit for bit (the same kernels and args run in both)."""
    t, _ = events
    lib = cq.load()
    import ctypes as C
    r0, c0 = C.c_uint64(0), C.c_uint64(0)
    lib.gpu_hip_graph_stats(C.byref(r0), C.byref(c0))
    for r in range(4):
        specs = _graph_specs(r)
        cq.set_setting("QE_BATCH_GRAPH", 1)
        graph = cq.aggregate_batch(t, specs)
        cq.set_setting("QE_BATCH_GRAPH", 0)
        eager = cq.aggregate_batch(t, specs)
        plain = t.aggregate(specs[1].funcs, specs[1].args, where=specs[1].where)
        for i, (sp, g, e) in enumerate(zip(specs, graph, eager)):
            assert np.array_equal(g.keys, e.keys), (r, i)
            assert np.array_equal(g.cells, e.cells), (r, i)
        _same(specs[1], graph[1].keys, _vals(graph[1]), plain.keys, _vals(plain), f"plain r{r}")
    r1, c1 = C.c_uint64(0), C.c_uint64(0)
    lib.gpu_hip_graph_stats(C.byref(r1), C.byref(c1))
    assert c1.value - c0.value >= 1
    assert r1.value - r0.value >= 3


def test_repeated_queries_run_once(events):
    """A batch of 8 distinct queries each sent 8 times, interleaved: every
    copy gets the result the distinct batch gives, with identical queries
    run once and also with that switched off."""
    t, _ = events
    distinct = wl.mixed_queries(8, seed=21)
    base = cq.aggregate_batch(t, distinct)
    specs = [distinct[i % 8] for i in range(64)]
    for dedupe in (1, 0):
        cq.set_setting("QE_BATCH_DEDUPE", dedupe)
        got = cq.aggregate_batch(t, specs)
        for i, (sp, r) in enumerate(zip(specs, got)):
            b = base[i % 8]
            _same(sp, r.keys, _vals(r), b.keys, _vals(b), f"dedupe={dedupe} q{i}")


def test_cost_estimate_picks_plain_or_shared(events):
    """Launches the estimate prices as cheaper plain run one plain operator
    per query, equal to calling it alone (f64 sums within 1e-12: the dense
    GROUP BY adds them with atomics in no fixed order); with the estimate
    off every query is fused into the shared scan."""
    t, _ = events
    specs = wl.mixed_queries(12, seed=23)
    cq.set_setting("QE_BATCH_COST_SHARED_MROWS", 1)  # shared scan priced far slower than plain
    plain_b, tr = cq.aggregate_batch(t, specs, with_trace=True)
    assert tr["fused"] == 0 and tr["fallback"] == 12
    for i, (sp, r) in enumerate(zip(specs, plain_b)):
        alone = t.aggregate(sp.funcs, sp.args, where=sp.where, group=sp.group)
        assert r.trace["fused"] == 0
        _same(sp, r.keys, _vals(r), alone.keys, _vals(alone), f"plain q{i}")
    cq.reset_settings()
    cq.set_setting("QE_BATCH_COST", 0)
    shared, tr = cq.aggregate_batch(t, specs, with_trace=True)
    assert tr["fused"] == 12 and tr["fallback"] == 0
    for i, (sp, a, b) in enumerate(zip(specs, shared, plain_b)):
        _same(sp, a.keys, _vals(a), b.keys, _vals(b), f"q{i}")


def test_few_queries_on_a_large_table_run_plain():
    """A launch of fewer queries than a block has waves leaves waves idle in
    the shared scan, so the estimate charges it at least eight queries'
    worth of query-rows: one 3-leaf AND/OR count over 16M rows runs as the
    plain operator, and with that minimum set to 1 it is fused instead.
    Both counts equal numpy's."""
    n = 16_000_000
    rng = np.random.default_rng(31)
    cols = [rng.integers(0, 1000, n, dtype=np.uint32) for _ in range(3)]
    t = cq.Table(n)
    try:
        for i, c in enumerate(cols):
            t.add_column(f"c{i}", cq.U32, c)
        sp = cq.AggSpec([cq.COUNT], [None], ((col(0) == 7) & (col(1) < 500)) | (col(2) > 990))
        expect = int(np.count_nonzero(((cols[0] == 7) & (cols[1] < 500)) | (cols[2] > 990)))
        res, tr = cq.aggregate_batch(t, [sp], with_trace=True)
        assert tr["fused"] == 0 and tr["fallback"] == 1
        assert int(res[0].cells[0, 0]) == expect
        cq.set_setting("QE_BATCH_COST_MIN_QUERIES", 1)
        res, tr = cq.aggregate_batch(t, [sp], with_trace=True)
        assert tr["fused"] == 1 and tr["fallback"] == 0
        assert int(res[0].cells[0, 0]) == expect
    finally:
        t.close()


def test_i64_leaves_compare_like_f64():
    """i64 WHERE leaves in the shared scan compare as f64, like the
    reference's CPU path: every count equals numpy's f64 comparison and the
    plain operator, at values around +-2^53 (where the conversion rounds)
    and +-2^62, for every comparison, with integral, fractional and
    unrepresentable-as-i64-exact literals."""
    rng = np.random.default_rng(31)
    big = 1 << 53
    edge = np.array([-(1 << 62), -big - 1, -big, -big + 1, -1, 0, 1, 12345, big - 2, big - 1, big, big + 1,
                     big + 2, 1 << 62], dtype=np.int64)
    n = 50_000
    x = np.concatenate([np.repeat(edge, 100), rng.integers(-(1 << 54), 1 << 54, n - 100 * len(edge))])
    t = cq.Table(n)
    t.add_column("x", cq.I64, x)
    t.add_column("g", cq.U32, (np.arange(n) % 7).astype(np.uint32))
    lits = [0, -1, 12345, big - 1, -(big - 1), big - 2, float(big), -float(big), 2.5, -7.5, float(big + 2)]
    ops = [lambda c: col(0) == c, lambda c: col(0) < c, lambda c: col(0) <= c, lambda c: col(0) > c,
           lambda c: col(0) >= c]
    npops = [np.equal, np.less, np.less_equal, np.greater, np.greater_equal]
    specs, expect = [], []
    for c in lits:
        for k in range(5):
            specs.append(cq.AggSpec([cq.COUNT], [None], ops[k](c), 1 if k % 2 else None))
            expect.append(npops[k](x.astype(np.float64), float(c)))
    cq.set_setting("QE_BATCH_COST", 0)
    res, tr = cq.aggregate_batch(t, specs, with_trace=True)
    assert tr["fused"] == len(specs)
    g = np.arange(n) % 7
    for i, (sp, r, mask) in enumerate(zip(specs, res, expect)):
        plain = t.aggregate(sp.funcs, sp.args, where=sp.where, group=sp.group)
        if sp.group is None:
            assert int(r.cells[0, 0]) == int(mask.sum()) == int(plain.cells[0, 0]), i
        else:
            want = np.bincount(g[mask], minlength=7)
            got = dict(zip(r.keys.astype(np.int64), r.cells[:, 0].astype(np.int64)))
            assert {k: v for k, v in got.items() if v} == {k: int(v) for k, v in enumerate(want) if v}, i
            _same(sp, r.keys, _vals(r), plain.keys, _vals(plain), f"plain q{i}")
    t.close()


def test_mixed_batch_with_nullable_timestamps(orc):
    """The 64-query mix over the event table with 2% NULLs in t0-t4: their
    WHERE leaves run in the shared scan with NULL rows failing, and every
    query equals the plain operator and the oracle's NULL-aware path."""
    rows = 300_007
    cols = wl.event_table(rows, seed=17)
    nulls = wl.event_nulls(rows, seed=17)
    t = wl.upload(cols, nulls)
    specs = wl.mixed_queries(64, seed=11)
    cq.set_setting("QE_BATCH_COST", 0)
    res, tr = cq.aggregate_batch(t, specs, with_trace=True)
    assert tr["fused"] == 64
    hit = 0
    for i, (sp, r) in enumerate(zip(specs, res)):
        plain = t.aggregate(sp.funcs, sp.args, where=sp.where, group=sp.group)
        _same(sp, r.keys, _vals(r), plain.keys, _vals(plain), f"plain q{i}")
        if i % 8 == 2:  # the shape whose leaves read t0-t4
            rk, rv = orc.group_aggregate(cols, sp.group, sp.funcs, sp.args, where=sp.where, nulls=nulls)
            _same(sp, r.keys, _vals(r), rk, rv, f"oracle q{i}")
            hit += 1
    assert hit == 8
    t.close()


@pytest.mark.parametrize("compact", [0, 1])
def test_sparse_and_dense_walks_match_oracle(events, orc, compact):
    """Grouped queries whose passing rows are sparse in a wave walk only
    those rows (QE_BATCH_COMPACT 1), denser waves the unrolled walk; with 0
    every wave takes the unrolled walk.  A very selective WHERE (every wave
    sparse) and a dense one, over f64 and i32-range u32 arguments with SUM,
    MIN and MAX, and keys outside the query's WHERE range, against the
    oracle either way."""
    t, cols = events
    m, u = wl.M0 + 2, wl.U0 + 1
    specs = [
        cq.AggSpec([cq.COUNT, cq.SUM, cq.MIN, cq.MAX], [None, m, m, m], col(u) < 2000, wl.REGION),          # ~0.2%
        cq.AggSpec([cq.COUNT, cq.SUM, cq.MIN, cq.MAX], [None, u, u, u], col(u) < 1500, wl.STATUS),          # sparse
        cq.AggSpec([cq.COUNT, cq.SUM, cq.MIN, cq.MAX], [None, m, m, m], col(u) >= 1000, wl.CATEGORY),       # dense
        cq.AggSpec([cq.SUM, cq.MAX, cq.MIN], [u, u, m], (col(wl.REGION) < 45) & (col(u) > 10), wl.STATUS),  # dense
    ]
    cq.set_setting("QE_BATCH_COST", 0)
    cq.set_setting("QE_BATCH_COMPACT", compact)
    res = cq.aggregate_batch(t, specs)
    for sp, r in zip(specs, res):
        assert r.trace["fused"] == 1
        rk, rv = orc.group_aggregate(cols, sp.group, sp.funcs, sp.args, where=sp.where)
        _same(sp, r.keys, _vals(r), rk, rv, (compact, sp.funcs))
