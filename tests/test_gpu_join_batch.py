"""qe_aggregate_join_batch: many joins over one fact table in one call.
Identical joins run once and joins that need the same fact side share it
(the fact rows in partition order on the partitioned path, the per-key
totals on the dense path).  Every result is checked against the oracle's
CPU join and against the same join run alone."""
import numpy as np
import pytest

import computeql_amd as cq
from computeql_amd.ir import col
from computeql_amd.serve import JoinSpec, Server, aggregate_join_batch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def star(orc):
    """fact (2M rows: id, region, amount, ts, fk) + dim A (131,072 spread
    keys, segment in [0, 100), tier in [0, 4)) + dim B (50 rows, key =
    region, attr)"""
    n = 2_000_003
    t = cq.Table(n)
    for name in ("id", "region", "amount", "ts", "fk"):
        t.add_column(name, cq.U32)
    t.gen_events(orc.SEED, id=0, region=1, amount=2, ts=3)
    dk, seg = orc.gen_dim(131_072, 100, seed=9)
    tier = (seg * 7 + 3) % 4
    dim = cq.Table(len(dk))
    dim.add_column("key", cq.U32, dk)
    dim.add_column("segment", cq.U32, seg)
    dim.add_column("tier", cq.U32, tier.astype(np.uint32))
    t.gen_fact_fk(4, orc.SEED, dim, 0)
    bk = np.arange(50, dtype=np.uint32)
    battr = (np.arange(50, dtype=np.uint32) * 7) % 10
    dimb = cq.Table(50)
    dimb.add_column("key", cq.U32, bk)
    dimb.add_column("attr", cq.U32, battr)
    host = orc.gen_events(n, seed=orc.SEED)
    fk = orc.gen_fact_fk(n, dk, seed=orc.SEED)
    yield t, dim, dimb, host, fk, (dk, seg, tier), (bk, battr)
    t.close()
    dim.close()
    dimb.close()


def _sorted(keys, cells):
    keys = np.asarray(keys, dtype=np.float64).reshape(len(cells), -1)
    order = np.lexsort(keys.T[::-1]) if keys.size else np.arange(len(cells))
    return keys[order], np.asarray(cells)[order]


def _oracle(orc, star, sp):
    """the oracle's answer for a JoinSpec over the star schema"""
    t, dim, dimb, host, fk, (dk, seg, tier), (bk, battr) = star
    cols_a = {1: seg, 2: tier}
    if sp.dim is dimb:
        fkey, dkey, dpass, groups = host[1], bk, None, None
        if sp.dim_where is not None:
            dpass = battr < sp.literal
    else:
        fkey, dkey = fk, dk
        dpass = None if sp.dim_where is None else (seg < sp.literal)
        g = sp.dim_group
        groups = None if g is None else ([cols_a[c] for c in g] if isinstance(g, list) else cols_a[g])
    args = [None if a is None else host[a] for a in sp.args]
    rk, rv = orc.join_aggregate(fkey, args, sp.funcs, dkey, dim_pass=dpass, dim_group=groups)
    return _sorted(rk, rv)


class Spec(JoinSpec):
    literal = None


def _spec(dim, fact_key, funcs, args, literal=None, group=None, where_col=1):
    sp = Spec(dim, fact_key, 0, funcs, args, dim_where=None if literal is None else col(where_col) < literal,
              dim_group=group)
    sp.literal = literal
    return sp


def _check(orc, star, specs, results):
    """each result against the oracle (groups ascending by their values)"""
    for sp, res in zip(specs, results):
        rk, rv = _oracle(orc, star, sp)
        assert res.groups == len(rv), (sp.funcs, sp.args, sp.literal)
        if sp.dim_group is not None:
            assert np.array_equal(np.asarray(res.keys, dtype=np.float64).reshape(res.groups, -1), rk)
        for e in range(len(sp.funcs)):
            assert np.array_equal(res.column(e), rv[:, e]), (sp.funcs, sp.args, sp.literal, e)


def test_partitioned_joins_share_the_fact_side(star, orc):
    """joins on dim A (partitioned path) with different dim WHEREs, group
    columns and functions over the same argument share one partition pass;
    identical ones are folded"""
    t, dim = star[0], star[1]
    specs = [
        _spec(dim, 4, [cq.COUNT, cq.SUM], [None, 2], group=1),
        _spec(dim, 4, [cq.SUM, cq.MIN, cq.MAX], [2, 2, 2], literal=60, group=1),
        _spec(dim, 4, [cq.COUNT, cq.SUM], [None, 2], group=1),  # identical to the first
        _spec(dim, 4, [cq.COUNT, cq.AVG], [None, 2], literal=90, group=[2, 1]),
        _spec(dim, 4, [cq.SUM], [2], literal=75),
    ]
    res, tr = aggregate_join_batch(t, specs, with_trace=True)
    assert tr["queries"] == 5 and tr["runs"] == 4 and tr["fact_passes"] == 1
    _check(orc, star, specs, res)
    # each equals the join run alone
    for sp, r in zip(specs, res):
        alone = cq.aggregate_join(t, 4, dim, 0, sp.funcs, sp.args, dim_where=sp.dim_where, dim_group=sp.dim_group)
        assert cq.AGG_PATHS[alone.trace["path"]] in ("join_part_group", "join_part_row")
        assert np.array_equal(alone.keys, r.keys) and np.array_equal(alone.cells, r.cells)


def test_dense_joins_share_per_key_totals(star, orc):
    """joins on the 50-row dim B (dense path: fact totals per key, combined
    on the host with each join's dim rows) share one per-key pass"""
    t, dimb = star[0], star[2]
    specs = [_spec(dimb, 1, [cq.COUNT], [None], literal=b) for b in (3, 7, 1, 9, 3)]
    res, tr = aggregate_join_batch(t, specs, with_trace=True)
    assert tr["queries"] == 5 and tr["runs"] == 4 and tr["fact_passes"] == 1
    _check(orc, star, specs, res)


def test_mixed_batch_and_no_folding(star, orc):
    """dims interleaved, argument sets differing, folding switched off: every
    join still runs, fact sides shared where the argument sets agree"""
    t, dim, dimb = star[0], star[1], star[2]
    specs = [
        _spec(dimb, 1, [cq.COUNT], [None], literal=4),
        _spec(dim, 4, [cq.COUNT, cq.SUM], [None, 2], literal=50, group=1),
        _spec(dimb, 1, [cq.COUNT, cq.SUM], [None, 2], literal=6),
        _spec(dim, 4, [cq.SUM], [1], group=2),
        _spec(dim, 4, [cq.COUNT, cq.SUM], [None, 2], literal=50, group=1),
        _spec(dimb, 1, [cq.COUNT], [None], literal=4),
    ]
    cq.set_setting("QE_BATCH_DEDUPE", 0)
    try:
        res, tr = aggregate_join_batch(t, specs, with_trace=True)
    finally:
        cq.reset_settings()
    assert tr["runs"] == 6
    # dim B: COUNT-only joins share, the SUM(amount) one has its own pass;
    # dim A: SUM(amount) joins share, SUM(region) has its own
    assert tr["fact_passes"] == 4
    _check(orc, star, specs, res)


def test_served_joins_batch(star, orc):
    """the served loop batches queued joins on one fact table; each result
    matches the oracle"""
    t, dim, dimb = star[0], star[1], star[2]
    specs = [_spec(dim, 4, [cq.COUNT, cq.SUM], [None, 2], literal=lim, group=1) for lim in (20, 40, 60, 80, 100)]
    specs += [_spec(dimb, 1, [cq.COUNT], [None], literal=b) for b in (2, 5, 8)]
    with Server(workers=1, max_batch=64, hbm_budget=8 << 30) as srv:
        tickets = [srv.submit(t, sp) for sp in specs]
        results = [tk.result() for tk in tickets]
        st = srv.stats()
    assert st["joins"] == len(specs)
    _check(orc, star, specs, results)


def test_join_batches_on_concurrent_threads(star, orc):
    """Join batches sorted and run on several host threads at once (as the
    served loop's workers run them): each thread's batch keeps its own run
    order and shared fact side, and every result matches the oracle."""
    import threading

    t, dim, dimb = star[0], star[1], star[2]
    mixes = [
        [_spec(dimb, 1, [cq.COUNT], [None], literal=b) for b in (2, 7, 4)]
        + [_spec(dim, 4, [cq.COUNT, cq.SUM], [None, 2], literal=lim, group=1) for lim in (30, 70)],
        [_spec(dim, 4, [cq.SUM], [2], group=1), _spec(dimb, 1, [cq.COUNT], [None], literal=5)],
        [_spec(dim, 4, [cq.COUNT, cq.SUM], [None, 2], literal=lim, group=1) for lim in (10, 90, 50, 20)],
    ]
    out, errors = {}, []

    def run(k):
        try:
            for rep in range(6):
                out[(k, rep)] = aggregate_join_batch(t, mixes[k])
        except Exception as e:  # reported below
            errors.append(repr(e))

    threads = [threading.Thread(target=run, args=(k,)) for k in range(len(mixes))]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors
    for k, specs in enumerate(mixes):
        first = out[(k, 0)]
        _check(orc, star, specs, first)
        for rep in range(1, 6):
            for a, b in zip(first, out[(k, rep)]):
                assert np.array_equal(a.keys, b.keys) and np.array_equal(a.cells, b.cells)
