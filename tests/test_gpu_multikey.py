"""GROUP BY over several columns, and the 2^53 exactness guard, through the
C ABI against the oracle.

Several keys (query_exec.c:10280-10299; the fused grouped kernel combines up
to 3 small-range keys as digits, :9462-9500):
- dense: 2 and 3 32-bit keys whose zone-map ranges multiply to <= 65536,
  with and without a fused WHERE, signed keys, args folded into counts;
- hash, digits: integer keys whose ranges multiply past the dense range;
- hash, tuple: f64 / i64 / u64 keys, hashed with per-slot key checks;
- a WHERE that is not an AND of comparisons (device row ids + hash);
- batches (a multi-key query runs on the plain operator inside the batch),
  host-backed tables (streamed per chunk), and joins grouped by 2 dim
  columns in every join form.
Groups are compared after sorting both sides by key tuple.
"""
import numpy as np
import pytest

import computeql_amd as cq
from computeql_amd.ir import col

pytestmark = pytest.mark.gpu

N = 2_000_003


def _sorted(keys, cells):
    keys = np.asarray(keys, dtype=np.float64)
    if keys.ndim == 1:
        keys = keys[:, None]
    o = np.lexsort(keys.T[::-1])
    return keys[o], np.asarray(cells)[o]


def _check(res, rk, rv, funcs):
    gk, gc = _sorted(res.keys, np.stack([res.column(e).astype(np.float64) for e in range(len(funcs))], axis=1))
    ok, ov = _sorted(rk, rv)
    assert gk.shape == ok.shape, (gk.shape, ok.shape)
    assert np.array_equal(gk, ok)
    for e, f in enumerate(funcs):
        if f == cq.AVG:
            np.testing.assert_allclose(gc[:, e], ov[:, e], rtol=1e-12)
        else:
            assert np.array_equal(gc[:, e], ov[:, e]), e


@pytest.fixture(scope="module")
def mk():
    r = np.random.default_rng(21)
    cols = {
        "a": r.integers(0, 50, N).astype(np.uint32),            # 50
        "b": r.integers(-10, 10, N).astype(np.int32),            # 20, signed
        "c": r.integers(3, 13, N).astype(np.uint32),             # 10
        "u": r.integers(0, 200_000, N).astype(np.uint32),        # wide
        "v": r.integers(0, 2**32, N, dtype=np.uint64).astype(np.uint32),
        "w": r.integers(-(2**31), 2**31, N).astype(np.int32),
        "f": (r.integers(0, 40, N) * 0.25 - 3.0).astype(np.float64),
        "i": (r.integers(-5, 5, N) * (2**40)).astype(np.int64),
        "g": r.integers(0, 7, N).astype(np.uint64),
    }
    types = {"a": cq.U32, "b": cq.I32, "c": cq.U32, "u": cq.U32, "v": cq.U32, "w": cq.I32, "f": cq.F64, "i": cq.I64,
             "g": cq.U64}
    t = cq.Table(N)
    for name, arr in cols.items():
        t.add_column(name, types[name], arr)
    host = list(cols.values())
    idx = {name: i for i, name in enumerate(cols)}
    yield t, host, idx
    t.close()


ALL7 = ([cq.COUNT, cq.SUM, cq.MIN, cq.MAX, cq.SUM, cq.MIN, cq.AVG], ["", "v", "v", "v", "w", "w", "w"])
DENSE_CASES = [
    # (keys, WHERE, functions and args): ranges 1000 and 200 take 2 args of
    # sums and min/max in LDS; 10000 groups take one kind per launch
    (["a", "b"], None) + ALL7,
    (["a", "b", "c"], None, [cq.COUNT, cq.SUM, cq.AVG], ["", "v", "v"]),
    (["b", "c"], col(0) < 25) + ALL7,
    (["c", "a", "b"], (col(4) > 2**31) & (col(0) >= 10), [cq.COUNT, cq.MIN, cq.MAX], ["", "w", "w"]),
]


@pytest.mark.parametrize("case", range(len(DENSE_CASES)))
def test_dense_multi_key(mk, orc, case):
    t, host, idx = mk
    names, where, funcs, anames = DENSE_CASES[case]
    keys = [idx[n] for n in names]
    args = [idx[a] if a else None for a in anames]
    res = t.aggregate(funcs, args, where=where, group=keys)
    assert cq.AGG_PATHS[res.trace["path"]] == ("fused_dense" if where is not None else "dense")
    rk, rv = orc.group_aggregate(host, keys, funcs, args, where=where)
    _check(res, rk, rv, funcs)
    # dense groups come out in key order, key 0 first
    k = res.keys
    assert all(tuple(k[i]) < tuple(k[i + 1]) for i in range(len(k) - 1))


def test_dense_multi_key_fold(mk, orc):
    """Arguments that are key columns fold into the counts."""
    t, host, idx = mk
    keys = [idx["a"], idx["b"]]
    funcs = [cq.COUNT, cq.SUM, cq.MIN, cq.SUM, cq.MAX]
    args = [None, idx["a"], idx["a"], idx["b"], idx["b"]]
    res = t.aggregate(funcs, args, group=keys)
    assert cq.AGG_PATHS[res.trace["path"]] == "dense_key_fold"
    rk, rv = orc.group_aggregate(host, keys, funcs, args)
    _check(res, rk, rv, funcs)


HASH_CASES = [
    (["u", "a"], None),                # digits: 200000 x 50
    (["a", "u", "c"], col(1) < 0),     # digits with a fused WHERE
    (["f", "a"], None),                # tuple: f64 key
    (["i", "g", "b"], None),           # tuple: i64, u64, i32
    (["a", "b", "c", "f"], None),      # four keys
    (["u", "b"], (col(0) < 3) | (col(2) == 7)),  # general WHERE: device row ids
]


@pytest.mark.parametrize("case", range(len(HASH_CASES)))
def test_hash_multi_key(mk, orc, case):
    t, host, idx = mk
    names, where = HASH_CASES[case]
    keys = [idx[n] for n in names]
    funcs = [cq.COUNT, cq.SUM, cq.MAX, cq.MIN]
    args = [None, idx["v"], idx["w"], idx["f"]]
    res = t.aggregate(funcs, args, where=where, group=keys)
    assert cq.AGG_PATHS[res.trace["path"]] == "hash"
    rk, rv = orc.group_aggregate(host, keys, funcs, args, where=where)
    _check(res, rk, rv, funcs)


def test_dense_disabled_runs_hash_digits(mk, orc):
    t, host, idx = mk
    keys = [idx["a"], idx["b"], idx["c"]]
    funcs = [cq.COUNT, cq.SUM]
    cq.set_setting("QE_AGG_DENSE", 0)
    try:
        res = t.aggregate(funcs, [None, idx["v"]], group=keys)
    finally:
        cq.reset_settings()
    assert cq.AGG_PATHS[res.trace["path"]] == "hash"
    rk, rv = orc.group_aggregate(host, keys, funcs, [None, idx["v"]])
    _check(res, rk, rv, funcs)


def test_batch_with_multi_key_queries(mk, orc):
    t, host, idx = mk
    specs = [cq.AggSpec([cq.COUNT, cq.SUM], [None, idx["v"]], col(0) < 20, idx["a"]),
             cq.AggSpec([cq.COUNT, cq.SUM], [None, idx["v"]], col(0) < 20, [idx["a"], idx["b"]]),
             cq.AggSpec([cq.COUNT, cq.MAX], [None, idx["w"]], None, [idx["c"], idx["u"]], cap=1 << 21),
             cq.AggSpec([cq.COUNT], [None], col(2) > 5)]
    cq.set_setting("QE_BATCH_COST", 0)
    try:
        res = cq.aggregate_batch(t, specs)
    finally:
        cq.reset_settings()
    assert res[0].trace["fused"] == 1 and res[1].trace["fused"] == 0 and res[2].trace["fused"] == 0
    for sp, r in zip(specs, res):
        rk, rv = orc.group_aggregate(host, sp.group, sp.funcs, sp.args, where=sp.where)
        _check(r, rk, rv, sp.funcs)


def test_streamed_multi_key(orc):
    """A host-backed table: each chunk's groups merged by key tuple."""
    r = np.random.default_rng(5)
    n = 3_000_001
    a = r.integers(0, 30, n).astype(np.uint32)
    b = r.integers(0, 5000, n).astype(np.uint32)
    v = r.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    t = cq.Table(n)
    try:
        for name, arr in (("a", a), ("b", b), ("v", v)):
            t.add_host_column(name, cq.U32, arr)
        cq.set_setting("QE_SCAN_STREAM_CHUNK_ROWS", 1 << 20)
        funcs = [cq.COUNT, cq.SUM, cq.MIN]
        for keys in ([0, 1], [1, 0], [0]):
            res = t.aggregate(funcs, [None, 2, 2], where=col(2) > 1000, group=keys)
            rk, rv = orc.group_aggregate([a, b, v], keys, funcs, [None, 2, 2], where=col(2) > 1000)
            _check(res, rk, rv, funcs)
        assert cq.stream_stats()["chunks"] == 3
    finally:
        cq.reset_settings()
        t.close()


JOIN_MODES = {
    "join_part_group": {},
    "join_part_row": {"QE_JOIN_GROUP_PAYLOAD": 0},
    "join_probe_group": {"QE_JOIN_PARTITION": 0},
    "join_probe_row": {"QE_JOIN_PARTITION": 0, "QE_JOIN_GROUP_PAYLOAD": 0},
}


@pytest.fixture(scope="module")
def star2(orc):
    n = 4_000_037
    t = cq.Table(n)
    for name in ("id", "region", "amount", "ts", "fk"):
        t.add_column(name, cq.U32)
    t.gen_events(orc.SEED, id=0, region=1, amount=2, ts=3)
    dk, seg = orc.gen_dim(100_000, 100, seed=9)
    r = np.random.default_rng(2)
    tier = r.integers(0, 4, len(dk)).astype(np.uint32)
    zone = (seg // 25).astype(np.int32) - 2          # signed
    dim = cq.Table(len(dk))
    dim.add_column("key", cq.U32, dk)
    dim.add_column("segment", cq.U32, seg)
    dim.add_column("tier", cq.U32, tier)
    dim.add_column("zone", cq.I32, zone)
    t.gen_fact_fk(4, orc.SEED, dim, 0)
    host = orc.gen_events(n, seed=orc.SEED)
    fk = orc.gen_fact_fk(n, dk, seed=orc.SEED)
    yield t, dim, host, fk, dk, seg, tier, zone
    t.close()
    dim.close()


@pytest.mark.parametrize("mode", list(JOIN_MODES))
def test_join_two_dim_group_columns(star2, orc, mode):
    t, dim, host, fk, dk, seg, tier, zone = star2
    funcs = [cq.COUNT, cq.SUM, cq.MIN, cq.MAX]
    for k, v in JOIN_MODES[mode].items():
        cq.set_setting(k, v)
    try:
        res = cq.aggregate_join(t, 4, dim, 0, funcs, [None, 2, 2, 2], dim_group=[2, 3])
        res3 = cq.aggregate_join(t, 4, dim, 0, [cq.COUNT, cq.SUM], [None, 2], dim_where=col(1) < 60,
                                 dim_group=[3, 1, 2])
    finally:
        cq.reset_settings()
    assert cq.AGG_PATHS[res.trace["path"]] == mode
    rk, rv = orc.join_aggregate(fk, [None, host[2], host[2], host[2]], funcs, dk, dim_group=[tier, zone.view(np.uint32)])
    rk = np.stack([rk[:, 0], rk[:, 1].astype(np.uint32).view(np.int32).astype(np.float64)], axis=1)
    _check(res, rk, rv, funcs)
    assert res.keys.shape == (16, 2)
    rk, rv = orc.join_aggregate(fk, [None, host[2]], [cq.COUNT, cq.SUM], dk, dim_pass=seg < 60,
                                dim_group=[zone.view(np.uint32), seg, tier])
    rk[:, 0] = rk[:, 0].astype(np.uint32).view(np.int32)
    _check(res3, rk, rv, [cq.COUNT, cq.SUM])


def test_join_multi_group_dense_fact_key(orc):
    """A small fact-key range: the dense per-key form, groups on the host."""
    r = np.random.default_rng(8)
    n = 1_000_003
    fk = r.integers(0, 300, n).astype(np.uint32)
    amt = r.integers(0, 1000, n).astype(np.uint32)
    dk = np.arange(0, 300, dtype=np.uint32)
    ga = (dk % 3).astype(np.uint32)
    gb = (dk % 7).astype(np.uint32)
    f, d = cq.Table(n), cq.Table(300)
    try:
        f.add_column("fk", cq.U32, fk)
        f.add_column("amt", cq.U32, amt)
        d.add_column("k", cq.U32, dk)
        d.add_column("ga", cq.U32, ga)
        d.add_column("gb", cq.U32, gb)
        res = cq.aggregate_join(f, 0, d, 0, [cq.COUNT, cq.SUM], [None, 1], dim_group=[1, 2])
        assert cq.AGG_PATHS[res.trace["path"]] == "join_dense"
        rk, rv = orc.join_aggregate(fk, [None, amt], [cq.COUNT, cq.SUM], dk, dim_group=[ga, gb])
        _check(res, rk, rv, [cq.COUNT, cq.SUM])
    finally:
        f.close()
        d.close()


# ---------------------------------------------------------------------------
# the 2^53 guard (qe_aggregate_check_sum_exact, query_exec.c:7028-7049)
# ---------------------------------------------------------------------------
def test_sum_reaching_2_pow_53_fails(orc):
    n = 5_000_011
    t = cq.Table(n)
    try:
        for name in ("id", "region", "amount", "ts"):
            t.add_column(name, cq.U32)
        t.gen_events(orc.SEED, id=0, region=1, amount=2, ts=3)
        ts = orc.gen_events(n, seed=orc.SEED)[3]
        assert int(ts.astype(np.uint64).sum()) >= 2**53
        with pytest.raises(cq.QueryError, match="2\\^53"):
            t.aggregate([cq.SUM], [3])
        with pytest.raises(orc.ExactnessError):
            orc.group_aggregate([ts], None, [cq.SUM], [0])
        # the first 4,000,000 rows stay below 2^53: exact, equal to the oracle
        head = ts[:4_000_000]
        assert int(head.astype(np.uint64).sum()) < 2**53
        res = t.aggregate([cq.SUM, cq.COUNT], [3, None], where=col(0) < 4_000_000)
        assert res.column(0)[0] == float(int(head.astype(np.uint64).sum()))
        ids = np.arange(n, dtype=np.uint32)
        _, rv = orc.group_aggregate([ids, ts], None, [cq.SUM], [1], where=col(0) < 4_000_000)
        assert res.column(0)[0] == rv[0, 0]
        # AVG / MIN / MAX / COUNT are not checked; grouped sums stay small
        t.aggregate([cq.AVG, cq.MAX, cq.COUNT], [3, 3, None])
        g = t.aggregate([cq.SUM], [3], group=1)
        assert g.groups == 50
        # the guard switched off returns the rounded double
        cq.set_setting("QE_WIDE_INT_CHECK", 0)
        try:
            res = t.aggregate([cq.SUM], [3])
        finally:
            cq.reset_settings()
        assert res.column(0)[0] == float(int(ts.astype(np.uint64).sum()))
    finally:
        t.close()


def test_grouped_batched_and_joined_sums_reaching_2_pow_53_fail(orc):
    """One group of 2,097,153 rows of 2^32 - 1 sums past 2^53 (its
    neighbours stay below): dense, hash, batched (fused and plain) and join
    forms all fail the query, as the CPU path does."""
    big = 2_097_153
    n = big + 1000
    key = np.zeros(n, dtype=np.uint32)
    key[big:] = 1 + np.arange(1000, dtype=np.uint32) % 10
    v = np.full(n, 0xFFFFFFFF, dtype=np.uint32)
    t, d = cq.Table(n), cq.Table(11)
    try:
        t.add_column("k", cq.U32, key)
        t.add_column("v", cq.U32, v)
        d.add_column("k", cq.U32, np.arange(11, dtype=np.uint32))
        d.add_column("g", cq.U32, np.arange(11, dtype=np.uint32))
        with pytest.raises(orc.ExactnessError):
            orc.group_aggregate([key, v], 0, [cq.COUNT, cq.SUM], [None, 1])
        for setting in ({}, {"QE_AGG_DENSE": 0}):
            for k, x in setting.items():
                cq.set_setting(k, x)
            try:
                with pytest.raises(cq.QueryError, match="2\\^53"):
                    t.aggregate([cq.COUNT, cq.SUM], [None, 1], group=0)
            finally:
                cq.reset_settings()
        # the small groups alone are exact
        res = t.aggregate([cq.COUNT, cq.SUM], [None, 1], where=col(0) > 0, group=0)
        rk, rv = orc.group_aggregate([key, v], 0, [cq.COUNT, cq.SUM], [None, 1], where=col(0) > 0)
        _check(res, rk, rv, [cq.COUNT, cq.SUM])
        for cost in (0, 1):
            cq.set_setting("QE_BATCH_COST", cost)
            try:
                with pytest.raises(cq.QueryError):
                    cq.aggregate_batch(t, [cq.AggSpec([cq.SUM], [1], None, 0), cq.AggSpec([cq.COUNT], [None])])
            finally:
                cq.reset_settings()
        with pytest.raises(cq.QueryError, match="2\\^53"):
            cq.aggregate_join(t, 0, d, 0, [cq.SUM], [1], dim_group=1)
        with pytest.raises(orc.ExactnessError):
            orc.join_aggregate(key, [v], [cq.SUM], np.arange(11, dtype=np.uint32), dim_group=np.arange(11, dtype=np.uint32))
    finally:
        t.close()
        d.close()
