"""GPU parity of the hash GROUP BY (keys beyond the dense range, any key
type, any WHERE) against the CPU oracle.  Counts, integer sums, min/max and
keys bit-exact; f64 sums within 1e-12 relative (the GPU adds a group's
values in arrival order through atomics, the CPU path in row order).
Groups come out of the hash path in table-slot order, so both sides are
sorted by key before comparing."""
import numpy as np
import pytest

import computeql_amd as cq
from computeql_amd.ir import col

pytestmark = pytest.mark.gpu

N = 300_007
F64_RTOL = 1e-12


@pytest.fixture(autouse=True)
def _settings():
    cq.reset_settings()
    yield
    cq.reset_settings()


@pytest.fixture(scope="module")
def ev(orc):
    t = cq.Table(N)
    for name in ("id", "region", "amount", "ts"):
        t.add_column(name, cq.U32)
    t.gen_events(orc.SEED, id=0, region=1, amount=2, ts=3)
    host = orc.gen_events(N, seed=orc.SEED)
    yield t, host
    t.close()


def _sorted(keys, vals):
    o = np.argsort(keys, kind="stable")
    return keys[o], vals[o]


def _check(res, rk, rv, funcs, f64_cols=()):
    rk, rv = _sorted(rk, rv)
    o = np.argsort(res.keys, kind="stable")
    keys = res.keys[o]
    assert len(keys) == len(rk)
    assert np.array_equal(keys, rk)
    for e, f in enumerate(funcs):
        got = res.column(e)[o]
        if e in f64_cols and f in (cq.SUM, cq.AVG):
            assert np.allclose(got, rv[:, e], rtol=F64_RTOL, atol=0.0), e
        else:
            assert np.array_equal(got, rv[:, e]), (e, f)


def test_hash_unique_u32_key(ev, orc):
    t, host = ev
    funcs = [cq.COUNT, cq.SUM, cq.MIN, cq.MAX, cq.AVG]
    args = [None, 2, 2, 2, 2]
    res = t.aggregate(funcs, args, group=0)
    assert res.trace["path"] == 30
    rk, rv = orc.group_aggregate(host, 0, funcs, args)
    _check(res, rk, rv, funcs)


def test_hash_fused_where(ev, orc):
    t, host = ev
    funcs = [cq.COUNT, cq.SUM, cq.MAX]
    where = (col(2) < 5000) & (col(1) > 9)
    res = t.aggregate(funcs, [None, 3, 2], where=where, group=3)
    rk, rv = orc.group_aggregate(host, 3, funcs, [None, 3, 2], where=where)
    _check(res, rk, rv, funcs)


@pytest.mark.parametrize("where", [(col(1) == 3) | (col(2) < 100), (col(2) >= 9000) | (col(3) < 2**30)])
def test_general_where_group_and_global(ev, orc, where):
    t, host = ev
    funcs = [cq.COUNT, cq.SUM, cq.MIN, cq.MAX]
    args = [None, 2, 2, 2]
    res = t.aggregate(funcs, args, where=where, group=1)
    rk, rv = orc.group_aggregate(host, 1, funcs, args, where=where)
    _check(res, rk, rv, funcs)
    g = t.aggregate(funcs, args, where=where)
    _, gv = orc.group_aggregate(host, None, funcs, args, where=where)
    assert g.groups == 1
    for e in range(len(funcs)):
        assert g.column(e)[0] == gv[0, e], e


def test_general_where_no_rows(ev, orc):
    t, host = ev
    where = (col(2) > 20000) | (col(1) > 1000)
    g = t.aggregate([cq.COUNT, cq.SUM], [None, 2], where=where)
    assert g.groups == 1 and int(g.cells[0, 0]) == 0 and g.column(1)[0] == 0.0
    r = t.aggregate([cq.COUNT, cq.SUM], [None, 2], where=where, group=1)
    assert r.groups == 0


@pytest.mark.parametrize("setting", [{"QE_AGG_DENSE": 0}, {"QE_AGG_DENSE": 0, "QE_AGG_HASH_LDS_SLOTS": 0},
                                     {"QE_AGG_HASH_LDS_SLOTS": 0}])
def test_hash_matches_oracle_under_settings(ev, orc, setting):
    t, host = ev
    for k, v in setting.items():
        cq.set_setting(k, v)
    funcs = [cq.COUNT, cq.SUM, cq.MIN, cq.MAX]
    for key in (1, 0):
        res = t.aggregate(funcs, [None, 2, 2, 2], group=key)
        if key == 1 and setting.get("QE_AGG_DENSE", 1) == 0:
            assert res.trace["path"] == 30
        rk, rv = orc.group_aggregate(host, key, funcs, [None, 2, 2, 2])
        _check(res, rk, rv, funcs)


def test_zipf_key_f64_measures(orc):
    rng = np.random.default_rng(11)
    n = 400_009
    user = ((rng.zipf(1.1, n) - 1) % 1_000_000).astype(np.uint32)
    m = rng.lognormal(3.0, 1.0, n)
    c = rng.integers(0, 1 << 20, n).astype(np.uint32)
    t = cq.Table(n)
    t.add_column("user", cq.U32, user)
    t.add_column("m", cq.F64, m)
    t.add_column("c", cq.U32, c)
    funcs = [cq.COUNT, cq.SUM, cq.AVG, cq.MIN, cq.MAX, cq.SUM]
    args = [None, 1, 1, 1, 1, 2]
    res = t.aggregate(funcs, args, group=0)
    assert res.trace["path"] == 30
    rk, rv = orc.group_aggregate([user, m, c], 0, funcs, args)
    _check(res, rk, rv, funcs, f64_cols=(1, 2))
    t.close()


def test_i64_and_f64_keys(orc):
    rng = np.random.default_rng(12)
    n = 200_003
    pool = rng.integers(-2**40, 2**40, 5000, dtype=np.int64)
    k64 = pool[rng.integers(0, 5000, n)]
    fpool = np.concatenate([rng.normal(0, 1e6, 997), [0.0, -0.0, 1.5]])
    kf = fpool[rng.integers(0, 1000, n)]
    v = rng.integers(-1000, 1000, n).astype(np.int32)
    t = cq.Table(n)
    t.add_column("k64", cq.I64, k64)
    t.add_column("kf", cq.F64, kf)
    t.add_column("v", cq.I32, v)
    funcs = [cq.COUNT, cq.SUM, cq.MIN, cq.MAX]
    for key in (0, 1):
        res = t.aggregate(funcs, [None, 2, 2, 2], group=key)
        assert res.trace["path"] == 30
        rk, rv = orc.group_aggregate([k64, kf, v], key, funcs, [None, 2, 2, 2])
        _check(res, rk, rv, funcs)
    t.close()


def test_table_growth_and_full_error(ev, orc):
    t, host = ev
    funcs = [cq.COUNT, cq.SUM]
    sub = col(0) < 1500
    cq.set_setting("QE_AGG_HASH_LOG2", 9)  # 512 slots for 1500 groups: grows twice
    res = t.aggregate(funcs, [None, 2], where=sub, group=0)
    rk, rv = orc.group_aggregate(host, 0, funcs, [None, 2], where=sub)
    _check(res, rk, rv, funcs)
    cq.set_setting("QE_AGG_HASH_LOG2", 4)  # 16 .. 128 slots: still full after three regrowths
    with pytest.raises(cq.QueryError, match="full"):
        t.aggregate(funcs, [None, 2], where=sub, group=0)
