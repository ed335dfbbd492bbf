"""Parity at the shapes the bench times.

- Join A with the bench's own 1,000,000-key dimension (same generator and
  seed), which takes the 256-partition form: the whole-line scatter
  (join_part_scatter_wc_1) and the 512-thread LDS probe.  Checked against
  the oracle's join_aggregate per segment (COUNT/SUM/MIN/MAX/AVG) on a
  fact table the CPU finishes in seconds, with uniform foreign keys, with
  skewed ones (one key holding 30% of the rows, the rest Zipf-like), with a
  dim WHERE, and in the per-dim-row form.
- The reference's closed-form table at 1,000,000,000 rows
  (bench_select_scale.c:39-54): every answer of :1123-1169 is a closed
  form of the row count (:98-168), so the GPU path is checked bit for bit
  at full size, beyond 2^32 bytes per column, including the 1M-row
  big_dim join on needle_key through the partitioned probe.
"""
import numpy as np
import pytest

import computeql_amd as cq
from computeql_amd.ir import col

pytestmark = pytest.mark.gpu

FACT_ROWS = 40_000_003  # ragged
DIM_A_ROWS = 1_000_000  # the bench's dimension A


def _norm(keys, vals):
    o = np.argsort(keys, kind="stable")
    return keys[o], vals[o]


@pytest.fixture(scope="module")
def join_a(orc):
    seed = orc.SEED
    t = cq.Table(FACT_ROWS)
    for name in ("id", "region", "amount", "ts", "fk"):
        t.add_column(name, cq.U32)
    t.gen_events(seed, id=0, region=1, amount=2, ts=3)
    dim = cq.Table(DIM_A_ROWS)
    dim.add_column("key", cq.U32)
    dim.add_column("segment", cq.U32)
    dim.gen_dim(seed ^ 0xA, 100, spread_keys=True, key=0, attr=1)
    t.gen_fact_fk(4, seed, dim, 0)
    dk, seg = orc.gen_dim(DIM_A_ROWS, 100, seed=seed ^ 0xA)
    amount = orc.gen_events(FACT_ROWS, seed=seed)[2]
    fk = orc.gen_fact_fk(FACT_ROWS, dk, seed=seed)
    yield t, dim, amount, fk, dk, seg
    t.close()
    dim.close()


def test_join_a_tables_match_oracle(join_a):
    t, dim, amount, fk, dk, seg = join_a
    assert np.array_equal(dim.read(0), dk) and np.array_equal(dim.read(1), seg)
    assert np.array_equal(t.read(4), fk) and np.array_equal(t.read(2), amount)


@pytest.mark.parametrize("funcs,probe", [
    ([cq.SUM], "join_part_probe_pk_1s_512"),                      # the bench's q6
    ([cq.COUNT, cq.SUM], "join_part_probe_pk_1s_512"),
    ([cq.COUNT, cq.SUM, cq.MIN, cq.MAX, cq.AVG], "join_part_probe_pk_1sm_512"),
])
def test_join_a_1m_keys_group_by_segment(join_a, orc, funcs, probe):
    t, dim, amount, fk, dk, seg = join_a
    args = [None if f == cq.COUNT else 2 for f in funcs]
    res = cq.aggregate_join(t, 4, dim, 0, funcs, args, dim_group=1)
    assert cq.AGG_PATHS[res.trace["path"]] == "join_part_group"
    assert res.trace["kernel"] == probe
    rk, rv = orc.join_aggregate(fk, [None if a is None else amount for a in args], funcs, dk, dim_group=seg)
    rk, rv = _norm(rk, rv)
    assert res.groups == 100 and np.array_equal(res.keys, rk)
    for e in range(len(funcs)):
        assert np.array_equal(res.column(e), rv[:, e]), (funcs, e)


@pytest.mark.parametrize("funcs,args,shape", [
    ([cq.COUNT], [None], "0"),
    ([cq.COUNT, cq.SUM], [None, 3], "1s"),           # SUM(ts): too wide for one count-and-sum atomic
    ([cq.MIN, cq.MAX], [2, 2], "1m"),
    ([cq.COUNT, cq.SUM, cq.SUM], [None, 2, 1], "2s"),
    ([cq.SUM, cq.MAX, cq.MIN], [2, 1, 2], "2sm"),
])
def test_join_a_packed_and_unpacked_tables(join_a, orc, funcs, args, shape):
    """The packed partition table (keys and group codes in one word, matched
    by subtraction) against the two-array table and the oracle."""
    t, dim, amount, fk, dk, seg = join_a
    res = cq.aggregate_join(t, 4, dim, 0, funcs, args, dim_group=1)
    assert res.trace["kernel"] == f"join_part_probe_pk_{shape}_512"
    cq.set_setting("QE_JOIN_PROBE_PACKED", 0)
    try:
        two = cq.aggregate_join(t, 4, dim, 0, funcs, args, dim_group=1)
    finally:
        cq.reset_settings()
    assert two.trace["kernel"] == f"join_part_probe_{shape}_512"
    assert np.array_equal(res.keys, two.keys) and np.array_equal(res.cells, two.cells)
    host = orc.gen_events(FACT_ROWS, seed=orc.SEED)
    rk, rv = orc.join_aggregate(fk, [None if a is None else host[a] for a in args], funcs, dk, dim_group=seg)
    rk, rv = _norm(rk, rv)
    assert np.array_equal(res.keys, rk)
    for e in range(len(funcs)):
        assert np.array_equal(res.column(e), rv[:, e]), (funcs, e)


def test_join_a_1m_keys_per_dim_row(join_a, orc):
    """The reference's form: totals per dim row, combined per segment."""
    t, dim, amount, fk, dk, seg = join_a
    funcs = [cq.COUNT, cq.SUM, cq.MIN, cq.MAX]
    cq.set_setting("QE_JOIN_GROUP_PAYLOAD", 0)
    try:
        res = cq.aggregate_join(t, 4, dim, 0, funcs, [None, 2, 2, 2], dim_group=1)
    finally:
        cq.reset_settings()
    assert cq.AGG_PATHS[res.trace["path"]] == "join_part_row"
    rk, rv = orc.join_aggregate(fk, [None, amount, amount, amount], funcs, dk, dim_group=seg)
    rk, rv = _norm(rk, rv)
    assert np.array_equal(res.keys, rk)
    for e in range(len(funcs)):
        assert np.array_equal(res.column(e), rv[:, e]), e


def test_join_a_1m_keys_dim_where(join_a, orc):
    t, dim, amount, fk, dk, seg = join_a
    funcs = [cq.COUNT, cq.SUM, cq.MAX]
    where = (col(1) < 37) | (col(1) == 90)
    res = cq.aggregate_join(t, 4, dim, 0, funcs, [None, 2, 2], dim_where=where, dim_group=1)
    assert cq.AGG_PATHS[res.trace["path"]] == "join_part_group"
    keep = (seg < 37) | (seg == 90)
    rk, rv = orc.join_aggregate(fk, [None, amount, amount], funcs, dk, dim_pass=keep, dim_group=seg)
    rk, rv = _norm(rk, rv)
    assert res.groups == 38 and np.array_equal(res.keys, rk)
    for e in range(len(funcs)):
        assert np.array_equal(res.column(e), rv[:, e]), e


def test_join_a_1m_keys_skewed_fk(join_a, orc):
    """Skewed foreign keys: 30% of the rows carry one key (one partition
    then holds ~30% of the fact rows), the rest follow a Zipf-like law over
    the dim keys, and 1% match no dim key at all."""
    t, dim, amount, fk, dk, seg = join_a
    n = 8_000_017
    r = np.random.default_rng(11)
    idx = np.minimum(r.zipf(1.3, n) - 1, DIM_A_ROWS - 1)
    idx[r.random(n) < 0.30] = 12345
    sk = dk[idx].astype(np.uint32)
    miss = r.random(n) < 0.01
    absent = np.setdiff1d(np.arange(1, 200_000, dtype=np.uint32), dk)[:1000]
    sk[miss] = absent[r.integers(0, len(absent), int(miss.sum()))]
    amt = r.integers(0, 10_000, n).astype(np.uint32)
    f = cq.Table(n)
    try:
        f.add_column("fk", cq.U32, sk)
        f.add_column("amount", cq.U32, amt)
        funcs = [cq.COUNT, cq.SUM, cq.MIN, cq.MAX]
        res = cq.aggregate_join(f, 0, dim, 0, funcs, [None, 1, 1, 1], dim_group=1)
        assert cq.AGG_PATHS[res.trace["path"]] == "join_part_group"
        rk, rv = orc.join_aggregate(sk, [None, amt, amt, amt], funcs, dk, dim_group=seg)
        rk, rv = _norm(rk, rv)
        assert np.array_equal(res.keys, rk)
        for e in range(len(funcs)):
            assert np.array_equal(res.column(e), rv[:, e]), e
        assert int(res.column(0).sum()) == n - int(miss.sum())
    finally:
        f.close()


# ---------------------------------------------------------------------------
# 1B rows: the reference's closed-form table (bench_select_scale.c:39-54)
# ---------------------------------------------------------------------------
BIG = 1_000_000_000
KEY_MUL = 2654435761  # SS_KEY_MULTIPLIER


@pytest.fixture(scope="module")
def scale_1b():
    t = cq.Table(BIG)
    for name in ("id", "region_id", "amount", "needle_key"):
        t.add_column(name, cq.U32)
    t.gen_scale_events(id=0, region=1, amount=2, key=3)
    yield t
    t.close()


def test_1b_global_and_filters(scale_1b, orc):
    t = scale_1b
    res = t.aggregate([cq.SUM, cq.MIN, cq.MAX], [2, 2, 2])
    assert (res.column(0)[0], res.column(1)[0], res.column(2)[0]) == (orc.ss_sum_amount(BIG), 0.0, 9999.0)
    assert t.count(col(2) > 9990) == orc.ss_count_amount_above(BIG, 9990)
    assert t.count(col(2) < 200) == 200 * (BIG // 10000)
    assert t.count(col(1) == 7) == orc.ss_count_region(BIG, 7)
    assert t.count((col(0) > 1000) & (col(0) <= 900_000_000)) == 900_000_000 - 1000
    for row in (0, 123_456_789, BIG - 1):
        n1, ids, _ = t.scan(col(3) == orc.ss_key_of_row(row))
        assert n1 == 1 and int(ids[0]) == row
    assert t.count(col(3) == orc.ss_key_of_row(BIG)) == 0  # a key of no row


def test_1b_group_by_region(scale_1b, orc):
    g = scale_1b.aggregate([cq.COUNT, cq.SUM], [None, 2], group=1)
    assert g.groups == 50
    for r in range(50):
        assert g.column(0)[r] == orc.ss_count_region(BIG, r) and g.column(1)[r] == orc.ss_region_amount_sum(BIG, r)


def test_1b_group_by_amount(scale_1b, orc):
    """The published 10k-group query (:1153-1156) and COUNT/SUM/MIN/MAX of
    a non-key column by amount: the rows of amount a are a + 10000 k."""
    t = scale_1b
    a = np.arange(10_000, dtype=np.float64)
    rows = np.array([orc.ss_amount_rows(BIG, int(x)) for x in a], dtype=np.float64)
    g = t.aggregate([cq.COUNT, cq.SUM], [None, 1], group=2)
    assert g.groups == 10_000 and np.array_equal(g.keys, a)
    assert np.array_equal(g.column(0), rows.astype(np.uint64))
    assert np.array_equal(g.column(1), rows * (a % 50))
    g = t.aggregate([cq.COUNT, cq.SUM, cq.MIN, cq.MAX], [None, 0, 0, 0], group=2)
    o = np.argsort(g.keys)
    assert np.array_equal(g.keys[o], a)
    # id = row + 1 over rows a + 10000 k, k < rows(a)
    assert np.array_equal(g.column(0)[o], rows.astype(np.uint64))
    assert np.array_equal(g.column(1)[o], rows * (a + 1) + 10_000 * rows * (rows - 1) / 2)
    assert np.array_equal(g.column(2)[o], a + 1)
    assert np.array_equal(g.column(3)[o], a + 1 + 10_000 * (rows - 1))


def test_1b_join_regions(scale_1b, orc):
    t = scale_1b
    regions = cq.Table(50)
    try:
        regions.add_column("region_id", cq.U32, np.arange(50, dtype=np.uint32))
        regions.add_column("zone", cq.U32, np.arange(50, dtype=np.uint32) // 10)
        j = cq.aggregate_join(t, 1, regions, 0, [cq.COUNT, cq.SUM], [None, 2], dim_where=col(1) == 2)
        assert int(j.cells[0, 0]) == sum(orc.ss_count_region(BIG, r) for r in range(20, 30))
        assert j.column(1)[0] == sum(orc.ss_region_amount_sum(BIG, r) for r in range(20, 30))
    finally:
        regions.close()


def test_1b_join_big_dim(scale_1b, orc):
    """big_dim (bench_select_scale.c:14-22): row j holds the needle_key of
    fact row 1000 j and weight j % 100.  Joined on needle_key and grouped
    by weight w: 10000 dim rows each, whose fact rows have amount
    (j % 10) * 1000 = (w % 10) * 1000.  1M dim keys: 256 partitions."""
    t = scale_1b
    j = np.arange(1_000_000, dtype=np.uint64)
    keys = ((j * (BIG // 1_000_000)) * KEY_MUL % (1 << 32)).astype(np.uint32)
    big = cq.Table(1_000_000)
    try:
        big.add_column("big_key", cq.U32, keys)
        big.add_column("weight", cq.U32, (j % 100).astype(np.uint32))
        funcs = [cq.COUNT, cq.SUM, cq.MIN, cq.MAX]
        res = cq.aggregate_join(t, 3, big, 0, funcs, [None, 2, 2, 2], dim_group=1)
        assert cq.AGG_PATHS[res.trace["path"]] == "join_part_group"
        w = np.arange(100, dtype=np.float64)
        amt = (w % 10) * 1000
        assert res.groups == 100 and np.array_equal(res.keys, w)
        assert np.array_equal(res.column(0), np.full(100, 10_000, dtype=np.uint64))
        assert np.array_equal(res.column(1), 10_000 * amt)
        assert np.array_equal(res.column(2), amt) and np.array_equal(res.column(3), amt)
        # COUNT(*), SUM(weight) of the reference's case (:1166-1168)
        assert int(res.column(0).sum()) == 1_000_000
        assert float((res.column(0) * w).sum()) == orc.ss_weight_sum(1_000_000)
    finally:
        big.close()
