"""The narrow dense GROUP BY (aggregate_dense_narrow_*): large key ranges
with one U32 argument in 16 bytes a group, counts and the sum's carries
sharing one LDS word and moved to the global totals as they fill.
Against the oracle, and against exact integer arithmetic on a skew that
makes one group's count and carries overflow their 16 bits many times in
every block."""
import numpy as np
import pytest

import computeql_amd as cq
from computeql_amd.ir import col

pytestmark = pytest.mark.gpu


def _exact_groups(keys, vals, ngroups):
    """per group: count, exact sum (16-bit halves summed in f64, exact
    below 2^53 per half), min, max"""
    k = keys.astype(np.int64)
    v = vals.astype(np.uint64)
    cnt = np.bincount(k, minlength=ngroups)
    lo = np.bincount(k, weights=(v & 0xFFFF).astype(np.float64), minlength=ngroups)
    hi = np.bincount(k, weights=(v >> 16).astype(np.float64), minlength=ngroups)
    mn = np.full(ngroups, 2**32, dtype=np.uint64)
    mx = np.zeros(ngroups, dtype=np.uint64)
    np.minimum.at(mn, k, v)
    np.maximum.at(mx, k, v)
    sums = [int(h) * 65536 + int(l) for h, l in zip(hi, lo)]
    return cnt, sums, mn, mx


@pytest.fixture(scope="module")
def wide():
    n = 20_000_011
    r = np.random.default_rng(31)
    cols = [r.integers(0, 10_000, n).astype(np.uint32),                         # key: 10k groups
            r.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32),           # full-range u32
            r.integers(0, 100, n).astype(np.uint32), r.integers(0, 100, n).astype(np.uint32),
            r.integers(0, 50, n).astype(np.uint32)]
    t = cq.Table(n)
    for i, c in enumerate(cols):
        t.add_column(f"c{i}", cq.U32, c)
    yield t, cols
    t.close()


@pytest.mark.parametrize("funcs", [[cq.COUNT, cq.SUM, cq.MIN, cq.MAX], [cq.COUNT, cq.SUM], [cq.MIN, cq.MAX, cq.COUNT]])
def test_10k_groups_against_oracle(wide, orc, funcs):
    t, cols = wide
    args = [None if f == cq.COUNT else 1 for f in funcs]
    res = t.aggregate(funcs, args, group=0)
    assert res.trace["kernel"] == "aggregate_dense_narrow" and cq.AGG_PATHS[res.trace["path"]] == "dense"
    rk, rv = orc.group_aggregate(cols, 0, funcs, args)
    o = np.argsort(rk)
    assert np.array_equal(res.keys, rk[o])
    for e in range(len(funcs)):
        assert np.array_equal(res.column(e).astype(np.float64), rv[o, e]), e


def test_multi_key_fused_and_encoded(wide, orc):
    t, cols = wide
    funcs = [cq.COUNT, cq.SUM, cq.MAX]
    w = col(4) < 25
    res = t.aggregate(funcs, [None, 1, 1], where=w, group=[2, 3])
    assert res.trace["kernel"] == "aggregate_dense_narrow"
    rk, rv = orc.group_aggregate(cols, [2, 3], funcs, [None, 1, 1], where=w)
    o = np.lexsort(rk.T[::-1])
    assert np.array_equal(res.keys, rk[o])
    for e in range(3):
        assert np.array_equal(res.column(e).astype(np.float64), rv[o, e])
    # narrow codes of a small-range argument (region-like, 1-byte codes)
    assert t.encode(4) == 1
    res = t.aggregate([cq.COUNT, cq.SUM, cq.MIN], [None, 4, 4], group=0)
    assert res.trace["kernel"] == "aggregate_dense_narrow"
    rk, rv = orc.group_aggregate(cols, 0, [cq.COUNT, cq.SUM, cq.MIN], [None, 4, 4])
    o = np.argsort(rk)
    for e in range(3):
        assert np.array_equal(res.column(e).astype(np.float64), rv[o, e])


def test_skew_overflows_the_16_bit_halves():
    """90% of 100M rows in one group with values near 2^32: in every block
    that group's count passes 0x8000 and its sum carries pass 0x8000 many
    times; the other groups stay small.  SUMs are past 2^53, so the check
    is off and each sum is the double nearest the exact integer."""
    n = 100_000_007
    r = np.random.default_rng(9)
    key = np.where(r.random(n) < 0.9, 0, r.integers(1, 10_000, n)).astype(np.uint32)
    key[:3] = 9_999  # the key range is 0..9999
    val = (0xFFFFFFFF - r.integers(0, 1 << 20, n)).astype(np.uint32)
    t = cq.Table(n)
    try:
        t.add_column("k", cq.U32, key)
        t.add_column("v", cq.U32, val)
        cq.set_setting("QE_WIDE_INT_CHECK", 0)
        res = t.aggregate([cq.COUNT, cq.SUM, cq.MIN, cq.MAX], [None, 1, 1, 1], group=0)
        assert res.trace["kernel"] == "aggregate_dense_narrow"
        cnt, sums, mn, mx = _exact_groups(key, val, 10_000)
        present = np.flatnonzero(cnt)
        assert np.array_equal(res.keys, present.astype(np.float64))
        assert np.array_equal(res.column(0), cnt[present].astype(np.uint64))
        assert res.column(1).tolist() == [float(sums[g]) for g in present]
        assert np.array_equal(res.column(2), mn[present].astype(np.float64))
        assert np.array_equal(res.column(3), mx[present].astype(np.float64))
        assert cnt[0] > 80_000_000 and sums[0] > 2**53
    finally:
        cq.reset_settings()
        t.close()


def test_narrow_off_takes_the_hash_path(wide, orc):
    t, cols = wide
    cq.set_setting("QE_AGG_DENSE_NARROW", 0)
    try:
        res = t.aggregate([cq.COUNT, cq.SUM, cq.MIN, cq.MAX], [None, 1, 1, 1], group=0)
    finally:
        cq.reset_settings()
    assert cq.AGG_PATHS[res.trace["path"]] == "hash"
    narrow = t.aggregate([cq.COUNT, cq.SUM, cq.MIN, cq.MAX], [None, 1, 1, 1], group=0)
    o = np.argsort(res.keys)
    assert np.array_equal(res.keys[o], narrow.keys) and np.array_equal(res.cells[o], narrow.cells)
