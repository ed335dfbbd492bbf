"""GPU parity of the row-pair join (qe_hash_join) against the CPU oracle
(orc_hash_join): pairs and positions bit-exact and in the same order (probe
position, then build position).  Cases: the four join types, duplicate
build keys (groups sorted by one thread up to 32 rows and by a block
beyond), row lists on both sides, mixed key types, f64 key semantics,
empty sides, truncation at the caller's capacity, and at 20M probe rows the
reference's known answer (every probe row p pairs with build row p % n)
through the order-free pair checksum of its tests."""
import ctypes as C

import numpy as np
import pytest

import computeql_amd as cq
from computeql_amd.native import QE_Trace

from test_hash_join_oracle import pair_checksum

pytestmark = pytest.mark.gpu

TYPES = {np.uint32: cq.U32, np.int32: cq.I32, np.float64: cq.F64, np.uint64: cq.U64, np.int64: cq.I64,
         np.float32: cq.F32}


def table(*cols):
    t = cq.Table(len(cols[0]))
    for i, c in enumerate(cols):
        t.add_column(f"c{i}", TYPES[c.dtype.type], c)
    return t


@pytest.mark.parametrize("jt", ["inner", "left", "semi", "anti"])
@pytest.mark.parametrize("seed", [1, 2])
def test_join_types_against_oracle(orc, jt, seed):
    r = np.random.default_rng(seed)
    pk = r.integers(0, 30_000, 200_003).astype(np.uint32)
    bk = r.integers(0, 40_000, 50_001).astype(np.uint32)
    got = cq.hash_join(table(pk), 0, table(bk), 0, jt)
    np.testing.assert_array_equal(got, orc.hash_join(pk, bk, jt))


@pytest.mark.parametrize("jt", ["inner", "left", "semi", "anti"])
def test_row_lists_and_mixed_types(orc, jt):
    r = np.random.default_rng(11)
    pk = r.integers(-500, 500, 60_000).astype(np.int32)
    bk = r.integers(-600, 600, 9_000).astype(np.float64)
    prow = r.permutation(60_000)[:40_000].astype(np.uint64)
    brow = r.permutation(9_000)[:5_000].astype(np.uint64)
    got = cq.hash_join(table(pk), 0, table(bk), 0, jt, probe_rows=prow, build_rows=brow)
    np.testing.assert_array_equal(got, orc.hash_join(pk, bk, jt, probe_rows=prow, build_rows=brow))


def test_large_duplicate_groups(orc):
    """Key 7 holds 5000 build rows and key 9 holds 100 (block-sorted groups),
    keys 0..999 two or three rows each (thread-sorted groups)."""
    r = np.random.default_rng(5)
    bk = np.concatenate([np.full(5000, 7), np.full(100, 9), r.integers(0, 1000, 2500)]).astype(np.uint32)
    bk = bk[r.permutation(len(bk))].copy()
    pk = r.integers(0, 1200, 3000).astype(np.uint32)
    pk[::100] = 7
    got = cq.hash_join(table(pk), 0, table(bk), 0, "inner")
    want = orc.hash_join(pk, bk, "inner")
    assert len(want) > 150_000
    np.testing.assert_array_equal(got, want)


def test_f64_key_semantics(orc):
    pk = np.array([0.0, -0.0, np.nan, 1.5, 2.0 ** 53, 3.25], dtype=np.float64)
    bk = np.array([-0.0, np.nan, 1.5, 0.0, 3.25, 3.25], dtype=np.float64)
    for jt in ("inner", "left", "semi", "anti"):
        np.testing.assert_array_equal(cq.hash_join(table(pk), 0, table(bk), 0, jt), orc.hash_join(pk, bk, jt))
    big = np.array([2 ** 53 + 1, 5], dtype=np.uint64)
    np.testing.assert_array_equal(cq.hash_join(table(big), 0, table(np.array([2.0 ** 53, 1.0])), 0, "inner"),
                                  [[0, 0]])


def test_empty_sides(orc):
    k = np.array([1, 2], dtype=np.uint32)
    e_rows = np.array([], dtype=np.uint64)
    assert cq.hash_join(table(k), 0, table(k), 0, "inner", probe_rows=e_rows).shape == (0, 2)
    assert cq.hash_join(table(k), 0, table(k), 0, "inner", build_rows=e_rows).shape == (0, 2)
    assert cq.hash_join(table(k), 0, table(k), 0, "left", build_rows=e_rows).tolist() == [
        [0, cq.BUILD_NULL], [1, cq.BUILD_NULL]]
    assert cq.hash_join(table(k), 0, table(k), 0, "anti", build_rows=e_rows).tolist() == [0, 1]


def test_capacity_truncates_and_reports_total(orc):
    r = np.random.default_rng(3)
    pk = r.integers(0, 100, 10_000).astype(np.uint32)
    bk = r.integers(0, 100, 300).astype(np.uint32)
    want = orc.hash_join(pk, bk, "inner")
    lib = cq.load()
    tp, tb = table(pk), table(bk)
    cap = 1234
    out = np.empty(cap * 2, dtype=np.uint64)
    n = C.c_uint64(0)
    assert lib.qe_hash_join(tp.handle, 0, None, 0, tb.handle, 0, None, 0, 0, out.ctypes.data, cap, C.byref(n),
                            C.byref(QE_Trace()))
    assert n.value == len(want)
    np.testing.assert_array_equal(out.reshape(cap, 2), want[:cap])


def test_non_numeric_or_bad_rows_fail_loudly():
    k = np.array([1, 2], dtype=np.uint32)
    with pytest.raises(cq.QueryError):
        cq.hash_join(table(k), 0, table(k), 0, "inner", probe_rows=np.array([5], dtype=np.uint64))


def test_calibrate_known_answer_20m():
    """bench_calibrate.c cal_measure_join at 20M probe rows over 1M build
    rows: probe row p pairs with build row p % n, exactly 20M pairs; checked
    through the reference's order-free checksum and in full."""
    nb, np_ = 1_000_000, 20_000_000
    bk = np.arange(nb, dtype=np.uint32)
    pk = (np.arange(np_, dtype=np.uint64) % nb).astype(np.uint32)
    got = cq.hash_join(table(pk), 0, table(bk), 0, "inner")
    assert got.shape == (np_, 2)
    p = np.arange(np_, dtype=np.uint64)
    assert pair_checksum(got) == pair_checksum(np.stack([p, p % nb], axis=1))
    assert np.array_equal(got[:, 0], p) and np.array_equal(got[:, 1], p % nb)
