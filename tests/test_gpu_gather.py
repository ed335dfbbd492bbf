"""GPU parity of row materialisation: qe_gather (values through row ids, 0
for a missing row; selection_gather.comp modes F32/F64) against numpy
indexing, also through the (probe, build) pairs of qe_hash_join at stride
2, and qe_scan_materialize (SELECT cols WHERE ...) against the oracle's
scan ids followed by indexing.  Bit-exact."""
import numpy as np
import pytest

import computeql_amd as cq
from computeql_amd.ir import col

pytestmark = pytest.mark.gpu

N = 200_003


@pytest.fixture(scope="module")
def tbl():
    r = np.random.default_rng(9)
    cols = {
        "u": r.integers(0, 2**32, N, dtype=np.uint64).astype(np.uint32),
        "i": r.integers(-2**31, 2**31, N).astype(np.int32),
        "f": r.normal(size=N),
        "g": r.integers(0, 2**64 - 1, N, dtype=np.uint64),
        "k": r.integers(0, 5000, N).astype(np.uint32),
    }
    t = cq.Table(N)
    for name, ct in (("u", cq.U32), ("i", cq.I32), ("f", cq.F64), ("g", cq.U64), ("k", cq.U32)):
        t.add_column(name, ct, cols[name])
    return t, cols


def take(a, rows):
    rows = np.asarray(rows, dtype=np.uint64)
    out = np.zeros(len(rows), dtype=a.dtype)
    ok = rows != np.uint64(2**64 - 1)
    out[ok] = a[rows[ok].astype(np.int64)]
    return out


def test_gather_every_width(tbl):
    t, cols = tbl
    r = np.random.default_rng(1)
    rows = r.integers(0, N, 70_001).astype(np.uint64)
    rows[::97] = 2**64 - 1
    got = t.gather(["u", "i", "f", "g", "k"], rows)
    for name, g in zip(["u", "i", "f", "g", "k"], got):
        np.testing.assert_array_equal(g.view(np.uint8), take(cols[name], rows).view(np.uint8))


def test_gather_through_join_pairs(orc, tbl):
    t, cols = tbl
    r = np.random.default_rng(2)
    dk = r.integers(0, 6000, 3000).astype(np.uint32)
    d = cq.Table(len(dk))
    d.add_column("key", cq.U32, dk)
    d.add_column("val", cq.F64, r.normal(size=len(dk)))
    dval = d.read("val")
    pairs = cq.hash_join(t, t.index("k"), d, 0, "left")
    np.testing.assert_array_equal(pairs, orc.hash_join(cols["k"], dk, "left"))
    probe_side = t.gather(["k", "f"], pairs, stride=2, offset=0)
    build_side = d.gather(["key", "val"], pairs, stride=2, offset=1)
    np.testing.assert_array_equal(probe_side[0], take(cols["k"], pairs[:, 0]))
    np.testing.assert_array_equal(probe_side[1], take(cols["f"], pairs[:, 0]))
    np.testing.assert_array_equal(build_side[0], take(dk, pairs[:, 1]))
    np.testing.assert_array_equal(build_side[1], take(dval, pairs[:, 1]))
    matched = pairs[:, 1] != cq.BUILD_NULL
    assert np.array_equal(build_side[0][matched], probe_side[0][matched])


@pytest.mark.parametrize("where", [None, col(4) < 100, (col(4) < 50) | (col(1) > 2**30)])
def test_select_against_oracle(orc, tbl, where):
    t, cols = tbl
    ids = np.arange(N, dtype=np.uint64) if where is None else orc.scan(
        [cols["u"], cols["i"], cols["f"], cols["g"], cols["k"]], where)
    got = t.select(["g", "k", "f"], where)
    for name, g in zip(["g", "k", "f"], got):
        np.testing.assert_array_equal(g.view(np.uint8), cols[name][ids.astype(np.int64)].view(np.uint8))


def test_select_cap_and_bad_rows(tbl):
    t, cols = tbl
    got = t.select(["k"], col(4) < 10, cap=7)
    assert len(got[0]) == 7 and np.all(got[0] < 10)
    with pytest.raises(cq.QueryError):
        t.gather(["k"], np.array([N], dtype=np.uint64))
