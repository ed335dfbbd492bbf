"""GPU parity of narrow column encoding (a3): u32/i32 columns kept as u8 or
u16 codes or recognised as sequences, read as codes by scan_simple, the
global aggregate and the dense GROUP BY.  Every result with QE_ENCODE on
equals the full-width result bit for bit and the CPU oracle's; the
encoding chosen follows the reference's rule (qe_column_encoding,
query_exec.c:1555-1589)."""
import numpy as np
import pytest

import computeql_amd as cq
from computeql_amd.ir import col

pytestmark = pytest.mark.gpu

N = 1_000_003


@pytest.fixture(autouse=True)
def _settings():
    cq.reset_settings()
    yield
    cq.reset_settings()


@pytest.fixture(scope="module")
def enc_table():
    r = np.random.default_rng(21)
    cols = [
        r.integers(0, 50, N).astype(np.uint32),                  # 0 region: u8
        r.integers(0, 10_000, N).astype(np.uint32),              # 1 amount: u16
        np.arange(N, dtype=np.uint32),                           # 2 id: sequence
        (np.arange(N) + 1000).astype(np.uint32),                 # 3 offset sequence
        r.integers(-100, 100, N).astype(np.int32),               # 4 i32 small: u8
        r.integers(-30_000, 30_000, N).astype(np.int32),         # 5 i32 wide: u16
        r.integers(0, 2**32, N, dtype=np.uint64).astype(np.uint32),  # 6 full range: none
        (r.integers(0, 200, N) + 4_000_000_000).astype(np.uint32),   # 7 u8 near the top of u32
    ]
    t = cq.Table(N)
    for i, c in enumerate(cols):
        t.add_column(f"c{i}", cq.I32 if c.dtype == np.int32 else cq.U32, c)
    encs = [t.encode(i) for i in range(len(cols))]
    return t, cols, encs


def test_encoding_choice(enc_table):
    _, _, encs = enc_table
    assert encs == [1, 2, 3, 3, 1, 2, 0, 1]


def both(fn):
    cq.set_setting("QE_ENCODE", 1)
    a = fn()
    cq.set_setting("QE_ENCODE", 0)
    b = fn()
    cq.set_setting("QE_ENCODE", 1)
    return a, b


WHERES = [
    col(0) < 7,
    (col(1) >= 123) & (col(1) <= 4567),
    col(2) == 777_777,
    (col(3) > 5000) & (col(3) < 200_000),
    (col(4) < -3) & (col(5) >= 1000),
    (col(5) == -12_345),
    col(7) > 4_000_000_150,
    (col(0) == 3.5),
    (col(4) > -1000.5) & (col(1) < 10_000.5),
]


@pytest.mark.parametrize("wi", range(len(WHERES)))
def test_scan_encoded_equals_full_and_oracle(orc, enc_table, wi):
    t, cols, _ = enc_table
    w = WHERES[wi]
    (na, ia, tra), (nb, ib, _) = both(lambda: t.scan(w))
    assert tra["kernel"] == "scan_simple"
    assert na == nb and np.array_equal(ia, ib)
    assert np.array_equal(ia, orc.scan(cols, w))


@pytest.mark.parametrize("wi", [None, 0, 1, 4, 6])
def test_global_aggregate_encoded(orc, enc_table, wi):
    t, cols, _ = enc_table
    w = None if wi is None else WHERES[wi]
    for funcs, args in (([cq.COUNT, cq.SUM, cq.MIN, cq.MAX, cq.SUM, cq.MIN, cq.MAX, cq.SUM],
                         [None, 1, 1, 1, 4, 4, 5, 3]),
                        ([cq.SUM, cq.MIN, cq.MAX, cq.SUM, cq.MAX, cq.AVG], [7, 7, 2, 2, 0, 5])):
        a, b = both(lambda: t.aggregate(funcs, args, where=w))
        assert np.array_equal(a.cells, b.cells)
        rk, rv = orc.group_aggregate(cols, None, funcs, args, where=w)
        for e in range(len(funcs)):
            assert float(a.column(e)[0]) == float(rv[0, e]), (funcs, e)


@pytest.mark.parametrize("key,wi", [(0, None), (4, None), (0, 1), (5, 0), (1, 4)])
def test_dense_group_by_encoded(orc, enc_table, key, wi):
    t, cols, _ = enc_table
    w = None if wi is None else WHERES[wi]
    funcs = [cq.COUNT, cq.SUM, cq.MIN, cq.MAX, cq.SUM, cq.MIN]
    args = [None, 1, 1, 4, 5, 7]
    a, b = both(lambda: t.aggregate(funcs, args, where=w, group=key))
    oa, ob = np.argsort(a.keys, kind="stable"), np.argsort(b.keys, kind="stable")
    assert np.array_equal(a.keys[oa], b.keys[ob]) and np.array_equal(a.cells[oa], b.cells[ob])
    rk, rv = orc.group_aggregate(cols, key, funcs, args, where=w)
    o, og = np.argsort(rk, kind="stable"), np.argsort(a.keys, kind="stable")   # the hash path is in slot order
    assert np.array_equal(a.keys[og], rk[o])
    for e in range(len(funcs)):
        assert np.array_equal(a.column(e).astype(np.float64)[og], rv[o, e]), e


def test_rewrite_drops_encoding(orc):
    t = cq.Table(10_000)
    t.add_column("a", cq.U32, np.arange(10_000, dtype=np.uint32) % 100)
    assert t.encode(0) == 1
    t.gen_events(5, region=0)     # rewrites column 0 (region = d(3r) % 50)
    new = t.read(0)
    n, ids, _ = t.scan(col(0) < 10)
    assert np.array_equal(ids, np.nonzero(new < 10)[0])
    assert t.encode(0) == 1
