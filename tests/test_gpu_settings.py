"""The reference's settings.cfg switches for this path, read under their
own names with the reference's meaning (query_exec.c line cited per
case); each switched path must give the default path's results."""
import numpy as np
import pytest

import computeql_amd as cq
from computeql_amd.ir import col

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _reset():
    cq.reset_settings()
    yield
    cq.reset_settings()


@pytest.fixture(scope="module")
def events(orc):
    n = 2_000_003
    t = cq.Table(n)
    for name in ("id", "region", "amount", "ts", "fk"):
        t.add_column(name, cq.U32)
    t.gen_events(orc.SEED, id=0, region=1, amount=2, ts=3)
    dk, seg = orc.gen_dim(100_000, 100, seed=4)
    dim = cq.Table(len(dk))
    dim.add_column("key", cq.U32, dk)
    dim.add_column("segment", cq.U32, seg)
    t.gen_fact_fk(4, orc.SEED, dim, 0)
    host = orc.gen_events(n, seed=orc.SEED)
    h = cq.Table(n)
    for i, name in enumerate(("id", "region", "amount", "ts")):
        h.add_host_column(name, cq.U32, host[i])
    yield t, dim, h, host
    t.close()
    dim.close()
    h.close()


def _sorted(res):
    o = np.argsort(res.keys, kind="stable")
    return res.keys[o], res.cells[o]


def test_fuse_aggregate_and_group_by_off(events):
    """QE_FUSE_AGGREGATE (:9892) / QE_FUSE_GROUP_BY (:9447) 0: the WHERE
    runs as its own scan feeding the aggregate."""
    t, _, _, _ = events
    w = (col(2) < 300) & (col(1) >= 10)
    g0 = t.aggregate([cq.COUNT, cq.SUM, cq.MAX], [None, 3, 3], where=w)
    k0 = t.aggregate([cq.COUNT, cq.SUM], [None, 3], where=w, group=1)
    assert cq.AGG_PATHS[g0.trace["path"]] == "fused_global" and cq.AGG_PATHS[k0.trace["path"]] == "fused_dense"
    cq.set_setting("QE_FUSE_AGGREGATE", 0)
    g1 = t.aggregate([cq.COUNT, cq.SUM, cq.MAX], [None, 3, 3], where=w)
    k1 = t.aggregate([cq.COUNT, cq.SUM], [None, 3], where=w, group=1)
    assert cq.AGG_PATHS[g1.trace["path"]] == "hash" and cq.AGG_PATHS[k1.trace["path"]] == "fused_dense"
    cq.set_setting("QE_FUSE_GROUP_BY", 0)
    k2 = t.aggregate([cq.COUNT, cq.SUM], [None, 3], where=w, group=1)
    assert cq.AGG_PATHS[k2.trace["path"]] == "hash"
    assert np.array_equal(g0.cells, g1.cells)
    for a, b in zip(_sorted(k0), _sorted(k2)):
        assert np.array_equal(a, b)


def test_generate_join_off_takes_generic_instances(events):
    """QE_GENERATE_JOIN (:8235) 0: the generic probe in place of the
    shape-specialised instance."""
    t, dim, _, _ = events
    a = cq.aggregate_join(t, 4, dim, 0, [cq.COUNT, cq.SUM], [None, 2], dim_group=1)
    cq.set_setting("QE_GENERATE_JOIN", 0)
    b = cq.aggregate_join(t, 4, dim, 0, [cq.COUNT, cq.SUM], [None, 2], dim_group=1)
    assert a.trace["kernel"] == "join_part_probe_pk_1s_512" and b.trace["kernel"] == "join_part_probe_512"
    assert np.array_equal(a.keys, b.keys) and np.array_equal(a.cells, b.cells)


def test_route_conjunctions_defers_the_first_compile(events, orc):
    """QE_ROUTE_CONJUNCTIONS (query_kernel.c:1735-1760) 1: a new WHERE
    shape runs once on the interpreter, and is generated when it comes back
    (with other literals: the shape, not the literal, keys the kernel)."""
    t, _, _, host = events
    cq.set_setting("QE_ROUTE_CONJUNCTIONS", 1)
    shape = lambda a, b: ((col(1) == a) | (col(3) > b)) & (col(2) != a + 17)
    n1, ids1, tr1 = t.scan(shape(3, 4_200_000_000))
    n2, ids2, tr2 = t.scan(shape(5, 4_100_000_000))
    assert cq.SCAN_PATHS[tr1["path"]] == "interpreter" and cq.SCAN_PATHS[tr2["path"]] == "generated"
    assert np.array_equal(ids1, orc.scan(host, shape(3, 4_200_000_000)))
    assert np.array_equal(ids2, orc.scan(host, shape(5, 4_100_000_000)))
    cq.set_setting("QE_ROUTE_CONJUNCTIONS", 0)
    n3, _, tr3 = t.scan(shape(7, 1))
    assert cq.SCAN_PATHS[tr3["path"]] == "generated"


def test_scan_and_agg_stream_off_read_in_place(events, orc):
    """QE_SCAN_STREAM (:2798) / QE_AGG_STREAM (:10379) 0: a host-backed
    table is read in place over the host link rather than streamed in
    chunks; the join pushdown needs QE_AGG_STREAM (:8446)."""
    t, dim, h, host = events
    w = (col(2) < 500) & (col(1) == 7)
    n0, ids0, _ = h.scan(w)
    st0 = cq.stream_stats()
    assert st0["chunks"] >= 1
    g0 = h.aggregate([cq.COUNT, cq.SUM], [None, 3], where=col(2) < 900, group=1)
    st1 = cq.stream_stats()
    cq.set_setting("QE_SCAN_STREAM", 0)
    n1, ids1, _ = h.scan(w)
    cq.set_setting("QE_AGG_STREAM", 0)
    g1 = h.aggregate([cq.COUNT, cq.SUM], [None, 3], where=col(2) < 900, group=1)
    assert cq.stream_stats() == st1  # neither streamed
    assert np.array_equal(ids0, ids1) and np.array_equal(ids1, orc.scan(host, w))
    for a, b in zip(_sorted(g0), _sorted(g1)):
        assert np.array_equal(a, b)
    with pytest.raises(cq.QueryError, match="QE_AGG_STREAM"):
        cq.aggregate_join(t, 4, dim, 0, [cq.COUNT], [None])


@pytest.mark.parametrize("slots", [2, 3, 8])
def test_stream_slots_and_chunk_rows(events, orc, slots):
    """QE_SCAN_STREAM_SLOTS (:2253, clamped to 2..8) chunks in flight,
    QE_SCAN_STREAM_CHUNK_ROWS (:2096-2099) rows a chunk, and
    GPU_MAX_BUFFER_SIZE (:2740-2756) bytes a chunk buffer at most."""
    _, _, h, host = events
    n = len(host[0])
    cq.set_setting("QE_RESIDENT", 0)  # stream every chunk
    cq.set_setting("QE_SCAN_STREAM_SLOTS", slots)
    cq.set_setting("QE_SCAN_STREAM_CHUNK_ROWS", 1 << 19)
    w = (col(2) > 9000) | (col(1) == 3)
    cnt, ids, _ = h.scan(w)
    assert np.array_equal(ids, orc.scan(host, w))
    assert cq.stream_stats()["chunks"] == -(-n // (1 << 19))
    cq.set_setting("GPU_MAX_BUFFER_SIZE", 1 << 20)  # 256K u32 rows
    g = h.aggregate([cq.COUNT, cq.SUM, cq.MIN], [None, 3, 3], group=1)
    assert cq.stream_stats()["chunks"] == -(-n // (1 << 18))
    rk, rv = orc.group_aggregate(host, 1, [cq.COUNT, cq.SUM, cq.MIN], [None, 3, 3])
    o = np.argsort(rk)
    assert np.array_equal(g.keys, rk[o]) and np.array_equal(g.column(1), rv[o, 1])
