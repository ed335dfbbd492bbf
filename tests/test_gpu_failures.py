"""A device-to-host read that fails must fail the query, never hand back
zero-filled or stale results (the backend's buffer_read zero-fills, as
gpu_vulkan.c:3107-3111 does; every operator copies through the checked
reads instead).  gpu_hip_fault_read_at(k) makes the k-th read from now
report failure, host side, with no device fault."""
import numpy as np
import pytest

import computeql_amd as cq
from computeql_amd.ir import col

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def events(orc):
    n = 1_000_003
    t = cq.Table(n)
    for name in ("id", "region", "amount", "ts", "fk"):
        t.add_column(name, cq.U32)
    t.gen_events(orc.SEED, id=0, region=1, amount=2, ts=3)
    dk, seg = orc.gen_dim(50_000, 100, seed=4)
    dim = cq.Table(len(dk))
    dim.add_column("key", cq.U32, dk)
    dim.add_column("segment", cq.U32, seg)
    t.gen_fact_fk(4, orc.SEED, dim, 0)
    yield t, dim
    t.close()
    dim.close()


def _rows(res):
    """(key values, cells) per group, sorted: hash groups come out in
    table-slot order, which may differ from run to run"""
    k = np.asarray(res.keys, dtype=np.float64).reshape(res.groups, -1).view(np.uint64)
    r = np.concatenate([k, res.cells], axis=1)
    return r[np.lexsort(r.T[::-1])]


OPS = {
    "global": lambda t, d: t.aggregate([cq.SUM, cq.MAX], [2, 2]).cells,
    "dense": lambda t, d: t.aggregate([cq.COUNT, cq.SUM], [None, 2], group=1).cells,
    "hash": lambda t, d: _rows(t.aggregate([cq.COUNT, cq.SUM], [None, 2], group=3, where=col(2) < 50)),
    "multi_key": lambda t, d: _rows(t.aggregate([cq.COUNT], [None], group=[1, 2])),
    "scan_ids": lambda t, d: t.scan(col(2) < 100)[1],
    "column_read": lambda t, d: t.read(2, 10, 1000),
    "join": lambda t, d: cq.aggregate_join(t, 4, d, 0, [cq.COUNT, cq.SUM], [None, 2], dim_group=1).cells,
    "batch": lambda t, d: np.concatenate([r.cells.reshape(-1) for r in cq.aggregate_batch(
        t, [cq.AggSpec([cq.COUNT], [None], col(1) < 9), cq.AggSpec([cq.SUM], [2], None, 1)])]),
}


@pytest.mark.parametrize("op", list(OPS))
def test_failed_read_fails_the_query(events, op):
    t, dim = events
    lib = cq.load()
    good = OPS[op](t, dim)
    lib.gpu_hip_fault_read_at(1)
    try:
        with pytest.raises(cq.QueryError):
            OPS[op](t, dim)
    finally:
        lib.gpu_hip_fault_read_at(0)
    again = OPS[op](t, dim)
    assert np.array_equal(again, good)


def test_checked_reads_report_failure(events):
    t, _ = events
    lib = cq.load()
    buf = lib.qe_table_column_buffer(t.handle, 2)
    out = np.empty(16, dtype=np.uint32)
    assert lib.gpu_hip_buffer_read_checked(buf, out.ctypes.data, 64) == 1
    lib.gpu_hip_fault_read_at(1)
    assert lib.gpu_hip_buffer_read_checked(buf, out.ctypes.data, 64) == 0
    assert b"injected" in lib.gpu_hip_last_error()
    # a read past the buffer is refused, not clipped
    assert lib.gpu_hip_buffer_read_checked(buf, None, 1 << 40) == 0


def test_dispatch_beyond_the_lds_limit_is_refused():
    """A kernel's own __shared__ bytes count against a workgroup's LDS beside
    the dynamic bytes the host asks for.  The shared scan keeps 8 static
    bytes (its query counters), so its dynamic stage may take that much less
    than gpu_hip_max_lds_bytes; a dispatch asking for more fails with an
    error on the host and never reaches the queue (there it would abort the
    queue with an invalid-allocation error)."""
    import ctypes as C

    from computeql_amd.native import String8

    lib = cq.load()
    name = C.create_string_buffer(b"shared_scan_1024_1")
    k = lib.gpu_hip_kernel_alloc(String8(C.cast(name, C.c_void_p), len(name.value)))
    assert k
    mx = lib.gpu_hip_max_lds_bytes()
    lim = lib.gpu_hip_kernel_max_lds_bytes(k)
    assert 0 < lim < mx
    try:
        lib.gpu_hip_kernel_set_lds_bytes(k, lim + 1)
        b = lib.gpu_hip_batch_begin(0, 0)
        lib.gpu_hip_batch_kernel_execute(b, k, 1024, 1024)
        assert lib.gpu_hip_batch_end(b) == 0
        assert b"exceed" in lib.gpu_hip_last_error()
    finally:
        lib.gpu_hip_kernel_set_lds_bytes(k, 0)
        lib.gpu_hip_kernel_release(k)
