"""Host-side marshalling of batches and WHERE trees (no GPU work): the
QE_BatchQuery records a batch hands to the C ABI point at the right
buffers, and a query's converted WHERE is built once and cannot be
changed afterwards."""
import ctypes as C

import numpy as np
import pytest

import computeql_amd as cq
from computeql_amd import engine, ir
from computeql_amd.ir import col
from computeql_amd.native import QE_BatchQuery


def test_record_layout_matches_the_struct():
    assert engine._BQ_DTYPE.itemsize == C.sizeof(QE_BatchQuery)
    for name, _ in QE_BatchQuery._fields_:
        assert engine._BQ_DTYPE.fields[name][1] == getattr(QE_BatchQuery, name).offset


def test_batch_records_point_at_their_outputs():
    specs = [cq.AggSpec([cq.COUNT], [None], col(0) < 5),
             cq.AggSpec([cq.COUNT, cq.SUM, cq.MAX], [None, 2, 2], (col(1) == 3) & (col(2) > 1), 1),
             cq.AggSpec([cq.SUM, cq.AVG], [4, 4], None, 2, cap=17),
             cq.AggSpec([cq.COUNT], [None], None, [1, 3, 2], cap=9)]
    recs, counts, outs = engine._batch_records(specs)
    for i, (sp, (funcs, keys, cells, cap)) in enumerate(zip(specs, outs)):
        r = recs[i]
        f, fa, aa, nodes, ptrs = sp.prepared()
        assert r["where"] == ptrs[0] and r["where_count"] == (len(nodes) if nodes is not None else 0)
        assert r["funcs"] == fa.ctypes.data and r["arg_cols"] == aa.ctypes.data and r["expr_count"] == len(funcs)
        gl = engine.group_list(sp.group)
        assert r["group_count"] == len(gl) and list(r["group_cols"][:len(gl)]) == gl
        assert r["cap"] == cap and keys.shape == (cap * max(len(gl), 1),) and cells.shape == (cap, len(funcs))
        assert r["out_keys"] == keys.ctypes.data and r["out_cells"] == cells.ctypes.data
        assert r["out_group_count"] == counts.ctypes.data + 8 * i
    assert outs[0][3] == 1 and outs[1][3] == 4096 and outs[2][3] == 17 and outs[3][3] == 9
    # the arena slices do not overlap
    spans = sorted((o[1].ctypes.data, o[2].ctypes.data + o[2].nbytes) for o in outs)
    assert all(a[1] <= b[0] for a, b in zip(spans, spans[1:]))


def test_where_nodes_are_built_once_and_read_only():
    w = (col(1) < 3) | col(2).isin([1, 2])
    a = w.nodes()
    assert a is w.nodes() and a is ir.to_nodes(w)
    with pytest.raises(ValueError):
        a[0]["value"] = 1.0
    sp = cq.AggSpec([cq.COUNT], [None], w)
    assert sp.prepared() is sp.prepared() and sp.prepared()[3] is a
