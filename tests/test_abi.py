"""The C-ABI library loads and exports every function include/cqlhip.h
declares (no GPU work: nothing here calls into the device)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cqlhip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"typedef struct GPU_Backend \{.*?\} GPU_Backend;", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\([^;{]*\)\s*;", text, flags=re.M)
    return sorted(set(n for n in names if not n.startswith("(")))


def test_header_declares_backend_and_operators():
    names = declared_functions()
    assert "gpu_hip_backend" in names and "qe_scan_filter" in names and "qe_aggregate_join" in names
    assert len(names) > 60


def test_library_exports_every_declared_symbol():
    from computeql_amd import native

    if not os.path.exists(native.LIB_PATH):
        pytest.fail("libcqlhip.so is not built (python -m computeql_amd.build)")
    lib = ctypes.CDLL(native.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from computeql_amd import native

    bound = {n for n, _, _ in native.SIGNATURES}
    assert set(declared_functions()) <= bound, sorted(set(declared_functions()) - bound)
    native.load()  # binds every signature (no device call)


def test_backend_table_layout():
    """struct GPU_Backend keeps the reference's member order: 45 function
    pointers after struct_size (gpu_backend.h:17-71)."""
    text = open(HEADER).read()
    body = re.search(r"typedef struct GPU_Backend \{(.*?)\} GPU_Backend;", text, flags=re.S).group(1)
    members = re.findall(r"\(\*(\w+)\)", body)
    assert members[:3] == ["init", "release", "wait"]
    assert members[-2:] == ["scratch_arena", "name"]
    assert len(members) == 45


def test_ir_layout_matches_oracle():
    import numpy as np

    from computeql_amd import ir
    from oracle import oracle

    w = (ir.col(1) == 7) & ir.col(2).isin([1, 2])
    a = w.nodes()
    assert a.dtype.itemsize == 24 and a.dtype == oracle.IR_DTYPE
    assert list(a["child_count"]) == [2, 2, 0, 0, 2, 0, 2, 0, 0]
    assert np.array_equal(a["type"], [1, 1, 2, 3, 1, 2, 4, 3, 3])
