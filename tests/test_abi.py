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


STRUCTS = {
    "QE_BatchQuery": "QE_BatchQuery",
    "QE_JoinQuery": "QE_JoinQuery",
    "QE_ScanQuery": "QE_ScanQuery",
    "QE_ServeRequest": "QE_ServeRequest",
    "QE_ServerConfig": "QE_ServerConfig",
    "QE_ServerStats": "QE_ServerStats",
}


def test_struct_layouts_match_ctypes(tmp_path):
    """Every field offset and struct size the Python binding assumes is the
    one a C compiler gives the header's structs."""
    import shutil
    import subprocess

    from computeql_amd import native

    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("no C compiler")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "cqlhip.h"', "int main(void) {"]
    for c_name, py_name in STRUCTS.items():
        cls = getattr(native, py_name)
        lines.append(f'  printf("{py_name} size %zu\\n", sizeof({c_name}));')
        for f, _ in cls._fields_:
            lines.append(f'  printf("{py_name} {f} %zu\\n", offsetof({c_name}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            s, f, v = line.split()
            got[(s, f)] = int(v)
    for py_name in STRUCTS.values():
        cls = getattr(native, py_name)
        assert got[(py_name, "size")] == ctypes.sizeof(cls), py_name
        for f, _ in cls._fields_:
            assert got[(py_name, f)] == getattr(cls, f).offset, (py_name, f)
