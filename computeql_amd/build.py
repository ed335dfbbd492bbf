"""Build recipe of libcqlhip.so (gfx950) and of the CPU oracle.

    python -m computeql_amd.build          # builds computeql_amd/lib/libcqlhip.so
    python -m computeql_amd.build --oracle # also oracle/_build/liboracle.so

HIP sources are compiled with hipcc for gfx950 only; the host operator layer
is C11 compiled with gcc; the library links libamdhip64 and libhiprtc.
Objects go to computeql_amd/_build/, the library to computeql_amd/lib/.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libcqlhip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["qk_library.hip", "gpu_hip.cpp"]
C_SOURCES = ["qe_core.c", "qe_scan.c", "qe_aggregate.c", "qe_join.c", "qe_hjoin.c", "qe_gather.c", "qe_stream.c", "qe_batch.c", "qe_serve.c",
             "query_kernel.c"]

# packed-fp32 instructions stay off in every HIP build
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wno-unused-result", "-Wno-unused-value",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-Wall", "-Wno-unused-function", f"-I{ROCM}/include",
           "-D__HIP_PLATFORM_AMD__"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{out}")
    if out.strip() and os.environ.get("CQLHIP_BUILD_VERBOSE"):
        print(out)


def write_prelude() -> str:
    """qk_abi.h + qk_device_common.h as one C string literal for the hiprtc
    generator (hiprtc sees no include path, so the include is inlined)."""
    abi = open(os.path.join(CSRC, "qk_abi.h")).read()
    src = open(os.path.join(CSRC, "qk_device_common.h")).read()
    src = abi + "\n" + src.replace('#include "qk_abi.h"\n', "")
    lines = []
    for line in src.splitlines():
        esc = line.replace("\\", "\\\\").replace('"', '\\"')
        lines.append(f'"{esc}\\n"')
    path = os.path.join(BUILD, "qk_prelude.inc")
    text = "\n".join(lines) + "\n"
    if not os.path.exists(path) or open(path).read() != text:
        with open(path, "w") as f:
            f.write(text)
    return path


def build_library() -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    write_prelude()
    jobs = []
    for s in HIP_SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(s)[0] + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", obj])
    for s in C_SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(s)[0] + ".o")
        jobs.append(["gcc", *C_FLAGS, f"-I{BUILD}", "-c", os.path.join(CSRC, s), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        list(ex.map(_run, jobs))
    objs = [os.path.join(BUILD, os.path.splitext(s)[0] + ".o") for s in HIP_SOURCES + C_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, f"-L{ROCM}/lib", "-lhiprtc",
          "-lpthread", f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> str:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "_build", "liboracle.so")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--oracle", action="store_true")
    ap.add_argument("--clean", action="store_true")
    a = ap.parse_args(argv)
    if a.clean:
        shutil.rmtree(BUILD, ignore_errors=True)
    print(build_library())
    if a.oracle:
        print(build_oracle())
    return 0


if __name__ == "__main__":
    sys.exit(main())
