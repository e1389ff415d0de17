"""Build libnullpath.so (gfx950) and the `render` CLI in-tree with hipcc.

Usage: python -m nullpath_amd.build [--incremental]   (or __graft_entry__.build())
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libnullpath.so")
BIN_DIR = os.path.join(PKG, "bin")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed FP32 instructions are kept out of generated code.
COMMON = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", f"-I{INCLUDE}", f"-I{CSRC}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument"]

LIB_SOURCES = ["np_frame.cpp", "kerr_render.hip", "geomu.hip", "geokerr.hip", "schw.hip", "adjoint.hip",
               "np_api.hip"]
OBJ_DIR = os.path.join(PKG, "_build")


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _jobs() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, int(os.environ.get("MAX_JOBS", "16")), len(LIB_SOURCES)))


def _newest_header() -> float:
    hdrs = [os.path.join(d, f) for d in (CSRC, INCLUDE) for f in os.listdir(d)
            if f.endswith((".h", ".hpp"))]
    return max(os.path.getmtime(h) for h in hdrs)


def build_lib(incremental: bool = False) -> str:
    """Every source compiled on its own (in parallel; no device code crosses
    translation units), then one link.  incremental=True skips objects newer
    than their source and every header."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    objs = [os.path.join(OBJ_DIR, s + ".o") for s in LIB_SOURCES]
    newest_hdr = _newest_header() if incremental else 0.0

    def compile_one(k: int) -> None:
        src = os.path.join(CSRC, LIB_SOURCES[k])
        if incremental and os.path.exists(objs[k]):
            t = os.path.getmtime(objs[k])
            if t > os.path.getmtime(src) and t > newest_hdr:
                return
        _run([HIPCC, *COMMON, "-x", "hip", "-c", src, "-o", objs[k]])

    with ThreadPoolExecutor(max_workers=_jobs()) as ex:
        list(ex.map(compile_one, range(len(LIB_SOURCES))))
    tmp = LIB + ".tmp"
    _run([HIPCC, *COMMON, "-shared", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_cli() -> str:
    """The `render` command line, linked against libnullpath.so next to it."""
    os.makedirs(BIN_DIR, exist_ok=True)
    exe = os.path.join(BIN_DIR, "render")
    tmp = exe + ".tmp"
    _run([HIPCC, *COMMON, "-x", "hip", os.path.join(CSRC, "render_main.cpp"), "-x", "none",
          f"-L{PKG}", "-lnullpath", "-Wl,-rpath,$ORIGIN/..", "-o", tmp])
    os.replace(tmp, exe)
    return exe


def build_all(incremental: bool = False) -> None:
    build_lib(incremental)
    build_cli()


if __name__ == "__main__":
    build_all(incremental="--incremental" in sys.argv[1:])
    sys.exit(0)
