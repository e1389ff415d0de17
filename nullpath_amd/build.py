"""Build libnullpath.so (gfx950) and the `render` CLI in-tree with hipcc.

Usage: python -m nullpath_amd.build   (or __graft_entry__.build())
"""
from __future__ import annotations

import os
import subprocess
import sys

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


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib() -> str:
    srcs = [os.path.join(CSRC, s) for s in LIB_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, *COMMON, "-shared", "-x", "hip", *srcs, "-o", tmp])
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


def build_all() -> None:
    build_lib()
    build_cli()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
