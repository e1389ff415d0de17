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


def device_functions(obj: str) -> tuple[list[str], list[str]]:
    """(kernels, other functions) in the gfx950 code object of a compiled
    object file: every kernel has a `<name>.kd` descriptor symbol; a function
    without one is a device function the compiler kept out of line."""
    import tempfile
    llvm = os.path.join(os.path.dirname(os.path.realpath(HIPCC)), "..", "llvm", "bin")
    with tempfile.TemporaryDirectory() as tmp:
        fb, elf = os.path.join(tmp, "fatbin"), os.path.join(tmp, "dev.elf")
        subprocess.run(["objcopy", f"--dump-section=.hip_fatbin={fb}", obj, os.path.join(tmp, "x.o")],
                       check=True, capture_output=True)
        subprocess.run([os.path.join(llvm, "clang-offload-bundler"), "--type=o", f"--input={fb}",
                        f"--output={elf}", f"--targets=hipv4-amdgcn-amd-amdhsa--{ARCH}", "--unbundle"],
                       check=True, capture_output=True)
        out = subprocess.run([os.path.join(llvm, "llvm-readelf"), "-s", "-W", elf], check=True,
                             capture_output=True, text=True).stdout
    funcs, descriptors = set(), set()
    for line in out.splitlines():
        f = line.split()
        if len(f) >= 8 and f[3] == "FUNC":
            funcs.add(f[7])
        elif len(f) >= 8 and f[3] == "OBJECT" and f[7].endswith(".kd"):
            descriptors.add(f[7][:-3])
    return sorted(funcs & descriptors), sorted(funcs - descriptors)


def check_no_device_calls() -> None:
    """No kernel of the library calls a device function: large numerics are
    force-inlined, because an out-of-line function larger than the branch
    range gets long-branch sequences that reuse its return-address registers
    (the cause of a memory-aperture fault this build has had).  Raises when
    a code object holds a non-kernel function."""
    bad = {}
    for src in LIB_SOURCES:
        if not src.endswith(".hip"):
            continue
        kernels, others = device_functions(os.path.join(OBJ_DIR, src + ".o"))
        if not kernels:
            raise RuntimeError(f"{src}: no kernels found in its gfx950 code object")
        if others:
            bad[src] = others
    if bad:
        raise RuntimeError(f"device functions kept out of line (kernels must inline them): {bad}")


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
    check_no_device_calls()
    build_cli()


if __name__ == "__main__":
    build_all(incremental="--incremental" in sys.argv[1:])
    sys.exit(0)
