"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute call: this runs without a GPU)."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECL = re.compile(r"^\s*(?:int|int64_t|double|void|const char\*|GeodesicLookupTable\*)\s+(\w+)\s*\(", re.M)


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in DECL.finditer(text):
            if not text[m.start():m.end()].lstrip().startswith("static"):
                names.add(m.group(1))
    return names


# C++-linkage declarations and their Itanium-mangled names
CXX_SYMBOLS = {"precomputeGeodesics": "_Z19precomputeGeodesicsfi"}


def test_library_exports_declared_symbols():
    from nullpath_amd import _lib
    L = _lib.lib()
    names = declared_symbols()
    missing = [n for n in sorted(names) if not hasattr(L, CXX_SYMBOLS.get(n, n))]
    assert not missing, missing
    assert "precomputeGeodesics" in names and len(names) >= 11


def test_status_strings():
    from nullpath_amd import _lib
    L = _lib.lib()
    assert L.np_status_string(0) == b"ok"
    assert L.np_status_string(-1) == b"invalid argument"
