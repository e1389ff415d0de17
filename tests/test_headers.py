"""The drop-in headers compile the way the reference's sources include them
(host C++, g++): bh_kerr.h brings GeodesicStateD with bh_common.h and
bh_types.h, and the types keep the reference's member order and sizes."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROBE = r'''
#include <cstddef>
#include <cstdio>
#include "bh_kerr.h"
#include "bh_schwarzschild.h"
#include "geokerr_ref_api.h"
#include "nullpath.h"
static_assert(sizeof(GeodesicStateD) == 64, "8 doubles");
static_assert(offsetof(GeodesicStateD, r) == 8 && offsetof(GeodesicStateD, pphi) == 56, "order");
static_assert(sizeof(GeodesicState) == 32, "8 floats");
static_assert(offsetof(GeodesicState, pt) == 16, "order");
int main() {
    GeodesicStateD s{0.0, 10.0, 1.5, 0.0, -1.0, 0.5, 0.1, 2.0};
    BhKleinSum<double> k;
    k.add(s.r); k.add(s.pr);
    std::printf("%.1f %d\n", k.result(), (int)RayStatus::Captured);
    return 0;
}
'''


def test_bh_kerr_header_compiles_as_reference_sources_use_it(tmp_path):
    src = tmp_path / "probe.cpp"
    src.write_text(PROBE)
    exe = tmp_path / "probe"
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", f"-I{ROOT}/include", str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert float(out[0]) == 10.5
