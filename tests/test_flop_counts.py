"""The adjoint benchmark's algorithmic flop figures come from counting the
operations of the Carter field and its vector-Jacobian product with a
counting scalar type (tests/support/flop_count.cpp)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_adjoint_flop_constants_match_the_op_count(tmp_path):
    exe = tmp_path / "flop_count"
    subprocess.run(["g++", "-std=c++17", "-O1", os.path.join(ROOT, "tests", "support", "flop_count.cpp"), "-o",
                    str(exe)], check=True)
    out = dict(line.split(None, 1) for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                              text=True).stdout.splitlines())
    rhs_flops, rhs_trans = (int(v) for v in out["rhs_r"].split())
    vjp_flops, vjp_trans = (int(v) for v in out["vjp_r"].split())
    sys.path.insert(0, ROOT)
    import bench
    assert (rhs_flops, vjp_flops) == (bench.ADJ_RHS, bench.ADJ_VJP)
    assert (rhs_trans, vjp_trans) == (2, 4)   # one sin/cos pair per field evaluation
    assert bench.ADJ_FWD_STEP == 4 * rhs_flops + 90
    assert bench.ADJ_BWD_STEP == bench.ADJ_FWD_STEP + 4 + 3 * rhs_flops + 4 * vjp_flops + 76
