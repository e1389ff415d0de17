"""The library's own sin/cos in FP64 and float (Cody-Waite reduction +
Taylor polynomials, nullpath_amd/csrc/np_math.hpp), built for the host,
against the C library's sin and cos (correctly rounded to within an ulp) and
against exact special values."""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    out = tmp_path_factory.mktemp("npmath") / "libmath_probe.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-ffp-contract=off",
                    os.path.join(HERE, "support", "math_probe.cpp"), "-o", str(out)], check=True)
    L = C.CDLL(str(out))
    L.probe_sincos.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    L.probe_sincos_f.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    return L


def _sc(probe, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    s, c = np.zeros_like(x), np.zeros_like(x)
    probe.probe_sincos(x.ctypes.data, len(x), s.ctypes.data, c.ctypes.data)
    return s, c


def test_sincos_against_libm(probe):
    rng = np.random.default_rng(81)
    x = np.concatenate([rng.uniform(0.0, math.pi, 200000), rng.uniform(-20.0, 20.0, 50000),
                        np.linspace(-1e-3, 1e-3, 1001)])
    s, c = _sc(probe, x)
    # 2 ulp of the result's magnitude, and a floor of 2 ulp of 1 near zeros
    tol_s = 2.0 * np.spacing(np.maximum(np.abs(np.sin(x)), 0.0)) + 2.0 * np.spacing(1.0) * np.abs(x) / 8
    tol_c = 2.0 * np.spacing(np.maximum(np.abs(np.cos(x)), 0.0)) + 2.0 * np.spacing(1.0) * np.abs(x) / 8
    assert np.all(np.abs(s - np.sin(x)) <= tol_s + 1e-300)
    assert np.all(np.abs(c - np.cos(x)) <= tol_c + 1e-300)


def test_sincos_special_values(probe):
    s, c = _sc(probe, [0.0, math.pi / 2, math.pi, -math.pi / 2])
    assert s[0] == 0.0 and c[0] == 1.0
    assert abs(s[1] - 1.0) <= 2.2e-16 and abs(c[1]) < 1e-16
    assert abs(s[2]) < 2e-16 and abs(c[2] + 1.0) <= 2.2e-16
    assert abs(s[3] + 1.0) <= 2.2e-16
    # sin^2 + cos^2 = 1 to rounding
    x = np.linspace(0, math.pi, 10001)
    s, c = _sc(probe, x)
    assert np.max(np.abs(s * s + c * c - 1.0)) < 4.5e-16


def _sc_f(probe, x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    s, c = np.zeros_like(x), np.zeros_like(x)
    probe.probe_sincos_f(x.ctypes.data, len(x), s.ctypes.data, c.ctypes.data)
    return s, c


def test_sincos_f_against_libm(probe):
    """Float version over the polar angles the Carter field sees and a wider
    range: within 2 float ulp of the exact value of the float input (plus 2
    ulp of 1 scaled by |x|/8 near the zeros, as for FP64)."""
    rng = np.random.default_rng(83)
    x = np.concatenate([rng.uniform(0.0, math.pi, 200000), rng.uniform(-20.0, 20.0, 50000),
                        np.linspace(-1e-3, 1e-3, 1001), [1e-6, math.pi / 2, math.pi - 1e-6]]).astype(np.float32)
    s, c = _sc_f(probe, x)
    xd = x.astype(np.float64)
    eps1 = float(np.spacing(np.float32(1.0)))
    tol_s = 2.0 * np.spacing(np.abs(np.sin(xd)).astype(np.float32)).astype(np.float64) + 2.0 * eps1 * np.abs(xd) / 8
    tol_c = 2.0 * np.spacing(np.abs(np.cos(xd)).astype(np.float32)).astype(np.float64) + 2.0 * eps1 * np.abs(xd) / 8
    es = np.abs(s.astype(np.float64) - np.sin(xd))
    ec = np.abs(c.astype(np.float64) - np.cos(xd))
    assert np.all(es <= tol_s), (x[np.argmax(es - tol_s)], es.max())
    assert np.all(ec <= tol_c), (x[np.argmax(ec - tol_c)], ec.max())
    # exact at zero; the pole-side sine keeps its relative accuracy
    s0, c0 = _sc_f(probe, [0.0, 1e-6, np.float32(math.pi) - np.float32(1e-6)])
    assert s0[0] == 0.0 and c0[0] == 1.0
    assert abs(s0[1] / 1e-6 - 1.0) < 1e-6
    xp = float(np.float32(math.pi) - np.float32(1e-6))
    assert abs(float(s0[2]) / math.sin(xp) - 1.0) < 1e-6

