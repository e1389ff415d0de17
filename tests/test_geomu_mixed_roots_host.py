"""The quartic root finders of np_geokerr.hpp built for the host
(tests/support/roots_probe.cpp): the FP64 ZROOTS against the oracle's
ZROOTS (geokerr_original.f:3343-3441, 3527-3606) bit for bit, and the
mixed-precision form of GEOMU's compared mode (float deflation, FP64 Newton
polish, FP64 ZROOTS where the float phase is ambiguous) against the same:
the same real/complex split, which fixes GEOMU's Table-1 case, and the
roots to 1e-12 relative, on U(u) = 1 + cc u^2 + dd u^3 + ee u^4 from the
GEOKERR batch distribution and from rays near the critical curve (near
double roots)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from oracle import np_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    out = tmp_path_factory.mktemp("roots") / "libroots_probe.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-ffp-contract=off",
                    "-I" + os.path.join(HERE, "support", "hostshim"),
                    os.path.join(HERE, "support", "roots_probe.cpp"), "-o", str(out)], check=True)
    L = C.CDLL(str(out))
    L.probe_roots.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    L.probe_roots.restype = None
    return L


def _coefs(n, seed, near_critical=False):
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.0, 0.998, n)
    mu0 = rng.uniform(0.05, 0.99, n)
    al, be = rng.uniform(-12, 12, (2, n))
    if near_critical:
        # Schwarzschild-like shadow edge b = 3 sqrt(3) scaled around a: impact
        # parameters close to the critical curve give nearly double roots
        r = 3.0 * np.sqrt(3.0) * (1.0 + rng.normal(0, 1e-3, n))
        ph = rng.uniform(0, 2 * np.pi, n)
        al, be = r * np.cos(ph), r * np.sin(ph)
        a = rng.uniform(0.0, 0.2, n)
    l = -al * np.sqrt(1 - mu0 ** 2)
    q2 = be ** 2 - (a * a - al * al) * mu0 ** 2
    cc = a * a - q2 - l * l
    dd = 2 * ((a - l) ** 2 + q2)
    ee = -a * a * q2
    keep = ee != 0.0
    return np.ascontiguousarray(np.stack([cc, dd, ee], axis=1)[keep])


def _run(probe, coef, mixed):
    n = len(coef)
    roots = np.empty((n, 8))
    fb = np.empty(n, np.int32)
    probe.probe_roots(coef.ctypes.data, n, 1 if mixed else 0, roots.ctypes.data, fb.ctypes.data)
    return roots[:, 0::2] + 1j * roots[:, 1::2], fb


def _pair_order(r):
    out = r.copy()
    for i in range(len(r)):
        out[i] = sorted(r[i], key=lambda z: (round(z.real, 9), z.imag))
    return out


def _oracle(coef):
    return np.array([O.zroots([1.0, 0.0, c[0], c[1], c[2]]) for c in coef])


@pytest.mark.parametrize("near", [False, True])
def test_zroots_fp64_and_mixed_against_oracle(probe, near):
    coef = _coefs(4000, 7 + near, near_critical=near)
    ref = _oracle(coef)
    r64, _ = _run(probe, coef, False)
    np.testing.assert_array_equal(r64, ref)             # the device restatement, bit for bit
    rm, fb = _run(probe, coef, True)
    # ZROOTS sorts by real part only: a conjugate pair may come in either
    # order, which GEOMU never looks at (it reads only the real roots)
    ref, rm = _pair_order(ref), _pair_order(rm)
    nreal_ref = (ref.imag == 0).sum(axis=1)
    nreal_mix = (rm.imag == 0).sum(axis=1)
    assert np.array_equal(nreal_mix, nreal_ref)
    rel = np.abs(rm - ref) / np.maximum(np.abs(ref), 1e-300)
    assert np.all(rel[fb == 0] <= 1e-12), rel[fb == 0].max()
    assert np.array_equal(rm[fb == 1], ref[fb == 1])
    frac = fb.mean()
    print(f"mixed roots: fallback {frac:.4%}, max rel {rel.max():.2e}")
    assert frac < (0.05 if near else 1e-3)
