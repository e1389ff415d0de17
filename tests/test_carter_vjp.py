"""The hand-written vector-Jacobian product of the Carter-form right-hand side
(nullpath_amd/csrc/np_carter.hpp), built for the host, against central finite
differences of the right-hand side itself; and the right-hand side against
Hamilton's equations of the oracle's metric on the constraint surface."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    out = tmp_path_factory.mktemp("carter") / "libcarter_probe.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-ffp-contract=off",
                    os.path.join(HERE, "support", "carter_probe.cpp"), "-o", str(out)], check=True)
    L = C.CDLL(str(out))
    pd = C.POINTER(C.c_double)
    for name in ("probe_rhs", "probe_rhs_r"):
        getattr(L, name).argtypes = [pd, C.c_double, C.c_double, C.c_double, pd]
    for name in ("probe_vjp", "probe_vjp_r"):
        getattr(L, name).argtypes = [pd, C.c_double, C.c_double, C.c_double, pd, pd]
    return L


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def rhs(probe, y, E, L, a, suffix=""):
    y = np.ascontiguousarray(y, dtype=np.float64)
    o = np.zeros(4)
    getattr(probe, "probe_rhs" + suffix)(_p(y), E, L, a, _p(o))
    return o


@pytest.mark.parametrize("suffix", ["", "_r"])
def test_vjp_matches_finite_differences(probe, suffix):
    """suffix "" is the affine-parameter field f, "_r" the r-scaled field r f."""
    rng = np.random.default_rng(61)
    worst = 0.0
    for _ in range(300):
        a = rng.uniform(0.0, 0.99)
        y = np.array([rng.uniform(2.5, 60.0), rng.uniform(0.2, 2.9), rng.normal(), rng.normal() * 3])
        E, L = rng.uniform(0.5, 1.5), rng.normal() * 4
        v = rng.normal(size=4)
        o = np.zeros(7)
        getattr(probe, "probe_vjp" + suffix)(_p(y), E, L, a, _p(np.ascontiguousarray(v)), _p(o))
        fd = np.zeros(7)
        for i in range(4):
            h = 1e-6 * max(1.0, abs(y[i]))
            yp, ym = y.copy(), y.copy()
            yp[i] += h
            ym[i] -= h
            fd[i] = v @ (rhs(probe, yp, E, L, a, suffix) - rhs(probe, ym, E, L, a, suffix)) / (2 * h)
        for j, (dE, dL, da) in enumerate([(1, 0, 0), (0, 1, 0), (0, 0, 1)]):
            h = 1e-6
            fp = rhs(probe, y, E + h * dE, L + h * dL, a + h * da, suffix)
            fm = rhs(probe, y, E - h * dE, L - h * dL, a - h * da, suffix)
            fd[4 + j] = v @ (fp - fm) / (2 * h)
        # error relative to the largest component of each vector (FD round-off
        # of a component near zero is ~1e-10 absolute)
        worst = max(worst, float(np.max(np.abs(o[:4] - fd[:4])) / np.max(np.abs(fd[:4]))),
                    float(np.max(np.abs(o[4:] - fd[4:])) / np.max(np.abs(fd[4:]))))
    assert worst < 1e-6, worst


def test_rhs_is_hamilton_flow_on_shell(probe, oracle):
    """On H = 0 the Carter-form field equals the reference Hamiltonian field."""
    rng = np.random.default_rng(62)
    for _ in range(100):
        a = rng.uniform(0.0, 0.99)
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        s = oracle.init_zamo(1.0, a, rng.uniform(4, 50), rng.uniform(0.3, 2.8), 0.0, n)
        d = oracle.rhs(s, 1.0, a)
        c = rhs(probe, s[[1, 2, 5, 6]], -s[4], s[7], a)
        np.testing.assert_allclose(c, d[[1, 2, 5, 6]], rtol=1e-9, atol=1e-12)
