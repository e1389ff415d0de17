"""Pins for the Schwarzschild oracle (oracle/schw_oracle.c).

* The reference's own sanity criteria (tests/bh_sanity_test.cu:60-163) on the
  FP32 restatement of precomputeGeodesicsKernel: b_crit from the
  escape/no-escape boundary within 5 %, r_min non-decreasing in b,
  |deflection| non-increasing for escaping rays;
* the analytic turning radius against the largest root of r^3 - b^2 r + b^2 Rs;
* the FP64 "plain method" sweep: capture iff b < 3 sqrt(3) M, and the
  deflection of escaping rays against the orbit quadrature
  2 * int_{u0}^{u_t} du / sqrt(1/b^2 - u^2 + 2 M u^3) (mpmath).
"""
import math

import numpy as np
import pytest


def test_reference_sanity_criteria_fp32(oracle):
    mass = 10.0
    Rs = 2 * mass
    bcrit = 0.5 * 3.0 * math.sqrt(3.0) * Rs
    n = 4000
    b_min, b_max = np.float32(0.30 * bcrit), np.float32(20.0 * bcrit)
    t = np.arange(n, dtype=np.float32) / np.float32(n - 1)
    b = (b_min * np.power(np.float32(b_max / b_min), t)).astype(np.float32)
    r_start = max(1000.0, 1.2 * float(b_max))
    defl, rmin, st = oracle.schw_precompute_f32(mass, r_start, b)
    esc = np.nonzero(st == 0)[0]
    assert esc.size > 0
    i_esc = esc[0]
    b_est = 0.5 * (b[i_esc - 1] + b[i_esc])
    assert abs(b_est - bcrit) / bcrit <= 0.05
    viol_r = np.sum(rmin[1:] + 1e-3 * Rs < rmin[:-1])
    assert viol_r <= int(0.005 * n) + 1
    both = (st[1:] == 0) & (st[:-1] == 0)
    ad = np.abs(defl)
    viol_d = np.sum(both & (ad[1:] > ad[:-1] + 1e-3))
    assert viol_d <= int(0.01 * max(1, both.sum())) + 1


@pytest.mark.parametrize("b_over_rs", [2.7, 3.0, 5.0, 12.0, 40.0])
def test_rmin_analytic_is_cubic_root(oracle, b_over_rs):
    Rs = 20.0
    b = b_over_rs * Rs
    ok, r = oracle.schw_rmin_analytic(Rs, b)
    assert ok
    roots = np.roots([1.0, 0.0, -b * b, b * b * Rs])
    r_big = max(x.real for x in roots if abs(x.imag) < 1e-9)
    assert abs(r - r_big) / r_big < 1e-6   # float32 output


def _deflection_quadrature(M, b, r0):
    import mpmath
    mpmath.mp.dps = 20
    # turning point: smallest positive root of 2 M u^3 - u^2 + 1/b^2
    roots = np.roots([2 * M, -1.0, 0.0, 1.0 / (b * b)])
    ut = min(x.real for x in roots if abs(x.imag) < 1e-12 and x.real > 0)
    ut = mpmath.findroot(lambda u: 2 * M * u ** 3 - u ** 2 + 1 / mpmath.mpf(b) ** 2, ut)
    f = lambda u: 1 / mpmath.sqrt(1 / mpmath.mpf(b) ** 2 - u * u + 2 * M * u ** 3)
    return float(2 * mpmath.quad(f, [1 / mpmath.mpf(r0), ut]))


def test_fp64_sweep_classes_and_deflection(oracle):
    M = 10.0
    r0 = 1000.0 * M
    bc = 3 * math.sqrt(3) * M
    b = np.array([0.6, 2.0, 4.5, 5.1, 5.3, 5.6, 7.0, 10.0, 14.5]) * M
    cls, dphi, steps = oracle.schw_sweep(M, r0, 2.01 * M, 1e-3 * M, 2_000_000, b)
    for bi, ci in zip(b, cls):
        assert ci == (1 if bi < bc else 0), (bi, ci)
    for bi, ci, di in zip(b, cls, dphi):
        if ci == 0:
            ref = _deflection_quadrature(M, bi, r0)
            # the ray stops within one step past r0; that tail adds < b/r0^2 per unit r
            assert abs(di - ref) < 2e-5, (bi, di, ref)
    # photon-sphere band
    cls2, _, _ = oracle.schw_sweep(M, r0, 2.01 * M, 1e-3 * M, 2_000_000, np.array([bc + 5e-4 * M]))
    assert cls2[0] == 2
    assert steps.max() > 10000   # rays of the FP64 sweep from r0 = 1000 M need more than the reference's 10000-step cap
