"""The reference's accuracy plan for Schwarzschild (docs/PLAN.md:38-43) on the
GPU FP64 plain-method sweep (np_schw_sweep, M = 10, r0 = 1000 M, the
photon-sphere band off so every ray is classified by its integration):

* b_crit: the capture/escape boundary found by three rounds of 4097-ray
  brackets, |b_num - 3 sqrt(3) M| / (3 sqrt(3) M) <= 1e-10;
* r_min(b): relative error <= 1e-12 for b >= 1.05 b_crit and <= 1e-6 down
  to 1.01 b_crit, against the root of r^3 - b^2 r + 2 M b^2 polished in
  mpmath;
* weak-field deflection (b >= 20 M): |alpha - alpha_PN| <= 1e-3, with
  alpha = delta-phi - pi + 2 arcsin(b / r0) (the straight-line angle between
  the launch and exit radii r0, which the plan's rays from infinity do not
  have) and alpha_PN = 4u + (15 pi/4) u^2 + (128/3) u^3 + (3465 pi/64) u^4,
  u = M/b (Keeton & Petters 2005); its truncation is 3e-4 at b = 20 M
  against an mpmath quadrature of the exact deflection;
* near-critical fit alpha = -c1 ln(b/b_crit - 1) + c2 over
  b/b_crit - 1 in [1e-6, 1e-3]: c1 within 1 % of 1 and c2 within 2 % of
  ln(216 (7 - 4 sqrt 3)) - pi (Bozza 2002, the strong-deflection limit);
* invariants: |H|max <= 1e-12 over the accepted steps of escaped rays; E
  and L are constants of the integration (p_t, p_phi are not stepped), so
  their drift is zero by construction.
"""
import math

import numpy as np
import pytest
import torch

import nullpath_amd as npa

pytestmark = pytest.mark.gpu

M = 10.0
BC = 3.0 * math.sqrt(3.0) * M


def _sweep(b, diag=False):
    t = torch.from_numpy(np.ascontiguousarray(b, dtype=np.float64)).cuda()
    out = npa.schw_sweep(t, M=M, ps_band=0.0, diagnostics=diag)
    torch.cuda.synchronize()
    return [x.cpu().numpy() for x in out]


def test_critical_impact_parameter(device):
    lo, hi = 5.0 * M, 5.4 * M
    for rnd in range(3):
        b = np.linspace(lo, hi, 4097)
        print(f"[gpu] np_schw_sweep b_crit bracket round {rnd}: [{lo!r}, {hi!r}]", flush=True)
        cls = _sweep(b)[0]
        assert set(np.unique(cls)) <= {0, 1}
        esc = cls == 0
        k = int(np.argmax(esc))
        assert 0 < k and esc[k:].all() and not esc[:k].any(), "capture/escape not monotone in b"
        lo, hi = b[k - 1], b[k]
    b_num = 0.5 * (lo + hi)
    print(f"[gpu] b_num {b_num!r}, rel err {abs(b_num - BC) / BC:.3g}", flush=True)
    assert abs(b_num - BC) / BC <= 1e-10


def _root(b):
    from test_oracle_geomu import mpmath
    mpmath.mp.dps = 40
    bb = mpmath.mpf(float(b))
    f = lambda r: r ** 3 - bb * bb * r + 2 * M * bb * bb
    r = 2 * bb / mpmath.sqrt(3) * mpmath.cos(mpmath.acos(-(3 * mpmath.sqrt(3) / 2) * 2 * M / bb) / 3)
    return float(mpmath.findroot(f, r))


def test_turning_radius_deflection_and_invariants(device):
    x_near = np.logspace(-6, -3, 16)
    b = BC * np.concatenate([1 + x_near, np.logspace(np.log10(1.01), np.log10(1.05), 10),
                             np.logspace(np.log10(1.05), np.log10(200 * M / BC), 60)])
    print(f"[gpu] np_schw_sweep with r_min and |H|max, {b.size} rays", flush=True)
    cls, dphi, steps, rmin, hmax = _sweep(b, diag=True)
    assert np.all(cls == 0)
    x = b / BC - 1.0
    # r_min(b)
    ref = np.array([_root(v) for v in b])
    rel = np.abs(rmin - ref) / ref
    assert rel[x >= 0.05 - 1e-12].max() <= 1e-12, rel[x >= 0.05 - 1e-12].max()
    assert rel[x >= 0.01 - 1e-12].max() <= 1e-6
    # weak field
    alpha = dphi - math.pi + 2.0 * np.arcsin(b / (1000.0 * M))
    wf = b >= 20 * M
    u = M / b[wf]
    pn = 4 * u + 15 * math.pi / 4 * u ** 2 + 128 / 3 * u ** 3 + 3465 * math.pi / 64 * u ** 4
    print(f"[gpu] weak field max |alpha - alpha_PN| {np.abs(alpha[wf] - pn).max():.3g}", flush=True)
    assert np.abs(alpha[wf] - pn).max() <= 1e-3
    # near-critical logarithmic divergence
    nc = x <= 1e-3 * (1 + 1e-9)
    A = np.stack([-np.log(x[nc]), np.ones(nc.sum())], axis=1)
    c1, c2 = np.linalg.lstsq(A, alpha[nc], rcond=None)[0]
    c2_ref = math.log(216.0 * (7.0 - 4.0 * math.sqrt(3.0))) - math.pi
    print(f"[gpu] strong-deflection fit c1 {c1:.6f}, c2 {c2:.6f} (ref {c2_ref:.6f})", flush=True)
    assert abs(c1 - 1.0) <= 0.01
    assert abs(c2 - c2_ref) <= 0.02 * abs(c2_ref)
    # invariant monitor
    print(f"[gpu] |H|max over escaped rays {hmax.max():.3g}", flush=True)
    assert hmax.max() <= 1e-12
