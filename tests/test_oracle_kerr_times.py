"""Pins for the FP64 Kerr oracle's end values of the coordinate time t, the
azimuth phi and the affine parameter lambda (oracle/kerr_oracle.c,
orc_render_ray: t integrated with the state as the reference's RK4 does,
bh_kerr.h:734-774 and 987-1032; lambda the sum of the accepted steps).

At a = 0 a photon moves in a plane through the hole with energy E = -p_t,
total angular momentum L^2 = p_theta^2 + p_phi^2 / sin^2(theta) and the
conserved value H of the Hamiltonian (the reference's camera launches rays
with H of order 1e-8 E^2, not exactly 0), and

    dr/dlambda = +-sqrt(E^2 - (1 - 2M/r) (L^2 / r^2 - 2H)),
    dt/dlambda = E / (1 - 2M/r),   dpsi/dlambda = L / r^2

(psi the angle swept in the orbital plane).  An escaping ray runs from the
camera in to the periapsis r_min (the largest root of dr/dlambda, near that
of r^3 - b^2 (r - 2M), b = L/E) and out to its last state r_end, so lambda,
t and psi are quadratures over r, evaluated here with scipy after r = r_min
+ u^2 removes the square-root singularity at the turning point.  phi follows from psi by
rotating the camera's position vector in the orbital plane.  None of this
shares code or formulas with the oracle's integrator.
"""
import math

import numpy as np
import pytest
from scipy.integrate import quad
from scipy.optimize import brentq


def _quad_leg(f, r_min, r_hi):
    """int_{r_min}^{r_hi} f(r) dr / |dr/dlambda| with r = r_min + u^2."""
    val, _ = quad(lambda u: 2.0 * u * f(r_min + u * u), 0.0, math.sqrt(r_hi - r_min),
                  epsabs=0.0, epsrel=1e-13, limit=400)
    return val


def _schwarzschild_ends(M, E, L, H, r_cam, r_end):
    b = L / E
    r0 = 2.0 * b / math.sqrt(3.0) * math.cos(math.acos(-3.0 * math.sqrt(3.0) * M / b) / 3.0)

    def rdot2(r):
        return E * E - (1.0 - 2.0 * M / r) * (L * L / (r * r) - 2.0 * H)

    r_min = brentq(rdot2, r0 * (1 - 1e-4), r0 * (1 + 1e-4), xtol=1e-13, rtol=1e-15)

    def rdot(r):
        return math.sqrt(max(rdot2(r), 0.0))

    def leg(g):
        # the integrand g / |rdot| is regular after the substitution: |rdot|
        # vanishes like sqrt(r - r_min)
        return (_quad_leg(lambda r: g(r) / rdot(r), r_min, r_cam) +
                _quad_leg(lambda r: g(r) / rdot(r), r_min, r_end))

    lam = leg(lambda r: 1.0)
    t = leg(lambda r: E / (1.0 - 2.0 * M / r))
    psi = leg(lambda r: L / (r * r))
    return lam, t, psi, r_min


@pytest.mark.parametrize("integrator", [1, 0])
def test_end_time_and_affine_parameter_by_quadrature(oracle, integrator):
    O = oracle
    p = O.default_params(width=9, height=5, fov_y_deg=30.0, spin=0.0, r_cam_Rs=50.0, rtol=1e-10,
                         atol=1e-12, integrator=integrator, max_steps=200000)
    ys, xs = np.mgrid[0:p.height, 0:p.width]
    _, info = O.render_pixels(p, xs.ravel(), ys.ravel(), 0, 1)
    M = float(np.float32(p.mass))
    r_cam = float(np.float32(p.r_cam_Rs) * np.float32(2.0 * np.float32(p.mass)))
    th0, ph0 = float(np.float32(p.theta_cam)), float(np.float32(p.phi_cam))
    esc = info["cls"] == 0  # NP_RAY_ESCAPED
    assert esc.sum() >= 20
    # measured 5e-12 (adaptive, rtol 1e-10) and 2e-11 (fixed steps of
    # 0.01 min(r / Rs, 100)); without H in dr/dlambda the gap is 8e-8
    tol = 5e-11 if integrator == 1 else 2e-10
    worst = [0.0, 0.0, 0.0]
    for k in np.flatnonzero(esc):
        pt, pr, pth, pph = info["p0"][k]
        E = -pt
        L = math.sqrt(pth * pth + (pph / math.sin(th0)) ** 2)
        f = 1.0 - 2.0 * M / r_cam
        H = 0.5 * (-E * E / f + f * pr * pr + L * L / (r_cam * r_cam))
        lam, t, psi, r_min = _schwarzschild_ends(M, E, L, H, r_cam, info["r_end"][k])
        assert info["min_r"][k] >= r_min * (1 - 1e-9)
        # phi: rotate the camera's position by psi in the orbital plane
        x0 = np.array([math.sin(th0) * math.cos(ph0), math.sin(th0) * math.sin(ph0), math.cos(th0)])
        e_th = np.array([math.cos(th0) * math.cos(ph0), math.cos(th0) * math.sin(ph0), -math.sin(th0)])
        e_ph = np.array([-math.sin(ph0), math.cos(ph0), 0.0])
        m = (pth * e_th + pph / math.sin(th0) * e_ph) / L
        x = math.cos(psi) * x0 + math.sin(psi) * m
        dphi = math.atan2(x[1], x[0]) - ph0
        target = math.copysign(psi, pph)
        dphi += 2.0 * math.pi * round((target - dphi) / (2.0 * math.pi))
        worst[0] = max(worst[0], abs(info["lam_end"][k] - lam) / lam)
        worst[1] = max(worst[1], abs(info["t_end"][k] - t) / t)
        worst[2] = max(worst[2], abs(info["phi_end"][k] - (ph0 + dphi)))
    assert worst[0] < tol and worst[1] < tol, worst
    assert worst[2] < tol * 10, worst


def test_end_values_fail_when_t_misses_its_lapse(oracle):
    """The quadrature tells t from the affine parameter: an integrand without
    the 1 / (1 - 2M/r) lapse factor (a plausible dropped term) misses the
    oracle's t by far more than the tolerance above."""
    O = oracle
    p = O.default_params(width=3, height=3, fov_y_deg=30.0, spin=0.0, r_cam_Rs=50.0, integrator=1,
                         max_steps=200000)
    ys, xs = np.mgrid[0:p.height, 0:p.width]
    _, info = O.render_pixels(p, xs.ravel(), ys.ravel(), 0, 1)
    k = int(np.flatnonzero(info["cls"] == 0)[0])
    M = float(np.float32(p.mass))
    r_cam = float(np.float32(p.r_cam_Rs) * np.float32(2.0 * np.float32(p.mass)))
    pt, pr, pth, pph = info["p0"][k]
    E = -pt
    L = math.sqrt(pth * pth + (pph / math.sin(float(np.float32(p.theta_cam)))) ** 2)
    f = 1.0 - 2.0 * M / r_cam
    H = 0.5 * (-E * E / f + f * pr * pr + L * L / (r_cam * r_cam))
    lam, _, _, _ = _schwarzschild_ends(M, E, L, H, r_cam, info["r_end"][k])
    t_no_lapse = E * lam
    assert abs(info["t_end"][k] - t_no_lapse) / t_no_lapse > 1e-3
