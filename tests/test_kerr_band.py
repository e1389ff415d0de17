"""Pins of the critical-curve margin the GPU parity tests use to state their
band (tests/kerr_band.py): the Schwarzschild closed form, Kerr's spherical
photon orbits (Teo 2003), and the oracle integrator's own capture/escape
classification."""
import numpy as np
import pytest

import kerr_band as KB


def test_schwarzschild_margin_is_one_minus_b2_over_27():
    b = np.linspace(0.5, 15.0, 301)
    for eta_frac in (0.0, 0.3, 1.0):       # a = 0: only lambda^2 + eta = b^2 matters
        lam = b * np.sqrt(1.0 - eta_frac)
        eta = b * b * eta_frac
        m = KB.critical_margin_lam_eta(lam, eta, 0.0, 1000.0)
        np.testing.assert_allclose(m, 1.0 - b * b / 27.0, rtol=0, atol=1e-13)


@pytest.mark.parametrize("a", [0.3, 0.7, 0.9, 0.998])
def test_spherical_photon_orbits_have_zero_margin(a):
    """On the critical curve U(u) has a double root at u = 1 / r_s, r_s the
    radius of the spherical photon orbit with those constants (Teo 2003,
    Eqs. 12-13 with M = 1)."""
    r_pro = 2.0 * (1.0 + np.cos(2.0 / 3.0 * np.arccos(-a)))
    r_ret = 2.0 * (1.0 + np.cos(2.0 / 3.0 * np.arccos(a)))
    r = np.linspace(r_pro, r_ret, 41)[1:-1]
    lam = -(r ** 3 - 3 * r ** 2 + a * a * r + a * a) / (a * (r - 1.0))
    eta = -r ** 3 * (r ** 3 - 6 * r ** 2 + 9 * r - 4 * a * a) / (a * a * (r - 1.0) ** 2)
    m = KB.critical_margin_lam_eta(lam, eta, a, 1000.0)
    np.testing.assert_allclose(m, 0.0, atol=1e-11)
    # a little inside the curve the ray is captured, a little outside it escapes
    assert np.all(KB.critical_margin_lam_eta(lam * 0.999, eta * 0.999, a, 1000.0) > 0)
    assert np.all(KB.critical_margin_lam_eta(lam * 1.001, eta * 1.001, a, 1000.0) < 0)


def test_margin_sign_matches_the_oracle_integrator(oracle):
    """Every traced ray with |m| > 1e-3 is captured iff m > 0."""
    import nullpath_amd as npa
    p = npa.kerr_disk_params(width=32, height=32, spin=0.9, theta_deg=60.0, fov=20.0, r_cam_Rs=25.0,
                             seed=1)
    q = oracle.default_params()
    for name, _ in npa.RenderParams._fields_:
        setattr(q, name, getattr(p, name))
    ys, xs = np.mgrid[0:p.height, 0:p.width]
    _, info = oracle.render_pixels(q, xs.ravel(), ys.ravel(), 0, 1)
    m = KB.critical_margin(info["p0"], p.mass, p.spin, p.theta_cam, 2.0 * p.r_cam_Rs * p.mass)
    np.testing.assert_array_equal(oracle.ray_p0_pixels(q, xs.ravel(), ys.ravel(), 0, 1), info["p0"])
    far = np.abs(m) > 1e-3
    assert far.sum() > 900 and (m[far] > 0).sum() > 100 and (m[far] < 0).sum() > 100
    assert np.all(info["cls"][far] != 3)
    np.testing.assert_array_equal(info["cls"][far] == 1, m[far] > 0)


def test_pole_distance():
    # a = 0: cos^2(theta_min) = eta / (eta + lambda^2)
    lam = np.array([0.0, 1.0, 3.0, 5.0])
    eta = np.array([10.0, 10.0, 2.0, 0.5])
    np.testing.assert_allclose(KB.pole_distance(lam, eta, 0.0) ** 2, lam ** 2 / (eta + lam ** 2), atol=1e-15)
    # Kerr: M(mu) vanishes at mu^2 = cos^2(theta_min)
    for a in (0.5, 0.9, 0.998):
        s = KB.pole_distance(lam, eta, a)
        x = 1.0 - s * s
        np.testing.assert_allclose(eta + (a * a - eta - lam ** 2) * x - a * a * x * x, 0.0, atol=1e-12)
        assert np.all((x > 0) & (x <= 1))
