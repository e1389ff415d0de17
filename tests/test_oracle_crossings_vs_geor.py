"""The renderer's equatorial-crossing radii against GEOR.

The oracle's FP64 renderer (oracle/kerr_oracle.c, the reference's
integrator and crossing rule, render.cu:118-363) integrates each ray
numerically and interpolates the radius where theta crosses pi/2.  GEOR
(oracle/geor_oracle.c, pinned in test_oracle_geor.py) gives the same radius
semi-analytically from the ray's constants of motion: E = -p_t, L = p_phi,
Carter Q = p_theta^2 + cos^2(theta) (L^2 / sin^2(theta) - a^2 E^2) at the
camera, mu_f = 0, and for the k-th crossing tpm = k - 1 mu turning points
(k if the ray first moves away from the equator).  The agreement pins the
crossing radii and their order (which crossing is which) independently of
the integrator.  The bench's Kerr thin-disk frame (a = 0.9, theta = 80 deg,
r_cam = 1000 M, rtol 1e-8), rays around the hole.
"""
import numpy as np
import pytest

from nullpath_amd import deg2rad_f32, isco_radius

A = 0.9
R_CAM = 1000.0


def _crossing_errors(oracle, crossing, n=48, seed=3):
    p = oracle.default_params(width=1920, height=1080, fov_y_deg=55.0, mass=1.0, spin=A, r_cam_Rs=R_CAM / 2,
                              theta_cam=deg2rad_f32(80.0), integrator=1, camera=1, rtol=1e-8, atol=1e-12,
                              disk_max_orders=3, emiss_q=3.0, emiss_m_units=1, max_steps=200000,
                              disk_rin=isco_radius(A), disk_rout=20.0, crossing=crossing)
    rng = np.random.default_rng(seed)
    xs = (960 + rng.integers(-30, 30, n)).astype(np.int32)
    ys = (540 + rng.integers(-15, 15, n)).astype(np.int32)
    _, info = oracle.render_pixels(p, xs, ys, 0, 1)
    th = float(p.theta_cam)
    mu0 = np.cos(th)
    errs, cases = [], set()
    for ri in info:
        pt, pr, pth, pph = ri["p0"]
        E = -pt
        L = pph / E
        Q = (pth ** 2 + np.cos(th) ** 2 * (pph ** 2 / np.sin(th) ** 2 - A * A * E * E)) / E ** 2
        su = 1.0 if pr < 0 else -1.0
        sm = -np.sign(pth)            # dmu/dlambda = -sin(theta) p_theta / Sigma
        t0 = 1 if sm > 0 else 0       # moving away from the equator: it turns first
        for k in range(ri["ncross"]):
            g = oracle.geor(1.0 / R_CAM, mu0, 0.0, A, L, Q, k + t0, su, sm)
            assert g["uf"] > 0, (k, g)
            r = 1.0 / g["uf"]
            errs.append(abs(ri["xr"][k] - r) / r)
            cases.add((g["ncase"], g["tpr"], k))
    return np.array(errs), cases


def test_hermite_crossing_radii_match_geor(oracle):
    e, cases = _crossing_errors(oracle, 1)
    assert e.size >= 40
    assert e.max() <= 1e-6, e.max()
    assert np.median(e) <= 1e-7, np.median(e)
    # first and second crossings, before and after the radial turning point,
    # on escaping (row 7) and captured (row 5) rays
    assert {k for _, _, k in cases} >= {0, 1}
    assert {t for _, t, _ in cases} == {0, 1}
    assert {c for c, _, _ in cases} >= {5, 7}


def test_linear_crossing_error_is_the_interpolant(oracle):
    # the reference's linear rule is off by the interpolant's h^2 r''/8
    e, _ = _crossing_errors(oracle, 0)
    assert 1e-5 < e.max() <= 1e-3, e.max()
