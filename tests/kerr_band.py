"""The band of rays whose classification or crossing count the last bits of
the arithmetic may decide: rays near the critical curve, and rays whose
equatorial crossing falls at an edge of the disk.

Critical margin.  A photon with E = -p_t, L = p_phi and Carter constant
Q = p_th^2 + cos^2(th) (L^2 / sin^2(th) - a^2 E^2) has, in units of M with
lambda = L/E, eta = Q/E^2, the radial potential
    R(r) = (r^2 + a^2 - a lambda)^2 - Delta (eta + (lambda - a)^2)
(Bardeen 1973; the potential GEOKERR writes as U(u) = R/r^4 in u = 1/r,
geokerr_original.f:836-860).  With B = a^2 - a lambda, K = eta + (lambda-a)^2,
    U(u) = 1 + c2 u^2 + c1 u^3 + c0 u^4,  c2 = 2B - K, c1 = 2K, c0 = B^2 - a^2 K.
A ray launched inward from u_cam = M / r_cam turns back at the first zero of
U; it is captured iff U > 0 on the whole path to the horizon u_+.  The
margin m = min of U over [u_cam, u_+] is negative for escaping rays,
positive for captured ones and 0 on the critical curve (a double root), so
|m| measures the distance to the critical curve (for a = 0,
m = 1 - b^2 / 27: twice the relative distance to b_crit).  U'(u) = 0 away
from u = 0 at the roots of 4 c0 u^2 + 3 c1 u + 2 c2 = 0, so m is exact.
"""
from __future__ import annotations

import numpy as np


def constants(p0: np.ndarray, mass: float, spin: float, theta_cam: float):
    """(lambda, eta) in units of M from initial covariant momenta
    (p_t, p_r, p_theta, p_phi) at polar angle theta_cam."""
    p0 = np.asarray(p0, dtype=np.float64).reshape(-1, 4)
    E = -p0[:, 0]
    L = p0[:, 3]
    pth = p0[:, 2]
    a = float(spin) * float(mass)
    s, c = np.sin(float(theta_cam)), np.cos(float(theta_cam))
    Q = pth * pth + c * c * (L * L / (s * s) - a * a * E * E)
    return L / (E * mass), Q / (E * E * mass * mass)


def u_potential(lam, eta, spin):
    """Coefficients (c2, c1, c0) of U(u) = 1 + c2 u^2 + c1 u^3 + c0 u^4."""
    a = float(spin)
    B = a * a - a * lam
    K = eta + (lam - a) ** 2
    return 2.0 * B - K, 2.0 * K, B * B - a * a * K


def critical_margin_lam_eta(lam, eta, spin, r_cam_over_m):
    """m = min over u in [M / r_cam, 1 / r_+] of U(u)."""
    lam = np.atleast_1d(np.asarray(lam, dtype=np.float64))
    eta = np.atleast_1d(np.asarray(eta, dtype=np.float64))
    a = float(spin)
    c2, c1, c0 = u_potential(lam, eta, a)
    u_lo = 1.0 / float(r_cam_over_m)
    u_hi = 1.0 / (1.0 + np.sqrt(1.0 - a * a))
    f = lambda u: 1.0 + c2 * u * u + c1 * u ** 3 + c0 * u ** 4
    m = np.minimum(f(np.full_like(lam, u_lo)), f(np.full_like(lam, u_hi)))
    # interior critical points: 4 c0 u^2 + 3 c1 u + 2 c2 = 0
    qa, qb, qc = 4.0 * c0, 3.0 * c1, 2.0 * c2
    with np.errstate(divide="ignore", invalid="ignore"):
        disc = qb * qb - 4.0 * qa * qc
        sq = np.sqrt(np.where(disc >= 0.0, disc, np.nan))
        lin = np.abs(qa) < 1e-300
        # stable quadratic roots; the linear case when c0 = 0 (a = 0 or K = 0)
        qq = -0.5 * (qb + np.copysign(sq, qb))
        r1 = np.where(lin, -qc / qb, qq / qa)
        r2 = np.where(lin, np.nan, qc / qq)
    for r in (r1, r2):
        ok = np.isfinite(r) & (r > u_lo) & (r < u_hi)
        m = np.where(ok, np.minimum(m, f(np.where(ok, r, u_lo))), m)
    return m


def critical_margin(p0, mass, spin, theta_cam, r_cam):
    lam, eta = constants(p0, mass, spin, theta_cam)
    return critical_margin_lam_eta(lam, eta, spin, float(r_cam) / float(mass))


def pole_distance(lam, eta, spin):
    """sin(theta_min) of the theta motion: the largest mu^2 with
    M(mu) = eta + (a^2 - eta - lambda^2) mu^2 - a^2 mu^4 = 0 (FINDMROOTS,
    geokerr_original.f:607-642) is cos^2(theta_min)."""
    lam = np.asarray(lam, dtype=np.float64)
    eta = np.asarray(eta, dtype=np.float64)
    a2 = float(spin) ** 2
    A = a2 - eta - lam * lam
    sq = np.sqrt(np.maximum(A * A + 4.0 * a2 * eta, 0.0))
    with np.errstate(divide="ignore", invalid="ignore"):
        x = np.where(A < 0.0, 2.0 * eta / (sq - A), (A + sq) / (2.0 * a2))
    return np.sqrt(np.clip(1.0 - x, 0.0, 1.0))


def edge_distance(xr, ncross, rin, rout):
    """Smallest relative distance of any recorded crossing radius (oracle
    ray info xr[:, :ncross]) to the disk edges; inf without crossings."""
    xr = np.asarray(xr, dtype=np.float64).reshape(-1, 4)
    k = np.arange(4)[None, :] < np.asarray(ncross)[:, None]
    d = np.minimum(np.abs(xr - rin) / rin, np.abs(xr - rout) / rout)
    return np.where(k, d, np.inf).min(axis=1)


# The stated band (tests and benchmark): |m| < CRIT_BAND, a crossing radius
# within EDGE_BAND (relative) of r_in or r_out, or a theta turning point
# within sin(theta_min) < POLE_BAND of the axis.
CRIT_BAND = 1e-4
EDGE_BAND = 1e-6
POLE_BAND = 1e-3


def in_band(margin, edge=None, pole=None):
    b = np.abs(margin) < CRIT_BAND
    if edge is not None:
        b |= edge < EDGE_BAND
    if pole is not None:
        b |= pole < POLE_BAND
    return b


def ray_band(p, p0, xr=None, ncross=None):
    """(in_band, margin, edge, pole) for rays of render parameters p given the
    oracle's initial momenta (and crossing radii when known)."""
    r_cam = 2.0 * float(p.r_cam_Rs) * float(p.mass)
    lam, eta = constants(p0, float(p.mass), float(p.spin), float(p.theta_cam))
    margin = critical_margin_lam_eta(lam, eta, float(p.spin), r_cam / float(p.mass))
    pole = pole_distance(lam, eta, float(p.spin))
    edge = None if xr is None else edge_distance(xr, ncross, float(p.disk_rin), float(p.disk_rout))
    return in_band(margin, edge, pole), margin, edge, pole
