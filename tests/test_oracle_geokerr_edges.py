"""Pins for the oracle's GEOKERR branches that the random physical inputs of
test_oracle_geor.py and test_oracle_geomu.py do not reach:

* the equal-root cubic (U(u) = dd (u - u1) (u - u2)^2, a = 0 and
  L^2 + Q2 = 27: the Schwarzschild photon orbit u2 = 1/3, u1 = -1/6) in GEOR
  (geokerr_original.f:2832-2860) and GEOMU (:905-933), including the clamp to
  u2 that both take from the caller's UF;
* GEOR's Table-2 row 2 with FIRSTPT off (:2795-2799), where the u0 piece of
  the u integral is set to zero whenever u0 is not u3;
* the WARNING edits of an initial u inside a forbidden region of U(u): GEOMU
  (:880-888 cubic, :1017-1025 quartic) and GEOR (:2807-2814, :2972-2980).

Each is checked through the Mino-time identity, which does not restate the
oracle: the u integral int du / sqrt|U(u)| from the (edited) start to the
returned u, by mpmath quadrature, equals the mu integral int dmu / sqrt M(mu)
along the mu path (quadrature, turning points found independently).  A wrong
root, sign or branch formula fails it.  The edits themselves (where u0 goes)
are the reference's rules and are checked for what they are.
"""
import math

import numpy as np
import pytest

from test_oracle_geomu import _imu_path, mpmath

EQUAL = [(3.0, 18.0), (0.0, 27.0), (1.0, 26.0), (-2.0, 23.0)]   # a = 0, L^2 + Q2 = 27


def _iu(a, l, q2, x, y):
    """|int_x^y du / sqrt|U(u)||, U(u) = 1 + cc u^2 + dd u^3 + ee u^4."""
    a, l, q2 = mpmath.mpf(a), mpmath.mpf(l), mpmath.mpf(q2)
    cc, dd, ee = a * a - q2 - l * l, 2 * ((a - l) ** 2 + q2), -a * a * q2
    f = lambda u: 1 / mpmath.sqrt(abs(1 + cc * u * u + dd * u ** 3 + ee * u ** 4))
    return float(abs(mpmath.quad(f, sorted([mpmath.mpf(x), mpmath.mpf(y)]))))


def _roots(a, l, q2):
    """Real roots of U(u), polished in mpmath (an integral that ends at a
    simple root turns an endpoint error d into an error ~sqrt(d))."""
    cc, dd, ee = a * a - q2 - l * l, 2 * ((a - l) ** 2 + q2), -a * a * q2
    c = [ee, dd, cc, 0.0, 1.0] if ee != 0.0 else [dd, cc, 0.0, 1.0]
    r = np.roots(c)
    r = np.sort(r[np.abs(r.imag) < 1e-12].real)
    A, L, Q = mpmath.mpf(a), mpmath.mpf(l), mpmath.mpf(q2)
    U = lambda u: 1 + (A * A - Q - L * L) * u * u + 2 * ((A - L) ** 2 + Q) * u ** 3 - A * A * Q * u ** 4
    return [mpmath.findroot(U, mpmath.mpf(float(x))) for x in r]


def _muf_a0(l, q2, mu0, T):
    """a = 0: M(mu) = Q2 - (Q2 + L^2) mu^2, so mu = mu+ sin(phi) with
    dphi = sqrt(Q2 + L^2) dMino; mu increasing (sm = +1), no turning point."""
    mup = math.sqrt(q2 / (q2 + l * l))
    arg = math.asin(mu0 / mup) + math.sqrt(q2 + l * l) * T
    assert arg < math.pi / 2
    return mup * math.sin(arg)


@pytest.fixture(autouse=True)
def _dps():
    mpmath.mp.dps = 30


@pytest.mark.parametrize("l,q2", EQUAL)
def test_geor_equal_root_cubic(oracle, l, q2):
    mup = math.sqrt(q2 / (q2 + l * l))
    mu0 = -0.5 * mup
    u1, u2 = -1.0 / 6.0, 1.0 / 3.0
    # (u0, su, uf_in, end): the Mino time T is a fraction of what lies between
    # u0 and `end` (the photon orbit is never reached: it takes infinite time)
    legs = [(0.05, 1.0, -1.0, u2), (0.2, 1.0, -1.0, u2), (0.2, -1.0, -1.0, 1e-3),
            (0.40, 1.0, 0.40, 0.499), (0.45, -1.0, 0.45, u2)]
    for u0, su, uf_in, end in legs:
        T = min(0.5 * _iu(0.0, l, q2, u0, end), 0.35)
        muf = _muf_a0(l, q2, mu0, T)
        r = oracle.geor(u0, mu0, muf, 0.0, l, q2, 0, su, 1.0, uf_in=uf_in)
        # the branch: one simple and one double root
        assert r["u2"] == r["u3"] and abs(r["u2"] - u2) < 1e-15 and abs(r["u1"] - u1) < 1e-15, r
        assert r["ncase"] == (1 if u0 <= u2 else 2) and r["tpr"] == 0
        imu_q = _imu_path(0.0, l, q2, mu0, muf, 0, 1.0)
        assert abs(r["imu"] - imu_q) <= 1e-10 * imu_q, (r, imu_q)
        lo, hi = min(u0, end), max(u0, end)
        assert lo < r["uf"] < hi, (u0, su, r)
        iu_q = _iu(0.0, l, q2, u0, r["uf"])
        assert abs(iu_q - r["imu"]) <= 1e-9 * iu_q, (u0, su, iu_q, r)
    # the clamp to the photon orbit, decided by the UF the caller passes in
    muf = _muf_a0(l, q2, mu0, 0.1)
    assert oracle.geor(0.2, mu0, muf, 0.0, l, q2, 0, 1.0, 1.0, uf_in=0.4)["uf"] == oracle.geor(
        0.2, mu0, muf, 0.0, l, q2, 0, 1.0, 1.0)["u2"]
    r = oracle.geor(0.45, mu0, muf, 0.0, l, q2, 0, -1.0, 1.0)   # uf_in = -1 < u2
    assert r["ncase"] == 2 and r["uf"] == r["u2"]


@pytest.mark.parametrize("l,q2", EQUAL)
def test_geomu_equal_root_cubic(oracle, l, q2):
    mup = math.sqrt(q2 / (q2 + l * l))
    for u0, uf, su in [(0.05, 0.3, 1.0), (0.2, 0.01, -1.0), (0.45, 0.49, 1.0), (0.45, 0.34, -1.0)]:
        for mu0, sm in [(-0.5 * mup, 1.0), (0.3 * mup, -1.0)]:
            g = oracle.geomu(u0, uf, mu0, 0.0, l, q2, 0, su, sm)
            assert g["u2"] == g["u3"] and abs(g["u2"] - 1.0 / 3.0) < 1e-15, g
            assert g["ncase"] == (1 if u0 <= 1.0 / 3.0 else 2) and g["tpr"] == 0 and g["uf_used"] == uf
            iu_q = _iu(0.0, l, q2, u0, uf)
            assert abs(g["iu"] - iu_q) <= 1e-10 * iu_q, (u0, uf, su, g, iu_q)
            imu_q = _imu_path(0.0, l, q2, mu0, g["muf"], g["tpm"], sm)
            assert abs(imu_q - iu_q) <= 1e-9 * iu_q, (u0, uf, su, mu0, sm, g, imu_q, iu_q)
    # a final u past the photon orbit is clamped to it, with the reference's
    # "infinite" Mino time
    g = oracle.geomu(0.1, 0.4, 0.0, 0.0, l, q2, 0, 1.0, 1.0)
    assert g["uf_used"] == g["u2"] and g["iu"] == 1e300


def test_geor_row2_firstpt_off_measures_from_u3(oracle):
    """Row 2 (u0 >= u3, a = 0, three real roots).  A first call with FIRSTPT
    on keeps the u0 piece IU0 of the u integral; a later call with FIRSTPT
    off, from the same u0, sets IU0 = 0 because u0 is not u3 (the ELSEIF
    tests u0 .ne. u3 where row 1 tests u0 .eq. u2), so its uf is the point
    Mino time IMU away from the turning point u3, not from u0."""
    l, q2 = 4.0, 20.0
    u1, u2, u3 = _roots(0.0, l, q2)
    u0, mu0 = 0.45, 0.0
    assert u3 < u0 < 0.5
    T1 = 0.3 * _iu(0.0, l, q2, u0, 0.4999)
    T2 = 0.6 * _iu(0.0, l, q2, u3, 0.4999)
    muf1, muf2 = _muf_a0(l, q2, mu0, T1), _muf_a0(l, q2, mu0, T2)
    for su in (1.0, -1.0):
        v = oracle.GkVars()
        first = oracle.gk_geor(v, u0, -1.0, mu0, muf1, 0.0, l, q2, 0.0, 0, 0, su, 1.0, firstpt=1)
        # IU0 is kept without the 1/sqrt(dd) of U(u) = dd (u - u1)(u - u2)(u - u3)
        dd = 2.0 * (l * l + q2)
        assert first["ncase"] == 2 and abs(first["iu0"] / math.sqrt(dd) - _iu(0.0, l, q2, u3, u0)) <= 1e-10
        later = oracle.gk_geor(v, u0, -1.0, mu0, muf2, 0.0, l, q2, 0.0, 0, 0, su, 1.0, firstpt=0)
        assert later["iu0"] == 0.0 and later["ncase"] == 2
        imu_q = _imu_path(0.0, l, q2, mu0, muf2, 0, 1.0)
        assert abs(later["imu"] - imu_q) <= 1e-10 * imu_q
        assert u3 < later["uf"] < 0.5
        iu_q = _iu(0.0, l, q2, u3, later["uf"])
        assert abs(iu_q - later["imu"]) <= 1e-9 * iu_q, (su, later, iu_q)
        # a fresh call measures from u0 instead, so the two differ
        fresh = oracle.geor(u0, mu0, muf2, 0.0, l, q2, 0, su, 1.0)
        assert abs(fresh["uf"] - later["uf"]) > 1e-6
    # row 1 with FIRSTPT off reuses the cached IU0: same result bit for bit
    v = oracle.GkVars()
    on = oracle.gk_geor(v, 0.1, -1.0, mu0, muf1, 0.0, l, q2, 0.0, 0, 0, 1.0, 1.0, firstpt=1)
    off = oracle.gk_geor(v, 0.1, -1.0, mu0, muf1, 0.0, l, q2, 0.0, 0, 0, 1.0, 1.0, firstpt=0)
    assert on["ncase"] == 1 and (on["uf"], on["imu"], on["tpr"]) == (off["uf"], off["imu"], off["tpr"])


# forbidden-region starts: (a, L, Q2, u0 inside (u2, u3), mu0, sm)
FORBIDDEN = [(0.0, 4.0, 20.0, 0.30, 0.2, 1.0), (0.9, 5.0, 20.0, 0.40, 0.1, -1.0)]


@pytest.mark.parametrize("case", FORBIDDEN)
def test_geomu_warning_edit(oracle, case):
    """GEOMU moves a start inside the forbidden region (U(u0) < 0) to u3
    (su = 1) or u2 for the cubic, to 0 (su = 1) or u+ for the quartic, and
    solves from there: the result satisfies the Mino-time identity from the
    edited u0."""
    a, l, q2, u0, mu0, sm = case
    r = _roots(a, l, q2)
    u2, u3 = (r[1], r[2])
    uplus = 1.0 / (1.0 + math.sqrt(1.0 - a * a))
    assert u2 < u0 < u3 < uplus
    quartic = a != 0.0
    for su, uf in [(1.0, 0.1 if quartic else 0.47), (-1.0, 0.68 if quartic else 0.1)]:
        g = oracle.geomu(u0, uf, mu0, a, l, q2, 0, su, sm)
        assert g["warned"] == 1
        want = (0.0 if su == 1.0 else uplus) if quartic else (u3 if su == 1.0 else u2)
        assert abs(g["u0_used"] - want) <= 1e-12, (su, g, want)
        # from the exact edited start (the solver treats u0 = u2, u3 as the root)
        iu_q = _iu(a, l, q2, want, g["uf_used"])
        assert abs(abs(g["iu"]) - iu_q) <= 1e-9 * iu_q, (su, g, iu_q)
        imu_q = _imu_path(a, l, q2, mu0, g["muf"], g["tpm"], sm)
        assert abs(imu_q - iu_q) <= 1e-8 * iu_q, (su, g, imu_q, iu_q)


@pytest.mark.parametrize("case", FORBIDDEN)
def test_geor_warning_edit(oracle, case):
    """GEOR moves a forbidden start to u3 (su = 1) or u2, cubic and quartic
    alike, and inverts from there: the u integral from the edited u0 to its
    uf equals its IMU, which equals the quadrature of the mu path."""
    a, l, q2, u0, mu0, sm = case
    r = _roots(a, l, q2)
    u2, u3 = r[1], r[2]
    for su in (1.0, -1.0):
        edited = u3 if su == 1.0 else u2
        end = (1.0 / (1.0 + math.sqrt(1.0 - a * a)) - 1e-3) if su == 1.0 else 1e-3
        # a reachable final mu: GEOMU from just inside the allowed side of the
        # edited start, a fraction of the way to `end`
        start = float(edited) + (1e-6 if su == 1.0 else -1e-6)
        g = oracle.geomu(start, start + 0.4 * (end - start), mu0, a, l, q2, 0, su, sm)
        assert g["warned"] == 0 and g["ncase"] != 0
        o = oracle.geor(u0, mu0, g["muf"], a, l, q2, g["tpm"], su, sm)
        assert o["warned"] == 1 and abs(o["u0_used"] - edited) <= 1e-12, (su, o)
        imu_q = _imu_path(a, l, q2, mu0, g["muf"], g["tpm"], sm)
        assert abs(abs(o["imu"]) - imu_q) <= 1e-9 * imu_q, (su, o, imu_q)
        iu_q = _iu(a, l, q2, edited, o["uf"])
        assert abs(iu_q - imu_q) <= 1e-8 * imu_q, (su, o, iu_q, imu_q)
        assert (o["uf"] - float(edited)) * su > 0, (su, o)


def test_geokerr_camera_invariants(oracle):
    """(alpha, beta) on the observer's sky at mu0 -> (L, Q2) with E = 1
    (Bardeen 1973): pinned by what a wrong mapping breaks rather than by its
    formula: at a = 0 spherical symmetry makes the total impact parameter
    L^2 + Q2 = alpha^2 + beta^2 for every mu0; an equatorial observer
    (mu0 = 0) sees L = -alpha, Q2 = beta^2; in general the polar potential
    at the observer, Theta(mu0) = Q2 + a^2 mu0^2 - L^2 mu0^2 / (1 - mu0^2),
    is beta^2 and |L| = |alpha| sqrt(1 - mu0^2).  sm follows the sign of
    beta; values below 1e-5 (1e-10 for Q2) are set to zero as the driver
    does."""
    rng = np.random.default_rng(7)
    for _ in range(500):
        al, be = rng.uniform(-12, 12, 2)
        mu0 = rng.uniform(0.01, 0.99)
        a = rng.uniform(0.05, 0.99)
        l, q2, aa, su, sm = oracle.geokerr_camera(al, be, mu0, 0.0)
        assert aa == 0.0 and su == 1.0
        assert abs(l * l + q2 - (al * al + be * be)) <= 1e-12 * (al * al + be * be)
        l, q2, aa, su, sm = oracle.geokerr_camera(al, be, 0.0, a)
        assert l == -al and abs(q2 - be * be) <= 1e-15 * be * be and aa == a
        l, q2, aa, su, sm = oracle.geokerr_camera(al, be, mu0, a)
        theta = q2 + a * a * mu0 ** 2 - l * l * mu0 ** 2 / (1 - mu0 ** 2)
        assert abs(theta - be * be) <= 1e-12 * max(be * be, 1.0)
        assert abs(abs(l) - abs(al) * math.sqrt(1 - mu0 ** 2)) <= 1e-14 * abs(al)
        assert np.sign(l) == -np.sign(al)
        assert sm == (1.0 if be >= 0 else -1.0)
    l, q2, aa, _, _ = oracle.geokerr_camera(5e-6, 1e-6, 0.3, 5e-6)
    assert l == 0.0 and q2 == 0.0 and aa == 0.0
