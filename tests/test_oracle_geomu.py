"""Pins for the oracle's GEOKERR GEOMU (oracle/geomu_oracle.c).

No Fortran build exists here (gfortran is absent), so GEOMU is pinned by
mathematics that does not restate it:

* Mino-time identity: GEOMU's I_u equals an mpmath quadrature of
  int du / sqrt(U(u)) from u0 to the (possibly clamped) final u, and the mu
  path it reports (tpm turning points, final mu) has the same Mino time
  int dmu / sqrt(M(mu)), with the turning points taken from the roots of
  M(mu) computed independently;
* the Carlson reductions (ellcubicreal, ellcubiccomplex, ellquarticreal,
  ellquarticcomplex, elldoublecomplex), every branch (p = 0, -2, -4), against
  direct quadrature of prod (a_i + b_i t)^(p_i/2);
* SNCNDN against mpmath's Jacobi elliptic functions;
* ZROOTS against numpy.roots.

The forbidden-region WARNING edits and the equal-root cubic are pinned in
test_oracle_geokerr_edges.py (Mino-time identity on the edited inputs).
Parity unpinned (only a Fortran build could pin it): ZROOTS root ordering
when two roots tie in real part.
"""
import math

import numpy as np
import pytest

mpmath = pytest.importorskip("mpmath")

U0 = min(1e-4, 1.0 / (100.0 * (144.0 ** 2 + 144.0 ** 2)))


def geokerr_batch_rays(n, seed):
    rng = np.random.default_rng(seed)
    alpha = rng.uniform(-12, 12, n)
    beta = rng.uniform(-12, 12, n)
    a = rng.uniform(0, 0.998, n)
    mu0 = rng.uniform(0.05, 0.99, n)
    uplus = 1 / (1 + np.sqrt(1 - a * a))
    uf = rng.uniform(0, 1, n) * uplus
    return alpha, beta, a, mu0, uf


def _iu_quad(a, l, q2, u0, uf, roots=None):
    """int du / sqrt(U(u)).  When the final u was clamped to a turning
    point, U is taken in product form over the solver's own (double) roots:
    the reduction integrates exactly up to that root, and U itself is
    slightly negative there after rounding."""
    cc = a * a - q2 - l * l
    dd = 2 * ((a - l) ** 2 + q2)
    ee = -a * a * q2
    if roots is None:
        f = lambda u: 1 / mpmath.sqrt(1 + cc * u * u + dd * u ** 3 + ee * u ** 4)
    else:
        u1, u2, u3, u4 = (mpmath.mpf(r) for r in roots)
        if ee == 0.0:   # cubic: U = dd (u - u1)(u2 - u)(u3 - u)
            f = lambda u: 1 / mpmath.sqrt(dd * (u - u1) * (u2 - u) * (u3 - u))
        else:
            f = lambda u: 1 / mpmath.sqrt(abs(ee) * (u - u1) * (u2 - u) * (u3 - u) * (u4 - u))
    return float(mpmath.quad(f, [u0, uf]))


def _mu_interval(a, l, q2, mu0):
    """Turning points of M(mu), as roots of a quadratic in mu^2, to the
    working precision of mpmath (the 1/sqrt endpoint singularity makes the
    path integral sensitive to the square root of any endpoint error)."""
    a, l, q2 = mpmath.mpf(a), mpmath.mpf(l), mpmath.mpf(q2)
    A = a * a
    B = a * a - q2 - l * l
    if A == 0:   # M(mu) = q2 - (q2 + l^2) mu^2
        mupl = mpmath.sqrt(q2 / (q2 + l * l))
        return -mupl, mupl
    disc = mpmath.sqrt(B * B + 4 * A * q2)
    xp = (B + disc) / (2 * A)
    xm = (B - disc) / (2 * A)
    mupl = mpmath.sqrt(max(xp, 0))
    if xm < 0:
        return -mupl, mupl
    s = 1 if mu0 >= 0 else -1
    lo, hi = sorted((s * mpmath.sqrt(xm), s * mupl))
    return lo, hi


def _imu_path(a, l, q2, mu0, muf, tpm, sm):
    a, l, q2 = mpmath.mpf(a), mpmath.mpf(l), mpmath.mpf(q2)
    Mf = lambda m: q2 + (a * a - q2 - l * l) * m * m - a * a * m ** 4

    def seg(x, y):
        if x == y:
            return 0.0
        lo, hi = min(x, y), max(x, y)
        # |M|: at a turning point known to double precision M is ~1e-17 either side
        return mpmath.quad(lambda m: 1 / mpmath.sqrt(abs(Mf(m))), [lo, hi])

    lo, hi = _mu_interval(a, l, q2, mu0)
    if tpm == 0:
        return float(seg(mu0, muf))
    first = hi if sm > 0 else lo
    other = lo if first == hi else hi
    total = seg(mu0, first)
    full = seg(lo, hi)
    total += full * (tpm - 1)
    last = first if tpm % 2 == 1 else other
    total += seg(last, muf)
    return float(total)


def test_mino_time_identity_geokerr_batch_rays(oracle):
    mpmath.mp.dps = 30
    alpha, beta, a, mu0, uf = geokerr_batch_rays(4000, 11)
    muf, tpm, ncase, iu = oracle.geomu_batch(alpha, beta, a, mu0, uf, U0)
    picked = []
    for c in (5, 6, 7):
        idx = np.nonzero(ncase == c)[0]
        picked += list(idx[:12])
    # rays with one and two mu turning points
    picked += list(np.nonzero(tpm == 2)[0][:4]) + list(np.nonzero(tpm == 1)[0][:4])
    checked = 0
    for i in sorted(set(picked)):
        l, q2, aa, su, sm = oracle.geokerr_camera(alpha[i], beta[i], mu0[i], a[i])
        o = oracle.geomu(U0, uf[i], mu0[i], aa, l, q2, 0, su, sm)
        assert o["muf"] == muf[i] and o["tpm"] == tpm[i]
        if o["warned"]:
            continue
        clamped = o["ncase"] in (1, 7) and o["uf_used"] == o["u2"]
        roots = (o["u1"], o["u2"], o["u3"], o["u4"]) if clamped else None
        iu_q = _iu_quad(aa, l, q2, o["u0_used"], o["uf_used"], roots)
        assert abs(o["iu"] - iu_q) <= 1e-9 * abs(iu_q) + 1e-12, (i, o["ncase"], o["iu"], iu_q)
        if clamped:   # the clamp point is a root of U to double precision
            cc = aa * aa - q2 - l * l
            dd = 2 * ((aa - l) ** 2 + q2)
            ee = -aa * aa * q2
            u2 = o["u2"]
            dU = 2 * cc * u2 + 3 * dd * u2 ** 2 + 4 * ee * u2 ** 3
            assert abs(1 + cc * u2 ** 2 + dd * u2 ** 3 + ee * u2 ** 4) < 1e-13 * max(1.0, abs(dU))
        imu = _imu_path(aa, l, q2, mu0[i], o["muf"], o["tpm"], sm)
        assert abs(imu - o["iu"]) <= 1e-10 * abs(o["iu"]) + 1e-12, (i, o["ncase"], o["tpm"], imu, o["iu"])
        checked += 1
    assert checked >= 30


def test_special_cases(oracle):
    mpmath.mp.dps = 20
    # a = 0 (Eq. 44 path) and q2 = 0 (Eq. 43 path): cubic U(u)
    for (a, l, q2, mu0, uf) in [(0.0, 3.0, 20.0, 0.4, 0.05), (0.0, -2.0, 30.0, 0.8, 0.1),
                                (0.6, 5.0, 0.0, 0.3, 0.08)]:
        o = oracle.geomu(U0, uf, mu0, a, l, q2, 0, 1.0, 1.0)
        assert o["ncase"] in (1, 2, 3)
        clamped = o["ncase"] == 1 and o["uf_used"] == o["u2"]
        roots = (o["u1"], o["u2"], o["u3"], o["u4"]) if clamped else None
        iu_q = _iu_quad(a, l, q2, o["u0_used"], o["uf_used"], roots)
        assert abs(o["iu"] - iu_q) <= 1e-9 * abs(iu_q)
        if q2 != 0.0:
            imu = _imu_path(a, l, q2, mu0, o["muf"], o["tpm"], 1.0)
            assert abs(imu - o["iu"]) <= 1e-7 * abs(o["iu"])
    # ncase 4: q2 = 0 and l = a gives U(u) = 1
    o = oracle.geomu(U0, 0.3, 0.5, 0.7, 0.7, 0.0, 0, 1.0, 1.0)
    assert o["ncase"] == 4 and abs(o["iu"] - (0.3 - U0)) < 1e-15 and o["muf"] == 0.0


@pytest.mark.parametrize("emmc", [0.9, 0.5, 0.1, 1e-3, -0.4])
def test_sncndn_vs_mpmath(oracle, emmc):
    mpmath.mp.dps = 25
    m = 1 - emmc
    for u in [-2.3, -0.2, 0.0, 0.4, 1.7, 5.5]:
        sn, cn, dn = oracle.sncndn(u, emmc)
        assert abs(sn - float(mpmath.re(mpmath.ellipfun("sn", u, m=m)))) < 1e-12
        assert abs(cn - float(mpmath.re(mpmath.ellipfun("cn", u, m=m)))) < 1e-12
        assert abs(dn - float(mpmath.re(mpmath.ellipfun("dn", u, m=m)))) < 1e-12


def test_zroots_vs_numpy(oracle):
    rng = np.random.default_rng(8)
    for _ in range(200):
        cc, dd, ee = rng.normal(size=3) * [10, 20, 5]
        coefs_low_first = [1.0, 0.0, cc, dd, ee]
        r = oracle.zroots(coefs_low_first)
        ref = np.roots([ee, dd, cc, 0.0, 1.0])
        ref = ref[np.argsort(ref.real, kind="stable")]
        # compare as multisets: match each oracle root to the nearest numpy root
        for z in r:
            assert np.min(np.abs(ref - z)) <= 1e-8 * max(1.0, abs(z))
        assert np.all(np.diff(r.real) >= 0)


# The [.., -4] branches of ellcubic*/ellquartic* drop the third-kind term
# whose coefficient is the linear coefficient of the polynomial under the
# root; U(u) = 1 + cc u^2 + dd u^3 + ee u^4 has none, so they are exact in
# the only form GEOPHITIME uses (fourth factor u, 1/u^2 integrand).  They are
# pinned in that form by test_minus4_reductions_on_U_polynomials; on the
# general polynomials of the tests below they would not match quadrature.


def _prod_integrand(factors):
    def f(t):
        v = mpmath.mpf(1)
        for (a, b, p) in factors:
            v *= (a + b * t) ** (mpmath.mpf(p) / 2)
        return v
    return f


@pytest.mark.parametrize("p4", [0, -2])
def test_ellcubicreal_branches(oracle, p4):
    mpmath.mp.dps = 25
    # roots u1 < 0 < y < x < u2 < u3, fourth factor a4 + b4 t positive on [y, x]
    u1, u2, u3 = -0.7, 0.6, 0.9
    y, x = 0.05, 0.45
    a4, b4 = 1.3, 0.5
    p = [-1, -1, -1, p4, 0]
    val, rff = oracle.ell("ellcubicreal", p, -u1, 1, u2, -1, u3, -1, a4, b4, y, x)
    if p4 != 0:
        # the -2/-4 branches reuse the RF piece of a p4 = 0 call, as GEOPHITIME does
        _, rff0 = oracle.ell("ellcubicreal", [-1, -1, -1, 0, 0], -u1, 1, u2, -1, u3, -1, a4, b4, y, x)
        val, _ = _ell_with_rff("ellcubicreal", oracle, p, rff0, -u1, 1, u2, -1, u3, -1, a4, b4, y, x)
    fac = [(-u1, 1, -1), (u2, -1, -1), (u3, -1, -1), (a4, b4, p4)]
    ref = float(mpmath.quad(_prod_integrand(fac), [y, x]))
    assert abs(val - ref) <= 1e-12 * abs(ref), (val, ref)


def _ell_with_rff(kind, oracle, p, rff_in, *args):
    import ctypes as C
    pp = (C.c_int * 5)(*p)
    rff = C.c_double(rff_in)
    f = getattr(oracle._glib(), "orc_" + kind)
    *coef, y, x = args
    return f(pp, *[float(c) for c in coef], C.byref(rff), float(y), float(x)), rff.value


@pytest.mark.parametrize("p5", [0, -2])
def test_ellquarticreal_branches(oracle, p5):
    mpmath.mp.dps = 25
    u1, u2, u3, u4 = -0.8, 0.5, 0.7, 1.4
    y, x = 0.02, 0.4
    a5, b5 = 0.9, 0.6
    p = [-1, -1, -1, -1, p5]
    _, rff0 = oracle.ell("ellquarticreal", [-1, -1, -1, -1, 0], -u1, 1, u2, -1, u3, -1, u4, -1, a5, b5, y, x)
    val, _ = _ell_with_rff("ellquarticreal", oracle, p, rff0, -u1, 1, u2, -1, u3, -1, u4, -1, a5, b5, y, x)
    fac = [(-u1, 1, -1), (u2, -1, -1), (u3, -1, -1), (u4, -1, -1), (a5, b5, p5)]
    ref = float(mpmath.quad(_prod_integrand(fac), [y, x]))
    assert abs(val - ref) <= 1e-12 * abs(ref), (val, ref)


def _quad_factor(f, g, h, p):
    return lambda t: (f + g * t + h * t * t) ** (mpmath.mpf(p) / 2)


@pytest.mark.parametrize("p4", [0, -2])
def test_ellcubiccomplex_branches(oracle, p4):
    mpmath.mp.dps = 25
    u1 = -0.6
    f, g, h = 1.2, -0.4, 1.0   # no real roots
    a4, b4 = 1.5, 0.7
    y, x = 0.1, 0.8
    p = [-1, -1, -1, p4, 0]
    _, rff0 = oracle.ell("ellcubiccomplex", [-1, -1, -1, 0, 0], -u1, 1, a4, b4, f, g, h, y, x)
    val, _ = _ell_with_rff("ellcubiccomplex", oracle, p, rff0, -u1, 1, a4, b4, f, g, h, y, x)
    q = _quad_factor(f, g, h, -1)
    integrand = lambda t: (-u1 + t) ** -0.5 * q(t) * (a4 + b4 * t) ** (mpmath.mpf(p4) / 2)
    ref = float(mpmath.quad(integrand, [y, x]))
    assert abs(val - ref) <= 1e-12 * abs(ref), (val, ref)


@pytest.mark.parametrize("p5", [0, -2])
def test_ellquarticcomplex_branches(oracle, p5):
    mpmath.mp.dps = 25
    u1, u4 = -0.6, 1.3
    f, g, h = 1.2, -0.4, 1.0
    a5, b5 = 1.5, 0.7
    y, x = 0.1, 0.8
    p = [-1, -1, -1, -1, p5]
    _, rff0 = oracle.ell("ellquarticcomplex", [-1, -1, -1, -1, 0], -u1, 1, u4, -1, a5, b5, f, g, h, y, x)
    val, _ = _ell_with_rff("ellquarticcomplex", oracle, p, rff0, -u1, 1, u4, -1, a5, b5, f, g, h, y, x)
    q = _quad_factor(f, g, h, -1)
    integrand = lambda t: ((-u1 + t) * (u4 - t)) ** -0.5 * q(t) * (a5 + b5 * t) ** (mpmath.mpf(p5) / 2)
    ref = float(mpmath.quad(integrand, [y, x]))
    assert abs(val - ref) <= 1e-12 * abs(ref), (val, ref)


@pytest.mark.parametrize("p5", [0, -2, -4])
def test_elldoublecomplex_branches(oracle, p5):
    mpmath.mp.dps = 25
    f1, g1, h1 = 1.0, 0.3, 0.8
    f2, g2, h2 = 0.7, -0.5, 1.1
    a5, b5 = 1.2, 0.4
    y, x = 0.05, 0.9
    p = [-1, -1, -1, -1, p5]
    _, rff0 = oracle.ell("elldoublecomplex", [-1, -1, -1, -1, 0], f1, g1, h1, f2, g2, h2, a5, b5, y, x)
    val, _ = _ell_with_rff("elldoublecomplex", oracle, p, rff0, f1, g1, h1, f2, g2, h2, a5, b5, y, x)
    q1, q2 = _quad_factor(f1, g1, h1, -1), _quad_factor(f2, g2, h2, -1)
    integrand = lambda t: q1(t) * q2(t) * (a5 + b5 * t) ** (mpmath.mpf(p5) / 2)
    ref = float(mpmath.quad(integrand, [y, x]))
    assert abs(val - ref) <= 1e-12 * abs(ref), (val, ref)


def _u_roots(a, l, q2):
    cc, dd, ee = a * a - q2 - l * l, 2 * ((a - l) ** 2 + q2), -a * a * q2
    coefs = [ee, dd, cc, 0.0, 1.0] if ee != 0.0 else [dd, cc, 0.0, 1.0]
    r = np.roots(coefs)
    return cc, dd, ee, r


@pytest.mark.parametrize("kind", ["cubicreal", "cubiccomplex", "quarticreal", "quarticcomplex"])
def test_minus4_reductions_on_U_polynomials(oracle, kind):
    """int dt / (t^2 sqrt(P(t))) with P the factored U(u) of a geodesic, the
    way GEOPHITIME calls the -4 branch (geokerr_original.f:1759-1761,
    1878-1880, 1989-1991, 2186-2188), against mpmath quadrature."""
    mpmath.mp.dps = 30
    if kind == "cubicreal":          # a = 0, three real roots (row 1)
        cc, dd, ee, r = _u_roots(0.0, 6.0, 10.0)
        u1, u2, u3 = np.sort(r.real)
        y, x = 0.02, u2
        args = (-u1, 1, u2, -1, u3, -1, 0.0, 1.0)
        P = lambda t: dd * (t - u1) * (u2 - t) * (u3 - t)
        p = [-1, -1, -1, -4, 0]
        scale = math.sqrt(dd)
    elif kind == "cubiccomplex":     # a = 0, one real root (row 3)
        cc, dd, ee, r = _u_roots(0.0, 2.0, 3.0)
        u1 = float(r[np.abs(r.imag) < 1e-12].real[0])
        f = -1 / dd / u1
        g = f / u1
        y, x = 0.02, 0.9
        args = (-u1, 1, 0.0, 1.0, f, g, 1.0)
        P = lambda t: dd * (t - u1) * (f + g * t + t * t)
        p = [-1, -1, -1, -4, 0]
        scale = math.sqrt(dd)
    elif kind == "quarticreal":      # four real roots (row 7)
        cc, dd, ee, r = _u_roots(0.9, 6.0, 10.0)
        u1, u2, u3, u4 = np.sort(r.real)
        y, x = 0.02, u2
        args = (-u1, 1, u2, -1, u3, -1, u4, -1, 0.0, 1.0)
        P = lambda t: abs(ee) * (t - u1) * (u2 - t) * (u3 - t) * (u4 - t)
        p = [-1, -1, -1, -1, -4]
        scale = math.sqrt(abs(ee))
    else:                            # two real roots (row 5)
        cc, dd, ee, r = _u_roots(0.9, 2.0, 3.0)
        real = np.sort(r[np.abs(r.imag) < 1e-12].real)
        u1, u4 = real[0], real[-1]
        qs = 1.0
        f = -qs / abs(ee) / u1 / u4
        g = (u4 + u1) / u1 / u4 * f
        y, x = 0.02, 0.6
        args = (-u1, 1, qs * u4, -qs, 0.0, 1.0, f, g, 1.0)
        P = lambda t: abs(ee) * (t - u1) * (u4 - t) * (f + g * t + t * t)
        p = [-1, -1, -1, -1, -4]
        scale = math.sqrt(abs(ee))
    name = "ell" + kind
    p0 = list(p)
    p0[3 if kind.startswith("cubic") else 4] = 0
    _, rff0 = oracle.ell(name, p0, *args, y, x)
    val, _ = _ell_with_rff(name, oracle, p, rff0, *args, y, x)
    ref = float(mpmath.quad(lambda t: 1 / (t * t * mpmath.sqrt(P(t))), [y, x])) * scale
    assert abs(val - ref) <= 1e-11 * abs(ref), (kind, val, ref)
