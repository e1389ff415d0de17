"""Pins for the oracle's GEOKERR GEOR (oracle/geor_oracle.c): u at a given
final mu, the inverse of GEOMU.

No Fortran build exists here, so GEOR is pinned by what does not restate it:

* round trip: GEOMU (Table 1, mu from u; itself pinned against quadrature in
  test_oracle_geomu.py) carries a geodesic from u0 to uf and reports muf and
  the mu turning points tpm; GEOR (Table 2, u from mu) from the same start
  must land on that uf, with the same case and radial turning count;
* Mino-time identity: GEOR's IMU equals an mpmath quadrature of the mu path
  (tpm turning points, the roots of M(mu) found independently), and the u
  integral from u0 to GEOR's uf (through the radial turning point when
  tpr = 1) equals IMU;
* the sentinel cases: an equatorial ray (Q2 = 0, L^2 >= a^2) and a final mu
  outside the allowed band give uf = -1, case 0.

Inputs are physical: M(mu0) >= 0, U(u0) >= 0, u0 below the horizon's
u+, and uf on the side of u0 that su points to (before the first radial
turning point, or after it for tpr = 1).

The equal-root cubic branch (a measure-zero set of constants), row 2 with
FIRSTPT off and the forbidden-region input edits, which these physical
inputs do not reach, are pinned in test_oracle_geokerr_edges.py.
"""
import numpy as np
import pytest

from test_oracle_geomu import _imu_path, mpmath

KINDS = ("general", "a0", "q20", "vortical")


def _u_poly(a, l, q2):
    cc = a * a - q2 - l * l
    dd = 2 * ((a - l) ** 2 + q2)
    ee = -a * a * q2
    return cc, dd, ee


def _turning_u(a, l, q2, u0, uplus):
    """Smallest real root of U(u) in (u0, u+), or None."""
    cc, dd, ee = _u_poly(a, l, q2)
    coefs = [ee, dd, cc, 0.0, 1.0] if ee != 0.0 else [dd, cc, 0.0, 1.0]
    r = np.roots(coefs)
    r = np.sort(r[np.abs(r.imag) < 1e-12].real)
    r = r[(r > u0) & (r < uplus)]
    return float(r[0]) if r.size else None


def physical_cases(n, seed, kinds=KINDS):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        kind = kinds[len(out) % len(kinds)]
        a = 0.0 if kind == "a0" else rng.uniform(0.05, 0.99)
        if kind == "q20":
            q2, l = 0.0, rng.uniform(-0.95, 0.95) * a
        elif kind == "vortical":
            q2, l = rng.uniform(-2.0, -0.01), rng.uniform(-1.0, 1.0)
        else:
            q2, l = rng.uniform(0.2, 40.0), rng.uniform(-7.0, 7.0)
        mu0 = rng.uniform(-0.99, 0.99)
        if q2 + (a * a - q2 - l * l) * mu0 ** 2 - a * a * mu0 ** 4 < 0.0:
            continue
        uplus = 1.0 / (1.0 + np.sqrt(1.0 - a * a))
        u0 = 10 ** rng.uniform(-3.0, -0.4)
        cc, dd, ee = _u_poly(a, l, q2)
        if u0 >= uplus or 1 + cc * u0 ** 2 + dd * u0 ** 3 + ee * u0 ** 4 < 0.0:
            continue
        ut = _turning_u(a, l, q2, u0, uplus)
        opts = [(1.0, 0), (-1.0, 0)] + ([(1.0, 1)] if ut is not None else [])
        su, tpr = opts[rng.integers(len(opts))]
        if (su, tpr) == (1.0, 0):
            hi = ut if ut is not None else uplus
            uf = u0 + (hi - u0) * rng.uniform(0.02, 0.98)
        elif tpr == 1:
            uf = ut * rng.uniform(0.02, 0.98)
        else:
            uf = u0 * rng.uniform(0.02, 0.98)
        sm = 1.0 if rng.uniform() < 0.5 else -1.0
        out.append((kind, u0, uf, mu0, a, l, q2, tpr, su, sm))
    return out


def _round_trip(oracle, case):
    kind, u0, uf, mu0, a, l, q2, tpr, su, sm = case
    g = oracle.geomu(u0, uf, mu0, a, l, q2, tpr, su, sm)
    r = oracle.geor(g["u0_used"], mu0, g["muf"], a, l, q2, g["tpm"], su, sm)
    return g, r


def test_round_trip_with_geomu(oracle):
    mpmath.mp.dps = 30
    cases = physical_cases(2400, 5)
    seen = set()
    geomu_tpm_quirk = 0
    for case in cases:
        kind, u0, uf, mu0, a, l, q2, tpr, su, sm = case
        g, r = _round_trip(oracle, case)
        assert g["ncase"] != 0 and not g["warned"], case
        # the target is GEOMU's final u: it clamps uf to a turning point it
        # would otherwise pass
        want = g["uf_used"]
        if not abs(r["uf"] - want) <= 1e-8 * want:
            # GEOMU's own mu turning-point count is not that of its mu path in
            # two places of the reference: with Q2 = 0 it never updates TPM,
            # and on a vortical orbit with mu0 < 0 its count takes SM = -1 as
            # moving away from the (negative) muplus.  Then GEOR, given that
            # count, follows a different mu path: its IMU must be that path's
            # quadrature, and GEOMU's I_u must not be.
            q = _imu_path(a, l, q2, mu0, g["muf"], g["tpm"], sm)
            assert kind in ("q20", "vortical") and mu0 * sm > 0, (case, g, r)
            assert abs(abs(r["imu"]) - q) <= 1e-9 * q and abs(g["iu"] - q) > 1e-6 * q, (case, g, r, q)
            geomu_tpm_quirk += 1
            continue
        assert r["ncase"] == g["ncase"], (case, g["ncase"], r["ncase"])
        if r["ncase"] in (1, 2, 7, 8):
            assert r["tpr"] == g["tpr"], (case, g["tpr"], r["tpr"])
        seen.add((r["ncase"], g["tpr"]))
    assert geomu_tpm_quirk <= 0.01 * len(cases), geomu_tpm_quirk
    # Table-1 rows with and without a radial turning point
    for need in [(1, 0), (1, 1), (2, 0), (3, 0), (5, 0), (6, 0), (7, 0), (7, 1), (8, 0)]:
        assert need in seen, (need, sorted(seen))


def _iu_path(a, l, q2, u0, uf, tpr):
    """Mino time along u: int du / sqrt(U) from u0 to uf, through the first
    turning point above u0 when tpr = 1 (found to mpmath precision)."""
    cc, dd, ee = (mpmath.mpf(x) for x in _u_poly(a, l, q2))
    U = lambda u: 1 + cc * u * u + dd * u ** 3 + ee * u ** 4
    f = lambda u: 1 / mpmath.sqrt(abs(U(u)))
    if tpr == 0:
        return float(abs(mpmath.quad(f, [u0, uf])))
    uplus = 1 / (1 + mpmath.sqrt(1 - mpmath.mpf(a) ** 2))
    ut = mpmath.findroot(U, _turning_u(a, l, q2, u0, float(uplus)))
    return float(mpmath.quad(f, [u0, ut]) + mpmath.quad(f, [uf, ut]))


def test_mino_time_identity(oracle):
    mpmath.mp.dps = 30
    cases = physical_cases(400, 9)
    checked = {}
    for case in cases:
        kind, u0, uf, mu0, a, l, q2, tpr, su, sm = case
        g, r = _round_trip(oracle, case)
        key = (r["ncase"], g["tpr"])
        if checked.get(key, 0) >= 3 or r["ncase"] == 0:
            continue
        imu_q = _imu_path(a, l, q2, mu0, g["muf"], g["tpm"], sm)
        assert abs(r["imu"] - imu_q) <= 1e-9 * abs(imu_q) + 1e-12, (case, r["imu"], imu_q)
        iu_q = _iu_path(a, l, q2, g["u0_used"], r["uf"], g["tpr"])
        assert abs(iu_q - r["imu"]) <= 1e-8 * abs(iu_q) + 1e-12, (case, iu_q, r["imu"])
        checked[key] = checked.get(key, 0) + 1
    assert sum(checked.values()) >= 15, checked


def test_sentinel_cases(oracle):
    # Q2 = 0 with L^2 >= a^2: mu stays on the equator, nothing to invert
    r = oracle.geor(1e-3, 0.0, 0.0, 0.9, 3.0, 0.0, 0, 1.0, 1.0)
    assert r["uf"] == -1.0 and r["ncase"] == 0 and r["imu"] == 0.0
    # vortical orbit (Q2 < 0): |mu| stays in [sqrt(mneg), sqrt(mpos)]; a
    # final mu inside the forbidden band has no solution
    a, l, q2, mu0 = 0.9, 0.1, -0.1, 0.8   # |mu| in [0.354, 0.993]
    assert q2 + (a * a - q2 - l * l) * mu0 ** 2 - a * a * mu0 ** 4 >= 0.0
    r = oracle.geor(1e-2, mu0, 0.05, a, l, q2, 0, 1.0, 1.0)
    assert r["uf"] == -1.0 and r["ncase"] == 0


def test_rows_return_minus_one_past_the_horizon(oracle):
    # A plunging ray (row 5, no radial turning point) asked for a mu it only
    # reaches after more Mino time than remains before the horizon
    case = next(c for c in physical_cases(200, 13, ("general",))
                if oracle.geomu(*c[1:]).get("ncase") == 5)
    kind, u0, uf, mu0, a, l, q2, tpr, su, sm = case
    g = oracle.geomu(u0, uf, mu0, a, l, q2, tpr, su, sm)
    r = oracle.geor(u0, mu0, g["muf"], a, l, q2, g["tpm"] + 200, su, sm)
    assert r["uf"] == -1.0 and r["ncase"] == 0


def test_batch_equals_single(oracle):
    cases = physical_cases(64, 21)
    g = [oracle.geomu(*c[1:]) for c in cases]
    cols = list(zip(*cases))
    uf, imu, tpr, ncase = oracle.geor_batch([x["u0_used"] for x in g], cols[3], [x["muf"] for x in g], cols[4],
                                            cols[5], cols[6], [x["tpm"] for x in g], cols[8], cols[9])
    for i, c in enumerate(cases):
        r = oracle.geor(g[i]["u0_used"], c[3], g[i]["muf"], c[4], c[5], c[6], g[i]["tpm"], c[8], c[9])
        assert (uf[i], imu[i], tpr[i], ncase[i]) == (r["uf"], r["imu"], r["tpr"], r["ncase"])
