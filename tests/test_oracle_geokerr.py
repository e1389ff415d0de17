"""Pins for the oracle's GEOPHITIME, GEOKERR driver and geokerr_ref_api
(oracle/geokerr_oracle.c).

* The reference's only stored Fortran output (tests/golden/
  geokerr_eval_lambda_fortran.npz, from gkval's GEOKERR cache) must be
  reproduced bit for bit: every one of its 4096 u, mu, t, phi and affine
  samples and both turning-point indices.  Its mu0 = 0.8 lies outside the
  allowed mu range of its constants (the reference clamps mu+ to mu0), so
  it pins the code path, not the physics.
* Physics, independently of the restatement: the driver's samples
  (u, mu, t, phi, affine lambda) against a DOP853 integration of the
  second-order Mino-time geodesic equations (E = 1, M = 1):
  r'' = R'(r)/2, mu'' = Theta'(mu)/2, t' and phi' from the Kerr potentials,
  lambda' = r^2 + a^2 mu^2, matched at the sample's u on the right side of
  the radial turning point.  Cases cover Table 1 rows 1, 3, 5, 7 and the
  q2 = 0 and a = 0 branches.
* geokerr_phi_omega against the inverse Kerr metric (M = 1, E = 1):
  dphi/dMino = Sigma (g^{phi phi} L - g^{t phi}).
* geokerr_estimate_periods: Ttheta against the affine distance between
  the first two mu turning points of the ODE, to the driver's sample
  spacing; Tr is always -1 because the driver only writes TPRI = 0 or 1.
"""
import math
import os

import numpy as np
import pytest

from oracle import np_oracle as oracle

scipy_integrate = pytest.importorskip("scipy.integrate")
scipy_optimize = pytest.importorskip("scipy.optimize")

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "geokerr_eval_lambda_fortran.npz")


def test_eval_lambda_reproduces_fortran_fixture_bitwise():
    f = np.load(GOLDEN, allow_pickle=False)
    u0, mu0, uf, a, lz, q2, t0 = (float(v) for v in f["inputs"])
    r = oracle.geokerr_eval_lambda(u0, mu0, uf, a, lz, q2, t0, float(f["lam0"]), float(f["dlam"]),
                                   int(f["s"]))
    assert r["info"] == 0
    for k in ("u", "mu", "t", "phi", "affine"):
        np.testing.assert_array_equal(r[k], f[k], err_msg=k)
    assert r["tpmi"] == int(f["tpmi"]) and r["tpri"] == int(f["tpri"])


def _potentials(a, l, q):
    R = lambda r: (r * r + a * a - a * l) ** 2 - (r * r - 2 * r + a * a) * (q + (l - a) ** 2)
    dR = lambda r: 4 * r * (r * r + a * a - a * l) - (2 * r - 2) * (q + (l - a) ** 2)
    Th = lambda m: q - (q + l * l - a * a) * m * m - a * a * m ** 4
    dTh = lambda m: -2 * (q + l * l - a * a) * m - 4 * a * a * m ** 3
    return R, dR, Th, dTh


def _ode(u0, mu0, a, l, q, su, sm, tau_end, r_max=None):
    R, dR, Th, dTh = _potentials(a, l, q)

    def f(tau, y):
        r, pr, m, pm, t, ph, lam = y
        D = r * r - 2 * r + a * a
        return [pr, 0.5 * dR(r), pm, 0.5 * dTh(m),
                (r * r + a * a) * (r * r + a * a - a * l) / D + a * (l - a * (1 - m * m)),
                a * (r * r + a * a - a * l) / D - a + l / (1 - m * m),
                r * r + a * a * m * m]

    r0 = 1 / u0
    y0 = [r0, -su * math.sqrt(max(R(r0), 0.0)), mu0, sm * math.sqrt(max(Th(mu0), 0.0)), 0.0, 0.0, 0.0]
    rplus = 1 + math.sqrt(1 - a * a)

    def near_horizon(tau, y):
        return y[0] - 1.2 * rplus
    near_horizon.terminal = True

    def far_out(tau, y):   # r reaches infinity at finite Mino time
        return y[0] - (r_max if r_max else 4.0 * r0)
    far_out.terminal = True
    return scipy_integrate.solve_ivp(f, (0, tau_end), y0, method="DOP853", rtol=1e-13, atol=1e-13,
                                     dense_output=True, events=[near_horizon, far_out])


# (u0, uf, mu0, a, l, q2); the ODE runs until the ray leaves 1.2 r+ .. 2 / min(u0, uf)
ODE_CASES = [
    (0.1, 0.01, 0.5, 0.5, 2.0, 4.0),     # quartic, outgoing (row 5)
    (0.02, 0.5, 0.3, 0.9, 6.0, 10.0),    # quartic real roots, radial turning point (row 7)
    (0.02, 0.9, 0.3, 0.9, 2.0, 3.0),     # quartic, plunging (row 5)
    (0.02, 0.5, 0.3, 0.0, 6.0, 10.0),    # a = 0 cubic real roots with turning point (row 1)
    (0.02, 0.9, 0.3, 0.0, 2.0, 3.0),     # a = 0 cubic, plunging (row 3)
    (0.02, 0.5, 0.0, 0.7, 6.0, 0.0),     # q2 = 0 equatorial scatter
    (0.05, 0.3, -0.4, 0.6, -3.0, 8.0),   # retrograde, mu0 < 0
]


@pytest.mark.parametrize("case", ODE_CASES)
def test_driver_samples_match_geodesic_ode(case):
    u0, uf, mu0, a, l, q2 = case
    d = oracle.geokerr_driver(u0, uf, mu0, a, l, q2, nup=256)
    n = d["n"]
    assert n > 16
    sol = _ode(u0, mu0, a, l, q2, d["su"], d["sm"], 5.0, r_max=2.0 / min(u0, uf))
    ts = np.linspace(0.0, sol.t[-1], 40001)
    Y = sol.sol(ts)
    uu = 1.0 / Y[0]
    flip = np.nonzero(np.sign(Y[1]) != np.sign(Y[1][1]))[0]
    itp = flip[0] if flip.size else ts.size
    checked = 0
    for k in range(n):
        uk, leg = d["ufi"][k], d["tpri"][k]
        seg = slice(0, itp) if leg == 0 else slice(itp, ts.size)
        g = uu[seg] - uk
        idx = np.nonzero(np.sign(g[:-1]) != np.sign(g[1:]))[0]
        if idx.size == 0:
            continue
        tt = ts[seg]
        tk = scipy_optimize.brentq(lambda x: 1.0 / sol.sol(x)[0] - uk, tt[idx[0]], tt[idx[0] + 1], xtol=1e-15)
        y = sol.sol(tk)
        for name, got, want in (("mu", d["mufi"][k], y[2]), ("t", d["dti"][k], y[4]),
                                ("phi", d["dphi"][k], y[5]), ("lambda", d["lami"][k], y[6])):
            assert abs(got - want) <= 1e-8 * max(1.0, abs(want)), (name, k, got, want)
        checked += 1
    # samples inside 1.2 r+ (where dphi/du grows like 1/Delta) are beyond the
    # integration's end
    assert checked >= n // 2, (checked, n)


def test_phi_omega_matches_inverse_metric():
    rng = np.random.default_rng(7)
    for _ in range(200):
        a = rng.uniform(-0.99, 0.99)
        r = rng.uniform(3.0, 500.0)
        mu = rng.uniform(-0.95, 0.95)
        lz = rng.uniform(-8, 8)
        # inverse Kerr metric, M = 1, theta from mu
        s2 = 1 - mu * mu
        sig = r * r + a * a * mu * mu
        delta = r * r - 2 * r + a * a
        g_tphi = -2 * a * r / (sig * delta)
        g_phiphi = (delta - a * a * s2) / (sig * delta * s2)
        want = sig * (g_phiphi * lz - g_tphi)   # p_t = -1, p_phi = lz
        got = oracle.geokerr_phi_omega(1.0 / r, mu, a, lz, 1.0)
        assert abs(got - want) <= 1e-11 * max(1.0, abs(want)), (a, r, mu, lz, got, want)


def test_estimate_periods_theta_matches_ode():
    u0, mu0, a, l, q2 = 0.02, 0.1, 0.3, 1.0, 25.0   # scatters with two mu turning points
    tr, tth, info = oracle.geokerr_estimate_periods(u0, mu0, a, l, q2)
    assert info == 0 and tr == -1.0
    assert tth > 0
    d = oracle.geokerr_driver(u0, u0, mu0, a, l, q2, nup=4096, su=1.0, uf_is_u0=True)
    spacing = np.max(np.abs(np.diff(d["lami"][: d["n"]])))
    sol = _ode(u0, mu0, a, l, q2, 1.0, -1.0 if mu0 > 0 else 1.0, 5.0, r_max=1.01 / u0)
    ts = np.linspace(0, sol.t[-1], 200001)
    Y = sol.sol(ts)
    flips = np.nonzero(np.sign(Y[3][:-1]) != np.sign(Y[3][1:]))[0]
    assert flips.size >= 2
    lam_tp = [Y[6][i] for i in flips[:2]]
    want = 2 * (lam_tp[1] - lam_tp[0])
    assert abs(tth - want) <= 4 * spacing, (tth, want, spacing)


def test_estimate_periods_plunging_orbit_has_no_samples():
    # no radial turning point with uf = u0: the driver's step is zero, it
    # writes nothing, and both periods keep the -1 sentinel
    tr, tth, info = oracle.geokerr_estimate_periods(0.02, 0.3, 0.9, 2.0, 3.0)
    assert (tr, tth, info) == (-1.0, -1.0, 0)


def test_eval_lambda_contract():
    r = oracle.geokerr_eval_lambda(0.02, 0.3, 0.5, 0.9, 6.0, 10.0, 5.0, 0.25, 0.5, 300)
    assert r["info"] == 0
    np.testing.assert_array_equal(r["affine"], 0.25 + 0.5 * np.arange(300))
    assert r["u"][0] == 0.02 and r["mu"][0] == 0.3
    assert np.all(np.isfinite(r["t"])) and np.all(r["t"] >= 5.0)
    # a radial turning point exists (row 7 scatter): its grid index is set
    assert r["tpri"] >= 0
    bad = oracle.geokerr_eval_lambda(0.02, 0.3, 0.5, 0.9, 6.0, 10.0, 0.0, 0.0, 0.1, 0)
    assert bad["info"] == 1


def test_eval_lambda_batch_equals_single_calls():
    rng = np.random.default_rng(3)
    n, s = 6, 200
    u0 = rng.uniform(0.01, 0.05, n)
    mu0 = rng.uniform(-0.4, 0.4, n)
    uf = rng.uniform(0.2, 0.6, n)
    a = rng.uniform(0.1, 0.95, n)
    lz = rng.uniform(-6, 6, n)
    q2 = rng.uniform(1, 30, n)
    t0 = np.zeros(n)
    b = oracle.geokerr_eval_lambda_batch(u0, mu0, uf, a, lz, q2, t0, 0.0, 0.2, s, nthreads=2)
    for i in range(n):
        r = oracle.geokerr_eval_lambda(u0[i], mu0[i], uf[i], a[i], lz[i], q2[i], 0.0, 0.0, 0.2, s)
        for k in ("u", "mu", "t", "phi", "affine"):
            np.testing.assert_array_equal(b[k][i], r[k])
        assert b["tpmi"][i] == r["tpmi"] and b["tpri"][i] == r["tpri"]
