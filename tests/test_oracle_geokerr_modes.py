"""The GEOKERR driver's USEGEOR and MUFILL modes in the oracle
(orc_gk_geokerr_full, geokerr_original.f:357-706).

No Fortran build pins these samples; what fixes them:
- FINDMROOTS against the roots of M(mu) = Q2 + (a^2 - Q2 - L^2) mu^2 - a^2 mu^4;
- INDEP_MUF against a brute-force walk of the mu path (bouncing between
  the turning points), which shares none of its turning-point algebra;
- every GEOR-placed sample (USEGEOR, MUFILL) against a GEOMU + GEOPHITIME
  solve at the same u: mu to 1e-9, the affine parameter, t and phi to 1e-9
  relative -- a wrong RF piece (RFFU1, RFFMU2), turning-point count or
  cached FIRSTPT value shows up here;
- the affine parameter increasing along every output, which fixes where
  the MUFILL samples sit between the two legs;
- the eval_lambda configuration (MUFILL on, NPTS = KEXT = 0) against the
  driver restatement behind geokerr_eval_lambda, sample for sample.
"""
import numpy as np
import pytest

from gk_cases import DRIVER_OPTS, fresh_point, geodesics
from oracle import np_oracle as O


def test_findmroots_are_roots_of_M():
    rng = np.random.default_rng(3)
    for _ in range(500):
        a, mu0 = rng.uniform(0.05, 0.998), rng.uniform(-0.95, 0.95)
        al, be = rng.uniform(-12, 12, 2)
        l = -al * np.sqrt(1 - mu0 ** 2)
        q2 = be ** 2 - (a * a - al * al) * mu0 ** 2
        mm, mp = O.findmroots(q2, l, a, mu0)
        r = np.roots([-a * a, a * a - q2 - l * l, q2])       # in x = mu^2
        xp = min(max(r.real), 1.0)
        if min(r.real) > 0:          # asymmetric: both turning points on mu0's side
            assert mp == pytest.approx(np.sign(mu0) * max(np.sqrt(xp), abs(mu0)), rel=1e-12)
            assert abs(mm) == pytest.approx(min(np.sqrt(min(r.real)), abs(mu0)), rel=1e-12)
        else:
            assert mp == pytest.approx(max(np.sqrt(xp), mu0), rel=1e-12)
            assert mm == -mp


def _walk(mm, mp, mu0, a1, n, muf, s):
    segs, cur, tgt = [], mu0, (mp if a1 == 1 else mm)
    for _ in range(n):
        segs.append((cur, tgt))
        cur, tgt = tgt, (mm if tgt == mp else mp)
    segs.append((cur, muf))
    tot = sum(abs(y - x) for x, y in segs)
    s = s * tot
    for j, (x, y) in enumerate(segs):
        if s <= abs(y - x) or j == len(segs) - 1:
            return x + np.sign(y - x) * s, j
        s -= abs(y - x)


def test_indep_muf_walks_the_mu_path():
    rng = np.random.default_rng(2)
    checked = 0
    for _ in range(3000):
        mp = rng.uniform(0.1, 1.0)
        mm = -mp if rng.random() < 0.5 else rng.uniform(0.01, mp)
        mu0, tpm0 = rng.uniform(mm, mp), int(rng.integers(0, 3))
        sm, n = float(rng.choice([-1.0, 1.0])), int(rng.integers(0, 4))
        a1 = int(sm * (-1) ** tpm0)
        first = mp if a1 == 1 else mm
        if n == 0:
            muf = mu0 + rng.uniform() * (first - mu0)
        else:
            last = first if n % 2 == 1 else (mm if first == mp else mp)
            muf = last + rng.uniform() * ((mm if last == mp else mp) - last)
        npts, off = int(rng.integers(1, 12)), float(rng.choice([0.0, 0.5]))
        for k in range(1, npts + 1):
            mun, tpmk = O.indep_muf(mm, mp, mu0, muf, tpm0, tpm0 + n, sm, k, npts, off)
            x, j = _walk(mm, mp, mu0, a1, n, muf, (k - off) / npts)
            assert mun == pytest.approx(x, abs=1e-12)
            assert tpmk == tpm0 + j
            checked += 1
    assert checked > 10000
    # outside [mu_-, mu_+] there is no solution: 0 and TPM0
    assert O.indep_muf(0.1, 0.5, 0.3, 0.7, 2, 3, 1.0, 1, 4, 0.5) == (0.0, 2)


def test_eval_lambda_configuration_matches_the_driver_restatement():
    g = geodesics("mufill", n=96, seed=11)
    g["uout"] = g["u0"]
    opts = dict(nup=48, npts=0, kext=0, phit=1, usegeor=0, mufill=1, offset=0.5)
    r = O.geokerr_full_batch(opts, g)
    for i in range(len(g)):
        ref = O.geokerr_driver(g["u0"][i], g["uf"][i], g["mu0"][i], g["a"][i], g["l"][i], g["q2"][i], 48,
                               su=g["su"][i], sm=g["sm"][i], tpr=1)
        m = ref["n"]
        for k in ("ufi", "mufi", "dti", "dphi", "lami", "tpmi", "tpri"):
            np.testing.assert_array_equal(r[k][i, :m], ref[k][:m], err_msg=k)
        assert r["nup_out"][i] == ref["nup_out"]


def _check_against_fresh(g, r, tot, min_checked):
    checked = 0
    for i in range(len(g)):
        uplus = 1.0 / (1.0 + np.sqrt(1.0 - g["a"][i] ** 2))
        for k in range(tot):
            u = r["ufi"][i, k]
            if not (0.0 < u < uplus) or r["tpri"][i, k] < 0:
                continue
            if abs(u - g["u0"][i]) <= 1e-9 * g["u0"][i]:
                continue   # back at the start radius: the one-point solve degenerates there
            f = fresh_point(g[i:i + 1], u, int(r["tpri"][i, k]))
            if not f["ufi"] == pytest.approx(u, rel=1e-12):
                continue
            assert f["mufi"] == pytest.approx(r["mufi"][i, k], abs=1e-9), (i, k)
            assert f["tpmi"] == r["tpmi"][i, k], (i, k)
            for key in ("lami", "dti"):
                assert f[key] == pytest.approx(r[key][i, k], rel=1e-9, abs=1e-9), (i, k, key)
            assert f["dphi"] == pytest.approx(r["dphi"][i, k], abs=1e-9), (i, k)
            checked += 1
    assert checked >= min_checked, checked


def _assert_monotone(lam):
    """The affine parameter moves one way along each geodesic (its sign
    follows su); with OFFSET = 0 and KEXT = 0 the last MUFILL sample and the
    first sample back from the turning point are the same point."""
    d = np.diff(lam, axis=1)
    assert np.all(np.all(d >= 0, axis=1) | np.all(d <= 0, axis=1))
    assert np.mean(d != 0) > 0.95


def test_usegeor_samples_lie_on_the_geodesic():
    g = geodesics("usegeor")
    o = DRIVER_OPTS["usegeor"]
    r = O.geokerr_full_batch(o, g)
    # mu evenly spaced from mu0 to MUF (:423, :452)
    k = np.arange(1, o["nup"] + 1)
    np.testing.assert_allclose(r["mufi"], g["mu0"][:, None] + (k - 0.5) * (g["muf"] - g["mu0"])[:, None] / o["nup"],
                               rtol=0, atol=1e-15)
    assert set(np.unique(r["ncase"])) >= {5, 7}
    _check_against_fresh(g, r, o["nup"], min_checked=400)
    ok = np.all((r["ufi"] > 0) & np.isfinite(r["lami"]), axis=1)
    assert ok.sum() >= 60
    _assert_monotone(r["lami"][ok])


@pytest.mark.parametrize("mode", ["mufill", "mufill_k0", "mufill_k2"])
def test_mufill_samples_lie_on_the_geodesic_in_order(mode):
    g = geodesics("mufill", seed=5)
    o = DRIVER_OPTS[mode]
    tot = o["nup"] + o["npts"] - 2 * o["kext"]
    r = O.geokerr_full_batch(o, g)
    filled = r["nup_out"] == tot
    assert filled.sum() >= 20
    assert np.all(r["npts_out"][filled] == o["npts"] - 2 * o["kext"])
    _check_against_fresh(g[filled], {k: v[filled] for k, v in r.items() if k != "args"}, tot, min_checked=400)
    lam = r["lami"][filled]
    assert np.all(np.isfinite(lam))
    _assert_monotone(lam)


def test_phit_off_leaves_times_at_zero_and_keeps_positions():
    g = geodesics("mufill", n=32, seed=9)
    on = O.geokerr_full_batch(dict(DRIVER_OPTS["nophit"], phit=1), g)
    off = O.geokerr_full_batch(DRIVER_OPTS["nophit"], g)
    for k in ("ufi", "mufi", "tpmi", "tpri"):
        np.testing.assert_array_equal(on[k], off[k])
    written = off["tpri"] >= 0
    assert np.all(off["lami"][written] == 0) and np.all(off["dti"][written] == 0)


def test_samples_outside_the_arrays_are_dropped():
    # KMAX < KEXT: the Fortran would write MUFILL samples before its arrays
    g = geodesics("mufill", n=16, seed=4)
    g["uf"] = 0.6 / (1.0 + np.sqrt(1.0 - g["a"] ** 2))
    o = dict(nup=2, npts=12, kext=5, phit=1, usegeor=0, mufill=1, offset=0.5)
    r = O.geokerr_full_batch(o, g)
    assert r["ufi"].shape == (16, 4)
    assert np.all((r["tpri"] >= -1) & (r["tpri"] <= 1))
