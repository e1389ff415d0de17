"""Pins of the FP64 forward-mode oracle of the differentiable disk render
(oracle/adjoint_oracle.c).

- ISCO: closed-form values (6 M at a = 0, the tabulated 4.2330025 M at
  a = 0.5, 1 M in the extremal limit) and its derivative against central
  differences.
- Flow: the r-scaled Carter field equals r times the reference Hamiltonian
  field (orc_kerr_rhs_d, bh_kerr.h:600-727) on the constraint surface.
- Initial data: the dual-number ZAMO tetrad equals orc_init_zamo_d
  (bh_kerr.h:897-964).
- Redshift: disk_term / (gain r^-q) equals orc_doppler_d^3
  (bh_kerr.h:833-862) for equatorial states inside the window plateau.
- Derivatives: per-pixel dI/d(a, theta_obs, q, gain) against central finite
  differences (h = 1e-5) of the oracle's own FP64 values.
"""
import math

import numpy as np
import pytest

from oracle import np_oracle as O


def test_isco_closed_forms():
    assert O.adj_isco(0.5)[0] == pytest.approx(4.2330025, abs=1e-7)
    assert O.adj_isco(1e-12)[0] == pytest.approx(6.0, abs=1e-9)
    assert O.adj_isco(1.0 - 1e-12)[0] == pytest.approx(1.0, abs=1e-3)
    for a in (0.1, 0.3, 0.6, 0.7, 0.9, 0.998):
        v, d = O.adj_isco(a)
        h = 1e-6
        fd = (O.adj_isco(a + h)[0] - O.adj_isco(a - h)[0]) / (2 * h)
        assert d == pytest.approx(fd, rel=1e-7)
        # r_isco solves the circular-orbit marginal-stability condition
        # r^2 - 6 r + 8 a sqrt(r) - 3 a^2 = 0 (prograde, M = 1)
        assert v * v - 6 * v + 8 * a * math.sqrt(v) - 3 * a * a == pytest.approx(0.0, abs=1e-9)


def test_flow_is_r_times_hamilton_flow(oracle):
    rng = np.random.default_rng(71)
    for _ in range(60):
        a = rng.uniform(0.0, 0.99)
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        s = oracle.init_zamo(1.0, a, rng.uniform(4, 300), rng.uniform(0.3, 2.8), 0.0, n)
        d = oracle.rhs(s, 1.0, a)
        f = O.adj_rhs(s[[1, 2, 5, 6]], -s[4], s[7], a)
        np.testing.assert_allclose(f, s[1] * d[[1, 2, 5, 6]], rtol=1e-9, atol=1e-12)


def test_zamo_initial_data(oracle):
    rng = np.random.default_rng(72)
    for _ in range(40):
        a = rng.uniform(0.0, 0.99)
        th = rng.uniform(0.2, 2.9)
        r = rng.uniform(10, 1000)
        n = rng.normal(size=3)
        n = (n / np.linalg.norm(n)).astype(np.float32)
        s = oracle.init_zamo(1.0, a, r, th, 0.0, n.astype(np.float64))
        E, L, pr, pth = O.adj_zamo_init(a, th, r, n)
        np.testing.assert_allclose([E, L, pr, pth], [-s[4], s[7], s[5], s[6]], rtol=1e-12, atol=1e-12)


def test_disk_term_is_doppler_cubed(oracle):
    rng = np.random.default_rng(73)
    for _ in range(50):
        a = rng.uniform(0.05, 0.95)
        rin = O.adj_isco(a)[0]
        r = rng.uniform(rin + 0.6, 19.4)
        b = rng.uniform(-8, 8)
        E, L = 1.0, b
        q, gain = rng.uniform(1.5, 3.5), rng.uniform(0.5, 2.0)
        st = np.array([0.0, r, np.pi / 2, 0.0, -E, 0.0, 0.0, L])
        g = oracle.doppler(st, 1.0, a)
        c = O.adj_disk_term(r, E, L, a, q, gain, 20.0, 0.5)
        assert c == pytest.approx(gain * r ** (-q) * g ** 3, rel=1e-12)


def test_pixel_gradients_match_finite_differences():
    p = O.adjoint_params(width=24, height=24)
    xs, ys = np.meshgrid(np.arange(24), np.arange(24))
    xs, ys = xs.ravel(), ys.ravel()
    rad, drad, nc = O.adj_pixels(p, xs, ys)
    sel = np.flatnonzero(rad > 1e-4)
    rng = np.random.default_rng(74)
    sel = rng.choice(sel, size=min(40, len(sel)), replace=False)
    h = 1e-5
    worst = 0.0
    for k, name in enumerate(("spin", "theta_obs", "emiss_q", "gain")):
        pp = O.adjoint_params(width=24, height=24, **{name: getattr(p, name) + h})
        pm = O.adjoint_params(width=24, height=24, **{name: getattr(p, name) - h})
        rp, _, ncp = O.adj_pixels(pp, xs[sel], ys[sel])
        rm, _, ncm = O.adj_pixels(pm, xs[sel], ys[sel])
        ok = (ncp == nc[sel]) & (ncm == nc[sel])
        assert ok.mean() > 0.9
        fd = (rp - rm) / (2 * h)
        an = drad[sel, k]
        err = np.abs(an - fd)[ok] / np.maximum(np.abs(fd[ok]), 1e-3 * rad[sel][ok])
        worst = max(worst, float(err.max()))
    assert worst < 1e-5, worst
