"""GPU parity of the GEOKERR driver and geokerr_ref_api (geokerr.hip):
the drop-in geokerr_eval_lambda / geokerr_estimate_periods /
geokerr_phi_omega and their np_geokerr_*_batch forms, against the
reference's stored Fortran output and the CPU oracle (oracle/
geokerr_oracle.c, itself bit-exact on that output and pinned by a geodesic
ODE in tests/test_oracle_geokerr.py).

Tolerances: the GPU contracts multiply-adds into FMAs and uses ROCm's
libm, the oracle neither, so FP64 results agree to rounding amplified by
the reductions, not bitwise: 1e-9 relative (1e-9 absolute near zero) on u,
mu, t, phi; the affine grid and the turning-point indices are exact.
"""
import os

import numpy as np
import pytest

import nullpath_amd as npa

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "geokerr_eval_lambda_fortran.npz")
TOL = 1e-9



def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda")


def _close(got, want, tol=TOL):
    got, want = np.asarray(got), np.asarray(want)
    both_nan = np.isnan(got) & np.isnan(want)
    err = np.abs(got - want) / np.maximum(1.0, np.abs(want))
    err[both_nan] = 0.0
    return err


def _cases(n, seed):
    """Mixed physical inputs: quartic rows 5 and 7 (plunge / scatter), a = 0
    cubic rows, q2 = 0, retrograde, outgoing rays."""
    rng = np.random.default_rng(seed)
    u0 = rng.uniform(0.005, 0.05, n)
    mu0 = rng.uniform(-0.6, 0.6, n)
    a = rng.uniform(0.05, 0.98, n)
    lz = rng.uniform(-7, 7, n)
    q2 = rng.uniform(0.5, 40, n)
    uf = rng.uniform(0.1, 0.9, n)
    a[::7] = 0.0
    q2[3::11] = 0.0
    mu0[3::11] = 0.0
    out = np.arange(n) % 5 == 4          # outgoing: uf < u0
    uf[out] = u0[out] * rng.uniform(0.1, 0.9, out.sum())
    # keep mu0 inside the allowed range: Theta(mu0) >= 0
    th = q2 - (q2 + lz * lz - a * a) * mu0 ** 2 - a * a * mu0 ** 4
    mu0[th < 0] = 0.0
    t0 = rng.uniform(0, 10, n)
    return u0, mu0, uf, a, lz, q2, t0


@pytest.mark.gpu
def test_drop_in_eval_lambda_matches_fortran_fixture(device):
    f = np.load(GOLDEN, allow_pickle=False)
    u0, mu0, uf, a, lz, q2, t0 = (float(v) for v in f["inputs"])
    print("[gpu] geokerr_eval_lambda on the stored Fortran case, S=4096", flush=True)
    r = npa.geokerr_eval_lambda(u0, mu0, uf, a, lz, q2, t0, float(f["lam0"]), float(f["dlam"]), int(f["s"]))
    assert r["info"] == 0
    np.testing.assert_array_equal(r["affine"], f["affine"])
    for k in ("u", "mu", "t", "phi"):
        e = _close(r[k], f[k])
        assert e.max() <= 1e-11, (k, e.max(), int(e.argmax()))
    assert r["tpmi"] == int(f["tpmi"]) and r["tpri"] == int(f["tpri"])


@pytest.mark.gpu
def test_port_solver_symbols_match_fortran_fixture(device):
    # the harness's port solver (gkval_adapters.c:661-695) passes the same
    # inputs without t0 and compares the five fields with the Fortran's
    f = np.load(GOLDEN, allow_pickle=False)
    u0, mu0, uf, a, lz, q2, t0 = (float(v) for v in f["inputs"])
    assert t0 == 0.0
    ref = npa.geokerr_eval_lambda(u0, mu0, uf, a, lz, q2, 0.0, float(f["lam0"]), float(f["dlam"]), int(f["s"]))
    for sym in ("geokerr_complete_stream_cuda", "geokerr_stream_cuda"):
        print(f"[gpu] {sym} on the stored Fortran case, S=4096", flush=True)
        code, r = npa.geokerr_port_stream(u0, uf, mu0, a, lz, q2, float(f["lam0"]), float(f["dlam"]),
                                          int(f["s"]), symbol=sym)
        assert code == 0
        for k in ("u", "mu", "t", "phi", "affine"):
            np.testing.assert_array_equal(r[k], ref[k])
            e = _close(r[k], f[k])
            assert e.max() <= 1e-11, (sym, k, e.max(), int(e.argmax()))


@pytest.mark.gpu
def test_eval_lambda_batch_matches_oracle(oracle, device):
    n, s = 96, 512
    u0, mu0, uf, a, lz, q2, t0 = _cases(n, 5)
    lam0 = np.zeros(n)
    dlam = np.full(n, 0.4)
    print(f"[gpu] np_geokerr_eval_lambda_batch n={n} s={s}", flush=True)
    g = npa.geokerr_eval_lambda_batch(_t(u0), _t(mu0), _t(uf), _t(a), _t(lz), _t(q2), _t(t0), _t(lam0),
                                      _t(dlam), s)
    g = {k: v.cpu().numpy() for k, v in g.items()}
    o = oracle.geokerr_eval_lambda_batch(u0, mu0, uf, a, lz, q2, t0, 0.0, 0.4, s, nthreads=8)
    np.testing.assert_array_equal(g["affine"], o["affine"])
    np.testing.assert_array_equal(g["info"], o["info"])
    for k in ("u", "mu", "t", "phi"):
        e = _close(g[k], o[k])
        assert e.max() <= TOL, (k, e.max(), np.unravel_index(e.argmax(), e.shape))
    np.testing.assert_array_equal(g["tpmi"], o["tpmi"])
    np.testing.assert_array_equal(g["tpri"], o["tpri"])
    # the case mix reaches both legs and both kinds of turning point
    assert (o["tpri"] >= 0).any() and (o["tpmi"] >= 0).any()


@pytest.mark.gpu
def test_eval_lambda_batch_grid_edge_cases(oracle, device):
    # negative step, zero step, grid starting past the end of the ray
    u0 = np.array([0.02, 0.02, 0.02])
    mu0 = np.array([0.3, 0.3, 0.3])
    uf = np.array([0.5, 0.5, 0.5])
    a = np.array([0.9, 0.9, 0.9])
    lz = np.array([6.0, 6.0, 6.0])
    q2 = np.array([10.0, 10.0, 10.0])
    t0 = np.zeros(3)
    lam0 = np.array([40.0, 7.0, 1e4])
    dlam = np.array([-0.5, 0.0, 1.0])
    s = 64
    print("[gpu] np_geokerr_eval_lambda_batch grid edge cases", flush=True)
    g = npa.geokerr_eval_lambda_batch(_t(u0), _t(mu0), _t(uf), _t(a), _t(lz), _t(q2), _t(t0), _t(lam0),
                                      _t(dlam), s)
    g = {k: v.cpu().numpy() for k, v in g.items()}
    for i in range(3):
        o = oracle.geokerr_eval_lambda(u0[i], mu0[i], uf[i], a[i], lz[i], q2[i], t0[i], lam0[i], dlam[i], s)
        for k in ("u", "mu", "t", "phi"):
            e = _close(g[k][i], o[k])
            assert e.max() <= TOL, (i, k, e.max())
        np.testing.assert_array_equal(g["affine"][i], o["affine"])
        assert g["tpmi"][i] == o["tpmi"] and g["tpri"][i] == o["tpri"]


@pytest.mark.gpu
def test_eval_lambda_full_size_properties(oracle, device):
    """256 geodesics x 4096 samples: the grid is exact, sample 0 is the
    start point, t never decreases along ingoing rays, and a spread of
    geodesics matches the oracle."""
    n, s = 256, 4096
    u0, mu0, uf, a, lz, q2, t0 = _cases(n, 9)
    lam0 = np.zeros(n)
    dlam = np.full(n, 0.1)
    print(f"[gpu] np_geokerr_eval_lambda_batch n={n} s={s}", flush=True)
    g = npa.geokerr_eval_lambda_batch(_t(u0), _t(mu0), _t(uf), _t(a), _t(lz), _t(q2), _t(t0), _t(lam0),
                                      _t(dlam), s)
    g = {k: v.cpu().numpy() for k, v in g.items()}
    np.testing.assert_array_equal(g["affine"], np.broadcast_to(0.1 * np.arange(s), (n, s)))
    np.testing.assert_array_equal(g["u"][:, 0], u0)
    np.testing.assert_array_equal(g["mu"][:, 0], mu0)
    ingoing = uf > u0
    dt = np.diff(g["t"][ingoing], axis=1)
    assert np.nanmin(dt) >= -1e-9
    for i in range(0, n, 37):
        o = oracle.geokerr_eval_lambda(u0[i], mu0[i], uf[i], a[i], lz[i], q2[i], t0[i], 0.0, 0.1, s)
        for k in ("u", "mu", "t", "phi"):
            e = _close(g[k][i], o[k])
            assert e.max() <= TOL, (i, k, e.max())
        assert g["tpmi"][i] == o["tpmi"] and g["tpri"][i] == o["tpri"]


@pytest.mark.gpu
def test_estimate_periods_matches_oracle(oracle, device):
    rng = np.random.default_rng(12)
    n = 24
    u0 = np.full(n, 0.02)
    mu0 = rng.uniform(-0.2, 0.2, n)
    a = rng.uniform(0.1, 0.9, n)
    lz = rng.uniform(-2, 2, n)
    q2 = rng.uniform(24, 32, n)        # near-critical scatterers with mu oscillations
    print(f"[gpu] np_geokerr_estimate_periods_batch n={n}", flush=True)
    tr, tth, info = (x.cpu().numpy() for x in
                     npa.geokerr_estimate_periods_batch(_t(u0), _t(mu0), _t(a), _t(lz), _t(q2)))
    hits = 0
    for i in range(n):
        o_tr, o_tth, o_info = oracle.geokerr_estimate_periods(u0[i], mu0[i], a[i], lz[i], q2[i])
        assert info[i] == o_info == 0
        assert tr[i] == o_tr == -1.0
        if o_tth < 0:
            assert tth[i] == o_tth
        else:
            hits += 1
            assert abs(tth[i] - o_tth) <= TOL * max(1.0, o_tth), (i, tth[i], o_tth)
    assert hits >= n // 4
    d_tr, d_tth, d_info = npa.geokerr_estimate_periods(u0[0], mu0[0], a[0], lz[0], q2[0])
    assert d_info == 0 and d_tr == tr[0] and d_tth == tth[0]


@pytest.mark.gpu
def test_phi_omega_batch_matches_oracle(oracle, device):
    rng = np.random.default_rng(13)
    n = 4096
    u0 = 10 ** rng.uniform(-4, -0.4, n)
    mu0 = rng.uniform(-0.999, 0.999, n)
    a = rng.uniform(-0.99, 0.99, n)
    lz = rng.uniform(-8, 8, n)
    q2 = rng.uniform(0, 30, n)
    u0[:8] = 0.0                     # r clamped to 1e12
    print(f"[gpu] np_geokerr_phi_omega_batch n={n}", flush=True)
    g = npa.geokerr_phi_omega_batch(_t(u0), _t(mu0), _t(a), _t(lz), _t(q2)).cpu().numpy()
    o = np.array([oracle.geokerr_phi_omega(*v) for v in zip(u0, mu0, a, lz, q2)])
    assert _close(g, o).max() <= 1e-13
    h = np.array([npa.geokerr_phi_omega(*v) for v in zip(u0[:64], mu0[:64], a[:64], lz[:64], q2[:64])])
    assert _close(h, o[:64]).max() <= 1e-13
