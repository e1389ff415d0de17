"""np_geokerr_batch (the split driver: a per-geodesic setup, then one lane
per sample from its leg's starting state, and a serial rerun of geodesics
that meet a sticky state) against np_geokerr_batch_serial (one lane walks
each geodesic's samples in the Fortran's order, itself checked against the
oracle in test_gpu_geokerr_modes).

Both forms run the same step functions on the same state, and the GEOKERR
numerics round every product and sum on its own (no fused multiply-add,
which the compiler would otherwise choose per kernel), so the outputs agree
bit for bit: integers, floats, and NaN where the driver writes nothing.  A
step run from the wrong state would differ, and flip turning-point counts.  The geodesic sets mix Kerr camera rays, a = 0
(cubic U), Q2 = 0 and outgoing rays; the wide sets move MUF across the
whole mu range and UF across (0, u+), so that GEOR leaves with NCASE = 0 on
some geodesics and the serial rerun of flagged geodesics is exercised."""
import numpy as np
import pytest
import torch

import nullpath_amd as npa
from gk_cases import DRIVER_OPTS, geodesics
from oracle import np_oracle as O

FLOATS = ("ufi", "mufi", "dti", "dphi", "lambdai", "u0_out", "uf_out", "muf_out")
INTS = ("tpmi", "tpri", "tpm_out", "tpr_out", "ncase_out", "nup_out", "npts_out")


def _rays(g):
    out = {}
    for k in ("u0", "uf", "uout", "mu0", "muf", "a", "l", "q2", "beta", "su", "sm"):
        out[k] = torch.tensor(np.ascontiguousarray(g[k]), dtype=torch.float64, device="cuda")
    for k in ("tpm", "tpr"):
        out[k] = torch.tensor(np.ascontiguousarray(g[k]), dtype=torch.int32, device="cuda")
    return out


def _widen(g, kind, seed):
    """MUF anywhere in [mu_-, mu_+] (and a little beyond) for USEGEOR; UF
    anywhere in (0, u+) for the u-sampled modes."""
    rng = np.random.default_rng(seed)
    g = g.copy()
    if kind == "usegeor":
        for i in range(len(g)):
            mm, mp = O.findmroots(g["q2"][i], g["l"][i], g["a"][i], g["mu0"][i])
            lo, hi = min(mm, mp), max(mm, mp)
            g["muf"][i] = rng.uniform(lo - 0.05, hi + 0.05)
    else:
        uplus = 1.0 / (1.0 + np.sqrt(1.0 - g["a"] ** 2))
        g["uf"] = rng.uniform(0.0, 1.0, len(g)) * uplus
    return g


def _compare(opts, g):
    rays = _rays(g)
    split = npa.geokerr_batch(opts, rays)
    serial = npa.geokerr_batch(opts, rays, serial=True)
    torch.cuda.synchronize()
    for k in INTS:
        a, b = split[k].cpu().numpy(), serial[k].cpu().numpy()
        bad = np.argwhere(a != b)
        assert bad.size == 0, (k, bad[:5].tolist(), a[tuple(bad[0])], b[tuple(bad[0])])
    nbits = nvals = 0
    worst, wkey = 0.0, None
    for k in FLOATS:
        a, b = split[k].cpu().numpy(), serial[k].cpu().numpy()
        assert np.array_equal(np.isnan(a), np.isnan(b)), k
        ok = ~np.isnan(b)
        nbits += int((a.view(np.int64) != b.view(np.int64)).sum())
        nvals += a.size
        err = np.abs(a[ok] - b[ok]) / np.maximum(np.maximum(np.abs(a[ok]), np.abs(b[ok])), 1.0)
        if err.size and float(err.max()) > worst:
            worst, wkey = float(err.max()), k
    print(f"[gpu]   {nbits} of {nvals} floats differ in their bits, max rel diff {worst:.2e}"
          + (f" ({wkey})" if worst > 0 else ""), flush=True)
    assert nbits == 0, (nbits, worst, wkey)
    return serial


@pytest.mark.gpu
@pytest.mark.parametrize("mode", sorted(DRIVER_OPTS))
def test_split_driver_equals_serial(device, mode):
    kind = "usegeor" if mode == "usegeor" else "mufill"
    opts = DRIVER_OPTS[mode]
    for seed in (3, 11):
        g = geodesics(kind, n=512, seed=seed)
        print(f"[gpu] split vs serial driver mode={mode} seed={seed} n={len(g)}", flush=True)
        _compare(opts, g)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["usegeor", "mufill", "mufill_k2", "single_point"])
def test_split_driver_equals_serial_wide_targets(device, mode):
    kind = "usegeor" if mode == "usegeor" else "mufill"
    opts = DRIVER_OPTS[mode]
    g = _widen(geodesics(kind, n=1024, seed=29), kind, 31)
    print(f"[gpu] split vs serial driver, wide targets, mode={mode} n={len(g)}", flush=True)
    ser = _compare(opts, g)
    if kind == "usegeor":
        # some geodesics leave GEOR with NCASE = 0 (uf = -1): the flagged path
        exits = int((ser["ufi"] == -1.0).any(dim=1).sum().item())
        print(f"[gpu] usegeor geodesics with a GEOR exit: {exits}", flush=True)
        assert exits > 0


@pytest.mark.gpu
def test_split_driver_larger_nup(device):
    """NUP 700, NPTS 90, KEXT 3 (MUFILL): sizes unlike the other modes'
    (several waves per geodesic), 256 geodesics."""
    opts = dict(nup=700, npts=90, kext=3, phit=1, usegeor=0, mufill=1, offset=0.5)
    g = geodesics("mufill", n=256, seed=5)
    print(f"[gpu] split vs serial driver nup=700 npts=90 n={len(g)}", flush=True)
    _compare(opts, g)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["usegeor", "mufill"])
def test_driver_near_polar_orbits_against_oracle(device, mode):
    """Camera rays with L ~ 0 and large Q2 (|alpha| < 0.1 M, |beta| 8-12 M):
    M(mu) has a root within ~1e-5 of mu = 1, where GEOPHITIME's phi mu-part
    cancels (n ~ 1e5 in rf + n RJ / 3) and amplifies rounding ~1e6-fold.
    Without multiply-add contraction, as in the oracle, both driver forms
    keep phi, t and lambda to 1e-10 of the oracle (a fused form there moved
    phi by 4e-9)."""
    kind = "usegeor" if mode == "usegeor" else "mufill"
    opts = DRIVER_OPTS[mode]
    g = geodesics(kind, n=128, seed=43, polar=True)
    print(f"[gpu] near-polar orbits, split and serial driver vs oracle, mode={mode} n={len(g)}", flush=True)
    ref = O.geokerr_full_batch(opts, g)
    rays = _rays(g)
    for serial in (False, True):
        got = npa.geokerr_batch(opts, rays, serial=serial)
        torch.cuda.synchronize()
        for gk, ok in (("dphi", "dphi"), ("dti", "dti"), ("lambdai", "lami"), ("mufi", "mufi"), ("ufi", "ufi")):
            a, b = got[gk].cpu().numpy(), ref[ok]
            assert np.array_equal(np.isnan(a), np.isnan(b)), gk
            m = ~np.isnan(b)
            err = np.abs(a[m] - b[m]) / np.maximum(np.abs(b[m]), 1.0)
            print(f"[gpu]   serial={serial} {gk}: max err {err.max():.2e}", flush=True)
            assert err.max() <= 1e-10, (serial, gk, float(err.max()))
        np.testing.assert_array_equal(got["tpmi"].cpu().numpy(), ref["tpmi"])
