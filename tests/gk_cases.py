"""Geodesic sets for the GEOKERR driver's USEGEOR and MUFILL modes
(geokerr_original.f:357-604), shared by the oracle and GPU tests.

Camera rays (u0 = 1e-4, constants from impact parameters at infinity as in
INITIALIZE_CAMERA_GEOKERR, geokerr_original.f:238-320), Schwarzschild rays
(a = 0, cubic U), Q2 = 0 rays and outgoing rays from r = 3.3 (rows 2/8)."""
import numpy as np

from oracle import np_oracle as O

DRIVER_OPTS = {
    "usegeor": dict(nup=16, npts=0, kext=0, phit=1, usegeor=1, mufill=0, offset=0.5),
    "mufill": dict(nup=32, npts=8, kext=1, phit=1, usegeor=0, mufill=1, offset=0.5),
    "mufill_k0": dict(nup=24, npts=6, kext=0, phit=1, usegeor=0, mufill=1, offset=0.0),
    "mufill_k2": dict(nup=20, npts=10, kext=2, phit=1, usegeor=0, mufill=1, offset=0.5),
    "nophit": dict(nup=12, npts=4, kext=1, phit=0, usegeor=0, mufill=1, offset=0.5),
    # the reference's validation programs (geokerr_single.f:60-67): one point at UF
    "single_point": dict(nup=1, npts=0, kext=0, phit=1, usegeor=0, mufill=0, offset=0.0),
}


def _camera(rng, n, a=None, polar=False):
    g = np.zeros(n, O.GK_ARGS_DTYPE)
    aa = rng.uniform(0.05, 0.998, n) if a is None else np.full(n, a)
    mu0 = rng.uniform(0.05, 0.95, n)
    al, be = rng.uniform(-12, 12, (2, n))
    if polar:
        # |alpha| < 0.1 M, |beta| 8-12 M: L ~ 0, large Q2, so a root of M(mu)
        # lies within ~1e-5 of mu = 1
        al = rng.uniform(-0.1, 0.1, n)
        be = rng.choice([-1.0, 1.0], n) * rng.uniform(8, 12, n)
    g["a"] = aa
    g["mu0"] = mu0
    g["l"] = -al * np.sqrt(1 - mu0 ** 2)
    g["q2"] = be ** 2 - (aa ** 2 - al ** 2) * mu0 ** 2
    g["beta"] = be
    g["u0"] = g["uout"] = 1e-4
    g["su"] = 1.0
    g["sm"] = np.where(be >= 0, 1.0, -1.0)
    return g


def _q2zero(rng, n):
    g = _camera(rng, n)
    a = rng.uniform(0.5, 0.998, n)
    mu0 = g["mu0"]
    al = rng.uniform(-0.4, 0.4, n) * a
    g["a"] = a
    g["l"] = -al * np.sqrt(1 - mu0 ** 2)
    g["beta"] = np.sqrt(a * a - al * al) * mu0
    g["q2"] = 0.0
    g["sm"] = rng.choice([-1.0, 1.0], n)
    return g


def geodesics(mode, n=96, seed=7, polar=False):
    """n geodesics with the per-mode targets: MUF between mu0 and the mu
    turning point ahead for USEGEOR, a final u in (0, u+) with TPR = 1 for
    the u-sampled modes.  polar: Kerr camera rays near the pole only."""
    rng = np.random.default_rng(seed)
    if polar:
        g = _camera(rng, n, polar=True)
    else:
        parts = [_camera(rng, n // 2), _camera(rng, n // 8, a=0.0), _q2zero(rng, n // 8)]
        out = _camera(rng, n - sum(len(p) for p in parts))
        out["u0"] = out["uout"] = 0.3
        out["su"] = -1.0
        parts.append(out)
        g = np.concatenate(parts)
    uplus = 1.0 / (1.0 + np.sqrt(1.0 - g["a"] ** 2))
    if mode == "usegeor":
        # GEOR counts mu turning points against mu_+ (geokerr_original.f:
        # 2911-2914) and the driver sets TPM = (sign(mu0) sm + 1) / 2 from
        # its second sample on (:451): with that TPM, MUF between mu0 and
        # mu_- (towards the equator for Q2 = 0) is reached without a further
        # turning point, from either direction sm.
        for i in range(len(g)):
            mm, mp = O.findmroots(g["q2"][i], g["l"][i], g["a"][i], g["mu0"][i])
            tgt = 0.0 if g["q2"][i] == 0.0 else mm
            g["muf"][i] = g["mu0"][i] + rng.uniform(0.3, 0.9) * (tgt - g["mu0"][i])
        g["tpm"] = ((np.sign(g["mu0"]) * g["sm"] + 1.0) / 2.0).astype(np.int32)
        g["tpr"] = 0
        g["uf"] = 0.0
    else:
        g["uf"] = np.where(g["su"] > 0, rng.uniform(0.0, 1.0, len(g)) * uplus, rng.uniform(0.0, 0.29, len(g)))
        g["tpm"] = 0
        g["tpr"] = 1
    return g


def fresh_point(g, u, tpr):
    """GEOMU + GEOPHITIME from (u0, mu0) to u with tpr u turning points, as
    the reference's own validation programs call the driver (NUP = 1,
    OFFSET = 0; for a point past the turning point NUP = 2 so that the
    driver's first sample fills GEOPHITIME's u0 pieces from U0)."""
    h = g.copy()
    h["uf"] = u
    h["tpr"] = tpr
    h["tpm"] = 0
    opts = dict(nup=1 if tpr == 0 else 2, npts=0, kext=0, phit=1, usegeor=0, mufill=0, offset=0.0)
    r = O.geokerr_full_batch(opts, h)
    j = 0 if tpr == 0 else 1
    return {k: r[k][0, j] for k in ("ufi", "mufi", "lami", "dti", "dphi", "tpmi")}
