"""The headline FP32-compensated tracer and the uneven multi-frame batch
against the FP64 oracle (the reference integrator restated, render.cu:
142-362 and bh_kerr.h:679-1105), with the Hermite crossing radius on both
sides.

Contract asserted, literally: every ray outside the band of kerr_band
(critical-curve margin |m| < 1e-4, a crossing radius within 1e-6 relative
of a disk edge, or a theta turning point with sin(theta_min) < 1e-3) has
the oracle's class and crossing count; every pixel whose samples are all
outside the band has the oracle's radiance to 1e-5 relative (max over the
three channels and all such pixels; captured pixels exactly 0).
"""
import numpy as np
import pytest
import torch

import kerr_band as KB
import nullpath_amd as npa
from nullpath_amd._lib import ray_record_dtype

RAD_TOL = 1e-5
_ORACLE_CACHE = {}


def _oracle_frame(oracle, p, spp, key):
    if key in _ORACLE_CACHE:
        return _ORACLE_CACHE[key]
    q = oracle.default_params()
    for name, _ in npa.RenderParams._fields_:
        setattr(q, name, getattr(p, name))
    q.precision = 0
    ys, xs = np.mgrid[0:p.height, 0:p.width]
    rgb, info = oracle.render_pixels(q, xs.ravel(), ys.ravel(), 0, spp)
    _ORACLE_CACHE[key] = (rgb, info)
    return rgb, info


def _gpu_frame(p, spp):
    n = p.width * p.height
    out = torch.zeros(n * 3, dtype=torch.float32, device="cuda")
    rec = torch.zeros(n * spp * 32, dtype=torch.uint8, device="cuda")
    print(f"[gpu] np_render_tile {p.width}x{p.height}x{spp} precision={p.precision} "
          f"fov={p.fov_y_deg} spin={p.spin}", flush=True)
    npa.render_tile(p, out, spp=spp, records=rec)
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(-1, 3), rec.cpu().numpy().view(ray_record_dtype())


def check_against_oracle(p, spp, g_rgb, g_rec, o_rgb, o_info, label):
    band, margin, edge, pole = KB.ray_band(p, o_info["p0"], o_info["xr"], o_info["ncross"])
    same = (g_rec["cls"] == o_info["cls"]) & (g_rec["orders"] == o_info["orders"])
    bad = ~same & ~band
    if (~same).any():
        mm = np.abs(margin[~same])
        print(f"[{label}] {(~same).sum()} ray(s) differ from the oracle, all in the band: "
              f"max |m| {mm.max():.3g}, min edge dist {edge[~same].min():.3g}", flush=True)
    assert not bad.any(), (
        f"{label}: {bad.sum()} ray(s) outside the band differ; |m| {np.abs(margin[bad])[:8]}, "
        f"edge {edge[bad][:8]}, pole {pole[bad][:8]}, gpu cls/orders "
        f"{g_rec['cls'][bad][:8]}/{g_rec['orders'][bad][:8]}, oracle {o_info['cls'][bad][:8]}/"
        f"{o_info['orders'][bad][:8]}")
    pix = ~band.reshape(-1, spp).any(axis=1)
    g = g_rgb[pix].astype(np.float64)
    o = o_rgb[pix].astype(np.float64)
    zero = o == 0.0
    assert np.all(g[zero] == 0.0), f"{label}: captured pixels not black"
    rel = np.abs(g - o) / np.where(zero, 1.0, np.abs(o))
    disk = (o_info["orders"].reshape(-1, spp) > 0).any(axis=1)[pix]
    rd = rel.max(axis=1)[disk]
    print(f"[{label}] {pix.sum()} of {pix.size} pixels outside the band, {disk.sum()} with disk; "
          f"rays in band {band.sum()}; disk rel err max {rd.max() if rd.size else 0:.3g} "
          f"p99 {np.quantile(rd, 0.99) if rd.size else 0:.3g}; all-pixel max {rel.max():.3g}", flush=True)
    assert rel.max() <= RAD_TOL, f"{label}: radiance rel err {rel.max():.3g} > {RAD_TOL}"
    return int(disk.sum()), int(band.sum())


# The 1920x1080x4 workload (a = 0.9, theta 80 deg, r = 1000 M, adaptive
# rtol 1e-8, disk ISCO..20 M, q = 3, <= 3 crossings) with the field of view
# narrowed onto the disk (1.6 deg) and onto the photon ring (0.8 deg).
ZOOMS = {"disk": (96, 54, 1.6), "photon_ring": (80, 45, 0.8)}


@pytest.mark.gpu
@pytest.mark.parametrize("zoom", sorted(ZOOMS))
@pytest.mark.parametrize("precision", [1, 0, 2])
def test_kerr_still_vs_oracle(oracle, device, zoom, precision):
    w, h, fov = ZOOMS[zoom]
    p = npa.kerr_disk_params(width=w, height=h, fov=fov, seed=1, precision=precision, crossing=1)
    g_rgb, g_rec = _gpu_frame(p, 1)
    o_rgb, o_info = _oracle_frame(oracle, p, 1, ("still", zoom))
    ndisk, nband = check_against_oracle(p, 1, g_rgb, g_rec, o_rgb, o_info,
                                        f"{['fp64', 'fp32c', 'fp64-carter'][precision]} {zoom}")
    assert ndisk > (1000 if zoom == "disk" else 300)
    if zoom == "photon_ring":
        assert (o_info["orders"] >= 2).sum() > 50


# The uneven batch's mix at small resolutions: spins {0, 0.5, 0.9, 0.998},
# inclinations 5..89 deg, pinhole at r = 50 M, fov 50 deg, 2 spp, adaptive.
BATCH = [((32, 32), 0.0, 5.0), ((24, 24), 0.5, 25.0), ((40, 24), 0.9, 45.0),
         ((16, 32), 0.998, 65.0), ((32, 16), 0.9, 80.0), ((24, 40), 0.5, 89.0)]


def _batch_frames(precision):
    return [npa.kerr_disk_params(width=w, height=h, spin=a, theta_deg=t, fov=50.0, seed=1 + i,
                                 precision=precision, crossing=1, r_cam_Rs=25.0)
            for i, ((w, h), a, t) in enumerate(BATCH)]


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tuning", [
    (0, None),
    (1, None),
    # an autotuner candidate: uncapped kernel, chunk 4, FP64 steps only inside 5 M
    (1, {"block_size": 128, "waves_per_simd": 2, "refill_chunk": 4, "fp64_radius": 5.0}),
])
def test_uneven_batch_vs_oracle(oracle, device, precision, tuning):
    frames = _batch_frames(precision)
    spp = 2
    total = sum(f.width * f.height for f in frames)
    out = torch.zeros(3 * total, dtype=torch.float32, device="cuda")
    rec = torch.zeros(total * spp * 32, dtype=torch.uint8, device="cuda")
    print(f"[gpu] np_render_batch {len(frames)} frames precision={precision} tuning={tuning}", flush=True)
    npa.render_batch(frames, out, spp=spp, records=rec, tuning=tuning)
    torch.cuda.synchronize()
    g_all = out.cpu().numpy().reshape(-1, 3)
    r_all = rec.cpu().numpy().view(ray_record_dtype())
    off = 0
    ndisk = 0
    for i, f in enumerate(frames):
        n = f.width * f.height
        o_rgb, o_info = _oracle_frame(oracle, f, spp, ("batch", i))
        d, _ = check_against_oracle(f, spp, g_all[off:off + n], r_all[spp * off:spp * (off + n)], o_rgb,
                                    o_info, f"batch frame {i} a={f.spin:.3f} prec={precision}")
        ndisk += d
        off += n
    assert ndisk > 200
