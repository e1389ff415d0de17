"""GPU parity of the FP64 Kerr renderer (np_render_tile) against the CPU oracle.

Both sides start every sample from bit-identical camera directions and
initial momenta; they then differ only by FMA contraction and device libm
ulps, so per-sample classification, crossing count and step count must match
except for rays whose fate the last bits decide (closest approach within a
thin band of the photon orbit), and colours agree to float rounding.
"""
import numpy as np
import pytest
import torch

import nullpath_amd as npa
from nullpath_amd._lib import ray_record_dtype


def _run_gpu(p, spp, sample_offset=0):
    n = p.width * p.height
    out = torch.zeros(n * 3, dtype=torch.float32, device="cuda")
    cnt = npa.new_counters()
    rec = torch.zeros(n * spp * 32, dtype=torch.uint8, device="cuda")
    print(f"[gpu] np_render_tile {p.width}x{p.height} spp={spp} integrator={p.integrator} "
          f"spin={p.spin}", flush=True)
    npa.render_tile(p, out, spp=spp, sample_offset=sample_offset, counters=cnt, records=rec)
    torch.cuda.synchronize()
    recs = rec.cpu().numpy().view(ray_record_dtype())
    return out.cpu().numpy().reshape(-1, 3), npa.counters_dict(cnt), recs


def _oracle(oracle, p, spp, sample_offset=0):
    q = oracle.default_params()
    for name, _ in npa.RenderParams._fields_:
        setattr(q, name, getattr(p, name))
    ys, xs = np.mgrid[0:p.height, 0:p.width]
    rgb, info = oracle.render_pixels(q, xs.ravel(), ys.ravel(), sample_offset, spp)
    return rgb, info


def _compare(p, spp, oracle, band=0.02):
    g_rgb, cnt, g_rec = _run_gpu(p, spp)
    o_rgb, o_info = _oracle(oracle, p, spp)
    Rs = 2.0 * p.mass
    r_ph = 1.5 * Rs
    # rays skimming the photon orbit are decided by the last bits: excluded
    near = np.abs(o_info["min_r"] - r_ph) < band * Rs * 10
    same_cls = g_rec["cls"] == o_info["cls"]
    assert np.all(same_cls | near), f"class mismatch away from photon orbit: {np.sum(~same_cls & ~near)}"
    same_ord = g_rec["orders"] == o_info["orders"]
    assert np.all(same_ord | near)
    assert cnt["rays"] == p.width * p.height * spp
    # pixel colours: pixels whose samples are all outside the band
    pix_near = near.reshape(-1, spp).any(axis=1)
    ok = ~pix_near
    err = np.abs(g_rgb[ok] - o_rgb[ok]) / np.maximum(np.abs(o_rgb[ok]), 1e-6)
    assert err.max() < 2e-5, err.max()
    # step counts match for almost every ray
    assert np.mean(g_rec["steps"] == o_info["steps"]) > 0.95
    return g_rgb, o_rgb, cnt, g_rec, o_info


@pytest.mark.gpu
def test_fp64_fixed_reference_defaults(oracle, device):
    """mode=balanced: fixed RK4, static camera, reference disk (render.cu defaults)."""
    p = npa.default_params(width=40, height=24, spin=0.0)
    _compare(p, 2, oracle)


@pytest.mark.gpu
def test_fp64_fixed_kerr_spin(oracle, device):
    p = npa.default_params(width=40, height=24, spin=0.9, theta_cam=npa.deg2rad_f32(80.0),
                           fov_y_deg=20.0, disk_max_orders=3)
    g_rgb, o_rgb, cnt, g_rec, o_info = _compare(p, 2, oracle)
    assert cnt["disk_hits"] > 0


@pytest.mark.gpu
def test_fp64_adaptive_zamo(oracle, device):
    """mode=accurate: adaptive step doubling + ZAMO camera (render.cu:472)."""
    p = npa.default_params(width=24, height=16, spin=0.7, theta_cam=npa.deg2rad_f32(70.0),
                           fov_y_deg=20.0, integrator=1, camera=1, check_inv=1,
                           disk_max_orders=2, rtol=1e-8)
    g_rgb, o_rgb, cnt, g_rec, o_info = _compare(p, 1, oracle)
    assert cnt["attempts"] >= cnt["steps"]


@pytest.mark.gpu
def test_fp64_hermite_crossing_matches_oracle(oracle, device):
    """The cubic Hermite crossing radius (np_crossing HERMITE) against the
    oracle's restatement of it, on the Kerr disk workload zoomed on the disk."""
    p = npa.kerr_disk_params(width=32, height=18, fov=3.0, crossing=1)
    g_rgb, o_rgb, cnt, g_rec, o_info = _compare(p, 1, oracle)
    assert cnt["disk_hits"] > 20


def _disk_pair(width, height, fov, spp=1, **kw):
    a = _run_gpu(npa.kerr_disk_params(width=width, height=height, fov=fov, seed=1, precision=0, **kw), spp)
    b = _run_gpu(npa.kerr_disk_params(width=width, height=height, fov=fov, seed=1, precision=1, **kw), spp)
    return a, b


@pytest.mark.gpu
def test_fp32c_matches_fp64_on_disk_workload(oracle, device):
    """FP32-compensated vs FP64 (both Hermite crossings) on the Kerr disk
    workload zoomed on the hole (160x90, fov 3 deg): every ray outside the
    band of kerr_band (critical margin |m| < 1e-4, sin(theta_min) < 1e-3, or
    an oracle crossing radius within 1e-6 relative of a disk edge) has the
    same class and crossing count in both; the radiance of every disk pixel
    whose samples are all outside the band agrees to 1e-5 relative (max);
    sky pixels are bitwise equal."""
    import kerr_band as KB
    (ia, ca, ra), (ib, cb, rb) = _disk_pair(160, 90, 3.0, crossing=1)
    p = npa.kerr_disk_params(width=160, height=90, fov=3.0, seed=1, precision=0, crossing=1)
    q = oracle.default_params()
    for name, _ in npa.RenderParams._fields_:
        setattr(q, name, getattr(p, name))
    ys, xs = np.mgrid[0:p.height, 0:p.width]
    xs, ys = xs.ravel(), ys.ravel()
    band, margin, _, pole = KB.ray_band(p, oracle.ray_p0_pixels(q, xs, ys, 0, 1))
    same = (ra["cls"] == rb["cls"]) & (ra["orders"] == rb["orders"])
    diff = np.flatnonzero(~same & ~band)
    if diff.size:   # the edge band needs the crossing radii: trace those rays
        assert diff.size < 50, diff.size
        _, info = oracle.render_pixels(q, xs[diff], ys[diff], 0, 1)
        edge = KB.edge_distance(info["xr"], info["ncross"], p.disk_rin, p.disk_rout)
        assert np.all(edge < KB.EDGE_BAND), (margin[diff], edge)
        band[diff] = True
    disk = (ra["orders"] > 0) & ~band
    assert disk.sum() > 1000
    rel = (np.abs(ib - ia) / np.maximum(np.abs(ia), 1e-30)).max(axis=1)[disk]
    print(f"[gpu] fp32c vs fp64 disk rel err: median {np.median(rel):.3g} p99 "
          f"{np.quantile(rel, 0.99):.3g} max {rel.max():.3g}; {(~same).sum()} rays differ, "
          f"{band.sum()} in the band", flush=True)
    assert rel.max() <= 1e-5
    sky = ~band & (ra["orders"] == 0)
    assert np.array_equal(ia[sky], ib[sky])


@pytest.mark.gpu
@pytest.mark.parametrize("fov", [55.0, 3.0])
def test_far_field_estimate_skip(device, fov):
    """Skipping the far-field error estimate (estimate_every_step 0, the
    default) against estimating on every step (1): the same image bits and
    per-ray class and crossing count; step counts differ only where a float
    attempt rejected on rounding noise was re-run in FP64; fewer right-hand
    sides."""
    runs = {}
    for every in (1, 0):
        p = npa.kerr_disk_params(width=96, height=54, fov=fov, seed=1, precision=1, crossing=1,
                                 estimate_every_step=every)
        runs[every] = _run_gpu(p, 1)
    (i1, c1, r1), (i0, c0, r0) = runs[1], runs[0]
    assert np.array_equal(i1, i0)
    assert np.array_equal(r1["cls"], r0["cls"]) and np.array_equal(r1["orders"], r0["orders"])
    assert np.mean(r1["steps"] == r0["steps"]) > 0.99
    for k in ("rays", "captured", "escaped", "disk_hits"):
        assert c0[k] == c1[k], k
    print(f"[gpu] far-field estimate skip: rhs {c0['rhs_evals']} vs {c1['rhs_evals']}", flush=True)
    assert c0["rhs_evals"] < 0.85 * c1["rhs_evals"]


@pytest.mark.gpu
def test_fp32c_kernel_variants_agree(device):
    """The 168-VGPR (default), 128-VGPR and compiler-budget kernel variants
    compute the same thing: outputs and records are bitwise identical."""
    p3 = npa.kerr_disk_params(width=64, height=36, fov=6.0, seed=1, precision=1, crossing=1)
    i3, c3, r3 = _run_gpu(p3, 1)
    for w in (4, 2):
        p = npa.kerr_disk_params(width=64, height=36, fov=6.0, seed=1, precision=1, crossing=1,
                                 waves_per_simd=w)
        i, c, r = _run_gpu(p, 1)
        assert np.array_equal(i3, i), w
        assert np.array_equal(r3, r), w
        assert c3 == c, w


@pytest.mark.gpu
@pytest.mark.parametrize("w", [4, 3, 2])
def test_deferred_fp64_attempts_change_nothing(device, w):
    """FP32C with the FP64 attempts deferred (defer_fp64 = 1: a lane whose
    next attempt is FP64 waits for more such lanes of its wave) against never
    deferred (-1), camera at 50 M where many rays run FP64 steps: every ray
    makes the same attempts, so outputs, records and counters are bitwise
    equal, in every kernel variant."""
    runs = {}
    for d in (-1, 1):
        p = npa.kerr_disk_params(width=96, height=64, fov=30.0, seed=2, precision=1, crossing=1,
                                 r_cam_Rs=25.0, waves_per_simd=w, defer_fp64=d)
        runs[d] = _run_gpu(p, 2)
    (i0, c0, r0), (i1, c1, r1) = runs[-1], runs[1]
    print(f"[gpu] waves {w}: FP64 attempts {c0['attempts_f64']} of {c0['attempts']}", flush=True)
    assert c0["attempts_f64"] > 0.01 * c0["attempts"]
    assert np.array_equal(i0, i1)
    assert np.array_equal(r0, r1)
    assert c0 == c1


def _batch_frames(precision):
    spins = [0.0, 0.5, 0.9, 0.998]
    sizes = [(40, 24), (32, 32), (48, 20), (24, 36)]
    thetas = [20.0, 45.0, 70.0, 88.0]
    return [npa.kerr_disk_params(width=w, height=h, spin=a, theta_deg=t, fov=30.0, seed=3,
                                 precision=precision, crossing=1, r_cam_Rs=25.0)
            for (w, h), a, t in zip(sizes, spins, thetas)]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("tuning", [None, {"block_size": 128, "waves_per_simd": 2,
                                           "refill_chunk": 4, "fp64_radius": 0.0},
                                    {"block_size": 256, "waves_per_simd": 0, "refill_chunk": 1,
                                     "fp64_radius": 0.0, "defer_fp64": 1}])
def test_batch_equals_per_frame_renders(device, precision, tuning):
    """One persistent launch over frames of different size, spin and
    inclination gives, frame by frame, exactly the single-frame render."""
    frames = _batch_frames(precision)
    total = sum(f.width * f.height for f in frames)
    out = torch.zeros(3 * total, dtype=torch.float32, device="cuda")
    cnt = npa.new_counters()
    print(f"[gpu] np_render_batch precision={precision} tuning={tuning}", flush=True)
    npa.render_batch(frames, out, spp=2, counters=cnt, tuning=tuning)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    off = 0
    for f in frames:
        ref, c, _ = _run_gpu(f, 2)
        n = f.width * f.height
        np.testing.assert_array_equal(got[3 * off:3 * (off + n)].reshape(-1, 3), ref)
        off += n
    assert npa.counters_dict(cnt)["rays"] == 2 * total


@pytest.mark.gpu
def test_autotune_batch_reports_candidates(device):
    frames = _batch_frames(1)
    best, trials = npa.autotune_batch(frames, spp=1, subsample=2)
    # block 2 x chunk 2 x kernel variant 3 (4/3/2 waves) x FP64 radius 3
    assert len(trials) == 36
    assert {t["waves_per_simd"] for t, _ in trials} == {4, 3, 2}
    assert all(ms > 0 for _, ms in trials)
    fastest = min(trials, key=lambda t: t[1])[0]
    assert best.as_dict() == fastest
    best2, trials2 = npa.autotune_batch(frames, spp=1, subsample=2)
    assert best2.as_dict() == best.as_dict() and trials2 == []
