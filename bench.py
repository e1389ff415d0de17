#!/usr/bin/env python3
"""Benchmark: Kerr thin-disk still on one MI355X, FP32-compensated vs FP64.

Workload (one "step" = one full frame): 1920x1080 pinhole camera at
r = 1000 M, a = 0.9 M, theta_obs = 80 deg, fov 55 deg, 4 jittered samples
per pixel (seed 1), adaptive step-doubling RK4 (rtol 1e-8, atol 1e-12),
disk from the ISCO to 20 M with emissivity (r/M)^-3 and g^3 redshift, up
to 3 equatorial crossings, float32 RGB output.  The frame constants are
resident before the timed region and the image stays in HBM.

Each step renders the frame with the FP64 path (reference formulas) and
with the FP32-compensated path, both with the cubic Hermite crossing radius;
`value` is the FP32-compensated rays/s.  Beside it: the FP64 rays/s, their
ratio, the FP32c-vs-FP64 error statistics, the same comparison with the
reference's linear crossing interpolation (960x540x1), the differentiable
512x512 render's gradients/s, and the CPU baseline (the FP64 oracle on the
host cores, a bounded sample of the same frame).  --gpus above 1 is
refused: this build measures one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Kerr rays/s @1080p a=0.9, FP32-comp vs FP64 at ≤1e-5 err; adjoint grads/s"
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}

# HBM bytes of one FP32C render launch on the 1920x1080x4 frame (default
# 4-wave variant, with ray records), measured with rocprofv3 FETCH_SIZE and
# WRITE_SIZE and scaled by what the counters read for a 256 MiB device copy
# (FETCH x2.00, WRITE x1.00; profiles/kerr_render_1920x1080x4_kernel_trace.txt):
# 96.0 MB read + 588.2 MB written.  The algorithmic bytes are the 24.9 MB
# image and 265.4 MB of ray records; the rest is register-spill scratch
# around the rare FP64 attempts.
FP32C_HBM_BYTES_1080P4 = (46868.6 * 2.0 + 574382.0 * 1.0) * 1024.0

# Algorithmic operation counts per unit (add, sub, mul, div, sqrt, rcp = 1;
# fma = 2; library sin/cos/pow counted apart), derived from the kernel source:
#   fp64 path: the reference right-hand side (bh_kerr.h:600-727) is 200
#     flops; per attempt three RK4 combinations with compensated sums and
#     updates (3 x 196) plus the error norm and step factor (29).
#   fp32c float steps: the Carter-form right-hand side with the split-theta
#     sin/cos correction is 70 flops; per attempt 10 stage inputs and stage
#     sums (10 x 18), the full step's increments with float-float r and
#     theta (26), the two half steps' float-float updates (2 x 56) and the
#     noise-floored error norm and step factor (90).
#   fp32c float steps without the error estimate (far field): 7 stage inputs
#     and sums (7 x 18), the two half steps' float-float updates (2 x 56) and
#     the pole bound that allows the skip (8).
#   fp32c FP64 steps: the FP64 Carter right-hand side with its polynomial
#     sincos is 105 flops; per attempt 12 stage inputs and sums (12 x 18),
#     3 updates (3 x 10) and the error norm (40).
#   fp64c (the FP32C path's formulas and step control in double): the FP64
#     Carter right-hand side, 105 flops; per attempt 10 stage inputs and sums
#     (10 x 18), 3 updates (3 x 10) and the error norm (40); without the
#     estimate 7 x 18 + 2 x 10 + 8.
FLOPS = {
    "fp64": {"rhs": 200, "attempt": 3 * 196 + 29},
    "fp64c": {"rhs": 105, "attempt": 10 * 18 + 3 * 10 + 40, "attempt_lazy": 7 * 18 + 2 * 10 + 8},
    "fp32c": {"rhs": 70, "attempt": 10 * 18 + 26 + 2 * 56 + 90, "attempt_lazy": 7 * 18 + 2 * 56 + 8,
              "rhs64": 105, "attempt64": 12 * 18 + 3 * 10 + 40},
}


# Schwarzschild FP64 sweep (schw.hip schw_sweep_kernel): the right-hand side
# is 33 flops (r^2, r^3, 1 - Rs/r, the five derivative terms and the force
# sum) plus one sincos; a step is 4 right-hand sides, 3 stage inputs (30),
# the step law (2), the stage combination and update (36) and the unwrapped
# delta-phi (2).
SCHW_FLOPS_PER_STEP = 4 * 33 + 30 + 2 + 36 + 2

# Differentiable render (adjoint.hip): the r-scaled Carter field is 53 flops
# and its vector-Jacobian product 258 (counted with a counting scalar on the
# host: tests/support/flop_count.cpp, pinned by tests/test_flop_counts.py).
# Forward FP32C step: 4 fields + 3 stage inputs (24) + the stage combination
# (20 + 2) + 4 float-float updates of 10 + reading the float-float state (4).
# Backward step: the segment's forward recompute (one forward step) + 4 float
# stores + the RK4 adjoint (3 fields, 4 products, 76 for stage inputs and
# adjoint updates).
ADJ_RHS, ADJ_VJP = 53, 258
ADJ_FWD_STEP = 4 * ADJ_RHS + 24 + 22 + 4 * 10 + 4
ADJ_BWD_STEP = ADJ_FWD_STEP + 4 + 3 * ADJ_RHS + 4 * ADJ_VJP + 76

# GEOMU batch (geomu.hip geomu_camera_kernel), data dependent (Laguerre and
# Carlson duplication counts): FP64 flops per ray measured on the 2^20-ray
# workload as SQ_INSTS_VALU_FLOPS_FP64 x 64 x the lane utilisation
# (SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU)), i.e. the flops the
# active lanes executed, division and square-root sequences included
# (profiles/geomu_2p20_counters.txt).
GEOMU_FLOPS_PER_RAY = 7092


def valu_flops(c: dict, prec: str):
    """(FP32, FP64) algorithmic flops of a launch from its counters."""
    fl = FLOPS[prec]
    if prec == "fp64":
        return 0, c["rhs_evals"] * fl["rhs"] + c["attempts"] * fl["attempt"]
    if prec == "fp64c":
        lazy = c.get("attempts_lazy", 0)
        return 0, c["rhs_evals"] * fl["rhs"] + (c["attempts"] - lazy) * fl["attempt"] + lazy * fl["attempt_lazy"]
    lazy = c.get("attempts_lazy", 0)
    f32 = (c["rhs_evals"] - c["rhs_evals_f64"]) * fl["rhs"] + \
        (c["attempts"] - c["attempts_f64"] - lazy) * fl["attempt"] + lazy * fl["attempt_lazy"]
    f64 = c["rhs_evals_f64"] * fl["rhs64"] + c["attempts_f64"] * fl["attempt64"]
    return f32, f64


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=4)
    ap.add_argument("--cpu-pixels", type=int, default=3000,
                    help="pixels (x spp rays) in the CPU-baseline sample; 0 skips it")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--precisions", default="fp64,fp64c,fp32c",
                    help="fp64 = the reference's formulas in double; fp64c = the FP32C path's formulas and "
                         "step control in double; fp32c = the headline")
    ap.add_argument("--fp64-steps", type=int, default=1,
                    help="timed steps of the FP64 comparison leg (no warmup; one 1080p x4 FP64 frame "
                         "takes ~18 s, so it does not follow --steps/--warmup)")
    ap.add_argument("--crossing", default="hermite", choices=["hermite", "linear"])
    ap.add_argument("--waves-per-simd", type=int, default=0, choices=[0, 2, 3, 4],
                    help="FP32C kernel variant (0 = the library default, 4 waves per SIMD)")
    ap.add_argument("--fp64c-waves-per-simd", type=int, default=0, choices=[0, 2, 3],
                    help="FP64 comparator kernel variant (0 = the library default, 3 waves per SIMD; "
                         "2 = the compiler's register budget)")
    ap.add_argument("--fp64-waves-per-simd", type=int, default=0, choices=[0, 1, 2],
                    help="reference-formula FP64 kernel variant (0 = the library default, 2 waves per "
                         "SIMD; 1 = the compiler's register budget)")
    ap.add_argument("--times", type=int, default=0,
                    help="1 = the render legs also integrate t and the affine parameter "
                         "(np_render_tile_times) -- a cost measurement, not the headline")
    ap.add_argument("--renorm-pr", type=int, default=0, choices=[0, 1],
                    help="1 = FP32C renormalises p_r from H = 0 after accepted steps")
    ap.add_argument("--full-contract", type=int, default=1,
                    help="1 = also time the FP32C frame with t and lambda integrated and p_r "
                         "renormalised (one step), beside the headline")
    ap.add_argument("--linear-check", default="960x540x1",
                    help="WxHxSPP of the linear-crossing comparison; 'none' skips it")
    ap.add_argument("--adjoint", type=int, default=1,
                    help="also time the differentiable 512x512 render (forward + adjoint)")
    ap.add_argument("--adjoint-size", type=int, default=512)
    ap.add_argument("--side-workloads", type=int, default=1,
                    help="also time the Schwarzschild 1e5 sweep and the GEOMU 2^20 batch")
    ap.add_argument("--geokerr-only", type=int, default=0,
                    help="only time geokerr_eval_lambda batches of this many geodesics x 4096 "
                         "samples (for profiling that path) and exit")
    ap.add_argument("--geomu-only", type=int, default=0,
                    help="only time the GEOMU batch (FP64 and the mixed-precision roots) on this many "
                         "rays and exit")
    ap.add_argument("--geokerr-driver-only", type=int, default=0,
                    help="only time the whole GEOKERR driver (MUFILL and USEGEOR runs) on this many "
                         "geodesics and exit")
    ap.add_argument("--only", default="", choices=["", "schw", "geomu", "adjoint", "batch"],
                    help="only time this side workload (with its CPU baseline unless --cpu-pixels 0) and exit")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run one device-to-device copy of this many MiB (a reference "
                         "kernel of known HBM traffic for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE) "
                         "and exit")
    return ap.parse_args()


def copy_calibration(mib: int) -> int:
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4}))
    return 0


def render_case(prec: str, crossing: str, width: int, height: int, spp: int, steps: int,
                warmup: int, waves_per_simd: int = 0, times: bool = False, renorm_pr: int = 0) -> dict:
    import torch

    import nullpath_amd as npa
    from nullpath_amd._lib import ray_record_dtype

    dev = torch.device("cuda:0")
    pid = {"fp64": 0, "fp32c": 1, "fp64c": 2}[prec]
    cr = 1 if crossing == "hermite" else 0
    npix = width * height
    rays = npix * spp
    p = npa.kerr_disk_params(width=width, height=height, spin=0.9, theta_deg=80.0, fov=55.0,
                             rtol=1e-8, max_orders=3, q=3.0, seed=1, precision=pid, crossing=cr,
                             waves_per_simd=waves_per_simd, renorm_pr=renorm_pr)
    out = torch.zeros(npix * 3, dtype=torch.float32, device=dev)
    cnt = npa.new_counters(dev)
    rec = torch.zeros(rays * 32, dtype=torch.uint8, device=dev)
    tm = torch.zeros(rays * 24 if times else 1, dtype=torch.uint8, device=dev)

    def launch(**kw):
        if times:
            npa.render_tile_times(p, out, tm, spp=spp, **kw)
        else:
            npa.render_tile(p, out, spp=spp, **kw)

    npa.render_tile(npa.kerr_disk_params(width=64, height=36, seed=1, precision=pid, crossing=cr),
                    out, spp=1)
    for _ in range(warmup):
        launch()
    torch.cuda.synchronize()
    # events on the stream the kernels are launched on (torch's current stream)
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    t0 = time.perf_counter()
    for k in range(steps):
        cnt.zero_()
        ev[k][0].record(stream)
        launch(counters=cnt, records=rec, stream=stream)
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    print(f"[bench] {prec}/{crossing} {width}x{height}x{spp}: {steps} step(s) in {wall:.2f} s",
          file=sys.stderr, flush=True)
    c = npa.counters_dict(cnt)
    assert c["rays"] == rays, c
    kern_ms = [a.elapsed_time(b) for a, b in ev]
    mean_s = sum(kern_ms) / len(kern_ms) / 1e3
    f32, f64 = valu_flops(c, prec)
    # fraction of the VALU peak: FP32 work at the FP32 rate, FP64 at the FP64 rate
    valu_s = f32 / (PEAK_TFLOPS["fp32"] * 1e12) + f64 / (PEAK_TFLOPS["fp64"] * 1e12)
    return {
        "prec": prec, "crossing": crossing, "wall": wall, "kern_ms": kern_ms, "mean_s": mean_s,
        "counters": c, "flops32": f32, "flops64": f64, "frac": valu_s / mean_s,
        "rays": rays * steps, "npix": npix, "spp": spp,
        "img": out.view(height, width, 3).cpu().numpy(),
        "rec": rec.cpu().numpy().view(ray_record_dtype()), "params": p,
    }


def case_summary(r: dict) -> dict:
    cc = r["counters"]
    return {
        "rays_per_s": r["rays"] / r["wall"],
        "kernel_ms": r["kern_ms"],
        "accepted_steps_per_ray": cc["steps"] / cc["rays"],
        "attempts_per_ray": cc["attempts"] / cc["rays"],
        "fp64_attempt_frac": cc["attempts_f64"] / max(1, cc["attempts"]),
        "lazy_attempt_frac": cc.get("attempts_lazy", 0) / max(1, cc["attempts"]),
        "rk4_steps_per_s": 3 * cc["attempts"] / r["mean_s"],
        "rhs_evals_per_s": cc["rhs_evals"] / r["mean_s"],
        "tflops_fp32": r["flops32"] / r["mean_s"] / 1e12,
        "tflops_fp64": r["flops64"] / r["mean_s"] / 1e12,
        "valu_peak_frac": r["frac"],
        "captured": cc["captured"], "escaped": cc["escaped"], "stepcap": cc["stepcap"],
        "disk_rays": cc["disk_hits"],
    }


def compare(a: dict, b: dict) -> dict:
    """FP32c (b) against FP64 (a) on the same frame: per-ray class and
    crossing count, per-pixel disk radiance relative error, sky pixels."""
    import numpy as np
    npix, spp = a["npix"], a["spp"]
    ra = a["rec"].reshape(npix, spp)
    rb = b["rec"].reshape(npix, spp)
    same_ray = (ra["cls"] == rb["cls"]) & (ra["orders"] == rb["orders"])
    same_pix = same_ray.all(axis=1)
    ia = a["img"].reshape(-1, 3).astype(np.float64)
    ib = b["img"].reshape(-1, 3).astype(np.float64)
    disk = (ra["orders"] > 0).any(axis=1) & same_pix
    rel = np.abs(ib - ia) / np.maximum(np.abs(ia), 1e-30)
    rel_disk = rel[disk].max(axis=1) if disk.any() else np.zeros(1)
    sky = same_pix & ~(ra["orders"] > 0).any(axis=1)
    return {
        "ray_class_mismatch": int((~same_ray).sum()),
        "ray_class_mismatch_frac": float((~same_ray).mean()),
        "mismatch_min_r_fp64_max": float(ra["min_r"][~same_ray].max()) if (~same_ray).any() else None,
        "disk_pixels_compared": int(disk.sum()),
        "disk_rel_err_max": float(rel_disk.max()),
        "disk_rel_err_p999": float(np.quantile(rel_disk, 0.999)),
        "disk_rel_err_p99": float(np.quantile(rel_disk, 0.99)),
        "disk_rel_err_median": float(np.median(rel_disk)),
        "disk_frac_within_1e-5": float(np.mean(rel_disk <= 1e-5)),
        "sky_pixels_max_abs_diff": float(np.abs(ib[sky] - ia[sky]).max()) if sky.any() else 0.0,
    }


def host_threads() -> int:
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    return min(16, ncpu)


# CPU baselines of the side workloads run about this long: whole passes over
# a sample (or over the workload itself) until the time is reached.
CPU_SAMPLE_S = 8.0


def cpu_for(seconds: float, fn):
    """Call fn(k) for k = 0, 1, ... (each call returns the units it did)
    until `seconds` have passed; returns (units, calls, elapsed s)."""
    t0 = time.perf_counter()
    units, k = 0, 0
    while True:
        units += fn(k)
        k += 1
        dt = time.perf_counter() - t0
        if dt >= seconds:
            return units, k, dt


def roofline(flops: float, seconds: float, prec: str, kernel: str) -> dict:
    peak = PEAK_TFLOPS[prec]
    achieved = flops / seconds / 1e12
    return {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
            "kernel": kernel, "flops_per_launch": flops}


def bench_adjoint(args) -> dict:
    """Gradients/s of the differentiable disk render: 512x512, 4096 fixed
    RK4 steps per ray (FP32-compensated), checkpoint every 64 steps; loss
    sum((I - I_target)^2) with the target at (a=0.7, theta=60 deg, q=3,
    gain=1) and the gradient taken at (0.6, 65 deg, 2.5, 0.8).  One gradient
    = one forward + one backward.  Beside it: the rooflines of the forward
    and adjoint kernels (algorithmic FP32 flops from the steps each ray ran,
    read back from the workspace) and the CPU baseline, the FP64
    forward-mode oracle gradient on a sample of pixels."""
    import math

    import numpy as np
    import torch

    from nullpath_amd import adjoint as A
    n = args.adjoint_size
    base = A.adjoint_params(width=n, height=n)
    ws = A.Workspace(base)
    target = A.with_params4(base, (0.7, math.radians(60.0), 3.0, 1.0))
    start = A.with_params4(base, (0.6, math.radians(65.0), 2.5, 0.8))
    tgt = A.forward(target, A.PREC_FP32C, ws=ws).clone()
    img = torch.empty_like(tgt)
    gimg = torch.empty_like(tgt)
    g4 = torch.empty(4, dtype=torch.float64, device=tgt.device)
    stream = torch.cuda.current_stream()

    def one():
        A.forward(start, A.PREC_FP32C, image=img, ws=ws)
        torch.sub(img, tgt, out=gimg).mul_(2.0)
        A.backward(start, gimg, ws, grad4=g4)

    one()
    torch.cuda.synchronize()
    steps = max(1, args.steps)
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(steps)]
    t0 = time.perf_counter()
    for k in range(steps):
        evs[k][0].record(stream)
        A.forward(start, A.PREC_FP32C, image=img, ws=ws)
        evs[k][1].record(stream)
        torch.sub(img, tgt, out=gimg).mul_(2.0)
        evs[k][2].record(stream)
        A.backward(start, gimg, ws, grad4=g4)
        evs[k][3].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    fwd = [a.elapsed_time(b) for a, b, _, _ in evs]
    bwd = [c.elapsed_time(d) for _, _, c, d in evs]
    tot = [a.elapsed_time(d) for a, _, _, d in evs]
    # steps each ray ran (forward) and the steps the adjoint walks back (from
    # the last counted crossing), from the workspace's per-ray records
    nrays = n * n
    nck = (base.n_steps + base.ckpt_every - 1) // base.ckpt_every
    # workspace rows: last step, crossings, crossing steps x3, work key, work
    # list, crossings up to the last contributing one (written by the backward)
    meta = ws.buf[nck * 8 * nrays * 4:(nck * 8 + 8) * nrays * 4].view(torch.int32).view(8, nrays).cpu().numpy()
    term, live = meta[0].astype(np.int64), meta[7]
    last = np.where(live > 0, meta[1 + np.clip(live, 1, 3), np.arange(nrays)], -1)
    bsteps = np.where((live > 0) & (meta[5] >= 0), last + 1, 0).sum()
    fl_f, fl_b = float(term.sum()) * ADJ_FWD_STEP, float(bsteps) * ADJ_BWD_STEP
    res = {
        "workload": f"{n}x{n}, 4096 fixed RK4 steps/ray FP32-comp, checkpoint every 64",
        "grads_per_s": steps / wall,
        "ms_per_grad": wall * 1e3 / steps,
        "forward_ms": fwd,
        "backward_ms": bwd,
        "forward_backward_ms": tot,
        "grad": g4.cpu().tolist(),
        "workspace_bytes": ws.nbytes,
        "forward_steps": int(term.sum()), "adjoint_steps": int(bsteps),
        "roofline_forward": roofline(fl_f, sum(fwd) / len(fwd) / 1e3, "fp32", "adj_forward_f32"),
        "roofline_backward": roofline(fl_b, sum(bwd) / len(bwd) / 1e3, "fp32", "adj_backward_f32"),
    }
    if args.cpu_pixels > 0:
        from oracle import np_oracle as O
        m = 65536
        rng = np.random.default_rng(7)
        sel = rng.choice(nrays, size=m, replace=False)
        q = O.adjoint_params()
        for name, _ in O.AdjointParams._fields_:
            setattr(q, name, getattr(start, name))
        th = host_threads()
        t1 = time.perf_counter()
        O.adj_pixels(q, sel % n, sel // n, nthreads=th)
        dt = time.perf_counter() - t1
        res["cpu_baseline"] = {"value": 1.0 / (dt * nrays / m), "unit": "gradients/s", "cores": th,
                               "kind": "port",
                               "sample": f"FP64 forward-mode oracle gradient of {m} seeded random pixels of the "
                                         f"{n}x{n} image ({dt:.1f} s), scaled to the whole image"}
    return res

def batch_frames(nframes: int = 64, seed: int = 4096):
    """The uneven multi-frame workload: nframes frames, resolution drawn from
    {256^2, 512^2, 768^2, 1024^2}, spin from {0, 0.5, 0.9, 0.998}, theta_obs
    ~ U[5, 89] deg (seeded); pinhole at r = 50 M, fov 50 deg, adaptive
    rtol 1e-8, disk ISCO..20 M, q = 3, <= 3 crossings, FP32C, Hermite
    crossing.  Sky rays take a few hundred steps, photon-ring rays ~10^4."""
    import numpy as np

    import nullpath_amd as npa
    rng = np.random.default_rng(seed)
    res = rng.choice([256, 512, 768, 1024], size=nframes)
    spins = rng.choice([0.0, 0.5, 0.9, 0.998], size=nframes)
    thetas = rng.uniform(5.0, 89.0, size=nframes)
    return [npa.kerr_disk_params(width=int(r), height=int(r), spin=float(a), theta_deg=float(t),
                                 fov=50.0, r_cam_Rs=25.0, seed=1 + i, precision=1, crossing=1)
            for i, (r, a, t) in enumerate(zip(res, spins, thetas))]


def bench_batch(spp: int = 2, cpu_pixels_per_frame: int = 0) -> dict:
    """One persistent launch over the 64-frame uneven batch, after the
    autotuner picked block size, refill chunk, FP32C kernel variant and
    FP64 radius on a 1/4-resolution copy; beside it the same frames as 64
    separate single-frame launches with the default launch shape."""
    import torch

    import nullpath_amd as npa
    frames = batch_frames()
    total = sum(f.width * f.height for f in frames)
    t0 = time.perf_counter()
    best, trials = npa.autotune_batch(frames, spp=spp, subsample=4)
    tune_s = time.perf_counter() - t0
    out = torch.zeros(3 * total, dtype=torch.float32, device="cuda")
    cnt = npa.new_counters()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    npa.render_batch(frames, out, spp=spp, counters=cnt, tuning=best, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    batch_ms = e0.elapsed_time(e1)
    c = npa.counters_dict(cnt)
    # the same frames one launch each, with the tuned kernel variant and FP64 radius
    for f in frames:
        f.waves_per_simd = best.waves_per_simd
        if best.fp64_radius > 0.0:
            f.fp64_radius = best.fp64_radius
    outs = [torch.zeros(3 * f.width * f.height, dtype=torch.float32, device="cuda") for f in frames]
    torch.cuda.synchronize()
    e0.record(stream)
    for f, o in zip(frames, outs):
        npa.render_tile(f, o, spp=spp, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    seq_ms = e0.elapsed_time(e1)
    offs = [sum(g.width * g.height for g in frames[:i]) for i in range(len(frames))]
    same = all(torch.equal(o, out[3 * off:3 * (off + f.width * f.height)]) for o, f, off in zip(outs, frames, offs))
    seq = torch.cat(outs)
    ndiff = int((seq != out).view(-1, 3).any(dim=1).sum().item())
    relpix = ((seq - out).abs() / seq.abs().clamp_min(1e-30)).view(-1, 3).max(dim=1).values
    rel = relpix.max().item()
    n_gt = int((relpix > 1e-5).sum().item())
    f32, f64 = valu_flops(c, "fp32c")
    achieved = (f32 + 2.0 * f64) / (batch_ms / 1e3) / 1e12
    res = {
        "workload": f"{len(frames)} frames (256^2..1024^2, a in {{0,0.5,0.9,0.998}}, theta U[5,89] deg), "
                    f"{spp} spp, r_cam=50M fov 50deg adaptive rtol 1e-8, FP32C, one persistent launch",
        "rays": total * spp,
        "rays_per_s": total * spp / (batch_ms / 1e3),
        "batch_kernel_ms": batch_ms,
        "per_frame_launches_ms": seq_ms,
        "speedup_over_per_frame_launches": seq_ms / batch_ms,
        "outputs_bitwise_equal_to_per_frame": bool(same),
        "pixels_differing_from_per_frame": ndiff,
        "max_rel_diff_vs_per_frame": rel,
        "pixels_rel_diff_above_1e-5": n_gt,
        "tuning": best.as_dict(),
        "autotune_s": tune_s,
        "autotune_trials": [{"tuning": t, "ms": ms} for t, ms in trials],
        "steps_per_ray_mean": c["steps"] / max(1, c["rays"]),
        "fp64_attempt_frac": c["attempts_f64"] / max(1, c["attempts"]),
        "valu_peak_frac": achieved / PEAK_TFLOPS["fp32"],
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_TFLOPS["fp32"], "unit": "TFLOP/s",
                     "frac": achieved / PEAK_TFLOPS["fp32"], "kernel": "kerr_render (FP32C, tuned variant)",
                     "note": "FP64 flops count twice (FP32-equivalent)"},
    }
    if cpu_pixels_per_frame > 0:
        import numpy as np

        from oracle import np_oracle as O
        rng = np.random.default_rng(99)
        th = host_threads()
        t1 = time.perf_counter()
        for f in frames:
            q = O.default_params()
            for name, _ in npa.RenderParams._fields_:
                setattr(q, name, getattr(f, name))
            sel = rng.choice(f.width * f.height, size=cpu_pixels_per_frame, replace=False)
            O.render_pixels(q, sel % f.width, sel // f.width, 0, spp, nthreads=th, with_info=False)
        dt = time.perf_counter() - t1
        m = cpu_pixels_per_frame * len(frames) * spp
        res["cpu_baseline"] = {"value": m / dt, "unit": "rays/s", "cores": th, "kind": "port",
                               "sample": f"{cpu_pixels_per_frame} seeded random pixels x {spp} spp of each of the "
                                         f"{len(frames)} frames ({m} rays, {dt:.1f} s), FP64 oracle, OpenMP"}
    return res


def bench_schwarzschild(n: int, cpu_rays: int = 0) -> dict:
    """Plain-method FP64 Schwarzschild sweep: M = 10, n equatorial photons
    from r0 = 1000 M, b uniform on [0.5 M, 15 M] (seed 2024), captured at
    r < 2.01 M, escaped at r > r0 outbound, photon sphere |b - 3 sqrt(3) M| <
    1e-3 M.  Timed with HIP events on the launch stream; the oracle's sweep
    on cpu_rays of the same impact parameters as the CPU baseline."""
    import numpy as np
    import torch

    import nullpath_amd as npa
    M = 10.0
    rng = np.random.default_rng(2024)
    bh = rng.uniform(0.5 * M, 15.0 * M, n)
    b = torch.tensor(bh, dtype=torch.float64, device="cuda")
    total = torch.zeros(1, dtype=torch.int64, device="cuda")
    npa.schw_sweep(b[:256], M=M)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    cls, dphi, steps = npa.schw_sweep(b, M=M, total_steps=total, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    c = np.bincount(cls.cpu().numpy(), minlength=4)
    nsteps = int(total.item())
    res = {
        "workload": f"Schwarzschild M=10, {n} equatorial photons from r0=1000M, b~U[0.5M,15M], FP64 RK4",
        "rays_per_s": n / (ms / 1e3), "kernel_ms": ms,
        "steps_per_s": nsteps / (ms / 1e3),
        "escaped": int(c[0]), "captured": int(c[1]), "photon_sphere": int(c[2]), "stepcap": int(c[3]),
        "roofline": roofline(float(nsteps) * SCHW_FLOPS_PER_STEP, ms / 1e3, "fp64", "schw_sweep_kernel"),
    }
    if cpu_rays > 0:
        from oracle import np_oracle as O
        th = host_threads()
        t1 = time.perf_counter()
        O.schw_sweep(M, 1000.0 * M, 2.01 * M, 1e-3 * M, 10 ** 7, bh[:cpu_rays], nthreads=th)
        dt = time.perf_counter() - t1
        res["cpu_baseline"] = {"value": cpu_rays / dt, "unit": "rays/s", "cores": th, "kind": "port",
                               "sample": f"the first {cpu_rays} of the {n} impact parameters ({dt:.1f} s), "
                                         "FP64 oracle sweep, OpenMP"}
    return res

def bench_geomu(n: int, cpu_rays: int = 0) -> dict:
    """GEOKERR GEOMU batch: n rays with alpha, beta ~ U[-12, 12] M, a ~
    U[0, 0.998], mu0 ~ U[0.05, 0.99], target u ~ U(0, u_+) (seed 2025), from
    u0 = 1e-4, FP64 with Carlson RF; timed with HIP events around 5 launches
    into preallocated outputs."""
    import numpy as np
    import torch

    import nullpath_amd as npa
    rng = np.random.default_rng(2025)
    t = lambda x: torch.tensor(x, dtype=torch.float64, device="cuda")
    a = rng.uniform(0.0, 0.998, n)
    uplus = 1.0 / (1.0 + np.sqrt(1.0 - a * a))
    args = [t(rng.uniform(-12, 12, n)), t(rng.uniform(-12, 12, n)), t(a),
            t(rng.uniform(0.05, 0.99, n)), t(rng.uniform(0.0, 1.0, n) * uplus)]
    stream = torch.cuda.current_stream()
    reps = 5

    def timed(fn):
        """Average kernel time of `reps` launches into outputs allocated by
        a full-size warm-up call (events bracket the launches only)."""
        out = fn(*args, 1e-4, stream=stream)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn(*args, 1e-4, stream=stream, out=out)
        e1.record(stream)
        torch.cuda.synchronize()
        return out, e0.elapsed_time(e1) / reps

    (muf, tpm, ncase, iu), ms = timed(npa.geomu_camera_batch)
    nc = np.bincount(ncase.cpu().numpy(), minlength=9)
    # the compared mode: the quartic's roots from a float deflation polished
    # in double, FP64 root finding for near-degenerate arguments
    (mm, tm, cm, im, fb), ms_mixed = timed(npa.geomu_camera_batch_mixed)
    same = (cm == ncase) & (tm == tpm)
    dmu = (mm - muf).abs()
    res = {
        "workload": f"GEOMU {n} rays, alpha,beta~U[-12,12]M, a~U[0,0.998], mu0~U[0.05,0.99], u~U(0,u+), FP64",
        "rays_per_s": n / (ms / 1e3), "kernel_ms": ms,
        "ncase_counts": [int(v) for v in nc],
        "nan_muf": int(torch.isnan(muf).sum().item()),
        "mixed_roots": {
            "rays_per_s": n / (ms_mixed / 1e3), "kernel_ms": ms_mixed, "speedup": ms / ms_mixed,
            "fp64_root_fallback_frac": float(fb.double().mean().item()),
            "case_or_tpm_differ_frac": float((~same).double().mean().item()),
            "muf_abs_diff_max": float(torch.nan_to_num(dmu, nan=0.0).max().item()),
            "muf_abs_diff_gt_1e-10_frac": float((dmu > 1e-10).double().mean().item()),
        },
    }
    if GEOMU_FLOPS_PER_RAY is not None:
        res["roofline"] = roofline(GEOMU_FLOPS_PER_RAY * n, ms / 1e3, "fp64", "geomu_camera_kernel<false>")
        res["roofline"]["note"] = "flops per ray measured by the FP64 flop counter on active lanes"
        # the same work priced the same way: the compared mode solves the same
        # rays, its float root phase standing in for FP64 work
        res["mixed_roots"]["roofline"] = roofline(GEOMU_FLOPS_PER_RAY * n, ms_mixed / 1e3, "fp64",
                                                  "geomu_camera_kernel<true>")
        res["mixed_roots"]["roofline"]["note"] = "the FP64 kernel's flops per ray, FP64 peak"
    if cpu_rays > 0:
        from oracle import np_oracle as O
        host = [x[:cpu_rays].cpu().numpy() for x in args]
        th = host_threads()

        def one(k):
            O.geomu_batch(*host, 1e-4, nthreads=th)
            return cpu_rays

        done, calls, dt = cpu_for(CPU_SAMPLE_S, one)
        res["cpu_baseline"] = {"value": done / dt, "unit": "rays/s", "cores": th, "kind": "port",
                               "sample": f"the first {cpu_rays} of the {n} rays, {calls} passes ({dt:.1f} s), "
                                         "FP64 oracle GEOMU restatement, OpenMP"}
    return res


def geokerr_cases(n: int, seed: int = 2026):
    """Mixed geodesics for the GEOKERR driver: camera-like starts u0 ~
    U[1e-3, 0.02] (r = 50..1000 M), toward uf ~ U(0.1, 0.9) (through the
    radial turning point when the ray scatters), a ~ U[0, 0.99],
    L ~ U[-7, 7], Q2 ~ U[0.5, 40], mu0 inside the allowed range."""
    import numpy as np
    rng = np.random.default_rng(seed)
    u0 = rng.uniform(1e-3, 0.02, n)
    uf = rng.uniform(0.1, 0.9, n)
    a = rng.uniform(0.0, 0.99, n)
    lz = rng.uniform(-7, 7, n)
    q2 = rng.uniform(0.5, 40, n)
    mu0 = rng.uniform(-0.6, 0.6, n)
    th = q2 - (q2 + lz * lz - a * a) * mu0 ** 2 - a * a * mu0 ** 4
    mu0[th < 0] = 0.0
    return u0, mu0, uf, a, lz, q2


def bench_geokerr(n: int, s: int, cpu_geodesics: int) -> dict:
    """geokerr_eval_lambda batch (np_geokerr_eval_lambda_batch): n geodesics
    x s driver samples resampled onto lambda = 0.1 k; HIP events around the
    three kernels on the launch stream; the oracle on cpu_geodesics of the
    same inputs on the host cores."""
    import numpy as np
    import torch

    import nullpath_amd as npa
    u0, mu0, uf, a, lz, q2 = geokerr_cases(n)
    t = lambda x: torch.tensor(x, dtype=torch.float64, device="cuda")
    args = [t(u0), t(mu0), t(uf), t(a), t(lz), t(q2), t(np.zeros(n)), t(np.zeros(n)), t(np.full(n, 0.1))]
    ws = npa._geokerr_workspace(s, n, "cuda")
    npa.geokerr_eval_lambda_batch(*[x[:64] for x in args], s, workspace=ws)   # warm-up
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 3
    e0.record(stream)
    for _ in range(reps):
        out = npa.geokerr_eval_lambda_batch(*args, s, stream=stream, workspace=ws)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    res = {
        "workload": f"geokerr_eval_lambda {n} geodesics x {s} samples (GEOKERR driver, NUP = S), "
                    "u0~U[1e-3,0.02], uf~U(0.1,0.9), a~U[0,0.99], L~U[-7,7], Q2~U[0.5,40], FP64",
        "samples_per_s": n * s / (ms / 1e3), "batch_ms": ms,
        "rays_with_u_turning_point": int((out["tpri"] >= 0).sum().item()),
        "nan_rows": int(torch.isnan(out["t"]).any(dim=1).sum().item()),
    }
    if cpu_geodesics > 0:
        import os
        import time

        from oracle import np_oracle as O
        m = min(cpu_geodesics, n)
        threads = min(16, len(os.sched_getaffinity(0)))

        def one(k):
            i0 = (k * m) % n
            sl = slice(i0, min(i0 + m, n))
            cnt = sl.stop - sl.start
            O.geokerr_eval_lambda_batch(u0[sl], mu0[sl], uf[sl], a[sl], lz[sl], q2[sl], np.zeros(cnt), 0.0, 0.1,
                                        s, nthreads=threads)
            return cnt * s

        done, calls, dt = cpu_for(CPU_SAMPLE_S, one)
        res["cpu_baseline"] = {"value": done / dt, "unit": "samples/s", "cores": threads, "kind": "port",
                               "sample": f"{calls} chunks of {m} of the {n} geodesics in turn ({done // s} "
                                         f"geodesics, {s} samples each, {dt:.1f} s)"}
    return res


def bench_geokerr_driver(n: int, cpu_geodesics: int) -> dict:
    """The whole GEOKERR driver (np_geokerr_batch, split into per-sample
    lanes; np_geokerr_batch_serial timed beside it) in its two other
    configurations: u-sampled with MUFILL (NUP 512, NPTS 64, KEXT 2) and
    mu-sampled (USEGEOR, NUP 512), both with PHIT, on n camera geodesics
    from u0 = 1e-4 (alpha, beta ~ U[-12, 12] M, a ~ U[0.05, 0.998],
    mu0 ~ U[0.05, 0.95], seed 2027); HIP events on the launch stream; the
    oracle on cpu_geodesics of the same geodesics on the host cores."""
    import numpy as np
    import torch

    import nullpath_amd as npa
    rng = np.random.default_rng(2027)
    a = rng.uniform(0.05, 0.998, n)
    mu0 = rng.uniform(0.05, 0.95, n)
    al, be = rng.uniform(-12, 12, (2, n))
    lz = -al * np.sqrt(1 - mu0 ** 2)
    q2 = be ** 2 - (a * a - al * al) * mu0 ** 2
    sm = np.where(be >= 0, 1.0, -1.0)
    uplus = 1.0 / (1.0 + np.sqrt(1.0 - a * a))
    # mu_- from M(mu) = Q2 + (a^2 - Q2 - L^2) mu^2 - a^2 mu^4 (FINDMROOTS)
    amq = a * a - q2 - lz * lz
    yy = -0.5 * (amq + np.sign(amq) * np.sqrt(amq * amq + 4 * q2 * a * a))
    mneg = np.where(amq < 0, -yy / (a * a), q2 / yy)
    mpos = np.minimum(np.where(amq < 0, q2 / yy, -yy / (a * a)), 1.0)
    muminus = np.where(mneg > 0, np.sqrt(np.abs(mneg)), -np.sqrt(mpos))
    cases = {
        "mufill": (dict(nup=512, npts=64, kext=2, phit=1, usegeor=0, mufill=1, offset=0.5),
                   rng.uniform(0, 1, n) * uplus, np.zeros(n), 1),
        "usegeor": (dict(nup=512, npts=0, kext=0, phit=1, usegeor=1, mufill=0, offset=0.5),
                    np.zeros(n), mu0 + 0.6 * (muminus - mu0), 0),
    }
    t = lambda x, dt=torch.float64: torch.tensor(np.ascontiguousarray(x), dtype=dt, device="cuda")
    out = {}
    for name, (opts, uf, muf, tpr) in cases.items():
        tpm = ((np.sign(mu0) * sm + 1) / 2).astype(np.int32) if opts["usegeor"] else np.zeros(n, np.int32)
        rays = {"u0": t(np.full(n, 1e-4)), "uf": t(uf), "uout": t(np.full(n, 1e-4)), "mu0": t(mu0), "muf": t(muf),
                "a": t(a), "l": t(lz), "q2": t(q2), "beta": t(be), "su": t(np.ones(n)), "sm": t(sm),
                "tpm": t(tpm, torch.int32), "tpr": t(np.full(n, tpr), torch.int32)}
        for serial in (False, True):   # warm-up
            npa.geokerr_batch(opts, {k: v[:64] for k, v in rays.items()}, serial=serial)
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

        def timed(serial):
            e0.record(stream)
            r = npa.geokerr_batch(opts, rays, stream=stream, serial=serial)
            e1.record(stream)
            torch.cuda.synchronize()
            return r, e0.elapsed_time(e1)

        res, ms = timed(False)
        ser, ms_serial = timed(True)
        int_mismatch, float_rel = 0, 0.0
        for k, v in res.items():
            w = ser[k]
            if v.dtype == torch.float64:
                both_nan = torch.isnan(v) & torch.isnan(w)
                den = torch.maximum(torch.maximum(v.abs(), w.abs()), torch.ones_like(v))
                rel = torch.where(both_nan, torch.zeros_like(v), (v - w).abs() / den)
                float_rel = max(float_rel, float(torch.nan_to_num(rel, nan=1.0).max().item()))
            else:
                int_mismatch += int((v != w).sum().item())
        written = int((res["tpri"] >= 0).sum().item())
        entry = {"opts": opts, "batch_ms": ms, "samples_written": written,
                 "samples_per_s": written / (ms / 1e3),
                 "geodesics_with_mufill": int((res["nup_out"] > opts["nup"]).sum().item()),
                 "serial_batch_ms": ms_serial, "speedup_over_serial": ms_serial / ms,
                 "vs_serial": {"integer_mismatches": int_mismatch, "float_max_rel_diff": float_rel}}
        del ser
        if cpu_geodesics > 0:
            import os
            import time

            from oracle import np_oracle as O
            m = min(cpu_geodesics, n)
            threads = min(16, len(os.sched_getaffinity(0)))
            host = {k: rays[k].cpu().numpy() for k in ("u0", "uf", "uout", "mu0", "muf", "a", "l", "q2", "beta",
                                                          "su", "sm", "tpm", "tpr")}

            def one(k, opts=opts, host=host):
                i0 = (k * m) % n
                g = np.zeros(min(m, n - i0), O.GK_ARGS_DTYPE)
                for key, v in host.items():
                    g[key] = v[i0:i0 + len(g)]
                ref = O.geokerr_full_batch(opts, g, nthreads=threads)
                return int((ref["tpri"] >= 0).sum())

            done, calls, dt = cpu_for(CPU_SAMPLE_S, one)
            entry["cpu_baseline"] = {"value": done / dt, "unit": "samples/s", "cores": threads, "kind": "port",
                                     "sample": f"{calls} chunks of {m} of the {n} geodesics in turn ({done} samples, "
                                               f"{dt:.1f} s)"}
        out[name] = entry
    out["workload"] = (f"GEOKERR driver (np_geokerr_batch), {n} camera geodesics from u0 = 1e-4: MUFILL "
                       "(NUP 512, NPTS 64, KEXT 2) and USEGEOR (NUP 512), PHIT on, FP64")
    return out


def cpu_baseline(args, fp64: dict) -> dict:
    """The FP64 oracle (a plain-C restatement of the reference integrator)
    on a bounded random sample of the same frame, on the host cores."""
    import numpy as np

    import nullpath_amd as npa
    from oracle import np_oracle as O
    p = fp64["params"]
    q = O.default_params()
    for name, _ in npa.RenderParams._fields_:
        setattr(q, name, getattr(p, name))
    npix = fp64["npix"]
    rng = np.random.default_rng(12345)
    sel = rng.choice(npix, size=min(args.cpu_pixels, npix), replace=False)
    xs = (sel % args.width).astype(np.int32)
    ys = (sel // args.width).astype(np.int32)
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    threads = args.cpu_threads or min(16, ncpu)
    t1 = time.perf_counter()
    o_rgb, _ = O.render_pixels(q, xs, ys, 0, args.spp, nthreads=threads, with_info=False)
    cpu_s = time.perf_counter() - t1
    g_rgb = fp64["img"].reshape(-1, 3)[sel]
    rel = np.abs(g_rgb - o_rgb) / np.maximum(np.abs(o_rgb), 1e-6)
    return {
        "value": len(sel) * args.spp / cpu_s,
        "unit": "rays/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{len(sel)} seeded random pixels x {args.spp} spp of the same frame "
                   f"({len(sel) * args.spp} rays, {cpu_s:.1f} s), FP64 oracle restatement, OpenMP"),
        "gpu_fp64_vs_oracle_max_rel_err_on_sample": float(rel.max()),
    }


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build",
              file=sys.stderr)
        return 2
    if args.copy_calibration_mib > 0:
        return copy_calibration(args.copy_calibration_mib)
    if args.geokerr_driver_only > 0:
        print(json.dumps(bench_geokerr_driver(args.geokerr_driver_only, 0)))
        return 0
    if args.geomu_only > 0:
        print(json.dumps(bench_geomu(args.geomu_only)))
        return 0
    if args.only:
        cpu = args.cpu_pixels > 0
        run = {"schw": lambda: bench_schwarzschild(100000, 100000 if cpu else 0),
               "geomu": lambda: bench_geomu(1 << 20, (1 << 20) if cpu else 0),
               "adjoint": lambda: bench_adjoint(args),
               "batch": lambda: bench_batch(2, 1024 if cpu else 0)}[args.only]
        print(json.dumps(run()))
        return 0
    if args.geokerr_only > 0:
        print(json.dumps(bench_geokerr(args.geokerr_only, 4096, 0)))
        return 0
    precs = args.precisions.split(",")
    res = {}
    for prec in precs:
        k, w = (args.steps, args.warmup) if prec == "fp32c" or len(precs) == 1 else (args.fp64_steps, 0)
        waves_p = {"fp32c": args.waves_per_simd, "fp64c": args.fp64c_waves_per_simd,
                   "fp64": args.fp64_waves_per_simd}[prec]
        res[prec] = render_case(prec, args.crossing, args.width, args.height, args.spp, k, w, waves_p,
                                bool(args.times), args.renorm_pr)
    head = "fp32c" if "fp32c" in res else precs[0]
    R = res[head]
    value = R["rays"] / R["wall"]
    waves = args.waves_per_simd or 4
    if head != "fp32c":
        kernel = "kerr_render_capped<StepF64, adaptive, 2 waves>"
    elif waves == 2:
        kernel = "kerr_render<StepF32C, adaptive>"
    else:
        kernel = f"kerr_render_capped<StepF32C, adaptive, {waves} waves>"
    achieved = (R["flops32"] + 2.0 * R["flops64"]) / R["mean_s"] / 1e12
    # the traffic constant was measured on the default kernel variant
    head_1080p4 = head == "fp32c" and waves == 4 and (args.width, args.height, args.spp) == (1920, 1080, 4)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": R["wall"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head == "fp32c" else "f64",
        "data": "synthetic (seeded jittered pinhole rays; no external data)",
        "config": {
            "workload": (f"Kerr thin-disk still {args.width}x{args.height}x{args.spp}spp a=0.9 "
                         f"theta=80deg fov=55deg r_cam=1000M adaptive rtol=1e-8 disk ISCO-20M q=3 "
                         f"<=3 crossings, {args.crossing} crossing radius"),
            "precision": head,
            "rays_per_step": R["rays"] // args.steps,
            "integrates_t_and_lambda": bool(args.times),
            "renorm_pr": args.renorm_pr,
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_TFLOPS["fp32"],
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_TFLOPS["fp32"],
            "traffic": FP32C_HBM_BYTES_1080P4 if head_1080p4 else None,
            "traffic_note": ("HBM bytes per launch, rocprofv3 FETCH_SIZE + WRITE_SIZE on this workload "
                             "(copy-calibrated); beyond the image and ray records, register-spill scratch around the rare FP64 attempts"
                             if head_1080p4 else "measured only for the 1920x1080x4 FP32C launch"),
            "bytes_algorithmic": 12 * R["npix"] + 32 * R["rays"] // args.steps,
            "kernel": kernel,
            "flops_per_launch": {"fp32": R["flops32"], "fp64": R["flops64"]},
            "note": "FP64 flops count twice (FP32-equivalent) since the FP64 rate is half",
        },
        "cpu_baseline": None,
    }
    for prec, r in res.items():
        line["config"][prec] = case_summary(r)
    if "fp64c" in res and "fp32c" in res:
        a, b = res["fp64c"], res["fp32c"]
        line["config"]["speedup_fp32c_over_fp64c"] = (b["rays"] / b["wall"]) / (a["rays"] / a["wall"])
        line["config"]["fp32c_vs_fp64c"] = compare(a, b)
    if "fp64" in res and "fp32c" in res:
        a, b = res["fp64"], res["fp32c"]
        line["config"]["speedup_fp32c_over_fp64"] = (b["rays"] / b["wall"]) / (a["rays"] / a["wall"])
        line["config"]["fp32c_vs_fp64"] = compare(a, b)
        if args.linear_check != "none":
            w, h, s = (int(v) for v in args.linear_check.split("x"))
            la = render_case("fp64", "linear", w, h, s, 1, 0)
            lb = render_case("fp32c", "linear", w, h, s, 1, 0)
            lin = compare(la, lb)
            lin["size"] = args.linear_check
            lin["speedup"] = (lb["rays"] / lb["wall"]) / (la["rays"] / la["wall"])
            line["config"]["fp32c_vs_fp64_linear_crossing"] = lin
    if args.full_contract and "fp32c" in res and not args.times and not args.renorm_pr:
        # the same frame with every item of the FP32C precision contract on:
        # t and lambda integrated (np_render_tile_times) and p_r renormalised
        # from H = 0; its image against the FP64 path's
        fc = render_case("fp32c", args.crossing, args.width, args.height, args.spp, 1, 0,
                         args.waves_per_simd, True, 1)
        full = {"rays_per_s": fc["rays"] / fc["wall"], "kernel_ms": fc["kern_ms"],
                "time_vs_headline": fc["mean_s"] / res["fp32c"]["mean_s"]}
        if "fp64" in res:
            full["speedup_over_fp64"] = full["rays_per_s"] / (res["fp64"]["rays"] / res["fp64"]["wall"])
            full["vs_fp64"] = compare(res["fp64"], fc)
        if "fp64c" in res:
            full["speedup_over_fp64c"] = full["rays_per_s"] / (res["fp64c"]["rays"] / res["fp64c"]["wall"])
        line["config"]["fp32c_full_contract"] = full

    if args.adjoint:
        line["config"]["adjoint"] = bench_adjoint(args)
    if args.side_workloads:
        cpu = args.cpu_pixels > 0
        line["config"]["uneven_batch"] = bench_batch(2, 1024 if cpu else 0)
        line["config"]["schwarzschild_sweep"] = bench_schwarzschild(100000, 100000 if cpu else 0)
        line["config"]["geomu_batch"] = bench_geomu(1 << 20, (1 << 20) if cpu else 0)
        line["config"]["geokerr_eval_lambda"] = bench_geokerr(256, 4096, 64 if args.cpu_pixels > 0 else 0)
        line["config"]["geokerr_driver"] = bench_geokerr_driver(131072, 64 if args.cpu_pixels > 0 else 0)

    if args.cpu_pixels > 0 and "fp64" in res:
        line["cpu_baseline"] = cpu_baseline(args, res["fp64"])
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
