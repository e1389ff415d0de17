#!/usr/bin/env python3
"""Benchmark: Kerr thin-disk still on one MI355X.

Workload (one "step" = one full frame): 1920x1080 pinhole camera at
r = 1000 M, a = 0.9 M, theta_obs = 80 deg, fov 55 deg, 4 jittered samples
per pixel, adaptive step-doubling RK4 (rtol 1e-8, atol 1e-12), disk from
the ISCO to 20 M with emissivity (r/M)^-3 and g^3 redshift, up to 3
equatorial crossings, float32 RGB output.  Inputs (frame constants) are
resident before the timed region; the output stays in HBM.

Prints one JSON line (see the keys below).  --gpus above 1 is refused: this
build measures one GPU.
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
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3

# Algorithmic FP64 operation counts of the reference formulas (add, sub,
# mul, div, sqrt = 1; fma = 2), derived line by line from
# bh_kerr.h:600-727 and 987-1105 and listed in DESIGN.md.  Transcendentals
# (one sincos per RHS, one pow per attempt) are not included.
FLOPS_RHS = 200          # one Hamilton right-hand side (metric, derivatives, two Klein sums)
FLOPS_RK4_COMBINE = 196  # stage states + compensated k1+2k2+2k3+k4 + update, per RK4 step
FLOPS_ATTEMPT_EXTRA = 29 # error norm and step-factor arithmetic per adaptive attempt


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=4)
    ap.add_argument("--cpu-pixels", type=int, default=160,
                    help="pixels (x spp rays) in the CPU-baseline sample; 0 skips it")
    ap.add_argument("--cpu-threads", type=int, default=0)
    return ap.parse_args()


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build",
              file=sys.stderr)
        return 2
    import numpy as np
    import torch

    import nullpath_amd as npa

    dev = torch.device("cuda:0")
    p = npa.kerr_disk_params(width=args.width, height=args.height, spin=0.9, theta_deg=80.0,
                             fov=55.0, rtol=1e-8, max_orders=3, q=3.0, seed=1)
    npix = p.width * p.height
    rays_per_step = npix * args.spp
    out = torch.zeros(npix * 3, dtype=torch.float32, device=dev)
    cnt = torch.zeros(8, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream()

    # load the code object and prime the queue slots on a tiny frame
    warm = npa.kerr_disk_params(width=64, height=36, spin=0.9, theta_deg=80.0, fov=55.0,
                                rtol=1e-8, max_orders=3, q=3.0, seed=1)
    npa.render_tile(warm, out, spp=1)
    for _ in range(args.warmup):
        npa.render_tile(p, out, spp=args.spp)
    torch.cuda.synchronize()

    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        cnt.zero_()
        ev[k][0].record(stream)
        npa.render_tile(p, out, spp=args.spp, counters=cnt)
        ev[k][1].record(stream)
        print(f"[bench] step {k + 1}/{args.steps} launched", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = [a.elapsed_time(b) for a, b in ev]
    c = npa.counters_dict(cnt)
    assert c["rays"] == rays_per_step, c

    ms_per_step = wall * 1e3 / args.steps
    value = rays_per_step * args.steps / wall
    mean_kern_s = sum(kern_ms) / len(kern_ms) / 1e3
    rk4_steps = 3 * c["attempts"]
    flops = c["rhs_evals"] * FLOPS_RHS + rk4_steps * FLOPS_RK4_COMBINE + c["attempts"] * FLOPS_ATTEMPT_EXTRA
    achieved = flops / mean_kern_s / 1e12

    img = out.view(p.height, p.width, 3).cpu().numpy()
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded jittered pinhole rays; no external data)",
        "config": {
            "workload": f"Kerr thin-disk still {p.width}x{p.height}x{args.spp}spp a=0.9 "
                        f"theta=80deg fov=55deg r_cam=1000M adaptive rtol=1e-8 disk ISCO-20M q=3 "
                        f"<=3 crossings, FP64 path",
            "precision": "fp64",
            "rays_per_step": rays_per_step,
            "accepted_steps_per_ray": c["steps"] / c["rays"],
            "attempts_per_ray": c["attempts"] / c["rays"],
            "rk4_steps_per_s": rk4_steps / mean_kern_s,
            "rhs_evals_per_s": c["rhs_evals"] / mean_kern_s,
            "captured": c["captured"], "escaped": c["escaped"], "stepcap": c["stepcap"],
            "disk_rays": c["disk_hits"],
            "kernel_ms": kern_ms,
            "image_mean_rgb": [float(v) for v in img.reshape(-1, 3).mean(axis=0)],
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": None,
            "kernel": "kerr_render_fp64<adaptive>",
            "flops_per_launch": flops,
        },
        "cpu_baseline": None,
    }

    if args.cpu_pixels > 0:
        from oracle import np_oracle as O
        q = O.default_params()
        for name, _ in npa.RenderParams._fields_:
            setattr(q, name, getattr(p, name))
        rng = np.random.default_rng(12345)
        sel = rng.choice(npix, size=args.cpu_pixels, replace=False)
        xs = (sel % p.width).astype(np.int32)
        ys = (sel // p.width).astype(np.int32)
        try:
            ncpu = len(os.sched_getaffinity(0))
        except AttributeError:
            ncpu = os.cpu_count() or 1
        threads = args.cpu_threads or min(16, ncpu)
        t1 = time.perf_counter()
        o_rgb, _ = O.render_pixels(q, xs, ys, 0, args.spp, nthreads=threads, with_info=False)
        cpu_s = time.perf_counter() - t1
        g_rgb = img.reshape(-1, 3)[sel]
        rel = np.abs(g_rgb - o_rgb) / np.maximum(np.abs(o_rgb), 1e-6)
        line["cpu_baseline"] = {
            "value": args.cpu_pixels * args.spp / cpu_s,
            "unit": "rays/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{args.cpu_pixels} seeded random pixels x {args.spp} spp of the same frame "
                      f"({args.cpu_pixels * args.spp} rays, {cpu_s:.1f} s), FP64 oracle restatement, OpenMP",
            "gpu_vs_oracle_max_rel_err_on_sample": float(rel.max()),
        }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
