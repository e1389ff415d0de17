"""The `render` command line (drop-in for render.cu's CLI): flag parsing,
validation messages and exit codes on the CPU; an end-to-end PPM against the
oracle on the GPU."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RENDER = os.path.join(ROOT, "nullpath_amd", "bin", "render")


def run(*args, cwd=None):
    return subprocess.run([RENDER, *args], capture_output=True, text=True, cwd=cwd, timeout=300)


@pytest.mark.parametrize("args,msg", [
    (["--w", "0"], "Invalid image dimensions."),
    (["-h", "-3"], "Invalid image dimensions."),
    (["--tile", "0"], "Invalid tile height."),
    (["-s", "0"], "Invalid samples per pixel."),
    (["--spp-total", "-1"], "Invalid total spp."),
    (["--frames", "0"], "Invalid frame count."),
    (["--mass", "-1"], "Invalid mass or camera radius."),
    (["--rcam", "0"], "Invalid mass or camera radius."),
    (["--spin", "1.0"], "Spin must satisfy |a/M| < 1."),
    (["--spin", "-1.5"], "Spin must satisfy |a/M| < 1."),
    (["--fov", "179"], "Invalid FOV; expected (0, 179) degrees."),
    (["--fov", "0"], "Invalid FOV; expected (0, 179) degrees."),
    (["--precision", "fp16"], "Invalid --precision; expected fp64 or fp32c."),
])
def test_validation_messages(args, msg):
    if not os.path.exists(RENDER):
        pytest.fail("render CLI not built")
    r = run(*args)
    assert r.returncode == 1
    assert msg in r.stderr


def test_unknown_flags_and_missing_values_are_ignored():
    # parsing quirks kept from render.cu:391-425: unknown flags are skipped and a
    # trailing flag without its value leaves the default; validation then passes
    # and the run stops only at the device check when no GPU is present.
    r = run("--bogus", "7", "--w")
    if "No HIP devices found." in r.stderr:
        assert r.returncode == 1
    else:
        pytest.skip("a HIP device is present; covered by the GPU test")


def _tonemap(accum, spp, exposure=0.0, gamma=2.2, reinhard=False):
    f = np.float32
    x = (accum.astype(f) * (f(1.0) / f(spp))) * f(2.0 ** exposure)
    if reinhard:
        x = x / (f(1.0) + x)
    if abs(gamma - 1.0) > 1e-6:
        x = np.power(np.maximum(x, f(0.0)), f(1.0) / f(max(1e-6, gamma))).astype(f)
    x = np.clip(x, 0.0, 1.0).astype(f)
    return (f(255.0) * x).astype(np.uint8)


@pytest.mark.gpu
def test_cli_end_to_end_against_oracle(oracle, device, tmp_path):
    import nullpath_amd as npa
    out = tmp_path / "frame.ppm"
    raw = tmp_path / "frame.f32"
    W, H = 40, 24
    print("[gpu] render CLI 40x24", flush=True)
    r = run("--w", str(W), "-h", str(H), "--spin", "0.6", "--theta", "75", "--fov", "20",
            "--samples", "1", "--spp-total", "3", "--checkpoint-every", "2",
            "--disk-max-orders", "2", "--out", str(out), "--raw-out", str(raw))
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert f"[checkpoint] {out} @ 2 spp" in lines and f"[checkpoint] {out} @ 3 spp" in lines
    assert lines[-1] == f"Wrote {out}"
    data = out.read_bytes()
    header = f"P6\n{W} {H}\n255\n".encode()
    assert data.startswith(header)
    img = np.frombuffer(data[len(header):], dtype=np.uint8).reshape(H, W, 3)
    p = npa.default_params(width=W, height=H, spin=0.6, theta_cam=npa.deg2rad_f32(75.0),
                           fov_y_deg=20.0, disk_max_orders=2)
    q = oracle.default_params()
    for name, _ in npa.RenderParams._fields_:
        setattr(q, name, getattr(p, name))
    ys, xs = np.mgrid[0:H, 0:W]
    acc = np.zeros((H * W, 3), np.float32)
    for s in range(3):   # three passes of one sample, accumulated in pass order
        rgb, _ = oracle.render_pixels(q, xs.ravel(), ys.ravel(), s, 1, with_info=False)
        acc += rgb
    ref = _tonemap(acc, 3).reshape(H, W, 3)
    diff = np.abs(img.astype(int) - ref.astype(int))
    assert np.mean(diff == 0) > 0.99 and diff.max() <= 2
    lin = np.fromfile(raw, dtype=np.float32).reshape(H * W, 3)
    rel = np.abs(lin - acc / np.float32(3)) / np.maximum(np.abs(acc / 3), 1e-6)
    assert np.median(rel) < 1e-6
