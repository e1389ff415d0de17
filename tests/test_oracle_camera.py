"""The oracle's pinhole camera and sample jitter (orc_camera_dir, restating
render.cu:88-99,137-159), pinned by the geometry a pinhole fixes rather than
by retyping its formula:

- every sample direction of pixel (x, y) projects back into that pixel's
  own footprint in normalised device coordinates, with the horizontal scale
  aspect * tan(fov / 2) and the vertical scale tan(fov / 2): a swapped axis,
  a wrong aspect ratio, a flipped row order or a dropped half-pixel offset
  moves samples out of their pixel;
- the forward axis is -z of the camera frame (n[0] = cam_z < 0 is the
  inward radial component) and the directions are unit vectors;
- the jitter covers the pixel: over many samples its offset from the pixel
  centre spans (-1/2, 1/2) in both axes with mean ~0 and mean |offset|
  ~1/4 (uniform), and different pixels and samples draw different offsets.

The hash keys themselves (which offset a given pixel and sample draw) stay
pinned by reading only; the GPU's directions equal these bit for bit
(tests/test_oracle_kerr.py and the GPU parity tests)."""
import numpy as np
import pytest

from oracle import np_oracle as O


def _ndc(p, n):
    """Normalised device coordinates of a camera-frame direction."""
    cam_z, cam_y, cam_x = float(n[0]), float(n[1]), float(n[2])
    tan_half = np.tan(0.5 * np.deg2rad(p.fov_y_deg))
    aspect = p.width / p.height
    return (cam_x / -cam_z) / (aspect * tan_half), -(cam_y / -cam_z) / tan_half


@pytest.mark.parametrize("w,h,fov,seed", [(64, 36, 50.0, 0), (33, 47, 20.0, 7), (128, 128, 80.0, 12345)])
def test_samples_land_in_their_pixel(oracle, w, h, fov, seed):
    p = O.default_params()
    p.width, p.height, p.fov_y_deg, p.seed = w, h, fov, seed
    rng = np.random.default_rng(seed + 1)
    for _ in range(400):
        x, y, s = int(rng.integers(0, w)), int(rng.integers(0, h)), int(rng.integers(0, 1 << 20))
        n, _ = O.camera_dir(p, x, y, s)
        assert abs(float(np.dot(n.astype(np.float64), n.astype(np.float64))) - 1.0) < 1e-6
        assert n[0] < 0.0
        nx, ny = _ndc(p, n)
        # pixel footprint in ndc, with float rounding slack
        lo_x, hi_x = 2.0 * x / w - 1.0, 2.0 * (x + 1) / w - 1.0
        lo_y, hi_y = 2.0 * y / h - 1.0, 2.0 * (y + 1) / h - 1.0
        eps = 1e-5
        assert lo_x - eps <= nx <= hi_x + eps, (x, y, s, nx, lo_x, hi_x)
        assert lo_y - eps <= ny <= hi_y + eps, (x, y, s, ny, lo_y, hi_y)


def test_image_orientation(oracle):
    """x grows to the right of the view (cam_x > 0) and rows grow downward
    (cam_y < 0); the corner pixels sit at the field of view's edges."""
    p = O.default_params()
    p.width, p.height, p.fov_y_deg = 200, 100, 60.0
    n_tl, _ = O.camera_dir(p, 0, 0, 0)
    n_br, _ = O.camera_dir(p, 199, 99, 0)
    assert n_tl[2] < 0 < n_br[2]          # cam_x: left negative, right positive
    assert n_tl[1] > 0 > n_br[1]          # cam_y: top positive, bottom negative
    tan_half = np.tan(np.deg2rad(30.0))
    for n in (n_tl, n_br):
        ny = abs(float(n[1]) / -float(n[0]))
        assert tan_half * (1 - 2.0 / 100) <= ny <= tan_half * (1 + 1e-5)


def test_jitter_is_uniform_over_the_pixel(oracle):
    p = O.default_params()
    p.width, p.height, p.fov_y_deg, p.seed = 16, 16, 40.0, 3
    offs = []
    for s in range(4096):
        n, _ = O.camera_dir(p, 5, 9, s)
        nx, ny = _ndc(p, n)
        # offset from the pixel centre, in pixels
        offs.append(((nx + 1.0) * 0.5 * p.width - 5.5, (ny + 1.0) * 0.5 * p.height - 9.5))
    offs = np.array(offs)
    assert offs.min() > -0.5 - 1e-4 and offs.max() < 0.5 + 1e-4
    assert offs.min(axis=0).max() < -0.49 and offs.max(axis=0).min() > 0.49
    assert np.all(np.abs(offs.mean(axis=0)) < 0.02)
    assert np.all(np.abs(np.abs(offs).mean(axis=0) - 0.25) < 0.02)
    # the two axes draw independent offsets
    assert abs(np.corrcoef(offs[:, 0], offs[:, 1])[0, 1]) < 0.05
    # different pixels draw different offsets for the same sample
    a, _ = O.camera_dir(p, 5, 9, 0)
    b, _ = O.camera_dir(p, 6, 9, 0)
    ax, _ = _ndc(p, a)
    bx, _ = _ndc(p, b)
    assert abs((bx - ax) * 0.5 * p.width - 1.0) > 1e-6
