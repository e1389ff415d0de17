"""np_render_batch_noalloc (the allocation-free batch render) against
np_render_batch through the C ABI: two calls in a row, the second on a
smaller batch, each bitwise equal to the allocating entry point's output
(same kernels, same frame table); capacity and argument errors."""
import ctypes as C

import numpy as np
import pytest
import torch

import nullpath_amd as npa
from nullpath_amd._lib import NullPathError, RenderParams, ray_record_dtype

NP_ECAPACITY = -6
NP_NOALLOC_MAX_FRAMES = 128


def _frames(sizes, spins, thetas, precision):
    return [npa.kerr_disk_params(width=w, height=h, spin=a, theta_deg=t, fov=30.0, seed=5,
                                 precision=precision, crossing=1, r_cam_Rs=25.0)
            for (w, h), a, t in zip(sizes, spins, thetas)]


def _render(frames, spp, noalloc):
    total = sum(f.width * f.height for f in frames)
    out = torch.full((3 * total,), -1.0, dtype=torch.float32, device="cuda")
    cnt = npa.new_counters()
    rec = torch.zeros(total * spp * 32, dtype=torch.uint8, device="cuda")
    print(f"[gpu] np_render_batch{'_noalloc' if noalloc else ''} frames={len(frames)} "
          f"px={total} spp={spp}", flush=True)
    npa.render_batch(frames, out, spp=spp, counters=cnt, records=rec, noalloc=noalloc)
    torch.cuda.synchronize()
    return out.cpu().numpy(), npa.counters_dict(cnt), rec.cpu().numpy().view(ray_record_dtype())


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [1, 0])
def test_noalloc_batch_equals_allocating_batch(device, precision):
    big = _frames([(40, 24), (32, 32), (48, 20), (24, 36), (16, 16)],
                  [0.0, 0.5, 0.9, 0.998, 0.7], [20.0, 45.0, 70.0, 88.0, 60.0], precision)
    small = _frames([(24, 16), (16, 20)], [0.9, 0.3], [80.0, 30.0], precision)
    for frames in (big, small):   # the second call is smaller: stale scratch must not leak in
        ref_img, ref_cnt, ref_rec = _render(frames, 2, noalloc=False)
        img, cnt, rec = _render(frames, 2, noalloc=True)
        np.testing.assert_array_equal(img, ref_img)
        np.testing.assert_array_equal(rec, ref_rec)
        assert cnt == ref_cnt
        assert cnt["rays"] == 2 * sum(f.width * f.height for f in frames)


@pytest.mark.gpu
def test_noalloc_batch_rejects_too_many_frames(device):
    frames = _frames([(8, 8)] * (NP_NOALLOC_MAX_FRAMES + 1), [0.5] * (NP_NOALLOC_MAX_FRAMES + 1),
                     [60.0] * (NP_NOALLOC_MAX_FRAMES + 1), 1)
    out = torch.zeros(3 * 64 * len(frames), dtype=torch.float32, device="cuda")
    arr = (RenderParams * len(frames))(*frames)
    code = npa.lib().np_render_batch_noalloc(arr, len(frames), 1, 0, out.data_ptr(), None, None, None,
                                             None)
    assert code == NP_ECAPACITY
    assert npa.lib().np_status_string(code) == b"request exceeds a fixed capacity of the entry point"
    with pytest.raises(NullPathError):
        npa.render_batch(frames[:3], out, spp=1, noalloc=True, tuning={"block_size": 100,
                                                                        "waves_per_simd": 3,
                                                                        "refill_chunk": 1,
                                                                        "fp64_radius": 0.0})
    # the full capacity is accepted
    npa.render_batch(frames[:NP_NOALLOC_MAX_FRAMES], out, spp=1, noalloc=True)
    torch.cuda.synchronize()
    assert np.isfinite(out[:3 * 64 * NP_NOALLOC_MAX_FRAMES].cpu().numpy()).all()
