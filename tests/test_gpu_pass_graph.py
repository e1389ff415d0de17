"""Progressive passes captured in a HIP graph (np_pass_graph_*) against the
same passes launched one by one (np_render_pass): every replay, with eager
work and synchronizations of the program between replays, equals the eager
accumulation bit for bit; replays continue the sample ids; seek restarts."""
import numpy as np
import pytest
import torch

import nullpath_amd as npa


def _eager(p, passes, spp, first=0):
    acc = torch.zeros(p.width * p.height * 3, dtype=torch.float32, device="cuda")
    for k in range(passes):
        npa.render_pass(p, acc, sample_offset=first + k * spp, spp=spp)
    torch.cuda.synchronize()
    return acc.cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [1, 0])
def test_graph_replays_equal_eager_passes(device, precision):
    p = npa.kerr_disk_params(width=40, height=24, fov=20.0, r_cam_Rs=15.0, seed=2, precision=precision,
                             crossing=1, spin=0.8)
    spp, passes = 2, 3
    acc = torch.zeros(p.width * p.height * 3, dtype=torch.float32, device="cuda")
    print(f"[gpu] np_pass_graph precision={precision}", flush=True)
    g = npa.PassGraph(p, acc, spp=spp, passes=passes, first_sample=0)
    other = torch.zeros_like(acc)
    for replay in range(1, 4):
        g.launch()
        # the program's own eager work between replays
        npa.render_pass(p, other, sample_offset=100 + replay, spp=1)
        torch.cuda.synchronize()
        want = _eager(p, passes * replay, spp)
        np.testing.assert_array_equal(acc.cpu().numpy(), want, err_msg=f"replay {replay}")
    # seek: a fresh buffer replayed from sample 0 again equals the first replay
    acc.zero_()
    g.seek(0)
    g.launch()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(acc.cpu().numpy(), _eager(p, passes, spp))
    # a pass graph started at sample 6 = eager passes from 6
    acc.zero_()
    g.seek(6)
    g.launch()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(acc.cpu().numpy(), _eager(p, passes, spp, first=6))
    g.close()


@pytest.mark.gpu
def test_graph_accumulates_invariant_maxima(device):
    p = npa.default_params(width=24, height=16, spin=0.7, theta_cam=npa.deg2rad_f32(70.0), fov_y_deg=20.0,
                           r_cam_Rs=15.0, check_inv=1, disk_max_orders=2)
    acc = torch.zeros(p.width * p.height * 3, dtype=torch.float32, device="cuda")
    inv_g = torch.zeros(3, dtype=torch.float32, device="cuda")
    g = npa.PassGraph(p, acc, spp=1, passes=2, inv=inv_g)
    g.launch()
    acc2 = torch.zeros_like(acc)
    inv_e = torch.zeros(3, dtype=torch.float32, device="cuda")
    for k in range(2):
        npa.render_pass(p, acc2, sample_offset=k, spp=1, inv=inv_e)
    torch.cuda.synchronize()
    assert torch.equal(acc, acc2) and torch.equal(inv_g, inv_e)
    assert inv_g[0].item() > 0.0
    g.close()
