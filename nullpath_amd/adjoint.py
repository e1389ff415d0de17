"""Differentiable Kerr disk render: the np_adjoint_* C ABI wrapped for torch.

`DiskRender.apply(params4, base)` renders the image (float32, H x W x 3, on
the HIP device) for params4 = (spin, theta_obs [rad], emissivity index q,
gain) with the FP32-compensated forward of libnullpath and returns it; its
backward runs the hand-written adjoint kernel (checkpoint segments
recomputed, discrete RK4 vector-Jacobian products) and returns dLoss/dparams4
as a float64 tensor.  params4 lives on the host (float64, 4 elements); the
image and every buffer live on the device.  There is no CPU path: without
libnullpath.so every call raises NullPathError.
"""
from __future__ import annotations

import ctypes as C
import math

from ._lib import AdjointParams, check, lib

PREC_FP64 = 0
PREC_FP32C = 1


def adjoint_params(**kw) -> AdjointParams:
    """np_adjoint_params_default (include/nullpath.h) with keyword overrides."""
    p = AdjointParams(width=512, height=512, fov_y_deg=2.8, seed=1, r_cam=1000.0,
                      theta_obs=60.0 * math.pi / 180.0, spin=0.7, emiss_q=3.0, gain=1.0,
                      disk_rout=20.0, edge_width=0.5, h=0.005, n_steps=4096, ckpt_every=64,
                      max_orders=3, reserved=0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def with_params4(base: AdjointParams, p4) -> AdjointParams:
    q = base.copy()
    q.spin, q.theta_obs, q.emiss_q, q.gain = (float(v) for v in p4)
    return q


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class Workspace:
    """Device workspace of one parameter shape (checkpoints, crossing
    steps, segment recompute)."""

    def __init__(self, p: AdjointParams, device="cuda"):
        import torch
        nbytes = lib().np_adjoint_workspace_bytes(C.byref(p))
        if nbytes < 0:
            check(-1, "np_adjoint_workspace_bytes")
        self.nbytes = int(nbytes)
        self.buf = torch.empty(self.nbytes, dtype=torch.uint8, device=device)
        self.shape = (p.width, p.height, p.n_steps, p.ckpt_every)

    def fits(self, p: AdjointParams) -> bool:
        return self.shape == (p.width, p.height, p.n_steps, p.ckpt_every)


def forward(p: AdjointParams, precision=PREC_FP32C, image=None, ws: Workspace | None = None,
            isum=None):
    """Render into `image` (float32 device tensor, H*W*3); FP32C needs `ws`,
    FP64 optionally fills `isum` (float64 device tensor, H*W)."""
    import torch
    if image is None:
        image = torch.empty(p.height * p.width * 3, dtype=torch.float32, device="cuda")
    wsp = ws.buf.data_ptr() if ws is not None else None
    code = lib().np_adjoint_forward(C.byref(p), precision, image.data_ptr(), wsp,
                                    isum.data_ptr() if isum is not None else None, _stream())
    check(code, "np_adjoint_forward")
    return image


def backward(p: AdjointParams, grad_image, ws: Workspace, grad4=None):
    """dLoss/d(spin, theta_obs, q, gain) as a float64 device tensor."""
    import torch
    if grad4 is None:
        grad4 = torch.empty(4, dtype=torch.float64, device=grad_image.device)
    code = lib().np_adjoint_backward(C.byref(p), grad_image.data_ptr(), ws.buf.data_ptr(),
                                     grad4.data_ptr(), _stream())
    check(code, "np_adjoint_backward")
    return grad4


def _diskrender_class():
    import torch

    class DiskRender(torch.autograd.Function):
        """image = render(params4); params4 = (spin, theta_obs, q, gain)."""

        @staticmethod
        def forward(ctx, params4, base: AdjointParams, ws: Workspace):
            p = with_params4(base, params4.detach().cpu().tolist())
            assert ws.fits(p)
            img = forward(p, PREC_FP32C, ws=ws)
            ctx.p = p
            ctx.ws = ws
            return img.view(p.height, p.width, 3)

        @staticmethod
        def backward(ctx, grad_img):
            g = backward(ctx.p, grad_img.contiguous().view(-1).float(), ctx.ws)
            return g.cpu(), None, None

    return DiskRender


_DR = None


def disk_render(params4, base: AdjointParams, ws: Workspace):
    """Differentiable render (torch.autograd.Function)."""
    global _DR
    if _DR is None:
        _DR = _diskrender_class()
    return _DR.apply(params4, base, ws)
