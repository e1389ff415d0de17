"""nullpath_amd -- MI355X (gfx950) null-geodesic hot path.

Host-side Python surface over libnullpath.so (C ABI in include/nullpath.h):
the Kerr renderer (`render_tile`, `render_accumulate`).  Device buffers are
torch tensors on a HIP device; torch is used only for memory and streams.
"""
from __future__ import annotations

import ctypes as C
import math

from ._lib import (COUNTER_FIELDS, NullPathError, RenderParams, TraceCounters, check, lib)

__all__ = ["RenderParams", "TraceCounters", "NullPathError", "default_params", "kerr_disk_params",
           "render_tile", "render_accumulate", "isco_radius", "lib"]

PI_F = 3.14159265358979323846


def _f32(x: float) -> float:
    import struct
    return struct.unpack("f", struct.pack("f", x))[0]


def default_params(**kw) -> RenderParams:
    """The reference renderer's defaults (render.cu:15-46, mode=balanced)."""
    p = RenderParams()
    p.width, p.height, p.fov_y_deg = 1280, 720, 50.0
    p.theta_cam = _f32(_f32(90.0 * _f32(PI_F)) / 180.0)
    p.phi_cam, p.mass, p.spin, p.r_cam_Rs = 0.0, 10.0, 0.0, 50.0
    p.integrator, p.camera, p.bg_mode = 0, 0, 0
    p.bg_sigma_rad = _f32(_f32(1.0 * _f32(PI_F)) / 180.0)
    p.bg_peak, p.disk_on, p.disk_gain, p.disk_max_orders = 3.0, 1, 1.0, 1
    p.check_inv, p.max_steps, p.precision, p.emiss_m_units = 0, 10000, 0, 0
    p.rtol, p.atol = 1e-10, 1e-12
    p.emiss_q, p.seed = 2.0, 0
    for k, v in kw.items():
        setattr(p, k, v)
    mass = float(p.mass)
    if "disk_rin" not in kw:
        p.disk_rin = 3.0 * 2.0 * mass
    if "disk_rout" not in kw:
        p.disk_rout = 20.0 * 2.0 * mass
    return p


def deg2rad_f32(d: float) -> float:
    """render.cu:483 in float arithmetic."""
    return _f32(_f32(_f32(d) * _f32(PI_F)) / 180.0)


def isco_radius(a: float) -> float:
    """Prograde ISCO in units of M (Bardeen, Press & Teukolsky 1972)."""
    z1 = 1.0 + (1.0 - a * a) ** (1.0 / 3.0) * ((1.0 + a) ** (1.0 / 3.0) + (1.0 - a) ** (1.0 / 3.0))
    z2 = math.sqrt(3.0 * a * a + z1 * z1)
    return 3.0 + z2 - math.sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2))


def kerr_disk_params(width=1920, height=1080, spin=0.9, theta_deg=80.0, fov=55.0,
                     rtol=1e-8, max_orders=3, q=3.0, gain=1.0, max_steps=200000,
                     adaptive=True, r_cam_Rs=500.0, **kw) -> RenderParams:
    """Kerr thin-disk still: pinhole at r = 2 r_cam_Rs M (default 1000 M;
    M = 1), disk from the ISCO to 20 M with emissivity (r/M)^-q and g^3
    redshift."""
    p = default_params(width=width, height=height, fov_y_deg=fov, mass=1.0, spin=spin,
                       r_cam_Rs=r_cam_Rs, theta_cam=deg2rad_f32(theta_deg),
                       integrator=1 if adaptive else 0, camera=1 if adaptive else 0,
                       rtol=rtol, atol=1e-12, disk_max_orders=max_orders, emiss_q=q,
                       emiss_m_units=1, disk_gain=gain, max_steps=max_steps,
                       disk_rin=isco_radius(spin), disk_rout=20.0, **kw)
    return p


def _stream_ptr(stream) -> int | None:
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    if hasattr(stream, "cuda_stream"):
        return stream.cuda_stream
    return int(stream)


def render_tile(p: RenderParams, out, y0=0, rows=None, spp=1, sample_offset=0, inv=None,
                counters=None, records=None, stream=None) -> None:
    """Launch the renderer on rows [y0, y0+rows) into `out` (torch float32
    device tensor of rows*width*3 elements)."""
    rows = p.height - y0 if rows is None else rows
    if out.numel() < rows * p.width * 3 or out.dtype.itemsize != 4 or not out.is_contiguous():
        raise NullPathError("out must be a contiguous float32 device tensor of rows*width*3")
    code = lib().np_render_tile(p, y0, rows, spp, sample_offset, out.data_ptr(),
                                None if inv is None else inv.data_ptr(),
                                None if counters is None else counters.data_ptr(),
                                None if records is None else records.data_ptr(),
                                _stream_ptr(stream))
    check(code, "np_render_tile")


def render_tile_times(p: RenderParams, out, times, y0=0, rows=None, spp=1, sample_offset=0,
                      counters=None, records=None, stream=None) -> None:
    """np_render_tile_times: render_tile that also writes each sample's end
    coordinate time, azimuth and affine parameter into `times` (device buffer
    of rows*width*spp np_ray_times, 24 bytes each; view it with
    ray_times_dtype())."""
    rows = p.height - y0 if rows is None else rows
    if out.numel() < rows * p.width * 3 or out.dtype.itemsize != 4 or not out.is_contiguous():
        raise NullPathError("out must be a contiguous float32 device tensor of rows*width*3")
    if times.numel() * times.element_size() < rows * p.width * spp * 24 or not times.is_contiguous():
        raise NullPathError("times must be a contiguous device buffer of rows*width*spp*24 bytes")
    code = lib().np_render_tile_times(p, y0, rows, spp, sample_offset, out.data_ptr(),
                                      None if counters is None else counters.data_ptr(),
                                      None if records is None else records.data_ptr(),
                                      times.data_ptr(), _stream_ptr(stream))
    check(code, "np_render_tile_times")


def new_counters(device="cuda"):
    """A zeroed device buffer for np_trace_counters (uint64 fields as int64)."""
    import torch
    return torch.zeros(len(COUNTER_FIELDS), dtype=torch.int64, device=device)


def render_pass(p: RenderParams, accum, sample_offset=0, spp=1, inv=None, stream=None) -> None:
    """np_render_pass: add one progressive pass of `spp` samples into the
    float32 device tensor `accum` (width*height*3)."""
    if accum.numel() < p.width * p.height * 3 or accum.dtype.itemsize != 4 or not accum.is_contiguous():
        raise NullPathError("accum must be a contiguous float32 device tensor of width*height*3")
    check(lib().np_render_pass(p, sample_offset, spp, accum.data_ptr(),
                               None if inv is None else inv.data_ptr(), _stream_ptr(stream)),
          "np_render_pass")


class PassGraph:
    """np_pass_graph: `passes` progressive passes of `spp` samples captured in
    a HIP graph; launch() replays them into `accum`, continuing the sample
    ids from the graph's device-side counter."""

    def __init__(self, p: RenderParams, accum, spp=1, passes=1, first_sample=0, inv=None):
        import ctypes as C
        if accum.numel() < p.width * p.height * 3 or accum.dtype.itemsize != 4:
            raise NullPathError("accum must be a float32 device tensor of width*height*3")
        self._accum = accum
        self._inv = inv
        h = C.c_void_p()
        check(lib().np_pass_graph_create(p, spp, passes, accum.data_ptr(),
                                         None if inv is None else inv.data_ptr(), first_sample, C.byref(h)),
              "np_pass_graph_create")
        self._h = h

    def launch(self, stream=None) -> None:
        check(lib().np_pass_graph_launch(self._h, _stream_ptr(stream)), "np_pass_graph_launch")

    def seek(self, next_sample: int, stream=None) -> None:
        check(lib().np_pass_graph_seek(self._h, next_sample, _stream_ptr(stream)), "np_pass_graph_seek")

    def close(self) -> None:
        if self._h:
            lib().np_pass_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def counters_dict(t) -> dict:
    """Decode a uint64 counter tensor (np_trace_counters, 10 elements) into a dict."""
    vals = t.cpu().tolist()
    return {k: int(v) & ((1 << 64) - 1) for k, v in zip(COUNTER_FIELDS, vals)}


CARLSON_KINDS = {"rf": 0, "rc": 1, "rd": 2, "rj": 3}


def _dptr(t):
    return None if t is None else t.data_ptr()


def carlson_batch(kind: str, x, y, z=None, p=None, out=None, stream=None):
    """Batched Carlson integral on float64 device tensors (np_carlson_batch)."""
    import torch
    out = torch.empty_like(x) if out is None else out
    code = lib().np_carlson_batch(CARLSON_KINDS[kind], _dptr(x), _dptr(y), _dptr(z), _dptr(p),
                                  x.numel(), out.data_ptr(), _stream_ptr(stream))
    check(code, "np_carlson_batch")
    return out


def _geomu_outputs(n, device, out, count):
    import torch
    if out is not None:
        assert len(out) == count and all(o.numel() == n for o in out)
        return out
    muf = torch.empty(n, dtype=torch.float64, device=device)
    tpm = torch.empty(n, dtype=torch.int32, device=device)
    rest = [torch.empty_like(tpm)] + [torch.empty_like(muf)] + [torch.empty_like(tpm)] * (count - 4)
    return (muf, tpm, *rest)


def geomu_camera_batch(alpha, beta, spin, mu0, uf, u0, stream=None, out=None):
    """GEOMU from screen coordinates (np_geomu_camera_batch); float64 device
    tensors in, (muf, tpm, ncase, iu) device tensors out (written into `out`
    when given: float64, int32, int32, float64 tensors of n)."""
    n = alpha.numel()
    muf, tpm, ncase, iu = _geomu_outputs(n, alpha.device, out, 4)
    code = lib().np_geomu_camera_batch(alpha.data_ptr(), beta.data_ptr(), spin.data_ptr(),
                                       mu0.data_ptr(), uf.data_ptr(), float(u0), n, muf.data_ptr(),
                                       tpm.data_ptr(), ncase.data_ptr(), iu.data_ptr(),
                                       _stream_ptr(stream))
    check(code, "np_geomu_camera_batch")
    return muf, tpm, ncase, iu


def geomu_camera_batch_mixed(alpha, beta, spin, mu0, uf, u0, stream=None, out=None):
    """GEOMU from screen coordinates with the quartic's roots in mixed
    precision (np_geomu_camera_batch_mixed); returns (muf, tpm, ncase, iu,
    fp64_roots) device tensors, fp64_roots = 1 where the FP64 root finder
    ran (written into `out` when given)."""
    n = alpha.numel()
    muf, tpm, ncase, iu, fb = _geomu_outputs(n, alpha.device, out, 5)
    code = lib().np_geomu_camera_batch_mixed(alpha.data_ptr(), beta.data_ptr(), spin.data_ptr(),
                                             mu0.data_ptr(), uf.data_ptr(), float(u0), n, muf.data_ptr(),
                                             tpm.data_ptr(), ncase.data_ptr(), iu.data_ptr(), fb.data_ptr(),
                                             _stream_ptr(stream))
    check(code, "np_geomu_camera_batch_mixed")
    return muf, tpm, ncase, iu, fb


def geomu_batch(u0, uf, mu0, spin, L, Q2, su, sm, tpr=None, stream=None):
    """GEOMU on constants of motion (np_geomu_batch)."""
    import torch
    n = u0.numel()
    muf = torch.empty(n, dtype=torch.float64, device=u0.device)
    iu = torch.empty_like(muf)
    tpm = torch.empty(n, dtype=torch.int32, device=u0.device)
    ncase = torch.empty_like(tpm)
    code = lib().np_geomu_batch(u0.data_ptr(), uf.data_ptr(), mu0.data_ptr(), spin.data_ptr(),
                                L.data_ptr(), Q2.data_ptr(), su.data_ptr(), sm.data_ptr(),
                                _dptr(tpr), n, muf.data_ptr(), tpm.data_ptr(), ncase.data_ptr(),
                                iu.data_ptr(), _stream_ptr(stream))
    check(code, "np_geomu_batch")
    return muf, tpm, ncase, iu


def geor_batch(u0, mu0, muf, spin, L, Q2, su, sm, tpm, uf_in=None, stream=None):
    """GEOR on constants of motion (np_geor_batch): float64 device tensors
    of n (tpm int32); returns (uf, imu, tpr, ncase) device tensors."""
    import torch
    n = u0.numel()
    uf = torch.empty(n, dtype=torch.float64, device=u0.device)
    imu = torch.empty_like(uf)
    tpr = torch.empty(n, dtype=torch.int32, device=u0.device)
    ncase = torch.empty_like(tpr)
    code = lib().np_geor_batch(u0.data_ptr(), mu0.data_ptr(), muf.data_ptr(), spin.data_ptr(), L.data_ptr(),
                               Q2.data_ptr(), su.data_ptr(), sm.data_ptr(), tpm.data_ptr(), _dptr(uf_in), n,
                               uf.data_ptr(), imu.data_ptr(), tpr.data_ptr(), ncase.data_ptr(),
                               _stream_ptr(stream))
    check(code, "np_geor_batch")
    return uf, imu, tpr, ncase


def geokerr_batch(opts, rays, stream=None, serial=False):
    """The whole GEOKERR driver (np_geokerr_batch) for n geodesics.

    opts: dict with nup, npts, kext, phit, usegeor, mufill, offset (the
    Fortran's switches, geokerr_original.f:357-360).  rays: dict of device
    tensors, n entries each -- float64 u0, uf, uout, mu0, muf, a, l, q2,
    beta, su, sm and int32 tpm, tpr.  Returns a dict of device tensors:
    ufi, mufi, dti, dphi, lambdai (float64) and tpmi, tpri (int32), each
    n x (nup + npts - 2 kext), and per geodesic u0_out, uf_out, muf_out,
    tpm_out, tpr_out, ncase_out, nup_out, npts_out.  serial=True runs
    np_geokerr_batch_serial (one lane walks each geodesic's samples)."""
    import torch
    from ._lib import GEOKERR_RAY_FIELDS, GeokerrOpts, GeokerrRays, GeokerrSamples
    o = GeokerrOpts(**{k: opts[k] for k in ("nup", "npts", "kext", "phit", "usegeor", "mufill", "offset")})
    n = rays["u0"].numel()
    dev = rays["u0"].device
    tot = max(int(o.nup) + int(o.npts) - 2 * int(o.kext), 0)
    for k in GEOKERR_RAY_FIELDS:
        want = torch.int32 if k in ("tpm", "tpr") else torch.float64
        t = rays[k]
        if t.dtype != want or t.numel() != n or not t.is_contiguous() or t.device != dev:
            raise ValueError(f"geokerr_batch: rays[{k!r}] must be a contiguous {want} tensor of {n} on {dev}")
    out = {k: torch.empty((n, tot), dtype=torch.float64, device=dev) for k in ("ufi", "mufi", "dti", "dphi", "lambdai")}
    out.update({k: torch.empty((n, tot), dtype=torch.int32, device=dev) for k in ("tpmi", "tpri")})
    out.update({k: torch.empty(n, dtype=torch.float64, device=dev) for k in ("u0_out", "uf_out", "muf_out")})
    out.update({k: torch.empty(n, dtype=torch.int32, device=dev)
                for k in ("tpm_out", "tpr_out", "ncase_out", "nup_out", "npts_out")})
    r = GeokerrRays(**{k: rays[k].data_ptr() for k in GEOKERR_RAY_FIELDS})
    s = GeokerrSamples(**{k: v.data_ptr() for k, v in out.items()})
    fn = lib().np_geokerr_batch_serial if serial else lib().np_geokerr_batch
    check(fn(C.byref(o), C.byref(r), n, C.byref(s), _stream_ptr(stream)),
          "np_geokerr_batch_serial" if serial else "np_geokerr_batch")
    return out


GEOKERR_PERIOD_SAMPLES = 4096


def _geokerr_workspace(nup, n, device):
    import torch
    nbytes = lib().np_geokerr_workspace_bytes(int(nup), int(n))
    if nbytes < 0:
        raise NullPathError("np_geokerr_workspace_bytes: invalid sizes")
    # torch's caching allocator returns blocks aligned to at least 512 bytes
    return torch.empty(max(nbytes, 1), dtype=torch.uint8, device=device)


def geokerr_eval_lambda_batch(u0, mu0, uf, a, lz, q2, t0, lam0, dlam, s, stream=None, workspace=None):
    """n geokerr_eval_lambda calls on float64 device tensors of n entries
    (np_geokerr_eval_lambda_batch).  Returns a dict of device tensors:
    u, mu, t, phi, affine (n x s) and tpmi, tpri, info (n, int32)."""
    import torch
    n = u0.numel()
    dev = u0.device
    out = {k: torch.empty((n, s), dtype=torch.float64, device=dev) for k in ("u", "mu", "t", "phi", "affine")}
    for k in ("tpmi", "tpri", "info"):
        out[k] = torch.empty(n, dtype=torch.int32, device=dev)
    ws = _geokerr_workspace(s, n, dev) if workspace is None else workspace
    code = lib().np_geokerr_eval_lambda_batch(
        u0.data_ptr(), mu0.data_ptr(), uf.data_ptr(), a.data_ptr(), lz.data_ptr(), q2.data_ptr(),
        t0.data_ptr(), lam0.data_ptr(), dlam.data_ptr(), int(s), n, out["u"].data_ptr(), out["mu"].data_ptr(),
        out["t"].data_ptr(), out["phi"].data_ptr(), out["affine"].data_ptr(), out["tpmi"].data_ptr(),
        out["tpri"].data_ptr(), out["info"].data_ptr(), ws.data_ptr(), _stream_ptr(stream))
    check(code, "np_geokerr_eval_lambda_batch")
    return out


def geokerr_estimate_periods_batch(u0, mu0, a, lz, q2, stream=None):
    """n geokerr_estimate_periods calls (np_geokerr_estimate_periods_batch):
    returns (Tr, Ttheta, info) device tensors."""
    import torch
    n = u0.numel()
    tr = torch.empty(n, dtype=torch.float64, device=u0.device)
    tth = torch.empty_like(tr)
    info = torch.empty(n, dtype=torch.int32, device=u0.device)
    ws = _geokerr_workspace(GEOKERR_PERIOD_SAMPLES, n, u0.device)
    code = lib().np_geokerr_estimate_periods_batch(u0.data_ptr(), mu0.data_ptr(), a.data_ptr(), lz.data_ptr(),
                                                   q2.data_ptr(), n, tr.data_ptr(), tth.data_ptr(),
                                                   info.data_ptr(), ws.data_ptr(), _stream_ptr(stream))
    check(code, "np_geokerr_estimate_periods_batch")
    return tr, tth, info


def geokerr_phi_omega_batch(u0, mu0, a, lz, q2, stream=None):
    """np_geokerr_phi_omega_batch on float64 device tensors."""
    import torch
    out = torch.empty_like(u0)
    code = lib().np_geokerr_phi_omega_batch(u0.data_ptr(), mu0.data_ptr(), a.data_ptr(), lz.data_ptr(),
                                            q2.data_ptr(), u0.numel(), out.data_ptr(), _stream_ptr(stream))
    check(code, "np_geokerr_phi_omega_batch")
    return out


def geokerr_eval_lambda(u0, mu0, uf, a, lz, q2, t0, lam0, dlam, s):
    """The drop-in geokerr_eval_lambda C entry point (host arrays)."""
    import ctypes as C
    import numpy as np
    arrs = [np.zeros(max(int(s), 0)) for _ in range(5)]
    tpmi, tpri, info = C.c_int(), C.c_int(), C.c_int()
    pd = C.POINTER(C.c_double)
    lib().geokerr_eval_lambda(u0, mu0, uf, a, lz, q2, t0, lam0, dlam, int(s),
                              *[x.ctypes.data_as(pd) for x in arrs], C.byref(tpmi), C.byref(tpri), C.byref(info))
    u, mu, t, phi, aff = arrs
    return dict(u=u, mu=mu, t=t, phi=phi, affine=aff, tpmi=tpmi.value, tpri=tpri.value, info=info.value)


def geokerr_estimate_periods(u0, mu0, a, lz, q2):
    """The drop-in geokerr_estimate_periods C entry point."""
    import ctypes as C
    tr, tth, info = C.c_double(), C.c_double(), C.c_int()
    lib().geokerr_estimate_periods(u0, mu0, a, lz, q2, C.byref(tr), C.byref(tth), C.byref(info))
    return tr.value, tth.value, info.value


def geokerr_phi_omega(u0, mu0, a, lz, q2):
    """The drop-in geokerr_phi_omega C entry point."""
    return lib().geokerr_phi_omega(u0, mu0, a, lz, q2)


def geokerr_port_stream(u0, uf, mu0, a, lz, q2, lam0, dlam, s, symbol="geokerr_complete_stream_cuda"):
    """The port-solver C entry points (include/geokerr_port.h):
    geokerr_complete_stream_cuda or geokerr_stream_cuda.  Returns
    (code, dict of u, mu, t, phi, affine host arrays); code 0 on success,
    -1 / -2 / -3 as the reference."""
    import ctypes as C
    import numpy as np
    arrs = [np.zeros(max(int(s), 0)) for _ in range(5)]
    pd = C.POINTER(C.c_double)
    code = getattr(lib(), symbol)(u0, uf, mu0, a, lz, q2, lam0, dlam, int(s),
                                  *[x.ctypes.data_as(pd) for x in arrs])
    return code, dict(zip(("u", "mu", "t", "phi", "affine"), arrs))


def schw_sweep(b, M=10.0, r0=None, r_capture=None, ps_band=None, max_steps=2_000_000,
               stream=None, total_steps=None, diagnostics=False):
    """FP64 plain-method Schwarzschild sweep (np_schw_sweep) over a float64
    device tensor of impact parameters; returns (cls, dphi, steps), and
    with diagnostics=True also (r_min, |H|max) per ray."""
    import torch
    r0 = 1000.0 * M if r0 is None else r0
    r_capture = 2.01 * M if r_capture is None else r_capture
    ps_band = 1e-3 * M if ps_band is None else ps_band
    n = b.numel()
    cls = torch.empty(n, dtype=torch.int32, device=b.device)
    dphi = torch.empty(n, dtype=torch.float64, device=b.device)
    steps = torch.empty(n, dtype=torch.int64, device=b.device)
    rmin = torch.empty(n, dtype=torch.float64, device=b.device) if diagnostics else None
    hmax = torch.empty(n, dtype=torch.float64, device=b.device) if diagnostics else None
    code = lib().np_schw_sweep(M, r0, r_capture, ps_band, max_steps, b.data_ptr(), n,
                               cls.data_ptr(), dphi.data_ptr(), steps.data_ptr(), _dptr(rmin), _dptr(hmax),
                               _dptr(total_steps), _stream_ptr(stream))
    check(code, "np_schw_sweep")
    return (cls, dphi, steps, rmin, hmax) if diagnostics else (cls, dphi, steps)


def schw_precompute(b, mass=10.0, r_start=1000.0, stream=None):
    """The reference's FP32 lookup-table kernel (np_schw_precompute)."""
    import torch
    n = b.numel()
    defl = torch.empty(n, dtype=torch.float32, device=b.device)
    rmin = torch.empty_like(defl)
    st = torch.empty(n, dtype=torch.int32, device=b.device)
    code = lib().np_schw_precompute(mass, r_start, b.data_ptr(), n, defl.data_ptr(),
                                    rmin.data_ptr(), st.data_ptr(), _stream_ptr(stream))
    check(code, "np_schw_precompute")
    return defl, rmin, st


def precompute_geodesics(mass=10.0, num_rays=20000):
    """precomputeGeodesics through its C twin; returns numpy arrays."""
    import ctypes as C

    import numpy as np

    from ._lib import LookupTable
    t = LookupTable()
    check(lib().np_precompute_geodesics(mass, num_rays, C.byref(t)), "np_precompute_geodesics")
    try:
        n = t.num_entries
        out = {k: np.ctypeslib.as_array(getattr(t, k), shape=(n,)).copy() for k in (
            "impact_parameters", "deflection_angles", "bending_angles", "closest_approaches",
            "ray_status")}
    finally:
        lib().np_free_lookup_table(C.byref(t))
    return out


def render_accumulate(p: RenderParams, accum, spp_done=0, samples_per_pass=1, passes=1,
                      tile_h=128, inv_max=None) -> None:
    """Progressive passes into a host numpy float32 accumulation buffer."""
    import ctypes as C

    import numpy as np
    assert accum.dtype == np.float32 and accum.flags.c_contiguous
    invp = None
    if inv_max is not None:
        assert inv_max.dtype == np.float32
        invp = inv_max.ctypes.data
    code = lib().np_render_accumulate(p, spp_done, samples_per_pass, passes, tile_h,
                                      accum.ctypes.data_as(C.c_void_p), invp)
    check(code, "np_render_accumulate")


def render_batch(frames, out, spp=1, sample_offset=0, records=None, counters=None, tuning=None,
                 stream=None, noalloc=False) -> None:
    """All frames (a list of RenderParams) in one persistent launch
    (np_render_batch); `out` is one float32 device tensor holding the frames'
    colour sums back to back.  noalloc=True calls np_render_batch_noalloc
    (no device allocation; at most NP_NOALLOC_MAX_FRAMES frames)."""
    import ctypes as C

    from ._lib import Tuning
    n = len(frames)
    arr = (RenderParams * n)(*frames)
    total = sum(f.width * f.height for f in frames)
    if out.numel() < 3 * total or out.dtype.itemsize != 4 or not out.is_contiguous():
        raise NullPathError("out must be a contiguous float32 device tensor of 3 * total pixels")
    t = None if tuning is None else C.byref(tuning if isinstance(tuning, Tuning) else Tuning(**tuning))
    fn = lib().np_render_batch_noalloc if noalloc else lib().np_render_batch
    code = fn(arr, n, spp, sample_offset, out.data_ptr(),
              None if records is None else records.data_ptr(),
              None if counters is None else counters.data_ptr(), t, _stream_ptr(stream))
    check(code, "np_render_batch_noalloc" if noalloc else "np_render_batch")


def autotune_batch(frames, spp=1, subsample=4, max_trials=64):
    """np_autotune_batch: returns (best Tuning, [(Tuning dict, ms), ...])."""
    import ctypes as C

    from ._lib import Tuning
    n = len(frames)
    arr = (RenderParams * n)(*frames)
    best = Tuning()
    trials = (Tuning * max_trials)()
    ms = (C.c_double * max_trials)()
    nt = C.c_int(0)
    code = lib().np_autotune_batch(arr, n, spp, subsample, C.byref(best), trials, ms, max_trials,
                                   C.byref(nt))
    check(code, "np_autotune_batch")
    return best, [(trials[i].as_dict(), ms[i]) for i in range(nt.value)]
