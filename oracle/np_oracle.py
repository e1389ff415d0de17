"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker of the HIP library,
never as a product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


class RenderParams(C.Structure):
    """Mirror of np_render_params (include/np_render.h)."""

    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("fov_y_deg", C.c_float),
        ("theta_cam", C.c_float), ("phi_cam", C.c_float), ("mass", C.c_float),
        ("spin", C.c_float), ("r_cam_Rs", C.c_float), ("integrator", C.c_int32),
        ("camera", C.c_int32), ("bg_mode", C.c_int32), ("bg_sigma_rad", C.c_float),
        ("bg_peak", C.c_float), ("disk_on", C.c_int32), ("disk_gain", C.c_float),
        ("disk_max_orders", C.c_int32), ("check_inv", C.c_int32),
        ("max_steps", C.c_int32), ("precision", C.c_int32), ("emiss_m_units", C.c_int32),
        ("rtol", C.c_double), ("atol", C.c_double), ("disk_rin", C.c_double),
        ("disk_rout", C.c_double), ("emiss_q", C.c_float), ("seed", C.c_uint32),
        ("fp64_radius", C.c_double), ("waves_per_simd", C.c_int32),
        ("crossing", C.c_int32), ("estimate_every_step", C.c_int32), ("renorm_pr", C.c_int32),
        ("defer_fp64", C.c_int32),
    ]


def default_params(**kw) -> RenderParams:
    """Reference renderer defaults (render.cu:15-46, mode=balanced)."""
    p = RenderParams()
    p.width, p.height, p.fov_y_deg = 1280, 720, 50.0
    p.theta_cam = np.float32(90.0) * np.float32(np.pi) / np.float32(180.0)
    p.phi_cam, p.mass, p.spin, p.r_cam_Rs = 0.0, 10.0, 0.0, 50.0
    p.integrator, p.camera, p.bg_mode = 0, 0, 0
    p.bg_sigma_rad = np.float32(1.0) * np.float32(np.pi) / np.float32(180.0)
    p.bg_peak, p.disk_on, p.disk_gain, p.disk_max_orders = 3.0, 1, 1.0, 1
    p.check_inv, p.max_steps, p.precision, p.emiss_m_units = 0, 10000, 0, 0
    p.rtol, p.atol = 1e-10, 1e-12
    p.disk_rin, p.disk_rout = 6.0 * p.mass, 40.0 * p.mass
    p.emiss_q, p.seed = 2.0, 0
    for k, v in kw.items():
        setattr(p, k, v)
    if "mass" in kw and "disk_rin" not in kw:
        p.disk_rin = 6.0 * float(np.float32(p.mass))
    if "mass" in kw and "disk_rout" not in kw:
        p.disk_rout = 40.0 * float(np.float32(p.mass))
    return p


class RayInfo(C.Structure):
    _fields_ = [
        ("cls", C.c_int32), ("orders", C.c_int32), ("steps", C.c_int64),
        ("attempts", C.c_int64), ("rhs_evals", C.c_int64), ("min_r", C.c_double),
        ("inv", C.c_float * 3), ("rgb", C.c_float * 3),
        ("ncross", C.c_int32), ("pad_", C.c_int32), ("xr", C.c_double * 4), ("p0", C.c_double * 4),
        ("t_end", C.c_double), ("phi_end", C.c_double), ("lam_end", C.c_double), ("r_end", C.c_double),
        ("th_end", C.c_double),
    ]


RAYINFO_DTYPE = np.dtype([
    ("cls", np.int32), ("orders", np.int32), ("steps", np.int64), ("attempts", np.int64),
    ("rhs_evals", np.int64), ("min_r", np.float64), ("inv", np.float32, 3),
    ("rgb", np.float32, 3), ("ncross", np.int32), ("pad_", np.int32), ("xr", np.float64, 4),
    ("p0", np.float64, 4), ("t_end", np.float64), ("phi_end", np.float64), ("lam_end", np.float64),
    ("r_end", np.float64), ("th_end", np.float64),
], align=True)


def build() -> None:
    """Compile liboracle.so from the C sources in this folder."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L: C.CDLL) -> None:
    d, i32, pd = C.c_double, C.c_int, C.POINTER(C.c_double)
    for name in ("orc_kerr_blocks_cov_d", "orc_kerr_blocks_contra_d",
                 "orc_kerr_blocks_contra_derivs_d"):
        f = getattr(L, name)
        f.argtypes = [d, d, d, d, pd]
        f.restype = None
    L.orc_kerr_rhs_d.argtypes = [pd, d, d, pd]
    L.orc_kerr_hamiltonian_d.argtypes = [pd, d, d]
    L.orc_kerr_hamiltonian_d.restype = d
    L.orc_kerr_rk4_fixed_d.argtypes = [pd, d, d, d]
    L.orc_kerr_rk4_fixed_d.restype = i32
    L.orc_kerr_rk4_step_d.argtypes = [pd, d, d, d, pd]
    L.orc_kerr_rk4_step_d.restype = i32
    L.orc_kerr_adaptive_step_d.argtypes = [pd, d, d, pd, d, d]
    L.orc_kerr_adaptive_step_d.restype = d
    for name in ("orc_init_static_d", "orc_init_zamo_d"):
        f = getattr(L, name)
        f.argtypes = [d, d, d, d, d, d, d, d, pd]
        f.restype = None
    L.orc_static_timelike_d.argtypes = [d, d, d, d]
    L.orc_static_timelike_d.restype = i32
    L.orc_doppler_d.argtypes = [pd, d, d]
    L.orc_doppler_d.restype = d
    L.orc_rk4_sum_d.argtypes = [d, d, d, d]
    L.orc_rk4_sum_d.restype = d
    L.orc_ray_p0_pixels.argtypes = [C.POINTER(RenderParams), i32, C.c_void_p, C.c_void_p, i32, i32,
                                    C.c_void_p]
    L.orc_ray_p0_pixels.restype = None
    L.orc_wanghash.argtypes = [C.c_uint32]
    L.orc_wanghash.restype = C.c_uint32
    L.orc_camera_dir.argtypes = [C.POINTER(RenderParams), i32, i32, C.c_uint32,
                                 C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.orc_render_ray.argtypes = [C.POINTER(RenderParams), i32, i32, C.c_uint32,
                                 C.POINTER(RayInfo)]
    L.orc_render_pixels.argtypes = [C.POINTER(RenderParams), i32, C.c_void_p, C.c_void_p,
                                    i32, i32, C.c_void_p, C.c_void_p, i32]
    L.orc_schw_precompute_f32.argtypes = [C.c_float, C.c_float, C.c_void_p, i32, C.c_void_p,
                                          C.c_void_p, C.c_void_p, i32]
    L.orc_schw_rmin_analytic.argtypes = [C.c_float, C.c_float, C.POINTER(C.c_float)]
    L.orc_schw_rmin_analytic.restype = i32
    L.orc_schw_sweep_d.argtypes = [d, d, d, d, C.c_int64, C.c_void_p, i32, C.c_void_p,
                                   C.c_void_p, C.c_void_p, i32]
    for name, n in (("orc_rf", 3), ("orc_rc", 2), ("orc_rd", 3), ("orc_rj", 4)):
        f = getattr(L, name)
        f.argtypes = [d] * n
        f.restype = d


class AdjointParams(C.Structure):
    """np_adjoint_params (include/nullpath.h)."""

    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("fov_y_deg", C.c_float),
        ("seed", C.c_uint32), ("r_cam", C.c_double), ("theta_obs", C.c_double),
        ("spin", C.c_double), ("emiss_q", C.c_double), ("gain", C.c_double),
        ("disk_rout", C.c_double), ("edge_width", C.c_double), ("h", C.c_double),
        ("n_steps", C.c_int32), ("ckpt_every", C.c_int32), ("max_orders", C.c_int32),
        ("reserved", C.c_int32),
    ]


def adjoint_params(**kw) -> AdjointParams:
    """np_adjoint_params_default (include/nullpath.h), then keyword overrides."""
    p = AdjointParams(width=512, height=512, fov_y_deg=2.8, seed=1, r_cam=1000.0,
                      theta_obs=60.0 * np.pi / 180.0, spin=0.7, emiss_q=3.0, gain=1.0,
                      disk_rout=20.0, edge_width=0.5, h=0.005, n_steps=4096, ckpt_every=64,
                      max_orders=3, reserved=0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


_adj_declared = False


def _alib():
    global _adj_declared
    L = lib()
    if not _adj_declared:
        d, pd = C.c_double, C.POINTER(C.c_double)
        L.orc_adj_pixels.argtypes = [C.POINTER(AdjointParams), C.c_int, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        L.orc_adj_pixels.restype = None
        L.orc_adj_isco.argtypes = [d, pd]
        L.orc_adj_isco.restype = d
        L.orc_adj_disk_term.argtypes = [d] * 8
        L.orc_adj_disk_term.restype = d
        L.orc_adj_zamo_init.argtypes = [d, d, d, C.POINTER(C.c_float), pd]
        L.orc_adj_zamo_init.restype = None
        L.orc_adj_rhs.argtypes = [pd, d, d, d, pd]
        L.orc_adj_rhs.restype = None
        _adj_declared = True
    return L


def adj_pixels(p: AdjointParams, xs, ys, nthreads=0):
    """FP64 radiance sum and its gradient d/d(spin, theta_obs, q, gain) per
    pixel, plus the number of crossings counted."""
    xs = _arr(xs, np.int32)
    ys = _arr(ys, np.int32)
    n = len(xs)
    rad = np.zeros(n)
    drad = np.zeros((n, 4))
    nc = np.zeros(n, dtype=np.int32)
    _alib().orc_adj_pixels(C.byref(p), n, xs.ctypes.data, ys.ctypes.data, rad.ctypes.data,
                           drad.ctypes.data, nc.ctypes.data, nthreads or (os.cpu_count() or 1))
    return rad, drad, nc


def adj_isco(a):
    g = C.c_double()
    v = _alib().orc_adj_isco(a, C.byref(g))
    return v, g.value


def adj_disk_term(r, E, L, a, q, gain, rout, ew):
    return _alib().orc_adj_disk_term(r, E, L, a, q, gain, rout, ew)


def adj_zamo_init(a, th, r_cam, n):
    nf = _arr(n, np.float32)
    out = np.zeros(4)
    _alib().orc_adj_zamo_init(a, th, r_cam, nf.ctypes.data_as(C.POINTER(C.c_float)), _ptr(out))
    return out


def adj_rhs(y, E, L, a):
    y = _arr(y)
    out = np.zeros(4)
    _alib().orc_adj_rhs(_ptr(y), E, L, a, _ptr(out))
    return out


def _arr(x, dtype=np.float64):
    return np.ascontiguousarray(x, dtype=dtype)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ---- thin numpy-facing helpers ----
def blocks_cov(M, a, r, th):
    out = np.zeros(5)
    lib().orc_kerr_blocks_cov_d(M, a, r, th, _ptr(out))
    return out


def blocks_contra(M, a, r, th):
    out = np.zeros(5)
    lib().orc_kerr_blocks_contra_d(M, a, r, th, _ptr(out))
    return out


def blocks_contra_derivs(M, a, r, th):
    out = np.zeros(10)
    lib().orc_kerr_blocks_contra_derivs_d(M, a, r, th, _ptr(out))
    return out


def rhs(s, M, a):
    s = _arr(s)
    d = np.zeros(8)
    lib().orc_kerr_rhs_d(_ptr(s), M, a, _ptr(d))
    return d


def hamiltonian(s, M, a):
    s = _arr(s)
    return lib().orc_kerr_hamiltonian_d(_ptr(s), M, a)


def rk4_fixed(s, M, a, h):
    s = _arr(s).copy()
    ok = lib().orc_kerr_rk4_fixed_d(_ptr(s), M, a, h)
    return s, bool(ok)


def rk4_step(s, M, a, h):
    s = _arr(s)
    out = np.zeros(8)
    ok = lib().orc_kerr_rk4_step_d(_ptr(s), M, a, h, _ptr(out))
    return out, bool(ok)


def adaptive_step(s, M, a, h, rtol=1e-10, atol=1e-12):
    s = _arr(s).copy()
    hh = np.array([h], dtype=np.float64)
    fac = lib().orc_kerr_adaptive_step_d(_ptr(s), M, a, _ptr(hh), rtol, atol)
    return s, float(hh[0]), fac


def init_static(M, a, r, th, ph, n):
    s = np.zeros(8)
    lib().orc_init_static_d(M, a, r, th, ph, float(n[0]), float(n[1]), float(n[2]), _ptr(s))
    return s


def init_zamo(M, a, r, th, ph, n):
    s = np.zeros(8)
    lib().orc_init_zamo_d(M, a, r, th, ph, float(n[0]), float(n[1]), float(n[2]), _ptr(s))
    return s


def doppler(s, M, a):
    s = _arr(s)
    return lib().orc_doppler_d(_ptr(s), M, a)


def camera_dir(p: RenderParams, x: int, y: int, sample_id: int):
    n = (C.c_float * 3)()
    cam = (C.c_float * 2)()
    lib().orc_camera_dir(C.byref(p), x, y, sample_id, n, cam)
    return np.array(n[:], dtype=np.float32), np.array(cam[:], dtype=np.float32)


def render_pixels(p: RenderParams, xs, ys, sample_offset=0, spp=1, nthreads=0, with_info=True):
    """Per-pixel colour sums (npix, 3) and per-ray info (npix*spp) records."""
    xs = _arr(xs, np.int32)
    ys = _arr(ys, np.int32)
    n = xs.size
    rgb = np.zeros((n, 3), dtype=np.float32)
    info = np.zeros(n * spp, dtype=RAYINFO_DTYPE) if with_info else None
    assert RAYINFO_DTYPE.itemsize == C.sizeof(RayInfo)
    lib().orc_render_pixels(C.byref(p), n, xs.ctypes.data, ys.ctypes.data, sample_offset, spp,
                            rgb.ctypes.data, info.ctypes.data if with_info else None, nthreads)
    return rgb, info


def ray_p0_pixels(p: RenderParams, xs, ys, sample_offset=0, spp=1):
    """Initial (p_t, p_r, p_theta, p_phi) of every sample, shape (npix*spp, 4)."""
    xs = _arr(xs, np.int32)
    ys = _arr(ys, np.int32)
    out = np.zeros((xs.size * spp, 4), dtype=np.float64)
    lib().orc_ray_p0_pixels(C.byref(p), xs.size, xs.ctypes.data, ys.ctypes.data, sample_offset, spp,
                            out.ctypes.data)
    return out


def render_image(p: RenderParams, spp=1, sample_offset=0, nthreads=0):
    ys, xs = np.mgrid[0:p.height, 0:p.width]
    rgb, info = render_pixels(p, xs.ravel(), ys.ravel(), sample_offset, spp, nthreads)
    return rgb.reshape(p.height, p.width, 3), info


def schw_precompute_f32(mass, r_start, b, nthreads=0):
    b = _arr(b, np.float32)
    n = b.size
    defl = np.zeros(n, np.float32)
    rmin = np.zeros(n, np.float32)
    st = np.zeros(n, np.int32)
    lib().orc_schw_precompute_f32(mass, r_start, b.ctypes.data, n, defl.ctypes.data,
                                  rmin.ctypes.data, st.ctypes.data, nthreads)
    return defl, rmin, st


def schw_rmin_analytic(Rs, b):
    out = C.c_float(0.0)
    ok = lib().orc_schw_rmin_analytic(Rs, b, C.byref(out))
    return bool(ok), out.value


def schw_sweep(M, r0, r_capture, ps_band, max_steps, b, nthreads=0):
    b = _arr(b)
    n = b.size
    cls = np.zeros(n, np.int32)
    dphi = np.zeros(n, np.float64)
    steps = np.zeros(n, np.int64)
    lib().orc_schw_sweep_d(M, r0, r_capture, ps_band, max_steps, b.ctypes.data, n,
                           cls.ctypes.data, dphi.ctypes.data, steps.ctypes.data, nthreads)
    return cls, dphi, steps


def rf(x, y, z):
    return lib().orc_rf(x, y, z)


def rc(x, y):
    return lib().orc_rc(x, y)


def rd(x, y, z):
    return lib().orc_rd(x, y, z)


def rj(x, y, z, p):
    return lib().orc_rj(x, y, z, p)


# ---- GEOKERR GEOMU (geomu_oracle.c) ----
class GeomuOut(C.Structure):
    _fields_ = [("muf", C.c_double), ("iu", C.c_double), ("u1", C.c_double), ("u2", C.c_double),
                ("u3", C.c_double), ("u4", C.c_double), ("u0_used", C.c_double),
                ("uf_used", C.c_double), ("i1mu", C.c_double), ("i3mu", C.c_double),
                ("h1", C.c_double), ("tpm", C.c_int32), ("ncase", C.c_int32), ("tpr", C.c_int32),
                ("warned", C.c_int32)]


class GeorOut(C.Structure):
    _fields_ = [("uf", C.c_double), ("imu", C.c_double), ("iu0", C.c_double), ("u1", C.c_double),
                ("u2", C.c_double), ("u3", C.c_double), ("u4", C.c_double), ("h1", C.c_double),
                ("u0_used", C.c_double), ("ncase", C.c_int32), ("tpr", C.c_int32),
                ("warned", C.c_int32), ("pad_", C.c_int32)]


def _declare_geomu(L):
    d, i32 = C.c_double, C.c_int
    L.orc_geomu.argtypes = [d, d, d, d, d, d, i32, d, d, C.POINTER(GeomuOut)]
    L.orc_geomu.restype = None
    L.orc_geor.argtypes = [d] * 7 + [C.c_int32, C.c_int32, d, d, C.POINTER(GeorOut)]
    L.orc_geor.restype = None
    L.orc_geor_batch.argtypes = [C.c_void_p] * 10 + [i32] + [C.c_void_p] * 4 + [i32]
    L.orc_geor_batch.restype = None
    L.orc_calcimusymf.argtypes = [d] * 6 + [C.POINTER(d)]
    L.orc_calcimusymf.restype = d
    L.orc_calcimuasymf.argtypes = [d] * 5 + [C.POINTER(d)]
    L.orc_calcimuasymf.restype = d
    L.orc_sncndn.argtypes = [d, d, C.POINTER(d), C.POINTER(d), C.POINTER(d)]
    L.orc_sncndn.restype = None
    L.orc_zroots.argtypes = [C.c_void_p, i32, C.c_void_p, i32]
    L.orc_zroots.restype = None
    L.orc_geokerr_camera.argtypes = [d, d, d, d] + [C.POINTER(d)] * 5
    L.orc_geokerr_camera.restype = None
    L.orc_geomu_batch.argtypes = [C.c_void_p] * 5 + [d, i32] + [C.c_void_p] * 4 + [i32]
    L.orc_geomu_batch.restype = None
    pi = C.POINTER(C.c_int)
    for name, nargs in (("orc_ellcubicreal", 8), ("orc_ellquarticreal", 10)):
        f = getattr(L, name)
        f.argtypes = [pi] + [d] * nargs + [C.POINTER(d), d, d]
        f.restype = d
    L.orc_ellcubiccomplex.argtypes = [pi] + [d] * 7 + [C.POINTER(d), d, d]
    L.orc_ellcubiccomplex.restype = d
    L.orc_ellquarticcomplex.argtypes = [pi] + [d] * 9 + [C.POINTER(d), d, d]
    L.orc_ellquarticcomplex.restype = d
    L.orc_elldoublecomplex.argtypes = [pi] + [d] * 8 + [C.POINTER(d), d, d]
    L.orc_elldoublecomplex.restype = d


_geomu_declared = False


def _glib():
    global _geomu_declared
    L = lib()
    if not _geomu_declared:
        _declare_geomu(L)
        _geomu_declared = True
    return L


def geomu(u0, uf, mu0, a, l, q2, tpr=0, su=1.0, sm=1.0) -> dict:
    o = GeomuOut()
    _glib().orc_geomu(u0, uf, mu0, a, l, q2, tpr, su, sm, C.byref(o))
    return {k: getattr(o, k) for k, _ in GeomuOut._fields_}


def geor(u0, mu0, muf, a, l, q2, tpm, su=1.0, sm=1.0, uf_in=-1.0, tpr_in=0) -> dict:
    """GEOR (geokerr_original.f:2672-3134) with FIRSTPT: uf at muf after
    tpm mu turning points; uf_in is the UF the caller passes in (read only
    by the equal-root cubic branch)."""
    o = GeorOut()
    _glib().orc_geor(u0, uf_in, mu0, muf, a, l, q2, int(tpm), int(tpr_in), su, sm, C.byref(o))
    return {k: getattr(o, k) for k, _ in GeorOut._fields_ if k != "pad_"}


class GkVars(C.Structure):
    """orc_gk_vars: GEOKERR's variables shared along one geodesic."""
    _fields_ = [("ncase", C.c_int32), ("pad_", C.c_int32)] + [
        (k, C.c_double) for k in ("u1", "u2", "u3", "u4", "h1", "iu0", "rffu0", "rffu1", "i1mu", "i3mu",
                                  "rffmu1", "rffmu2", "rffmu3", "rdc", "rjc", "tu01", "tu02", "tu03", "tu04",
                                  "tmu1", "tmu3", "phimu1", "phimu3")]


def gk_geor(v: GkVars, u0, uf, mu0, muf, a, l, q2, imu, tpm, tpr, su, sm, pht=0, firstpt=1) -> dict:
    """orc_gk_geor: GEOR with the Fortran's by-reference arguments and its
    FIRSTPT/PHT switches; v carries the shared variables between calls."""
    L = _glib()
    if not getattr(L, "_gk_geor_declared", False):
        d, pd, pi = C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_int32)
        L.orc_gk_geor.argtypes = [C.POINTER(GkVars), pd, pd, d, d, d, d, d, pd, C.c_int32, pi, d, d,
                                  C.c_int, C.c_int, pi]
        L.orc_gk_geor.restype = None
        L._gk_geor_declared = True
    U0, UF, IMU = C.c_double(u0), C.c_double(uf), C.c_double(imu)
    TPR, W = C.c_int32(tpr), C.c_int32(0)
    L.orc_gk_geor(C.byref(v), C.byref(U0), C.byref(UF), mu0, muf, a, l, q2, C.byref(IMU), int(tpm), C.byref(TPR),
                  su, sm, int(pht), int(firstpt), C.byref(W))
    return {"u0": U0.value, "uf": UF.value, "imu": IMU.value, "tpr": TPR.value, "warned": W.value,
            "ncase": v.ncase, "iu0": v.iu0}


def geor_batch(u0, mu0, muf, a, l, q2, tpm, su, sm, uf_in=None, nthreads=0):
    """orc_geor_batch over arrays of n: returns (uf, imu, tpr, ncase)."""
    arr = [np.ascontiguousarray(x, dtype=np.float64) for x in (u0, mu0, muf, a, l, q2, su, sm)]
    n = arr[0].size
    tpm = np.ascontiguousarray(tpm, dtype=np.int32)
    ufi = None if uf_in is None else np.ascontiguousarray(uf_in, dtype=np.float64)
    uf, imu = np.zeros(n), np.zeros(n)
    tpr, ncase = np.zeros(n, np.int32), np.zeros(n, np.int32)
    _glib().orc_geor_batch(arr[0].ctypes.data, None if ufi is None else ufi.ctypes.data, arr[1].ctypes.data,
                           arr[2].ctypes.data, arr[3].ctypes.data, arr[4].ctypes.data, arr[5].ctypes.data,
                           tpm.ctypes.data, arr[6].ctypes.data, arr[7].ctypes.data, n, uf.ctypes.data,
                           imu.ctypes.data, tpr.ctypes.data, ncase.ctypes.data, nthreads)
    return uf, imu, tpr, ncase


def sncndn(u, emmc):
    sn, cn, dn = C.c_double(), C.c_double(), C.c_double()
    _glib().orc_sncndn(u, emmc, C.byref(sn), C.byref(cn), C.byref(dn))
    return sn.value, cn.value, dn.value


def zroots(coefs_low_first, polish=True):
    a = np.ascontiguousarray(coefs_low_first, dtype=np.complex128)
    m = a.size - 1
    roots = np.zeros(m, dtype=np.complex128)
    _glib().orc_zroots(a.ctypes.data, m, roots.ctypes.data, 1 if polish else 0)
    return roots


def geokerr_camera(alpha, beta, mu0, a):
    L, Q2, aa, su, sm = (C.c_double() for _ in range(5))
    _glib().orc_geokerr_camera(alpha, beta, mu0, a, C.byref(L), C.byref(Q2), C.byref(aa),
                               C.byref(su), C.byref(sm))
    return L.value, Q2.value, aa.value, su.value, sm.value


def geomu_batch(alpha, beta, a, mu0, uf, u0, nthreads=0):
    arrs = [_arr(v) for v in (alpha, beta, a, mu0, uf)]
    n = arrs[0].size
    muf = np.zeros(n)
    tpm = np.zeros(n, np.int32)
    ncase = np.zeros(n, np.int32)
    iu = np.zeros(n)
    _glib().orc_geomu_batch(*[x.ctypes.data for x in arrs], u0, n, muf.ctypes.data,
                            tpm.ctypes.data, ncase.ctypes.data, iu.ctypes.data, nthreads)
    return muf, tpm, ncase, iu


def ell(kind, p, *args):
    """Call one of the Carlson reductions: args are the coefficient list
    followed by (y, x); returns (value, rff)."""
    pp = (C.c_int * 5)(*p)
    rff = C.c_double(0.0)
    f = getattr(_glib(), "orc_" + kind)
    *coef, y, x = args
    val = f(pp, *[float(c) for c in coef], C.byref(rff), float(y), float(x))
    return val, rff.value


# ---- GEOKERR GEOPHITIME, driver and geokerr_ref_api (geokerr_oracle.c) ----
_gk_declared = False


def _gklib():
    global _gk_declared
    L = _glib()
    if not _gk_declared:
        d, i32 = C.c_double, C.c_int
        pd, pi32 = C.POINTER(C.c_double), C.POINTER(C.c_int32)
        L.orc_gk_geokerr.argtypes = [pd, pd, d, d, d, d, d, C.c_int32, C.c_int32, d, d, i32, d, i32,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]
        L.orc_gk_geokerr.restype = i32
        L.orc_geokerr_eval_lambda.argtypes = [d] * 9 + [i32] + [C.c_void_p] * 5 + [C.POINTER(C.c_int)] * 3
        L.orc_geokerr_eval_lambda.restype = None
        L.orc_geokerr_estimate_periods.argtypes = [d] * 5 + [pd, pd, C.POINTER(C.c_int)]
        L.orc_geokerr_estimate_periods.restype = None
        L.orc_geokerr_phi_omega.argtypes = [d] * 5
        L.orc_geokerr_phi_omega.restype = d
        L.orc_geokerr_eval_lambda_batch.argtypes = [C.c_void_p] * 7 + [d, d, i32, i32] + [C.c_void_p] * 8 + [i32]
        L.orc_geokerr_eval_lambda_batch.restype = None
        _gk_declared = True
    return L


def geokerr_driver(u0, uf, mu0, a, l, q2, nup, su=None, sm=None, tpr=1, offset=0.5, uf_is_u0=False):
    """GEOKERR's sample arrays (USEGEOR off, PHIT on) for one geodesic:
    returns dict with n (entries written), nup_out, u0/uf as modified, and
    ufi, mufi, dti, dphi, tpmi, tpri, lami (length nup)."""
    if su is None:
        su = 1.0 if uf - u0 >= 0 else -1.0
    if sm is None:
        sm = 1.0 if mu0 <= 0 else -1.0
    out = {k: np.zeros(nup) for k in ("ufi", "mufi", "dti", "dphi", "lami")}
    out["tpmi"] = np.zeros(nup, np.int32)
    out["tpri"] = np.zeros(nup, np.int32)
    U0, UF, nn = C.c_double(u0), C.c_double(uf), C.c_int(0)
    n = _gklib().orc_gk_geokerr(C.byref(U0), C.byref(UF), u0, mu0, a, l, q2, 0, tpr, su, sm, nup, offset,
                                1 if uf_is_u0 else 0, out["ufi"].ctypes.data, out["mufi"].ctypes.data,
                                out["dti"].ctypes.data, out["dphi"].ctypes.data, out["tpmi"].ctypes.data,
                                out["tpri"].ctypes.data, out["lami"].ctypes.data, C.byref(nn))
    out.update(n=n, nup_out=nn.value, u0=U0.value, uf=UF.value, su=su, sm=sm)
    return out


class GkOpts(C.Structure):
    """orc_gk_opts: the GEOKERR driver's switches (geokerr_original.f:357-360)."""
    _fields_ = [("nup", C.c_int32), ("npts", C.c_int32), ("kext", C.c_int32), ("phit", C.c_int32),
                ("usegeor", C.c_int32), ("mufill", C.c_int32), ("offset", C.c_double)]


GK_ARGS_DTYPE = np.dtype([(k, np.float64) for k in ("u0", "uf", "uout", "mu0", "muf", "a", "l", "q2",
                                                     "beta", "su", "sm")] +
                         [("tpm", np.int32), ("tpr", np.int32)], align=True)


def findmroots(q2, l, a, mu0):
    L = _gklib()
    L.orc_findmroots.argtypes = [C.c_double] * 4 + [C.POINTER(C.c_double)] * 2
    L.orc_findmroots.restype = None
    mm, mp = C.c_double(), C.c_double()
    L.orc_findmroots(q2, l, a, mu0, C.byref(mm), C.byref(mp))
    return mm.value, mp.value


def indep_muf(muminus, muplus, mu0, muf, tpm0, tpmf, sm, k, npts, offset):
    L = _gklib()
    L.orc_indep_muf.argtypes = [C.c_double] * 4 + [C.c_int32, C.c_int32, C.c_double, C.c_int32, C.c_int32,
                                                   C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_int32)]
    L.orc_indep_muf.restype = None
    mun, tpmk = C.c_double(), C.c_int32()
    L.orc_indep_muf(muminus, muplus, mu0, muf, tpm0, tpmf, sm, k, npts, offset, C.byref(mun), C.byref(tpmk))
    return mun.value, tpmk.value


def geokerr_full_batch(opts: dict, args: np.ndarray, nthreads=0) -> dict:
    """The whole GEOKERR driver (orc_gk_geokerr_full) over n geodesics.
    opts: nup, npts, kext, phit, usegeor, mufill, offset.  args: a
    GK_ARGS_DTYPE array (copied; the returned 'args' holds the by-reference
    arguments as the driver leaves them).  Sample arrays are n x (nup +
    npts - 2 kext)."""
    L = _gklib()
    L.orc_gk_geokerr_full_batch.argtypes = [C.POINTER(GkOpts), C.c_void_p, C.c_int] + [C.c_void_p] * 10 + [C.c_int]
    L.orc_gk_geokerr_full_batch.restype = None
    o = GkOpts(**{k: opts[k] for k in ("nup", "npts", "kext", "phit", "usegeor", "mufill", "offset")})
    g = np.ascontiguousarray(args, dtype=GK_ARGS_DTYPE).copy()
    n = len(g)
    tot = o.nup + o.npts - 2 * o.kext
    out = {k: np.empty((n, tot)) for k in ("ufi", "mufi", "dti", "dphi", "lami")}
    out["tpmi"] = np.empty((n, tot), np.int32)
    out["tpri"] = np.empty((n, tot), np.int32)
    for k in ("ncase", "nup_out", "npts_out"):
        out[k] = np.empty(n, np.int32)
    L.orc_gk_geokerr_full_batch(C.byref(o), g.ctypes.data, n, out["ufi"].ctypes.data, out["mufi"].ctypes.data,
                                out["dti"].ctypes.data, out["dphi"].ctypes.data, out["tpmi"].ctypes.data,
                                out["tpri"].ctypes.data, out["lami"].ctypes.data, out["ncase"].ctypes.data,
                                out["nup_out"].ctypes.data, out["npts_out"].ctypes.data, nthreads)
    out["args"] = g
    return out


def geokerr_eval_lambda(u0, mu0, uf, a, lz, q2, t0, lam0, dlam, s):
    arrs = [np.zeros(max(s, 0)) for _ in range(5)]
    tpmi, tpri, info = C.c_int(), C.c_int(), C.c_int()
    _gklib().orc_geokerr_eval_lambda(u0, mu0, uf, a, lz, q2, t0, lam0, dlam, s,
                                     *[x.ctypes.data for x in arrs], C.byref(tpmi), C.byref(tpri),
                                     C.byref(info))
    u, mu, t, phi, aff = arrs
    return dict(u=u, mu=mu, t=t, phi=phi, affine=aff, tpmi=tpmi.value, tpri=tpri.value, info=info.value)


def geokerr_estimate_periods(u0, mu0, a, lz, q2):
    tr, tth, info = C.c_double(), C.c_double(), C.c_int()
    _gklib().orc_geokerr_estimate_periods(u0, mu0, a, lz, q2, C.byref(tr), C.byref(tth), C.byref(info))
    return tr.value, tth.value, info.value


def geokerr_phi_omega(u0, mu0, a, lz, q2):
    return _gklib().orc_geokerr_phi_omega(u0, mu0, a, lz, q2)


def geokerr_eval_lambda_batch(u0, mu0, uf, a, lz, q2, t0, lam0, dlam, s, nthreads=0):
    ins = [_arr(v) for v in (u0, mu0, uf, a, lz, q2, t0)]
    n = ins[0].size
    outs = [np.zeros((n, s)) for _ in range(5)]
    tpmi = np.zeros(n, np.int32)
    tpri = np.zeros(n, np.int32)
    info = np.zeros(n, np.int32)
    _gklib().orc_geokerr_eval_lambda_batch(*[x.ctypes.data for x in ins], lam0, dlam, s, n,
                                           *[x.ctypes.data for x in outs], tpmi.ctypes.data,
                                           tpri.ctypes.data, info.ctypes.data, nthreads)
    u, mu, t, phi, aff = outs
    return dict(u=u, mu=mu, t=t, phi=phi, affine=aff, tpmi=tpmi, tpri=tpri, info=info)
