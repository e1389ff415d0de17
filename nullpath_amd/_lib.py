"""ctypes binding of libnullpath.so (include/nullpath.h).

The library is loaded from this package folder.  There is no fallback: if
libnullpath.so is missing or cannot be loaded, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libnullpath.so")

_lib: C.CDLL | None = None


class NullPathError(RuntimeError):
    pass


class RenderParams(C.Structure):
    """np_render_params (include/np_render.h)."""

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

    def copy(self) -> "RenderParams":
        q = RenderParams()
        C.pointer(q)[0] = self
        return q


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

    def copy(self) -> "AdjointParams":
        q = AdjointParams()
        C.pointer(q)[0] = self
        return q


class Tuning(C.Structure):
    """np_tuning (include/nullpath.h)."""

    _fields_ = [("block_size", C.c_int32), ("waves_per_simd", C.c_int32),
                ("refill_chunk", C.c_int32), ("fp64_radius", C.c_double),
                ("defer_fp64", C.c_int32)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class TraceCounters(C.Structure):
    """np_trace_counters (include/nullpath.h)."""

    _fields_ = [(n, C.c_uint64) for n in (
        "rays", "steps", "attempts", "rhs_evals", "captured", "escaped", "stepcap", "disk_hits",
        "attempts_f64", "rhs_evals_f64", "attempts_lazy")]


COUNTER_FIELDS = [f[0] for f in TraceCounters._fields_]

# np_ray_record as a numpy dtype (32 bytes)
def ray_record_dtype():
    import numpy as np
    return np.dtype([("cls", np.int32), ("orders", np.int32), ("steps", np.int32),
                     ("attempts", np.int32), ("min_r", np.float64), ("r_end", np.float64)])


# np_ray_times as a numpy dtype (24 bytes)
def ray_times_dtype():
    import numpy as np
    return np.dtype([("t", np.float64), ("phi", np.float64), ("lam", np.float64)])


def lib() -> C.CDLL:
    """Load libnullpath.so (raises NullPathError when it is absent)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NullPathError(
                f"{LIB_PATH} is missing: build it with `python -m nullpath_amd.build`")
        try:
            L = C.CDLL(LIB_PATH)
        except OSError as e:  # pragma: no cover - depends on the machine
            raise NullPathError(f"cannot load {LIB_PATH}: {e}") from e
        _declare(L)
        _lib = L
    return _lib


class GeokerrOpts(C.Structure):
    """np_geokerr_opts (nullpath.h): the GEOKERR driver's switches."""
    _fields_ = [("nup", C.c_int32), ("npts", C.c_int32), ("kext", C.c_int32), ("phit", C.c_int32),
                ("usegeor", C.c_int32), ("mufill", C.c_int32), ("offset", C.c_double)]


GEOKERR_RAY_FIELDS = ("u0", "uf", "uout", "mu0", "muf", "a", "l", "q2", "beta", "su", "sm", "tpm", "tpr")
GEOKERR_SAMPLE_FIELDS = ("ufi", "mufi", "dti", "dphi", "lambdai", "tpmi", "tpri", "u0_out", "uf_out", "muf_out",
                         "tpm_out", "tpr_out", "ncase_out", "nup_out", "npts_out")


class GeokerrRays(C.Structure):
    """np_geokerr_rays: device pointers, n entries each."""
    _fields_ = [(k, C.c_void_p) for k in GEOKERR_RAY_FIELDS]


class GeokerrSamples(C.Structure):
    """np_geokerr_samples: device pointers."""
    _fields_ = [(k, C.c_void_p) for k in GEOKERR_SAMPLE_FIELDS]


def _declare(L: C.CDLL) -> None:
    i32, vp = C.c_int, C.c_void_p
    L.np_status_string.argtypes = [i32]
    L.np_status_string.restype = C.c_char_p
    L.np_render_tile.argtypes = [C.POINTER(RenderParams), i32, i32, i32, i32, vp, vp, vp, vp, vp]
    L.np_render_tile.restype = i32
    L.np_render_tile_times.argtypes = [C.POINTER(RenderParams), i32, i32, i32, i32, vp, vp, vp, vp, vp]
    L.np_render_tile_times.restype = i32
    L.np_render_accumulate.argtypes = [C.POINTER(RenderParams), i32, i32, i32, i32, vp, vp]
    L.np_render_accumulate.restype = i32
    i64, d = C.c_int64, C.c_double
    L.np_carlson_batch.argtypes = [i32, vp, vp, vp, vp, i64, vp, vp]
    L.np_carlson_batch.restype = i32
    L.np_geomu_batch.argtypes = [vp] * 9 + [i64] + [vp] * 5
    L.np_geomu_batch.restype = i32
    L.np_geor_batch.argtypes = [vp] * 10 + [i64] + [vp] * 5
    L.np_geor_batch.restype = i32
    L.np_geokerr_batch.argtypes = [C.POINTER(GeokerrOpts), C.POINTER(GeokerrRays), i64, C.POINTER(GeokerrSamples), vp]
    L.np_geokerr_batch.restype = i32
    L.np_geokerr_batch_serial.argtypes = L.np_geokerr_batch.argtypes
    L.np_geokerr_batch_serial.restype = i32
    L.np_geomu_camera_batch.argtypes = [vp] * 5 + [d, i64] + [vp] * 5
    L.np_geomu_camera_batch.restype = i32
    L.np_geomu_camera_batch_mixed.argtypes = [vp] * 5 + [d, i64] + [vp] * 6
    L.np_geomu_camera_batch_mixed.restype = i32
    f = C.c_float
    L.np_schw_precompute.argtypes = [f, f, vp, i64, vp, vp, vp, vp]
    L.np_schw_precompute.restype = i32
    L.np_schw_sweep.argtypes = [d, d, d, d, i64, vp, i64, vp, vp, vp, vp, vp, vp, vp]
    L.np_schw_sweep.restype = i32
    L.np_precompute_geodesics.argtypes = [f, i32, C.POINTER(LookupTable)]
    L.np_precompute_geodesics.restype = i32
    L.np_free_lookup_table.argtypes = [C.POINTER(LookupTable)]
    L.np_free_lookup_table.restype = None
    prp, pt = C.POINTER(RenderParams), C.POINTER(Tuning)
    L.np_render_batch.argtypes = [prp, i32, i32, i32, vp, vp, vp, pt, vp]
    L.np_render_batch.restype = i32
    L.np_render_pass.argtypes = [prp, i32, i32, vp, vp, vp]
    L.np_render_pass.restype = i32
    L.np_pass_graph_create.argtypes = [prp, i32, i32, vp, vp, i32, C.POINTER(vp)]
    L.np_pass_graph_create.restype = i32
    L.np_pass_graph_launch.argtypes = [vp, vp]
    L.np_pass_graph_launch.restype = i32
    L.np_pass_graph_seek.argtypes = [vp, i32, vp]
    L.np_pass_graph_seek.restype = i32
    L.np_pass_graph_destroy.argtypes = [vp]
    L.np_pass_graph_destroy.restype = None
    L.np_render_batch_noalloc.argtypes = [prp, i32, i32, i32, vp, vp, vp, pt, vp]
    L.np_render_batch_noalloc.restype = i32
    L.np_autotune_batch.argtypes = [prp, i32, i32, i32, pt, pt, C.POINTER(C.c_double), i32,
                                    C.POINTER(C.c_int)]
    L.np_autotune_batch.restype = i32
    pa = C.POINTER(AdjointParams)
    L.np_adjoint_workspace_bytes.argtypes = [pa]
    L.np_adjoint_workspace_bytes.restype = i64
    L.np_adjoint_forward.argtypes = [pa, i32, vp, vp, vp, vp]
    L.np_adjoint_forward.restype = i32
    L.np_adjoint_backward.argtypes = [pa, vp, vp, vp, vp]
    L.np_adjoint_backward.restype = i32
    L.np_geokerr_workspace_bytes.argtypes = [i32, i64]
    L.np_geokerr_workspace_bytes.restype = i64
    L.np_geokerr_eval_lambda_batch.argtypes = [vp] * 9 + [i32, i64] + [vp] * 10
    L.np_geokerr_eval_lambda_batch.restype = i32
    L.np_geokerr_estimate_periods_batch.argtypes = [vp] * 5 + [i64] + [vp] * 5
    L.np_geokerr_estimate_periods_batch.restype = i32
    L.np_geokerr_phi_omega_batch.argtypes = [vp] * 5 + [i64, vp, vp]
    L.np_geokerr_phi_omega_batch.restype = i32
    pd, pi = C.POINTER(C.c_double), C.POINTER(C.c_int)
    L.geokerr_eval_lambda.argtypes = [d] * 9 + [C.c_int] + [pd] * 5 + [pi] * 3
    L.geokerr_eval_lambda.restype = None
    L.geokerr_estimate_periods.argtypes = [d] * 5 + [pd, pd, pi]
    L.geokerr_estimate_periods.restype = None
    L.geokerr_phi_omega.argtypes = [d] * 5
    L.geokerr_phi_omega.restype = d
    for name in ("geokerr_complete_stream_cuda", "geokerr_stream_cuda"):
        getattr(L, name).argtypes = [d] * 8 + [C.c_int] + [pd] * 5
        getattr(L, name).restype = C.c_int


class LookupTable(C.Structure):
    """GeodesicLookupTable (include/bh_types.h)."""

    _fields_ = [("impact_parameters", C.POINTER(C.c_float)),
                ("deflection_angles", C.POINTER(C.c_float)),
                ("bending_angles", C.POINTER(C.c_float)),
                ("closest_approaches", C.POINTER(C.c_float)),
                ("ray_status", C.POINTER(C.c_int)),
                ("num_entries", C.c_int), ("mass", C.c_float)]


def check(code: int, what: str) -> None:
    if code != 0:
        msg = lib().np_status_string(code).decode()
        raise NullPathError(f"{what} failed: {msg} ({code})")
