/*
 * np_render.h -- plain-C description of one render job (one frame, one pass).
 *
 * The first block of fields is the reference renderer's RenderParams
 * (render.cu:15-46) after main() has turned the CLI values into kernel
 * arguments (render.cu:467-521, 545-550): the camera polar angle is in
 * radians, the background sigma in radians, the mode string already split
 * into integrator/camera switches.  The second block holds the additive
 * extensions the Kerr disk workloads need; their defaults
 * (np_render_params_default) reproduce the reference renderer exactly.
 *
 * Shared by the HIP library (libnullpath.so), the `render` CLI and the CPU
 * oracle, so every field is a fixed-width scalar and the layout is identical
 * in C and C++.
 */
#ifndef NP_RENDER_H
#define NP_RENDER_H

#include <stdint.h>
#include <math.h>

#ifdef __cplusplus
extern "C" {
#endif

enum np_integrator { NP_INTEGRATOR_FIXED = 0, NP_INTEGRATOR_ADAPTIVE = 1 };
enum np_camera     { NP_CAMERA_STATIC = 0, NP_CAMERA_ZAMO = 1 };
enum np_bg         { NP_BG_GRADIENT = 0, NP_BG_POINT = 1 };
/* FP64: the reference's formulas in double.  FP32C: float-compensated
 * stepping with the Carter-form field (FP64 where float cannot decide).
 * FP64_CARTER: a comparator, the FP32C path's formulas and step control all
 * in double, so that FP32C against it measures the arithmetic type alone. */
enum np_precision  { NP_PREC_FP64 = 0, NP_PREC_FP32C = 1, NP_PREC_FP64_CARTER = 2 };
/* Radius of an equatorial crossing between two accepted states: LINEAR is
 * the reference's interpolation in theta (render.cu:211-226), whose error
 * (~h^2 r'') depends on where the adaptive steps fall; HERMITE solves the
 * cubic Hermite interpolant of theta(lambda) through both states and their
 * derivatives, so the radius no longer depends on the step sequence at the
 * ~1e-4 level. */
enum np_crossing   { NP_CROSSING_LINEAR = 0, NP_CROSSING_HERMITE = 1 };

/* Per-ray termination class written by the tracers (same values as
 * RayStatus, bh_common.h:295-331, plus the renderer's own outcomes). */
enum np_ray_class {
    NP_RAY_ESCAPED   = 0,  /* r > r_cam + 10 Rs after step 10 (render.cu:182,250) */
    NP_RAY_CAPTURED  = 1,  /* horizon, pole, or adaptive-rejection exhaustion (render.cu:209) */
    NP_RAY_STEPCAP   = 3   /* step budget exhausted, treated as not captured */
};

typedef struct np_render_params {
    /* ---- reference fields (render.cu:118-123 kernel arguments) ---- */
    int32_t width;            /* pixels (render.cu:16) */
    int32_t height;           /* pixels (render.cu:17) */
    float   fov_y_deg;        /* vertical field of view, degrees (render.cu:18) */
    float   theta_cam;        /* camera polar angle, radians (render.cu:484) */
    float   phi_cam;          /* camera azimuth, radians (render.cu:485,632) */
    float   mass;             /* M in absolute length units (render.cu:20) */
    float   spin;             /* a/M (render.cu:29) */
    float   r_cam_Rs;         /* camera radius in units of Rs = 2M (render.cu:21) */
    int32_t integrator;       /* np_integrator (render.cu:470-474) */
    int32_t camera;           /* np_camera (render.cu:471-474) */
    int32_t bg_mode;          /* np_bg (render.cu:516) */
    float   bg_sigma_rad;     /* point-source width, radians (render.cu:517) */
    float   bg_peak;          /* point-source brightness (render.cu:518) */
    int32_t disk_on;          /* 1 = thin disk on (render.cu:519) */
    float   disk_gain;        /* disk brightness multiplier (render.cu:520) */
    int32_t disk_max_orders;  /* crossings accumulated, >= 1 (render.cu:521) */
    int32_t check_inv;        /* 1 = invariant monitor on (render.cu:472) */
    /* ---- extensions; defaults reproduce the reference ---- */
    int32_t max_steps;        /* accepted-step budget per ray (reference: 10000) */
    int32_t precision;        /* np_precision (reference: FP64) */
    int32_t emiss_m_units;    /* 0: j = max(r,1)^-q with absolute r (render.cu:67); 1: r in units of M */
    double  rtol;             /* adaptive relative tolerance (reference: 1e-10) */
    double  atol;             /* adaptive absolute tolerance (reference: 1e-12) */
    double  disk_rin;         /* disk inner edge, absolute units (reference: 3 Rs) */
    double  disk_rout;        /* disk outer edge, absolute units (reference: 20 Rs) */
    float   emiss_q;          /* emissivity index q (reference: 2) */
    uint32_t seed;            /* jitter seed; 0 gives the reference's hash keys */
    double  fp64_radius;      /* NP_PREC_FP32C only: steps that start inside this radius (absolute
                                 units) run in FP64; 0 = auto, max(30 M, 1.5 disk_rout) */
    int32_t waves_per_simd;   /* kernel variant: the waves per SIMD it runs at, 0 = the default
                                 (the fastest measured).  NP_PREC_FP32C: 0 or 4 = capped at 128
                                 VGPRs (the float attempt runs without spills, the rare FP64
                                 branch spills), 3 = capped at 168 VGPRs, 2 = the compiler's own
                                 register budget.  NP_PREC_FP64_CARTER: 0 or 3 = capped at 168
                                 VGPRs, 2 = the compiler's budget.  NP_PREC_FP64: 0 or 2 = capped
                                 at 256 VGPRs, 1 = the compiler's budget (256 VGPRs + AGPRs).
                                 Other values are rejected (NP_EINVAL) */
    int32_t crossing;         /* np_crossing: disk-crossing radius between accepted steps
                                 (reference: linear in theta) */
    int32_t estimate_every_step; /* NP_PREC_FP32C, adaptive: 0 = a ray's steps at the maximum
                                 step size beyond 100 M skip the step-doubling error estimate
                                 (the full step) once one there came out below 1e-3, while the
                                 pole stays clear of the step; there the truncation error of a
                                 step is ~1e-5 of the tolerance, far from the 0.59 that would
                                 shrink the next step; 1 = every step estimates its error */
    int32_t renorm_pr;        /* NP_PREC_FP32C: 1 = after every 16th accepted step, set p_r from
                                 the constraint H = 0 (Delta^2 p_r^2 = P^2 - Delta (p_theta^2 +
                                 (L/sin - a E sin)^2), in FP64 on the float-float state) where
                                 radial motion dominates (Delta p_r^2 >= P^2 / (4 Delta)), so the
                                 square root is well conditioned; 0 = off (the reference never
                                 renormalises) */
    int32_t defer_fp64;       /* NP_PREC_FP32C, adaptive: whether a lane whose next attempt runs in
                                 FP64 (inside fp64_radius, or the re-run of a float attempt) waits
                                 until enough lanes of its wave need one, so the wave runs its FP64
                                 branch for several lanes at once.  Changes no result (each ray's
                                 attempts are the same), only the time.  1 = defer, -1 = never,
                                 0 = the library's choice: defer when the FP64 sphere seen from the
                                 camera spans more than a tenth of the vertical half field of view
                                 (asin(fp64 radius / camera radius) > 0.1 fov_y/2), where many
                                 rays run FP64 steps, and the launch has at least 8 rays per lane
                                 of its persistent grid (with fewer, the deferred rays, a few
                                 times slower, decide when the launch ends) */
} np_render_params;

/* Fill *p with the reference renderer's defaults (render.cu:15-46 with
 * mode=balanced, i.e. fixed-step RK4 and the static camera). */
static inline void np_render_params_default(np_render_params* p) {
    p->width = 1280; p->height = 720; p->fov_y_deg = 50.0f;
    p->theta_cam = 90.0f * 3.14159265358979323846f / 180.0f;
    p->phi_cam = 0.0f; p->mass = 10.0f; p->spin = 0.0f; p->r_cam_Rs = 50.0f;
    p->integrator = NP_INTEGRATOR_FIXED; p->camera = NP_CAMERA_STATIC;
    p->bg_mode = NP_BG_GRADIENT; p->bg_sigma_rad = 1.0f * 3.14159265358979323846f / 180.0f;
    p->bg_peak = 3.0f; p->disk_on = 1; p->disk_gain = 1.0f; p->disk_max_orders = 1;
    p->check_inv = 0; p->max_steps = 10000; p->precision = NP_PREC_FP64;
    p->emiss_m_units = 0; p->rtol = 1e-10; p->atol = 1e-12;
    p->disk_rin = 3.0 * 2.0 * (double)p->mass; p->disk_rout = 20.0 * 2.0 * (double)p->mass;
    p->emiss_q = 2.0f; p->seed = 0u; p->fp64_radius = 0.0;
    p->waves_per_simd = 0;
    p->crossing = NP_CROSSING_LINEAR;
    p->estimate_every_step = 0;
    p->renorm_pr = 0;
    p->defer_fp64 = 0;
}

/* Kerr prograde ISCO radius in units of M (Bardeen, Press & Teukolsky 1972,
 * Eq. 2.21).  The reference never computes it; the Kerr disk workloads put
 * the inner disk edge there. */
static inline double np_isco_radius(double a) {
    double z1 = 1.0 + cbrt(1.0 - a * a) * (cbrt(1.0 + a) + cbrt(1.0 - a));
    double z2 = sqrt(3.0 * a * a + z1 * z1);
    return 3.0 + z2 - sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2));
}

#ifdef __cplusplus
}
#endif

#endif /* NP_RENDER_H */
