// np_frame.hpp -- per-frame constants of a Kerr render, computed once on the
// host in FP64 and handed to the kernels by value (they are staged into LDS
// or kept in SGPRs).  Everything that depends only on the camera position
// -- the observer tetrad, metric blocks at the camera, the pinhole scale --
// is folded here so the per-ray camera set-up is a handful of multiplies.
#pragma once

#include <stdint.h>
#include "np_render.h"

#define NP_PI_F_HOST 3.14159265358979323846f

namespace np {

struct FrameConst {
    // geometry (absolute units)
    double M, a, r_plus, r_cam, th_cam, ph_cam, Rs_d, r_far, rin, rout;
    double rtol, atol;
    double r_isl;        // FP32C: steps starting at r < r_isl run in FP64
    double s_eq, c_eq;   // sin, cos of pi/2 as the host libm returns them
    double s2_cam, c2_cam;  // sin^2, cos^2 of th_cam (per-ray Carter-constant bound)
    // observer tetrad: p^mu components as functions of the local direction n
    //   p^t   = ct0 + n_phi * ct_ph
    //   p^r   = n_r * cr
    //   p^th  = n_th * cth
    //   p^phi = cph0 + n_phi * cph_ph
    double ct0, ct_ph, cr, cth, cph0, cph_ph;
    // covariant metric at the camera, for lowering p^mu
    double g_tt, g_tphi, g_phiphi, g_rr, g_thth;
    // FP32 copies for the compensated stepper
    float Mf, af, a2f, rpf, rmf, rpf_guard, rtolf, atolf;
    // domain check of a float attempt on the high words of r and theta:
    // r_hi > r_okf and th_okf_lo < th_hi < th_okf_hi imply the FP64 check
    // r > r_+ + 1e-12, 1e-9 < theta < pi - 1e-9 of the full value
    float r_okf, th_okf_lo, th_okf_hi;
    // float pinhole camera (render.cu:137-139)
    float aspect, tan_half, Rs, mass, disk_gain, emiss_q, bg_sigma_rad, bg_peak;
    uint32_t salt;        // jitter salt (0 = reference hash keys)
    int32_t width, height;
    int32_t zamo;         // 1 when the ZAMO tetrad is used
    int32_t bg_mode, disk_on, disk_max_orders, check_inv, max_steps, emiss_m_units;
    int32_t variant;      // kernel variant (FP32C: waves-per-SIMD cap 3 or 4; 0 = none)
    int32_t crossing;     // 0 linear crossing radius (reference), 1 cubic Hermite
    // FP32C: 1 = steps at the maximum step size beyond r_lazyf may skip
    // their error estimate
    float r_lazyf;
    int32_t lazy;
    int32_t renorm;       // FP32C: p_r from H = 0 after accepted steps
    // FP32C, adaptive: FP64 attempts deferred in the wave.  0 = no; 2 = yes
    // (asked for); 1 = yes (the library's choice), which the launch turns
    // off when the persistent grid has few rays per lane (settle_defer)
    int32_t defer64;
};

// Returns 0 on success, NP_EINVAL when the parameters are outside what the
// reference accepts (render.cu:431-465).
int make_frame(const np_render_params* p, FrameConst* f);

}  // namespace np
