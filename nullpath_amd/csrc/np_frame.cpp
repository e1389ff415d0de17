// np_frame.cpp -- host-side per-frame constants (see np_frame.hpp).
//
// The tetrad formulas are those of init_ray_from_camera_general_d
// (bh_kerr.h:781-826) and init_ray_from_camera_zamo_d (bh_kerr.h:910-967),
// evaluated once per frame.  Contraction is off so each product and sum is
// rounded separately, as the reference's source reads.
#include <cmath>
#include "np_frame.hpp"
#include "nullpath.h"

#pragma clang fp contract(off)

namespace np {

namespace {

void cov_blocks(double M, double a, double r, double th, double g[5]) {
    double ct = std::cos(th), st = std::sin(th);
    double st2 = st * st, r2 = r * r, a2 = a * a;
    double Sigma = r2 + a2 * ct * ct;
    double Delta = r2 - 2.0 * M * r + a2;
    double A = (r2 + a2) * (r2 + a2) - a2 * Delta * st2;
    g[0] = -(1.0 - 2.0 * M * r / std::fmax(1e-12, Sigma));
    g[1] = -2.0 * M * a * r * st2 / std::fmax(1e-12, Sigma);
    g[2] = A * st2 / std::fmax(1e-12, Sigma);
    g[3] = Sigma / std::fmax(1e-12, Delta);
    g[4] = Sigma;
}

// Second-order compensated sum (BhKleinSum, bh_common.h:451-483).
struct KleinH {
    double sum = 0, cs = 0, ccs = 0;
    void add(double x) {
        double t = sum + x;
        double c = (std::fabs(sum) >= std::fabs(x)) ? (sum - t) + x : (x - t) + sum;
        sum = t;
        t = cs + c;
        double cc = (std::fabs(cs) >= std::fabs(c)) ? (cs - t) + c : (c - t) + cs;
        cs = t;
        ccs += cc;
    }
    double result() const { return sum + cs + ccs; }
};

uint32_t wanghash(uint32_t a) {
    a = (a ^ 61u) ^ (a >> 16);
    a *= 9u;
    a = a ^ (a >> 4);
    a *= 0x27d4eb2du;
    a = a ^ (a >> 15);
    return a;
}

}  // namespace

int make_frame(const np_render_params* p, FrameConst* f) {
    if (!p || !f) return NP_EINVAL;
    if (p->width <= 0 || p->height <= 0) return NP_EINVAL;
    if (!(p->mass > 0.0f) || !(p->r_cam_Rs > 0.0f)) return NP_EINVAL;
    if (!(std::fabs(p->spin) < 1.0f)) return NP_EINVAL;
    if (!(p->fov_y_deg > 0.0f && p->fov_y_deg < 179.0f)) return NP_EINVAL;
    if (p->max_steps < 0 || p->disk_max_orders < 1) return NP_EINVAL;
    // kernel variants (waves per SIMD) that exist for each precision
    {
        const int w = p->waves_per_simd;
        const bool ok = w == 0 || (p->precision == NP_PREC_FP32C && w >= 2 && w <= 4) ||
                        (p->precision == NP_PREC_FP64_CARTER && (w == 2 || w == 3)) ||
                        (p->precision == NP_PREC_FP64 && (w == 1 || w == 2));
        if (!ok) return NP_EINVAL;
    }
    if (p->integrator == NP_INTEGRATOR_ADAPTIVE && !(p->rtol > 0.0 && p->atol >= 0.0)) return NP_EINVAL;

    // render.cu:131-139 (float arithmetic)
    const float Rs = 2.0f * p->mass;
    const float Mf = p->mass;
    const float af = p->spin * Mf;
    const float r_cam = p->r_cam_Rs * Rs;
    f->aspect = (float)p->width / (float)p->height;
    const float fov_y = p->fov_y_deg * NP_PI_F_HOST / 180.0f;
    f->tan_half = std::tan(0.5f * fov_y);
    f->Rs = Rs;
    f->mass = Mf;
    f->disk_gain = p->disk_gain;
    f->emiss_q = p->emiss_q;
    f->bg_sigma_rad = p->bg_sigma_rad;
    f->bg_peak = p->bg_peak;
    f->salt = p->seed ? wanghash(p->seed) : 0u;
    f->width = p->width;
    f->height = p->height;
    f->bg_mode = p->bg_mode;
    f->disk_on = p->disk_on;
    f->disk_max_orders = p->disk_max_orders;
    f->check_inv = p->check_inv;
    f->max_steps = p->max_steps;
    f->emiss_m_units = p->emiss_m_units;

    const double M = (double)Mf, a = (double)af;
    f->M = M;
    f->a = a;
    double disc = M * M - a * a;
    if (disc < 0.0) disc = 0.0;
    f->r_plus = M + std::sqrt(disc);
    f->r_cam = (double)r_cam;
    f->th_cam = (double)p->theta_cam;
    f->ph_cam = (double)p->phi_cam;
    f->Rs_d = (double)Rs;
    f->r_far = (double)r_cam + 10.0 * (double)Rs;   // render.cu:182
    f->rin = p->disk_rin;
    f->rout = p->disk_rout;
    f->rtol = p->rtol;
    f->atol = p->atol;
    f->crossing = p->crossing == NP_CROSSING_HERMITE ? 1 : 0;
    // Kernel variant = the waves per SIMD the kernel runs at; 0 picks the
    // fastest measured.  FP32C: 4 (default) or 3, capped; 2 is the compiler's
    // own register budget.  FP64 comparator: 3 (default), capped; 2 the
    // compiler's budget.  Reference-formula FP64: 2 (default), capped; 1 the
    // compiler's budget (256 VGPRs + AGPRs).
    if (p->precision == NP_PREC_FP32C)
        f->variant = (p->waves_per_simd == 2) ? 0 : (p->waves_per_simd == 3) ? 3 : 4;
    else if (p->precision == NP_PREC_FP64_CARTER)
        f->variant = (p->waves_per_simd == 2) ? 0 : 3;
    else
        f->variant = (p->waves_per_simd == 1) ? 0 : 2;
    f->r_isl = p->fp64_radius > 0.0 ? p->fp64_radius : std::fmax(30.0 * M, 1.5 * p->disk_rout);
    if (p->estimate_every_step != 0 && p->estimate_every_step != 1) return NP_EINVAL;
    f->lazy = p->estimate_every_step == 0 ? 1 : 0;
    if (p->renorm_pr != 0 && p->renorm_pr != 1) return NP_EINVAL;
    f->renorm = p->precision == NP_PREC_FP32C ? p->renorm_pr : 0;
    if (p->defer_fp64 < -1 || p->defer_fp64 > 1) return NP_EINVAL;
    {
        // the library's choice: the FP64 sphere's angular radius from the
        // camera against the vertical half field of view
        const double half_fov = 0.5 * (double)p->fov_y_deg * 3.141592653589793238462643383279502884 / 180.0;
        const double ang = std::asin(std::fmin(1.0, f->r_isl / (double)r_cam));
        const bool wide = ang > 0.1 * half_fov;
        const int on = p->defer_fp64 == 1 ? 2 : (p->defer_fp64 == 0 && wide) ? 1 : 0;
        f->defer64 = (p->precision == NP_PREC_FP32C && p->integrator == NP_INTEGRATOR_ADAPTIVE) ? on : 0;
    }
    f->r_lazyf = (float)(100.0 * M);
    f->s_eq = std::sin(3.141592653589793238462643383279502884 * 0.5);
    f->c_eq = std::cos(3.141592653589793238462643383279502884 * 0.5);
    f->s2_cam = std::sin(f->th_cam) * std::sin(f->th_cam);
    f->c2_cam = std::cos(f->th_cam) * std::cos(f->th_cam);
    f->Mf = (float)M;
    f->af = (float)a;
    f->a2f = f->af * f->af;
    f->rpf = (float)f->r_plus;
    f->rmf = (float)(M - std::sqrt(disc));
    f->rpf_guard = (float)(f->r_plus + 1e-12);
    f->rtolf = (float)p->rtol;
    f->atolf = (float)p->atol;
    // A float-float value differs from its high word by at most 2^-24 of it
    // (6e-8 relative in r; 1.9e-7 absolute in theta < pi): the thresholds
    // sit 1e-6 (relative, absolute) inside the FP64 bounds and are rounded
    // outward-safe, so passing them implies passing the FP64 check.
    const auto up = [](double x) {
        float v = (float)x;
        return ((double)v < x) ? std::nextafter(v, HUGE_VALF) : v;
    };
    const auto down = [](double x) {
        float v = (float)x;
        return ((double)v > x) ? std::nextafter(v, -HUGE_VALF) : v;
    };
    const double pi = 3.141592653589793238462643383279502884;
    f->r_okf = up((f->r_plus + 1e-12) * (1.0 + 1e-6));
    f->th_okf_lo = up(1e-9 + 1e-6);
    f->th_okf_hi = down(pi - 1e-9 - 1e-6);

    double g[5];
    cov_blocks(M, a, f->r_cam, f->th_cam, g);
    f->g_tt = g[0]; f->g_tphi = g[1]; f->g_phiphi = g[2]; f->g_rr = g[3]; f->g_thth = g[4];
    const bool static_ok = g[0] < 0.0;   // kerr_static_timelike (bh_kerr.h:880-884)
    f->zamo = (p->camera == NP_CAMERA_ZAMO || !static_ok) ? 1 : 0;   // render.cu:163-164

    if (!f->zamo) {
        // Static observer tetrad (bh_kerr.h:792-815).
        double ut = 1.0 / std::sqrt(std::fmax(1e-12, -g[0]));
        double er_r = 1.0 / std::sqrt(std::fmax(1e-12, g[3]));
        double eth_th = 1.0 / std::sqrt(std::fmax(1e-12, g[4]));
        double ephi_t, ephi_phi;
        if (std::fabs(g[1]) > 1e-12) {
            double beta = -g[0] / g[1];
            KleinH n2;
            n2.add(g[0]);
            n2.add(2.0 * g[1] * beta);
            n2.add(g[2] * beta * beta);
            double inv = 1.0 / std::sqrt(std::fmax(1e-12, std::fabs(n2.result())));
            ephi_t = inv;
            ephi_phi = beta * inv;
        } else {
            ephi_t = 0.0;
            ephi_phi = 1.0 / std::sqrt(std::fmax(1e-12, g[2]));
        }
        f->ct0 = ut; f->ct_ph = ephi_t; f->cr = er_r; f->cth = eth_th;
        f->cph0 = 0.0; f->cph_ph = ephi_phi;
    } else {
        // ZAMO tetrad (bh_kerr.h:921-955).
        double ct = std::cos(f->th_cam), st = std::sin(f->th_cam);
        double r2 = f->r_cam * f->r_cam, a2 = a * a;
        double Sigma = std::fmax(1e-12, r2 + a2 * ct * ct);
        double Delta = r2 - 2.0 * M * f->r_cam + a2;
        double A = std::fmax(1e-12, (r2 + a2) * (r2 + a2) - a2 * Delta * st * st);
        double alpha = std::sqrt(std::fmax(1e-12, (Delta * Sigma) / A));
        double omega = (2.0 * M * a * f->r_cam) / A;
        f->ct0 = 1.0 / std::fmax(1e-12, alpha);
        f->ct_ph = 0.0;
        f->cph0 = omega / std::fmax(1e-12, alpha);
        f->cr = std::sqrt(std::fmax(1e-12, Delta / Sigma));
        f->cth = 1.0 / std::sqrt(Sigma);
        f->cph_ph = std::sqrt(std::fmax(1e-12, Sigma / A)) / std::fmax(1e-12, std::fabs(st));
    }
    return NP_OK;
}

}  // namespace np
