// kerr_render.hip -- gfx950 Kerr null-geodesic renderer, FP64 path.
//
// One ray per lane, persistent grid, wave64 work queue:
//   * work item = one pixel; the lane traces that pixel's spp samples one
//     after the other (so the per-pixel colour sum has the reference's order,
//     render.cu:141-362) and refills from the queue the moment its pixel is
//     done -- a long photon-ring ray never holds a whole wave hostage;
//   * the integrator loop is flattened to one adaptive *attempt* per
//     iteration, so lanes that reject and lanes that accept stay converged;
//   * ray state (r, theta, phi, p_r, p_theta; p_t and p_phi are exactly
//     conserved) lives in VGPRs; HBM sees only the 12 B/pixel result.
//
// Numerics follow bh_kerr.h (metric blocks 600-617, derivatives 624-672,
// right-hand side 679-727, RK4 734-774 and 987-1032, step doubling
// 1058-1105) and render.cu:124-363 (camera, step law, disk crossing,
// shading).  Two exact savings over the reference: the full step and the
// first half step share k1 = f(y) (the reference evaluates it twice), and a
// retried attempt reuses it again; t is not integrated because nothing in
// the renderer reads it.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "np_device.hpp"
#include "np_frame.hpp"
#include "np_kernels.hpp"
#include "nullpath.h"

namespace np {

namespace {

constexpr double HMIN = 1e-6;     // BH_RK_HMIN (bh_common.h:245)
constexpr double HMAX = 0.2;      // BH_RK_HMAX (bh_common.h:253)
constexpr double SAFETY = 0.9;    // BH_RK_SAFETY (bh_common.h:236)
constexpr float STEP0 = 0.01f;    // BH_INTEGRATION_STEP_DEFAULT (bh_common.h:213)

struct Deriv {
    double r, th, ph, pr, pth;
};

// computeGeodesicDerivativesKerrD (bh_kerr.h:679-727) with the metric
// blocks of bh_kerr.h:600-617 and 624-672 inlined, one sincos shared.
// g^tt only feeds dt/dlambda, which the renderer never reads.
__device__ __forceinline__ Deriv kerr_rhs(const double M, const double a, const double r_plus,
                                          const double r, const double th, const double pt,
                                          const double pr, const double pth, const double pph) {
    Deriv d;
    if (r <= r_plus + 1e-12 || r <= 1e-12) {
        d.r = d.th = d.ph = d.pr = d.pth = 0.0;
        return d;
    }
    double s, c;
    sincos(th, &s, &c);
    const double r2 = r * r;
    const double a2 = a * a;
    const double Delta = r2 - 2.0 * M * r + a2;
    // kerr_blocks_contra_d
    const double st2c = fmax(1e-12, s * s);
    const double SigmaC = fmax(1e-12, r2 + a2 * c * c);
    const double grr = fmax(1e-12, Delta) / SigmaC;
    const double gthth = 1.0 / SigmaC;
    const double gtphi = -2.0 * M * a * r / fmax(1e-12, SigmaC * Delta);
    const double gphiphi = (Delta - a2 * st2c) / fmax(1e-12, SigmaC * Delta * st2c);
    // kerr_blocks_contra_derivs_d
    const double s2 = (s * s > 1e-30 ? s * s : 1e-30);
    const double B = r2 + a2;
    const double Sigma = r2 + a2 * c * c;
    const double A = (B * B) - a2 * Delta * s2;
    const double D = Sigma * Delta;
    const double D2 = D * D + 1e-48;
    const double s2_th = 2.0 * s * c;
    const double Sigma_r = 2.0 * r;
    const double Sigma_th = -2.0 * a2 * s * c;
    const double Delta_r = 2.0 * (r - M);
    const double A_r = 4.0 * r * B - a2 * Delta_r * s2;
    const double A_th = -a2 * Delta * s2_th;
    const double D_r = Sigma_r * Delta + Sigma * Delta_r;
    const double D_th = Sigma_th * Delta;
    const double gtt_r = -(A_r * D - A * D_r) / D2;
    const double gtt_th = -(A_th * D - A * D_th) / D2;
    const double Nt = 2.0 * M * a * r;
    const double Nt_r = 2.0 * M * a;
    const double gtphi_r = -(Nt_r * D - Nt * D_r) / D2;
    const double gtphi_th = -(0.0 * D - Nt * D_th) / D2;
    const double SS = Sigma * Sigma;
    const double grr_r = (Delta_r * Sigma - Delta * Sigma_r) / SS;
    const double grr_th = -(Delta * Sigma_th) / SS;
    const double gthth_r = -Sigma_r / SS;
    const double gthth_th = -Sigma_th / SS;
    const double N = Delta - a2 * s2;
    const double N_r = Delta_r;
    const double N_th = -a2 * s2_th;
    const double Den = Delta * Sigma * s2 + 1e-48;
    const double Den_r = Delta_r * Sigma * s2 + Delta * Sigma_r * s2;
    const double Den_th = Delta * Sigma_th * s2 + Delta * Sigma * s2_th;
    const double Den2 = Den * Den + 1e-48;
    const double gphiphi_r = (N_r * Den - N * Den_r) / Den2;
    const double gphiphi_th = (N_th * Den - N * Den_th) / Den2;

    d.r = grr * pr;
    d.th = gthth * pth;
    d.ph = gtphi * pt + gphiphi * pph;
    Klein<double> sr;
    sr.add(gtt_r * pt * pt);
    sr.add(2.0 * gtphi_r * pt * pph);
    sr.add(gphiphi_r * pph * pph);
    sr.add(grr_r * pr * pr);
    sr.add(gthth_r * pth * pth);
    Klein<double> sth;
    sth.add(gtt_th * pt * pt);
    sth.add(2.0 * gtphi_th * pt * pph);
    sth.add(gphiphi_th * pph * pph);
    sth.add(grr_th * pr * pr);
    sth.add(gthth_th * pth * pth);
    d.pr = -0.5 * sr.result();
    d.pth = -0.5 * sth.result();
    return d;
}

struct Y5 {
    double r, th, ph, pr, pth;
};

// Classic RK4 step from y with a precomputed k1 = f(y); the stage
// combination k1 + 2k2 + 2k3 + k4 is accumulated with the compensated sum
// in the reference's order (bh_rk4_sum, bh_common.h:485-493) as the stages
// are produced, so only the running sums stay live.
__device__ __forceinline__ Y5 rk4_from_k1(const FrameConst& F, const Y5& y, const Deriv& k1,
                                          const double pt, const double pph, const double h,
                                          int& rhs_count) {
    Klein<double> sr, sth, sph, spr, spth;
    sr.add(k1.r); sth.add(k1.th); sph.add(k1.ph); spr.add(k1.pr); spth.add(k1.pth);
    const double hh = 0.5 * h;
    Deriv k = kerr_rhs(F.M, F.a, F.r_plus, y.r + hh * k1.r, y.th + hh * k1.th, pt,
                       y.pr + hh * k1.pr, y.pth + hh * k1.pth, pph);
    sr.add(2.0 * k.r); sth.add(2.0 * k.th); sph.add(2.0 * k.ph); spr.add(2.0 * k.pr); spth.add(2.0 * k.pth);
    k = kerr_rhs(F.M, F.a, F.r_plus, y.r + hh * k.r, y.th + hh * k.th, pt,
                 y.pr + hh * k.pr, y.pth + hh * k.pth, pph);
    sr.add(2.0 * k.r); sth.add(2.0 * k.th); sph.add(2.0 * k.ph); spr.add(2.0 * k.pr); spth.add(2.0 * k.pth);
    k = kerr_rhs(F.M, F.a, F.r_plus, y.r + h * k.r, y.th + h * k.th, pt,
                 y.pr + h * k.pr, y.pth + h * k.pth, pph);
    sr.add(k.r); sth.add(k.th); sph.add(k.ph); spr.add(k.pr); spth.add(k.pth);
    rhs_count += 3;
    const double h6 = h / 6.0;
    Y5 o;
    o.r = y.r + h6 * sr.result();
    o.th = y.th + h6 * sth.result();
    o.ph = y.ph + h6 * sph.result();
    o.pr = y.pr + h6 * spr.result();
    o.pth = y.pth + h6 * spth.result();
    return o;
}

__device__ __forceinline__ double err1(double x, double y, double scale) {
    return fabs(x - y) / fmax(scale, 1e-30);
}

// doppler_g_general_d (bh_kerr.h:833-862) at the disk plane.
__device__ __forceinline__ double doppler_g(const FrameConst& F, double r, double pt, double pph) {
    const double M = F.M, a = F.a;
    // kerr_blocks_cov_d at theta = pi/2 (bh_kerr.h:575-592)
    double s, c;
    sincos(NP_PI_D * 0.5, &s, &c);
    const double st2 = s * s, r2 = r * r, a2 = a * a;
    const double Sigma = r2 + a2 * c * c;
    const double Delta = r2 - 2.0 * M * r + a2;
    const double Aq = (r2 + a2) * (r2 + a2) - a2 * Delta * st2;
    const double g_tt = -(1.0 - 2.0 * M * r / fmax(1e-12, Sigma));
    const double g_tphi = -2.0 * M * a * r * st2 / fmax(1e-12, Sigma);
    const double g_phiphi = Aq * st2 / fmax(1e-12, Sigma);
    const double M_safe = fmax(1e-12, M);
    const double r_over_M = r / M_safe;
    const double Omega = 1.0 / (a / M_safe + pow(fmax(1.0, r_over_M), 1.5)) / M_safe;
    Klein<double> ds;
    ds.add(g_tt);
    ds.add(2.0 * g_tphi * Omega);
    ds.add(g_phiphi * Omega * Omega);
    const double denom = -ds.result();
    const double u_t = 1.0 / sqrt(denom > 1e-24 ? denom : 1e-24);
    const double u_phi = Omega * u_t;
    const double E_inf = -pt;
    Klein<double> ks;
    ks.add(pt * u_t);
    ks.add(pph * u_phi);
    const double k_u = -ks.result();
    const double g = E_inf / (k_u > 1e-18 ? k_u : 1e-18);
    return g > 0.0 ? g : 0.0;
}

__device__ __forceinline__ uint32_t wanghash(uint32_t a) {
    a = (a ^ 61u) ^ (a >> 16);
    a *= 9u;
    a = a ^ (a >> 4);
    a *= 0x27d4eb2du;
    a = a ^ (a >> 15);
    return a;
}

__device__ __forceinline__ float rnd01(uint32_t seed) {
    return (float)(wanghash(seed) & 0x00ffffffu) * (1.0f / 16777216.0f);
}

// Camera direction of one sample (render.cu:143-159), float arithmetic
// with each operation rounded on its own.
__device__ void camera_dir(const FrameConst& F, int x, int y, uint32_t sid, float& n_r,
                           float& n_th, float& n_ph, float& cam_x_out, float& cam_y_out) {
#pragma clang fp contract(off)
    const uint32_t ux = (uint32_t)x, uy = (uint32_t)y;
    const float jx = rnd01((ux * 73856093u ^ uy * 19349663u ^ sid * 83492791u) ^ F.salt) - 0.5f;
    const float jy = rnd01((ux * 83492791u ^ uy * 2654435761u ^ sid * 19349663u) ^ F.salt) - 0.5f;
    const float ndc_x = (((float)x + 0.5f + jx) / (float)F.width) * 2.0f - 1.0f;
    const float ndc_y = (((float)y + 0.5f + jy) / (float)F.height) * 2.0f - 1.0f;
    float cam_x = ndc_x * F.aspect * F.tan_half;
    float cam_y = -ndc_y * F.tan_half;
    float cam_z = -1.0f;
    const float norm = 1.0f / sqrtf(cam_x * cam_x + cam_y * cam_y + cam_z * cam_z);
    cam_x *= norm;
    cam_y *= norm;
    cam_z *= norm;
    n_r = cam_z;
    n_th = cam_y;
    n_ph = cam_x;
    cam_x_out = cam_x;
    cam_y_out = cam_y;
}

// Initial covariant momentum from the frame tetrad (bh_kerr.h:808-825 and
// 945-964 with the position-only factors precomputed on the host).
__device__ void init_momentum(const FrameConst& F, float n_r, float n_th, float n_ph, double& pt,
                              double& pr, double& pth, double& pph) {
#pragma clang fp contract(off)
    const double nr = (double)n_r, nth = (double)n_th, nph = (double)n_ph;
    double pt_c, pph_c;
    if (F.zamo) {
        pt_c = F.ct0;
        pph_c = F.cph0 + nph * F.cph_ph;
    } else {
        pt_c = F.ct0 + nph * F.ct_ph;
        pph_c = nph * F.cph_ph;
    }
    const double pr_c = nr * F.cr;
    const double pth_c = nth * F.cth;
    pt = F.g_tt * pt_c + F.g_tphi * pph_c;
    pr = F.g_rr * pr_c;
    pth = F.g_thth * pth_c;
    pph = F.g_tphi * pt_c + F.g_phiphi * pph_c;
}

// H, dE/E, dL/L for the invariant monitor (render.cu:234-243, 329-335).
__device__ void invariants(const FrameConst& F, double r, double th, double pt, double pr,
                           double pth, double pph, double pt0, double pph0, float inv[3]) {
    double s, c;
    sincos(th, &s, &c);
    const double M = F.M, a = F.a;
    const double st2 = fmax(1e-12, s * s);
    const double r2 = r * r, a2 = a * a;
    const double Sigma = fmax(1e-12, r2 + a2 * c * c);
    const double Delta = r2 - 2.0 * M * r + a2;
    const double A = (r2 + a2) * (r2 + a2) - a2 * Delta * st2;
    const double grr = fmax(1e-12, Delta) / Sigma;
    const double gthth = 1.0 / Sigma;
    const double gtt = -A / fmax(1e-12, Sigma * Delta);
    const double gtphi = -2.0 * M * a * r / fmax(1e-12, Sigma * Delta);
    const double gphiphi = (Delta - a2 * st2) / fmax(1e-12, Sigma * Delta * st2);
    const double H = 0.5 * (gtt * pt * pt + 2.0 * gtphi * pt * pph + gphiphi * pph * pph +
                            grr * pr * pr + gthth * pth * pth);
    inv[0] = fmaxf(inv[0], (float)fabs(H));
    inv[1] = fmaxf(inv[1], (float)(fabs(pt - pt0) / fmax(1e-12, fabs(pt0))));
    inv[2] = fmaxf(inv[2], (float)(fabs(pph - pph0) / fmax(1e-12, fabs(pph0))));
}

struct Ray {
    Y5 y;
    Deriv k1;
    double pt, pph, pt0, pph0;
    double prev_r, prev_th, min_r, h_adapt, h;
    float disk[3];
    float inv[3];
    float cam_x, cam_y;
    int step, attempt, orders, attempts_total;
    bool k1_valid;
};

__device__ void start_ray(const FrameConst& F, int x, int yy, uint32_t sid, Ray& R) {
    float n_r, n_th, n_ph;
    camera_dir(F, x, yy, sid, n_r, n_th, n_ph, R.cam_x, R.cam_y);
    R.y.r = F.r_cam;
    R.y.th = F.th_cam;
    R.y.ph = F.ph_cam;
    init_momentum(F, n_r, n_th, n_ph, R.pt, R.y.pr, R.y.pth, R.pph);
    R.pt0 = R.pt;
    R.pph0 = R.pph;
    R.prev_r = R.y.r;
    R.prev_th = R.y.th;
    R.min_r = R.y.r;
    R.h_adapt = 0.0;
    R.h = 0.0;
    R.disk[0] = R.disk[1] = R.disk[2] = 0.0f;
    R.inv[0] = R.inv[1] = R.inv[2] = 0.0f;
    R.step = 0;
    R.attempt = 0;
    R.orders = 0;
    R.attempts_total = 0;
    R.k1_valid = false;
}

// Colour of a finished sample (render.cu:312-359).
__device__ void shade_final(const FrameConst& F, const Ray& R, int cls, float col[3]) {
    const float Rs = F.Rs;
    if (cls == NP_RAY_CAPTURED) {
        if (F.bg_mode == NP_BG_POINT) {
            const float r_ph = 1.5f * Rs;
            const float width = fmaxf(0.02f * Rs, 0.005f * Rs + 0.02f * F.bg_sigma_rad * Rs);
            const float dr = ((float)R.min_r - r_ph);
            const float w = expf(-0.5f * (dr / width) * (dr / width));
            const float I = F.bg_peak * w + 0.0f;
            col[0] = col[1] = col[2] = I;
        } else {
            col[0] = col[1] = col[2] = 0.0f;
        }
    } else if (R.orders > 0 && F.disk_on) {
        col[0] = R.disk[0];
        col[1] = R.disk[1];
        col[2] = R.disk[2];
    } else if (F.bg_mode == NP_BG_POINT) {
        const float r_ph = 1.5f * Rs;
        const float width = fmaxf(0.02f * Rs, 0.005f * Rs + 0.02f * F.bg_sigma_rad * Rs);
        const float dr = ((float)R.min_r - r_ph);
        const float w = expf(-0.5f * (dr / width) * (dr / width));
        const float I = F.bg_peak * w + 0.02f;
        col[0] = col[1] = col[2] = I;
    } else {
        // shade_background (render.cu:74-86) with v from the camera direction
        const float v = R.cam_y * 0.5f + 0.5f;
        const float t = fminf(fmaxf(v, 0.f), 1.f);
        float c0 = 0.02f * (1 - t) + 0.08f * t;
        float c1 = 0.03f * (1 - t) + 0.09f * t;
        float c2 = 0.05f * (1 - t) + 0.12f * t;
        const float band = expf(-50.0f * (v - 0.5f) * (v - 0.5f));
        c0 += 0.4f * band;
        c1 += 0.35f * band;
        c2 += 0.3f * band;
        col[0] = fminf(fmaxf(c0, 0.f), 1.f);
        col[1] = fminf(fmaxf(c1, 0.f), 1.f);
        col[2] = fminf(fmaxf(c2, 0.f), 1.f);
    }
}

// One flattened integrator iteration for one lane.  Returns -1 while the
// ray is still running, else its np_ray_class.
template <bool ADAPTIVE>
__device__ __forceinline__ int advance(const FrameConst& F, Ray& R, int& rhs_count, int& attempts) {
    const double Rs = F.Rs_d;
    bool ok = false;
    attempts += 1;
    R.attempts_total += 1;
    if (ADAPTIVE) {
        if (R.attempt == 0) {
            // step-size choice at the start of a step (render.cu:185-191)
            const double scale = fmin(fmax(1.0, R.y.r / (0.5 * Rs)), 100.0);
            double h_base = (double)STEP0 * scale;
            h_base = fmin(fmax(h_base, HMIN), HMAX);
            double h = (R.h_adapt > 0.0) ? R.h_adapt : h_base;
            R.h = fmin(fmax(h, HMIN), HMAX);
        }
        const double h = R.h;
        if (!R.k1_valid) {
            R.k1 = kerr_rhs(F.M, F.a, F.r_plus, R.y.r, R.y.th, R.pt, R.y.pr, R.y.pth, R.pph);
            rhs_count += 1;
            R.k1_valid = true;
        }
        // rk4_adaptive_step_d (bh_kerr.h:1058-1105)
        const Y5 full = rk4_from_k1(F, R.y, R.k1, R.pt, R.pph, h, rhs_count);
        const double h2 = 0.5 * h;
        const Y5 half1 = rk4_from_k1(F, R.y, R.k1, R.pt, R.pph, h2, rhs_count);
        const Deriv k1b = kerr_rhs(F.M, F.a, F.r_plus, half1.r, half1.th, R.pt, half1.pr, half1.pth, R.pph);
        rhs_count += 1;
        const Y5 half2 = rk4_from_k1(F, half1, k1b, R.pt, R.pph, h2, rhs_count);
        const double lim = NP_PI_D - 1e-9;
        const bool ok1 = full.r > F.r_plus + 1e-12 && full.th > 1e-9 && full.th < lim;
        const bool ok2 = half1.r > F.r_plus + 1e-12 && half1.th > 1e-9 && half1.th < lim;
        const bool ok3 = half2.r > F.r_plus + 1e-12 && half2.th > 1e-9 && half2.th < lim;
        double h_try = h;
        double fac = 0.0;
        if (!(ok1 && ok2 && ok3)) {
            h_try *= 0.5;
        } else {
            const double rt = F.rtol, at = F.atol;
            const double e_r = err1(full.r, half2.r, at + rt * fmax(fabs(full.r), fabs(half2.r)));
            const double e_th = err1(full.th, half2.th, at + rt * fmax(fabs(full.th), fabs(half2.th)));
            const double e_ph = err1(full.ph, half2.ph, at + rt * fmax(fabs(full.ph), fabs(half2.ph)));
            const double e_pr = err1(full.pr, half2.pr, at + rt * fmax(fabs(full.pr), fabs(half2.pr)));
            const double e_pth = err1(full.pth, half2.pth, at + rt * fmax(fabs(full.pth), fabs(half2.pth)));
            // p_phi is conserved exactly, so its error term is zero.
            const double e = fmax(fmax(e_r, e_th), fmax(fmax(e_ph, e_pr), fmax(e_pth, 0.0)));
            if (e <= 1.0) {
                fac = fmin(2.0, SAFETY * pow(fmax(1e-12, e), -0.2));
                R.y = half2;
                R.k1_valid = false;
            } else {
                h_try *= fmax(0.1, SAFETY * pow(fmax(1e-12, e), -0.25));
            }
        }
        if (fac > 0.0) {
            ok = true;
            R.h_adapt = fmin(fmax(h_try * fac, HMIN), HMAX);
            R.attempt = 0;
        } else {
            // retry loop of render.cu:193-205
            R.h = fmax(h_try, HMIN);
            R.attempt += 1;
            if (R.h <= HMIN * 1.01 || R.attempt >= 8) return NP_RAY_CAPTURED;
            return -1;
        }
    } else {
        // integrateGeodesicKerrD (bh_kerr.h:734-774), h = 0.01 * scale
        const double scale = fmin(fmax(1.0, R.y.r / (0.5 * Rs)), 100.0);
        const double h = (double)STEP0 * scale;
        const Deriv k1 = kerr_rhs(F.M, F.a, F.r_plus, R.y.r, R.y.th, R.pt, R.y.pr, R.y.pth, R.pph);
        rhs_count += 1;
        R.y = rk4_from_k1(F, R.y, k1, R.pt, R.pph, h, rhs_count);
        ok = R.y.r > F.r_plus + 1e-9 && R.y.th > 1e-9 && R.y.th < NP_PI_D - 1e-9;
    }
    if (!ok) return NP_RAY_CAPTURED;

    // accepted step: bookkeeping of render.cu:210-250
    R.min_r = fmin(R.min_r, R.y.r);
    const double d_prev = R.prev_th - NP_PI_D * 0.5;
    const double d_curr = R.y.th - NP_PI_D * 0.5;
    const double eps = 1e-8;
    const bool sign_change = (d_prev > 0 && d_curr < 0) || (d_prev < 0 && d_curr > 0);
    const bool near_zero = (fabs(d_prev) < eps && fabs(d_curr) >= eps) ||
                           (fabs(d_curr) < eps && fabs(d_prev) >= eps);
    if ((sign_change || near_zero) && R.orders < F.disk_max_orders) {
        const double denom = fabs(d_prev) + fabs(d_curr) + 1e-30;
        const double alpha = fabs(d_prev) / denom;
        const double r_cross = R.prev_r + alpha * (R.y.r - R.prev_r);
        if (r_cross >= F.rin && r_cross <= F.rout) {
            // p_t and p_phi are constant, so their interpolants equal them
            const float g = (float)doppler_g(F, r_cross, R.pt, R.pph);
            const float rr = F.emiss_m_units ? (float)(r_cross / F.M) : (float)r_cross;
            const float j = powf(fmaxf(rr, 1.0f), -F.emiss_q);
            const float I = j * g * g * g;
            R.disk[0] += (1.0f * I) * F.disk_gain;
            R.disk[1] += (0.75f * I) * F.disk_gain;
            R.disk[2] += (0.45f * I) * F.disk_gain;
            R.orders += 1;
            if (F.check_inv)
                invariants(F, r_cross, NP_PI_D * 0.5, R.pt, R.y.pr, R.y.pth, R.pph, R.pt0, R.pph0, R.inv);
        }
    }
    R.prev_r = R.y.r;
    R.prev_th = R.y.th;
    const int step = R.step;
    R.step += 1;
    if (R.y.r > F.r_far && step > 10) return NP_RAY_ESCAPED;
    if (R.step >= F.max_steps) return NP_RAY_STEPCAP;
    return -1;
}

template <bool ADAPTIVE>
__global__ void __launch_bounds__(256)
kerr_render_fp64(const FrameConst F, const int y0, const int rows, const int spp,
                 const int sample_offset, float* __restrict__ out, const int accumulate,
                 float* __restrict__ inv_max3, np_trace_counters* __restrict__ counters,
                 np_ray_record* __restrict__ records, unsigned long long* __restrict__ work) {
    const int W = F.width;
    const int y_end = min(y0 + rows, F.height);
    const uint64_t npix = (y_end > y0) ? (uint64_t)(y_end - y0) * (uint64_t)W : 0ull;

    Ray R;
    bool has = false, done = false;
    uint64_t pix = 0;
    int sample = 0;
    float acc[3] = {0.0f, 0.0f, 0.0f};
    float inv_lane[3] = {0.0f, 0.0f, 0.0f};
    int rhs_count = 0, attempts = 0;
    uint64_t c_rhs = 0, c_att = 0, c_steps = 0, c_rays = 0, c_cap = 0, c_esc = 0, c_cap_steps = 0,
             c_disk = 0;
    if (spp <= 0 || F.max_steps <= 0) done = true;

    while (true) {
        const bool need = !has && !done;
        if (__ballot(need)) {
            const uint64_t idx = wave_fetch(need, work);
            if (need) {
                if (idx < npix) {
                    pix = idx;
                    sample = 0;
                    acc[0] = acc[1] = acc[2] = 0.0f;
                    const int x = (int)(pix % (uint64_t)W);
                    const int yy = y0 + (int)(pix / (uint64_t)W);
                    start_ray(F, x, yy, (uint32_t)(sample_offset + sample), R);
                    has = true;
                } else {
                    done = true;
                }
            }
        }
        if (__ballot(has) == 0) break;
        if (has) {
            rhs_count = 0;
            attempts = 0;
            const int cls = advance<ADAPTIVE>(F, R, rhs_count, attempts);
            c_rhs += (uint64_t)rhs_count;
            c_att += (uint64_t)attempts;
            if (cls >= 0) {
                float col[3];
                shade_final(F, R, cls, col);
                if (F.check_inv && cls != NP_RAY_CAPTURED && !(R.orders > 0 && F.disk_on))
                    invariants(F, R.y.r, R.y.th, R.pt, R.y.pr, R.y.pth, R.pph, R.pt0, R.pph0, R.inv);
                acc[0] += col[0];
                acc[1] += col[1];
                acc[2] += col[2];
                inv_lane[0] = fmaxf(inv_lane[0], R.inv[0]);
                inv_lane[1] = fmaxf(inv_lane[1], R.inv[1]);
                inv_lane[2] = fmaxf(inv_lane[2], R.inv[2]);
                c_rays += 1;
                c_steps += (uint64_t)R.step;
                c_cap += (cls == NP_RAY_CAPTURED);
                c_esc += (cls == NP_RAY_ESCAPED);
                c_cap_steps += (cls == NP_RAY_STEPCAP);
                c_disk += (R.orders > 0);
                if (records) {
                    np_ray_record rec;
                    rec.cls = cls;
                    rec.orders = R.orders;
                    rec.steps = R.step;
                    rec.attempts = R.attempts_total;
                    rec.min_r = R.min_r;
                    rec.r_end = R.y.r;
                    records[pix * (uint64_t)spp + (uint64_t)sample] = rec;
                }
                sample += 1;
                const int x = (int)(pix % (uint64_t)W);
                const int yy = y0 + (int)(pix / (uint64_t)W);
                if (sample < spp) {
                    start_ray(F, x, yy, (uint32_t)(sample_offset + sample), R);
                } else {
                    float* o = out + 3 * pix;
                    if (accumulate) {
                        o[0] += acc[0];
                        o[1] += acc[1];
                        o[2] += acc[2];
                    } else {
                        o[0] = acc[0];
                        o[1] = acc[1];
                        o[2] = acc[2];
                    }
                    has = false;
                }
            }
        }
    }
    if (counters) {
        wave_add_u64(&counters->rays, c_rays);
        wave_add_u64(&counters->steps, c_steps);
        wave_add_u64(&counters->attempts, c_att);
        wave_add_u64(&counters->rhs_evals, c_rhs);
        wave_add_u64(&counters->captured, c_cap);
        wave_add_u64(&counters->escaped, c_esc);
        wave_add_u64(&counters->stepcap, c_cap_steps);
        wave_add_u64(&counters->disk_hits, c_disk);
    }
    if (F.check_inv && inv_max3) {
        atomic_max_nonneg(&inv_max3[0], inv_lane[0]);
        atomic_max_nonneg(&inv_max3[1], inv_lane[1]);
        atomic_max_nonneg(&inv_max3[2], inv_lane[2]);
    }
}

}  // namespace

int launch_kerr_fp64(const FrameConst& F, bool adaptive, int y0, int rows, int spp,
                     int sample_offset, float* d_out, int accumulate, float* d_inv,
                     np_trace_counters* d_counters, np_ray_record* d_records,
                     unsigned long long* d_work, hipStream_t stream, int grid) {
    const dim3 block(256);
    if (adaptive) {
        hipLaunchKernelGGL(kerr_render_fp64<true>, dim3(grid), block, 0, stream, F, y0, rows, spp,
                           sample_offset, d_out, accumulate, d_inv, d_counters, d_records, d_work);
    } else {
        hipLaunchKernelGGL(kerr_render_fp64<false>, dim3(grid), block, 0, stream, F, y0, rows, spp,
                           sample_offset, d_out, accumulate, d_inv, d_counters, d_records, d_work);
    }
    return hipGetLastError() == hipSuccess ? NP_OK : NP_ELAUNCH;
}

int occupancy_kerr_fp64(bool adaptive, int* blocks_per_cu) {
    hipError_t e;
    if (adaptive)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, kerr_render_fp64<true>, 256, 0);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, kerr_render_fp64<false>, 256, 0);
    return e == hipSuccess ? NP_OK : NP_ELAUNCH;
}

}  // namespace np
