// kerr_render.hip -- gfx950 Kerr null-geodesic renderer: the FP64 path
// (reference formulas) and the FP32-compensated path, behind one persistent
// work-queue kernel.
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

// Right-hand side of the integrated coordinates and momenta; t (dt/dlambda)
// is formed by every field but only the kernels that report the end time
// (TIME) integrate it -- elsewhere it is dead code.
struct Deriv {
    double r, th, ph, pr, pth, t;
};

// computeGeodesicDerivativesKerrD (bh_kerr.h:679-727) with the metric
// blocks of bh_kerr.h:600-617 and 624-672 inlined, one sincos shared.
__device__ __forceinline__ Deriv kerr_rhs(const double M, const double a, const double r_plus,
                                          const double r, const double th, const double pt,
                                          const double pr, const double pth, const double pph) {
    Deriv d;
    if (r <= r_plus + 1e-12 || r <= 1e-12) {
        d.r = d.th = d.ph = d.pr = d.pth = d.t = 0.0;
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
    const double Ac = (r2 + a2) * (r2 + a2) - a2 * Delta * st2c;
    const double gtt = -Ac / fmax(1e-12, SigmaC * Delta);
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

    d.t = gtt * pt + gtphi * pph;
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
    double r, th, ph, pr, pth, t;
};

// Classic RK4 step from y with a precomputed k1 = f(y); the stage
// combination k1 + 2k2 + 2k3 + k4 is accumulated with the compensated sum
// in the reference's order (bh_rk4_sum, bh_common.h:485-493) as the stages
// are produced, so only the running sums stay live.
template <bool TIME>
__device__ __forceinline__ Y5 rk4_from_k1(const FrameConst& F, const Y5& y, const Deriv& k1,
                                          const double pt, const double pph, const double h,
                                          int& rhs_count) {
    Klein<double> sr, sth, sph, spr, spth, st;
    sr.add(k1.r); sth.add(k1.th); sph.add(k1.ph); spr.add(k1.pr); spth.add(k1.pth);
    if (TIME) st.add(k1.t);
    const double hh = 0.5 * h;
    Deriv k = kerr_rhs(F.M, F.a, F.r_plus, y.r + hh * k1.r, y.th + hh * k1.th, pt,
                       y.pr + hh * k1.pr, y.pth + hh * k1.pth, pph);
    sr.add(2.0 * k.r); sth.add(2.0 * k.th); sph.add(2.0 * k.ph); spr.add(2.0 * k.pr); spth.add(2.0 * k.pth);
    if (TIME) st.add(2.0 * k.t);
    k = kerr_rhs(F.M, F.a, F.r_plus, y.r + hh * k.r, y.th + hh * k.th, pt,
                 y.pr + hh * k.pr, y.pth + hh * k.pth, pph);
    sr.add(2.0 * k.r); sth.add(2.0 * k.th); sph.add(2.0 * k.ph); spr.add(2.0 * k.pr); spth.add(2.0 * k.pth);
    if (TIME) st.add(2.0 * k.t);
    k = kerr_rhs(F.M, F.a, F.r_plus, y.r + h * k.r, y.th + h * k.th, pt,
                 y.pr + h * k.pr, y.pth + h * k.pth, pph);
    sr.add(k.r); sth.add(k.th); sph.add(k.ph); spr.add(k.pr); spth.add(k.pth);
    if (TIME) st.add(k.t);
    rhs_count += 3;
    const double h6 = h / 6.0;
    Y5 o;
    o.r = y.r + h6 * sr.result();
    o.th = y.th + h6 * sth.result();
    o.ph = y.ph + h6 * sph.result();
    o.pr = y.pr + h6 * spr.result();
    o.pth = y.pth + h6 * spth.result();
    o.t = TIME ? y.t + h6 * st.result() : 0.0;
    return o;
}

// Step doubling of rk4_adaptive_step_d (bh_kerr.h:1058-1070): one RK4
// step of size h and two of size h/2 from y, given k1 = f(y).  The ten
// remaining right-hand sides run in a uniform, non-unrolled stage loop so
// the kernel holds one copy of the RHS code and one stage's temporaries;
// every product, compensated sum and update is the one rk4_step_d makes.
template <bool TIME>
__device__ __forceinline__ void step_doubling_f64(const FrameConst& F, const Y5& y, const Deriv& k1,
                                                  const double pt, const double pph, const double h,
                                                  Y5& full, Y5& half1, Y5& half2) {
    const double h2 = 0.5 * h;
    Klein<double> sr, sth, sph, spr, spth, st;
    sr.add(k1.r); sth.add(k1.th); sph.add(k1.ph); spr.add(k1.pr); spth.add(k1.pth);
    if (TIME) st.add(k1.t);
    Deriv kp = k1;
    Y5 base = y;
#pragma clang loop unroll(disable)
    for (int j = 0; j < 10; ++j) {
        // stage coefficients: full step (j 0-2), first half step (3-5),
        // second half step (6-9; stage 6 is its k1)
        const double c = (j <= 1) ? 0.5 * h : (j == 2) ? h : (j == 5 || j == 9) ? h2 : 0.5 * h2;
        Y5 in = base;
        if (j != 6) {
            in.r = base.r + c * kp.r;
            in.th = base.th + c * kp.th;
            in.pr = base.pr + c * kp.pr;
            in.pth = base.pth + c * kp.pth;
        }
        const Deriv k = kerr_rhs(F.M, F.a, F.r_plus, in.r, in.th, pt, in.pr, in.pth, pph);
        if (j == 6) {
            sr = Klein<double>(); sth = Klein<double>(); sph = Klein<double>();
            spr = Klein<double>(); spth = Klein<double>();
            sr.add(k.r); sth.add(k.th); sph.add(k.ph); spr.add(k.pr); spth.add(k.pth);
            if (TIME) {
                st = Klein<double>();
                st.add(k.t);
            }
        } else {
            const double w = (j == 2 || j == 5 || j == 9) ? 1.0 : 2.0;
            sr.add(w * k.r); sth.add(w * k.th); sph.add(w * k.ph); spr.add(w * k.pr); spth.add(w * k.pth);
            if (TIME) st.add(w * k.t);
        }
        kp = k;
        if (j == 2 || j == 5 || j == 9) {
            const double h6 = (j == 2 ? h : h2) / 6.0;
            const Y5& b = (j == 9) ? half1 : y;
            Y5 o;
            o.r = b.r + h6 * sr.result();
            o.th = b.th + h6 * sth.result();
            o.ph = b.ph + h6 * sph.result();
            o.pr = b.pr + h6 * spr.result();
            o.pth = b.pth + h6 * spth.result();
            o.t = TIME ? b.t + h6 * st.result() : 0.0;
            if (j == 2) {
                full = o;
                sr = Klein<double>(); sth = Klein<double>(); sph = Klein<double>();
                spr = Klein<double>(); spth = Klein<double>();
                sr.add(k1.r); sth.add(k1.th); sph.add(k1.ph); spr.add(k1.pr); spth.add(k1.pth);
                if (TIME) {
                    st = Klein<double>();
                    st.add(k1.t);
                }
                kp = k1;
            } else if (j == 5) {
                half1 = o;
                base = o;
            } else {
                half2 = o;
            }
        }
    }
}

__device__ __forceinline__ double err1(double x, double y, double scale) {
    return fabs(x - y) / fmax(scale, 1e-30);
}

// doppler_g_general_d (bh_kerr.h:833-862) at the disk plane.  sin and
// cos of pi/2 come from the host (F.s_eq, F.c_eq); (r/M)^1.5 is x*sqrt(x).
__device__ __forceinline__ double doppler_g(const FrameConst& F, double r, double pt, double pph) {
    const double M = F.M, a = F.a;
    // kerr_blocks_cov_d at theta = pi/2 (bh_kerr.h:575-592)
    const double s = F.s_eq, c = F.c_eq;
    const double st2 = s * s, r2 = r * r, a2 = a * a;
    const double Sigma = r2 + a2 * c * c;
    const double Delta = r2 - 2.0 * M * r + a2;
    const double Aq = (r2 + a2) * (r2 + a2) - a2 * Delta * st2;
    const double g_tt = -(1.0 - 2.0 * M * r / fmax(1e-12, Sigma));
    const double g_tphi = -2.0 * M * a * r * st2 / fmax(1e-12, Sigma);
    const double g_phiphi = Aq * st2 / fmax(1e-12, Sigma);
    const double M_safe = fmax(1e-12, M);
    const double x = fmax(1.0, r / M_safe);
    const double Omega = 1.0 / (a / M_safe + x * sqrt(x)) / M_safe;
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

// ---------------------------------------------------------------------------
// Lane rows in LDS.
//
// What a lane touches only when a ray starts, crosses the disk plane, runs
// an FP64 attempt or ends (the pixel and sample it works on, the colour
// sums, the crossing count, p_t and p_phi in double, the per-lane totals)
// lives in a row of LDS words instead of VGPRs, so the integrator attempt
// has the register file to itself.  Rows are kLaneSlots words long, an odd
// stride, so the 64 lanes of a wave reading the same slot hit distinct
// banks; every slot is addressed from one base register with an immediate
// offset.  Doubles and u64 take two words (rows are only 4-byte aligned).
// The invariant monitor's values sit in a second region after the rows of
// the whole block.  Launches size the dynamic LDS with lane_lds_bytes.
// ---------------------------------------------------------------------------
enum LaneSlot : int {
    LS_PIX = 0,        // u64: pixel of the window the lane works on
    LS_CEND = 2,       // u64: end of the lane's refill chunk
    LS_SAMPLE = 4,     // sample index within the pixel
    LS_ACC = 5,        // 3 floats: colour sum of the pixel's samples
    LS_N_RAYS = 8,     // per-lane totals (u32), reduced at the end
    LS_N_STEPS = 9,
    LS_N_CAP = 10,
    LS_N_ESC = 11,
    LS_N_STEPCAP = 12,
    LS_N_DISK = 13,
    LS_N_ATT = 14,
    LS_ATT_RAY = 15,   // attempts of the current ray
    LS_DISK = 16,      // 3 floats: disk radiance of the current ray
    LS_CAMY = 19,      // camera-frame y of the ray direction (sky shading)
    LS_ORDERS = 20,    // disk crossings counted so far
    LS_PT = 21,        // double: covariant p_t (exactly conserved)
    LS_PPH = 23,       // double: covariant p_phi
    LS_FULL = 25,      // 5 doubles: the FP64 re-run's full step (FP32C)
};
constexpr int kLaneSlots = 35;
// invariant monitor region (INV kernels): p_t, p_phi at launch, the ray's
// H / dE / dL maxima and the lane's maxima over its rays
enum InvSlot : int { IS_PT0 = 0, IS_PPH0 = 2, IS_INV = 4, IS_INVL = 7 };
constexpr int kInvSlots = 11;

extern __shared__ uint32_t np_lane_lds[];

__device__ __forceinline__ uint32_t* lane_row() { return np_lane_lds + threadIdx.x * kLaneSlots; }
__device__ __forceinline__ uint32_t* inv_row() {
    return np_lane_lds + blockDim.x * kLaneSlots + threadIdx.x * kInvSlots;
}
__device__ __forceinline__ float ld_f(const uint32_t* row, int k) { return __uint_as_float(row[k]); }
__device__ __forceinline__ void st_f(uint32_t* row, int k, float v) { row[k] = __float_as_uint(v); }
__device__ __forceinline__ double ld_d(const uint32_t* row, int k) {
    return __hiloint2double((int)row[k + 1], (int)row[k]);
}
__device__ __forceinline__ void st_d(uint32_t* row, int k, double v) {
    row[k] = (uint32_t)__double2loint(v);
    row[k + 1] = (uint32_t)__double2hiint(v);
}
__device__ __forceinline__ uint64_t ld_u64(const uint32_t* row, int k) {
    return (uint64_t)row[k] | ((uint64_t)row[k + 1] << 32);
}
__device__ __forceinline__ void st_u64(uint32_t* row, int k, uint64_t v) {
    row[k] = (uint32_t)v;
    row[k + 1] = (uint32_t)(v >> 32);
}

// ---------------------------------------------------------------------------
// Per-lane ray bookkeeping shared by the steppers (render.cu:171-183): the
// values every accepted step reads stay in registers, the rest in the lane
// row above.
// ---------------------------------------------------------------------------
struct RayBase {
    double prev_r, prev_th, min_r;   // previous accepted state, closest approach
    int step, attempt;
};

template <bool INV>
__device__ __forceinline__ void base_start(const FrameConst& F, RayBase& R, float cam_y, double pt,
                                           double pph) {
    R.prev_r = F.r_cam;
    R.prev_th = F.th_cam;
    R.min_r = F.r_cam;
    R.step = 0;
    R.attempt = 0;
    uint32_t* row = lane_row();
    st_f(row, LS_DISK + 0, 0.0f);
    st_f(row, LS_DISK + 1, 0.0f);
    st_f(row, LS_DISK + 2, 0.0f);
    st_f(row, LS_CAMY, cam_y);
    row[LS_ORDERS] = 0;
    row[LS_ATT_RAY] = 0;
    st_d(row, LS_PT, pt);
    st_d(row, LS_PPH, pph);
    if (INV) {
        uint32_t* iv = inv_row();
        st_d(iv, IS_PT0, pt);
        st_d(iv, IS_PPH0, pph);
        st_f(iv, IS_INV + 0, 0.0f);
        st_f(iv, IS_INV + 1, 0.0f);
        st_f(iv, IS_INV + 2, 0.0f);
    }
}

// invariants() on the lane's own monitor values
__device__ __forceinline__ void invariants_lane(const FrameConst& F, double r, double th, double pt,
                                                double pr, double pth, double pph) {
    uint32_t* iv = inv_row();
    float inv[3] = {ld_f(iv, IS_INV + 0), ld_f(iv, IS_INV + 1), ld_f(iv, IS_INV + 2)};
    invariants(F, r, th, pt, pr, pth, pph, ld_d(iv, IS_PT0), ld_d(iv, IS_PPH0), inv);
    st_f(iv, IS_INV + 0, inv[0]);
    st_f(iv, IS_INV + 1, inv[1]);
    st_f(iv, IS_INV + 2, inv[2]);
}

// Crossing radius from the cubic Hermite interpolants of theta(lambda) and
// r(lambda) over the accepted step [0, h], given the end values and end
// derivatives: Newton on theta(s) = pi/2 in s = lambda/h from the linear
// estimate s0, kept inside [0, 1]; returns the linear radius when the
// iteration leaves the interval or stalls.
__device__ __forceinline__ double hermite_crossing(double h, double r0, double th0, double dr0,
                                                   double dth0, double r1, double th1, double dr1,
                                                   double dth1, double s0, double r_lin) {
    const double m0 = h * dth0, m1 = h * dth1, n0 = h * dr0, n1 = h * dr1;
    const double g0 = th0 - NP_PI_D * 0.5, g1 = th1 - NP_PI_D * 0.5;
    double s = s0;
#pragma unroll
    for (int it = 0; it < 4; ++it) {
        const double s2 = s * s, s3 = s2 * s;
        const double f = (2 * s3 - 3 * s2 + 1) * g0 + (s3 - 2 * s2 + s) * m0 + (3 * s2 - 2 * s3) * g1 +
                         (s3 - s2) * m1;
        const double fp = (6 * s2 - 6 * s) * g0 + (3 * s2 - 4 * s + 1) * m0 + (6 * s - 6 * s2) * g1 +
                          (3 * s2 - 2 * s) * m1;
        if (!(fabs(fp) > 0.0)) return r_lin;
        s -= f / fp;
        if (!(s >= 0.0 && s <= 1.0)) return r_lin;
    }
    const double s2 = s * s, s3 = s2 * s;
    return (2 * s3 - 3 * s2 + 1) * r0 + (s3 - 2 * s2 + s) * n0 + (3 * s2 - 2 * s3) * r1 + (s3 - s2) * n1;
}

// Accepted-step bookkeeping of render.cu:210-250: closest approach,
// equatorial crossing with linear interpolation of r (or, with
// F.crossing = 1, the cubic Hermite refinement above), disk shading, escape
// and step budget.  Returns -1 while the ray keeps going.
template <class S, bool INV>
__device__ __forceinline__ int after_accept(const FrameConst& F, typename S::Ray& R, double r, double th,
                                            double pr, double pth) {
    R.min_r = fmin(R.min_r, r);
    const double d_prev = R.prev_th - NP_PI_D * 0.5;
    const double d_curr = th - NP_PI_D * 0.5;
    const double eps = 1e-8;
    const bool sign_change = (d_prev > 0 && d_curr < 0) || (d_prev < 0 && d_curr > 0);
    const bool near_zero = (fabs(d_prev) < eps && fabs(d_curr) >= eps) ||
                           (fabs(d_curr) < eps && fabs(d_prev) >= eps);
    uint32_t* row = lane_row();
    if ((sign_change || near_zero) && (int)row[LS_ORDERS] < F.disk_max_orders) {
        const double denom = fabs(d_prev) + fabs(d_curr) + 1e-30;
        const double alpha = fabs(d_prev) / denom;
        double r_cross = R.prev_r + alpha * (r - R.prev_r);
        // The cubic Hermite radius differs from the chord between the two
        // accepted states by at most 4/27 (|n0 - dr| + |n1 - dr|), n = h
        // dr/dlambda at either end, dr = r - prev_r.  Beyond the disk a null
        // ray has |dr/dlambda| <= 1.3 E and h <= HMAX, so when both ends lie
        // more than 2 HMAX (1 + 2E) outside r_out the refined radius cannot
        // reach the disk either: the crossing adds nothing and is left at the
        // chord (no FP64 end derivative).
        const double r_refine = F.rout + 2.0 * HMAX * (1.0 + 2.0 * fabs(S::pt(R)));
        if (F.crossing == 1 && sign_change && fmin(R.prev_r, r) <= r_refine) {
            double h, dr0, dth0, dr1, dth1;
            S::cross_derivs(F, R, h, dr0, dth0, dr1, dth1);
            r_cross = hermite_crossing(h, R.prev_r, R.prev_th, dr0, dth0, r, th, dr1, dth1, alpha, r_cross);
        }
        if (r_cross >= F.rin && r_cross <= F.rout) {
            // p_t and p_phi are constant, so their interpolants equal them
            const double pt = S::pt(R), pph = S::pph(R);
            const float g = (float)doppler_g(F, r_cross, pt, pph);
            const float rr = F.emiss_m_units ? (float)(r_cross / F.M) : (float)r_cross;
            const float j = powf(fmaxf(rr, 1.0f), -F.emiss_q);
            const float I = j * g * g * g;
            st_f(row, LS_DISK + 0, ld_f(row, LS_DISK + 0) + (1.0f * I) * F.disk_gain);
            st_f(row, LS_DISK + 1, ld_f(row, LS_DISK + 1) + (0.75f * I) * F.disk_gain);
            st_f(row, LS_DISK + 2, ld_f(row, LS_DISK + 2) + (0.45f * I) * F.disk_gain);
            row[LS_ORDERS] += 1;
            if (INV) invariants_lane(F, r_cross, NP_PI_D * 0.5, pt, pr, pth, pph);
        }
    }
    R.prev_r = r;
    R.prev_th = th;
    const int step = R.step;
    R.step += 1;
    if (r > F.r_far && step > 10) return NP_RAY_ESCAPED;
    if (R.step >= F.max_steps) return NP_RAY_STEPCAP;
    return -1;
}

// Colour of a finished sample (render.cu:312-359).
__device__ void shade_final(const FrameConst& F, const RayBase& R, int cls, float col[3]) {
    const float Rs = F.Rs;
    const uint32_t* row = lane_row();
    const int orders = (int)row[LS_ORDERS];
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
    } else if (orders > 0 && F.disk_on) {
        col[0] = ld_f(row, LS_DISK + 0);
        col[1] = ld_f(row, LS_DISK + 1);
        col[2] = ld_f(row, LS_DISK + 2);
    } else if (F.bg_mode == NP_BG_POINT) {
        const float r_ph = 1.5f * Rs;
        const float width = fmaxf(0.02f * Rs, 0.005f * Rs + 0.02f * F.bg_sigma_rad * Rs);
        const float dr = ((float)R.min_r - r_ph);
        const float w = expf(-0.5f * (dr / width) * (dr / width));
        const float I = F.bg_peak * w + 0.02f;
        col[0] = col[1] = col[2] = I;
    } else {
        // shade_background (render.cu:74-86) with v from the camera direction
        const float v = ld_f(row, LS_CAMY) * 0.5f + 0.5f;
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

// ---------------------------------------------------------------------------
// FP64 stepper: the reference formulas (bh_kerr.h) in double precision.
// ---------------------------------------------------------------------------
struct StepF64 {
    struct Ray : RayBase {
        Y5 y;
        Deriv k1;
        double pt, pph;   // read by every right-hand side
        double h_adapt, h;
        double lam;       // affine parameter (TIME kernels)
        bool k1_valid;
    };

    template <bool INV>
    __device__ static void start(const FrameConst& F, Ray& R, float n_r, float n_th, float n_ph,
                                 float cam_y) {
        init_momentum(F, n_r, n_th, n_ph, R.pt, R.y.pr, R.y.pth, R.pph);
        R.y.r = F.r_cam;
        R.y.th = F.th_cam;
        R.y.ph = F.ph_cam;
        R.y.t = 0.0;
        R.lam = 0.0;
        R.h_adapt = 0.0;
        R.h = 0.0;
        R.k1_valid = false;
        base_start<INV>(F, R, cam_y, R.pt, R.pph);
    }

    __device__ static double pt(const Ray& R) { return R.pt; }
    __device__ static double pph(const Ray& R) { return R.pph; }
    __device__ static double t(const Ray& R) { return R.y.t; }
    __device__ static double phi(const Ray& R) { return R.y.ph; }
    __device__ static double lam(const Ray& R) { return R.lam; }
    // an accepted step of size R.h advances the affine parameter
    __device__ static void advance(Ray& R) { R.lam += R.h; }
    static constexpr unsigned kRenormEvery = 16;
    static constexpr bool kDefer = false;
    __device__ static void post_accept(const FrameConst&, Ray&, bool) {}

    __device__ static double r(const Ray& R) { return R.y.r; }
    __device__ static double th(const Ray& R) { return R.y.th; }
    __device__ static double pr(const Ray& R) { return R.y.pr; }
    __device__ static double pth(const Ray& R) { return R.y.pth; }

    // Step size and (r, theta) derivatives at both ends of the step just
    // accepted, for the Hermite crossing: R.k1 still holds f at the step's
    // start; f at its end is evaluated here and kept as the next k1.
    __device__ static void cross_derivs(const FrameConst& F, Ray& R, double& h, double& dr0,
                                        double& dth0, double& dr1, double& dth1) {
        h = R.h;
        dr0 = R.k1.r;
        dth0 = R.k1.th;
        R.k1 = kerr_rhs(F.M, F.a, F.r_plus, R.y.r, R.y.th, R.pt, R.y.pr, R.y.pth, R.pph);
        R.k1_valid = true;
        dr1 = R.k1.r;
        dth1 = R.k1.th;
    }

    // One integrator attempt.  Returns 1 accepted, 0 rejected (retry),
    // 2 failed (the ray is captured: horizon, pole or rejection budget).
    template <bool ADAPTIVE, bool TIME>
    __device__ static int attempt(const FrameConst& F, Ray& R, int& rhs_count, int& /*rhs64*/,
                                  int& /*lazy*/) {
        const double Rs = F.Rs_d;
        if (!ADAPTIVE) {
            // integrateGeodesicKerrD (bh_kerr.h:734-774), h = 0.01 * scale
            const double scale = fmin(fmax(1.0, R.y.r / (0.5 * Rs)), 100.0);
            const double h = (double)STEP0 * scale;
            const Deriv k1 = (R.k1_valid) ? R.k1
                                          : kerr_rhs(F.M, F.a, F.r_plus, R.y.r, R.y.th, R.pt, R.y.pr,
                                                     R.y.pth, R.pph);
            rhs_count += R.k1_valid ? 0 : 1;
            R.k1 = k1;
            R.k1_valid = false;
            R.h = h;
            R.y = rk4_from_k1<TIME>(F, R.y, k1, R.pt, R.pph, h, rhs_count);
            const bool ok = R.y.r > F.r_plus + 1e-9 && R.y.th > 1e-9 && R.y.th < NP_PI_D - 1e-9;
            return ok ? 1 : 2;
        }
        if (R.attempt == 0) {
            // step-size choice at the start of a step (render.cu:185-191)
            const double scale = fmin(fmax(1.0, R.y.r / (0.5 * Rs)), 100.0);
            double h_base = (double)STEP0 * scale;
            h_base = fmin(fmax(h_base, HMIN), HMAX);
            const double h = (R.h_adapt > 0.0) ? R.h_adapt : h_base;
            R.h = fmin(fmax(h, HMIN), HMAX);
        }
        const double h = R.h;
        if (!R.k1_valid) {
            R.k1 = kerr_rhs(F.M, F.a, F.r_plus, R.y.r, R.y.th, R.pt, R.y.pr, R.y.pth, R.pph);
            rhs_count += 1;
            R.k1_valid = true;
        }
        // rk4_adaptive_step_d (bh_kerr.h:1058-1105)
        Y5 full, half1, half2;
        step_doubling_f64<TIME>(F, R.y, R.k1, R.pt, R.pph, h, full, half1, half2);
        rhs_count += 10;
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
            R.h_adapt = fmin(fmax(h_try * fac, HMIN), HMAX);
            R.attempt = 0;
            return 1;
        }
        // retry loop of render.cu:193-205
        R.h = fmax(h_try, HMIN);
        R.attempt += 1;
        if (R.h <= HMIN * 1.01 || R.attempt >= 8) return 2;
        return 0;
    }
};

// ---------------------------------------------------------------------------
// FP32-compensated stepper.
//
// * State (r, theta, phi, p_r, p_theta) is a float-float pair per
//   component; each RK4 update is added with an error-free two-sum, so the
//   long far-field legs (thousands of steps of a few 0.1 M) lose nothing to
//   rounding of the state itself.
// * The right-hand side is evaluated in FP32 from the separable (Carter)
//   form of the Kerr Hamiltonian, 2 Sigma H = Delta p_r^2 - P^2/Delta + p_th^2
//   + (L/sin - a E sin)^2 with P = (r^2+a^2)E - aL, and the radial force
//   taken on the constraint surface H = 0, which removes the r/M-fold
//   cancellation of the quotient-rule derivatives:
//     dr/dl   = Delta p_r / Sigma
//     dth/dl  = p_th / Sigma
//     dph/dl  = (a (2MrE - aL)/Delta + L/sin^2) / Sigma
//     dpr/dl  = ((r-M) Q/Delta - P (2EM(r^2-a^2) - 2aL(r-M))/Delta^2) / Sigma
//     dpth/dl = cos sin w (L/sin^2 + aE) / Sigma,  w = L/sin^2 - aE,
//   with Q = p_th^2 + sin^2 w^2 and Delta = (r - r+)(r - r-).
// * Step control is the reference's step doubling (same tolerances, safety
//   factor, growth caps and retry budget) on the compensated state.
// ---------------------------------------------------------------------------
struct ff {
    float hi, lo;
};

__device__ __forceinline__ ff ff_from_d(double d) {
    ff x;
    x.hi = (float)d;
    x.lo = (float)(d - (double)x.hi);
    return x;
}
__device__ __forceinline__ double ff_to_d(ff x) { return (double)x.hi + (double)x.lo; }
__device__ __forceinline__ float ff_f(ff x) { return x.hi + x.lo; }

// x + y with an error-free transformation (Knuth two-sum) and a final
// renormalisation (Dekker fast two-sum).
__device__ __forceinline__ ff ff_add(ff x, float y) {
#pragma clang fp contract(off)
    const float s = x.hi + y;
    const float bb = s - x.hi;
    const float err = (x.hi - (s - bb)) + (y - bb);
    const float lo = err + x.lo;
    ff o;
    o.hi = s + lo;
    o.lo = lo - (o.hi - s);
    return o;
}

// (x - y) of two float-float numbers, rounded to float.
__device__ __forceinline__ float ff_diff(ff x, ff y) {
#pragma clang fp contract(off)
    return (x.hi - y.hi) + (x.lo - y.lo);
}

struct DerivF {
    float r, th, ph, pr, pth, t;
};

struct FF5 {
    ff r, th, ph, pr, pth, t;
};

__device__ __forceinline__ float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }

// theta enters as th + thd with th a float and thd a small float offset (the
// low word of the float-float state plus the stage increment): sin and cos
// of th (s0, c0) are corrected by the rotation through thd (Taylor to
// thd^4).  With theta rounded to one float instead, cos(theta) near the
// equator carries a relative error of tan(theta) * 6e-8, which is what
// limits dp_theta.  The stages of one RK4 run share th (only thd moves), so
// callers evaluate sincos(th) once per run and pass it in.
__device__ __forceinline__ DerivF kerr_rhs_f32_sc(const FrameConst& F, float r, float s0, float c0,
                                                  float thd, float pr, float pth, float E, float L) {
    DerivF d;
    if (r <= F.rpf_guard) {
        d.r = d.th = d.ph = d.pr = d.pth = d.t = 0.0f;
        return d;
    }
    const float d2 = thd * thd;
    const float cd = 1.0f - 0.5f * d2 * (1.0f - d2 * (1.0f / 12.0f));
    const float sd = thd * (1.0f - d2 * (1.0f / 6.0f));
    const float s = s0 * cd + c0 * sd, c = c0 * cd - s0 * sd;
    const float M = F.Mf, a = F.af, a2 = F.a2f;
    const float r2 = r * r;
    const float s2 = fmaxf(s * s, 1e-12f);
    const float Sigma = r2 + a2 * (c * c);
    const float Delta = (r - F.rpf) * (r - F.rmf);
    const float iS = rcpf(Sigma);
    const float iD = rcpf(Delta);
    const float is2 = rcpf(s2);
    const float P = (r2 + a2) * E - a * L;
    const float w = L * is2 - a * E;
    const float Q = pth * pth + s2 * (w * w);
    const float X = 2.0f * E * M * (r2 - a2) - 2.0f * a * L * (r - M);
    d.r = Delta * pr * iS;
    d.th = pth * iS;
    d.ph = (a * (2.0f * M * r * E - a * L) * iD + L * is2) * iS;
    d.pr = ((r - M) * Q * iD - P * X * (iD * iD)) * iS;
    d.pth = c * s * w * (L * is2 + a * E) * iS;
    // dt/dlambda = ((r^2 + a^2) P / Delta + a (L - a E sin^2)) / Sigma
    d.t = ((r2 + a2) * P * iD + a * (L - a * E * s2)) * iS;
    return d;
}

__device__ __forceinline__ DerivF kerr_rhs_f32(const FrameConst& F, float r, float th, float thd,
                                               float pr, float pth, float E, float L) {
    float s0, c0;
    sincos_f(th, s0, c0);
    return kerr_rhs_f32_sc(F, r, s0, c0, thd, pr, pth, E, L);
}

// The same Carter-form field in FP64 (the FP64 steps of the FP32C path:
// inside the FP64 radius and re-runs of float attempts).  Delta = r^2 -
// 2Mr + a^2 as in the reference's FP64 metric; one division for the three
// reciprocals; sincos_d instead of the library's large-argument sincos.
__device__ __forceinline__ Deriv kerr_rhs_c64(const FrameConst& F, double r, double th, double pr,
                                              double pth, double E, double L) {
    Deriv d;
    if (r <= F.r_plus + 1e-12) {
        d.r = d.th = d.ph = d.pr = d.pth = d.t = 0.0;
        return d;
    }
    double s, c;
    sincos_d(th, s, c);
    const double M = F.M, a = F.a, a2 = a * a, r2 = r * r;
    const double s2 = fmax(s * s, 1e-12);
    const double Sigma = r2 + a2 * (c * c);
    const double Delta = r2 - 2.0 * M * r + a2;
    const double SD = Sigma * Delta;
    const double inv = 1.0 / (SD * s2);
    const double iS = Delta * s2 * inv, iD = Sigma * s2 * inv, is2 = SD * inv;
    const double P = (r2 + a2) * E - a * L;
    const double w = L * is2 - a * E;
    const double Q = pth * pth + s2 * (w * w);
    const double X = 2.0 * E * M * (r2 - a2) - 2.0 * a * L * (r - M);
    d.r = Delta * pr * iS;
    d.th = pth * iS;
    d.ph = (a * (2.0 * M * r * E - a * L) * iD + L * is2) * iS;
    d.pr = ((r - M) * Q * iD - P * X * (iD * iD)) * iS;
    d.pth = c * s * w * (L * is2 + a * E) * iS;
    d.t = ((r2 + a2) * P * iD + a * (L - a * E * s2)) * iS;
    return d;
}

// FP64 image of a float-float state.
__device__ __forceinline__ Y5 ff5_to_y5(const FF5& y) {
    return Y5{ff_to_d(y.r), ff_to_d(y.th), ff_to_d(y.ph), ff_to_d(y.pr), ff_to_d(y.pth), ff_to_d(y.t)};
}

__device__ __forceinline__ bool in_domain_d(const FrameConst& F, const Y5& y) {
    return y.r > F.r_plus + 1e-12 && y.th > 1e-9 && y.th < NP_PI_D - 1e-9;
}

// Step doubling with kerr_rhs_c64: one RK4 step of h and two of h/2 from y,
// plain FP64 stage sums, as three RK4 runs of four stages in one
// non-unrolled loop.  This path is rare in a wave, so its live set is kept
// small: k1 = f(y) is evaluated twice (runs 0 and 1) instead of being held,
// the full step waits in the lane row (slots LS_FULL + 2c, c = r, th, ph,
// pr, pth) and the intermediate half step is only checked against the
// domain.
// Returns true when all three results are inside the domain; k1r, k1th are
// the r and theta components of f(y).
template <bool TIME>
__device__ __forceinline__ bool step_doubling_c64(const FrameConst& F, const FF5& y, const double E,
                                                  const double L, const double h, Y5& half2,
                                                  double& k1r, double& k1th) {
    Deriv S, kp;
    Y5 base = ff5_to_y5(y);
    S.r = S.th = S.ph = S.pr = S.pth = S.t = 0.0;
    kp = S;
    bool ok = true;
    uint32_t* row = lane_row();
#pragma clang loop unroll(disable)
    for (int j = 0; j < 12; ++j) {
        const int st = j & 3, run = j >> 2;
        const double hr = (run == 0) ? h : 0.5 * h;
        const double c = (st == 0) ? 0.0 : (st == 3) ? hr : 0.5 * hr;
        const Deriv k = kerr_rhs_c64(F, base.r + c * kp.r, base.th + c * kp.th, base.pr + c * kp.pr,
                                     base.pth + c * kp.pth, E, L);
        const double w = (st == 0 || st == 3) ? 1.0 : 2.0;
        if (j == 0) {
            k1r = k.r;
            k1th = k.th;
        }
        if (st == 0) {
            S = k;
        } else {
            S.r += w * k.r; S.th += w * k.th; S.ph += w * k.ph; S.pr += w * k.pr; S.pth += w * k.pth;
            if (TIME) S.t += w * k.t;
        }
        kp = k;
        if (st == 3) {
            const double h6 = hr / 6.0;
            Y5 o;
            o.r = base.r + h6 * S.r;
            o.th = base.th + h6 * S.th;
            o.ph = base.ph + h6 * S.ph;
            o.pr = base.pr + h6 * S.pr;
            o.pth = base.pth + h6 * S.pth;
            o.t = TIME ? base.t + h6 * S.t : 0.0;
            ok = ok && in_domain_d(F, o);
            if (run == 0) {
                st_d(row, LS_FULL + 0, o.r);
                st_d(row, LS_FULL + 2, o.th);
                st_d(row, LS_FULL + 4, o.ph);
                st_d(row, LS_FULL + 6, o.pr);
                st_d(row, LS_FULL + 8, o.pth);
                base = ff5_to_y5(y);
            } else if (run == 1) {
                base = o;
            } else {
                half2 = o;
            }
        }
    }
    return ok;
}

template <bool TIME>
__device__ __forceinline__ Y5 rk4_c64(const FrameConst& F, const Y5& y, const double E, const double L,
                                      const double h, double& k1r, double& k1th) {
    const double hh = 0.5 * h;
    const Deriv k1 = kerr_rhs_c64(F, y.r, y.th, y.pr, y.pth, E, L);
    k1r = k1.r;
    k1th = k1.th;
    const Deriv k2 = kerr_rhs_c64(F, y.r + hh * k1.r, y.th + hh * k1.th, y.pr + hh * k1.pr,
                                  y.pth + hh * k1.pth, E, L);
    const Deriv k3 = kerr_rhs_c64(F, y.r + hh * k2.r, y.th + hh * k2.th, y.pr + hh * k2.pr,
                                  y.pth + hh * k2.pth, E, L);
    const Deriv k4 = kerr_rhs_c64(F, y.r + h * k3.r, y.th + h * k3.th, y.pr + h * k3.pr,
                                  y.pth + h * k3.pth, E, L);
    const double h6 = h / 6.0;
    Y5 o;
    o.r = y.r + h6 * (k1.r + 2.0 * (k2.r + k3.r) + k4.r);
    o.th = y.th + h6 * (k1.th + 2.0 * (k2.th + k3.th) + k4.th);
    o.ph = y.ph + h6 * (k1.ph + 2.0 * (k2.ph + k3.ph) + k4.ph);
    o.pr = y.pr + h6 * (k1.pr + 2.0 * (k2.pr + k3.pr) + k4.pr);
    o.pth = y.pth + h6 * (k1.pth + 2.0 * (k2.pth + k3.pth) + k4.pth);
    o.t = TIME ? y.t + h6 * (k1.t + 2.0 * (k2.t + k3.t) + k4.t) : 0.0;
    return o;
}

template <bool TIME>
__device__ __forceinline__ FF5 rk4_f32(const FrameConst& F, const FF5& y, const DerivF& k1, float E,
                                       float L, float h, int& rhs_count) {
    const float hh = 0.5f * h;
    const float yr = ff_f(y.r), ypr = ff_f(y.pr), ypth = ff_f(y.pth);
    float s0, c0;
    sincos_f(y.th.hi, s0, c0);
    const DerivF k2 = kerr_rhs_f32_sc(F, yr + hh * k1.r, s0, c0, y.th.lo + hh * k1.th, ypr + hh * k1.pr,
                                      ypth + hh * k1.pth, E, L);
    const DerivF k3 = kerr_rhs_f32_sc(F, yr + hh * k2.r, s0, c0, y.th.lo + hh * k2.th, ypr + hh * k2.pr,
                                      ypth + hh * k2.pth, E, L);
    const DerivF k4 = kerr_rhs_f32_sc(F, yr + h * k3.r, s0, c0, y.th.lo + h * k3.th, ypr + h * k3.pr,
                                      ypth + h * k3.pth, E, L);
    rhs_count += 3;
    const float h6 = h * (1.0f / 6.0f);
    FF5 o;
    o.r = ff_add(y.r, h6 * (k1.r + 2.0f * (k2.r + k3.r) + k4.r));
    o.th = ff_add(y.th, h6 * (k1.th + 2.0f * (k2.th + k3.th) + k4.th));
    o.ph = ff_add(y.ph, h6 * (k1.ph + 2.0f * (k2.ph + k3.ph) + k4.ph));
    o.pr = ff_add(y.pr, h6 * (k1.pr + 2.0f * (k2.pr + k3.pr) + k4.pr));
    o.pth = ff_add(y.pth, h6 * (k1.pth + 2.0f * (k2.pth + k3.pth) + k4.pth));
    if (TIME) o.t = ff_add(y.t, h6 * (k1.t + 2.0f * (k2.t + k3.t) + k4.t));
    return o;
}

// FP32 counterpart of step_doubling_f64: same stage sequence, float RHS,
// float stage sums, float-float updates of the state.  Stages 0-5 (the full
// step's k2..k4 and the first half step's k2..k4) share the base theta of y,
// whose sin and cos (s0, c0) the caller passes in; stages 6-9 share that of
// half1.  The full step is returned as its float increments inc = full - y
// (all the error norm needs) and as float-float r and theta (for the domain
// check); the half steps as float-float states.
// full_step (wave-uniform) = false skips stages 0-2 (the full step), for a
// wave whose lanes all skip the error estimate; inc, full_r and full_th are
// then left unset.
template <bool TIME>
__device__ __forceinline__ void step_doubling_f32(const FrameConst& F, const FF5& y, const DerivF& k1,
                                                  const float s0, const float c0, const float E,
                                                  const float L, const float h, const bool full_step,
                                                  DerivF& inc, ff& full_r, ff& full_th, FF5& half1,
                                                  FF5& half2) {
    const float h2 = 0.5f * h;
    DerivF S = k1, kp = k1;
    float br = ff_f(y.r), bthl = y.th.lo, bpr = ff_f(y.pr), bpth = ff_f(y.pth);
    float bs = s0, bc = c0;
    // unrolled: the stage coefficients below fold to constants
#pragma unroll
    for (int j = 0; j < 10; ++j) {
        if (j < 3 && !full_step) continue;
        const float c = (j <= 1) ? 0.5f * h : (j == 2) ? h : (j == 5 || j == 9) ? h2 : 0.5f * h2;
        const float cc = (j == 6) ? 0.0f : c;
        const DerivF k = kerr_rhs_f32_sc(F, br + cc * kp.r, bs, bc, bthl + cc * kp.th, bpr + cc * kp.pr,
                                         bpth + cc * kp.pth, E, L);
        if (j == 6) {
            S = k;
        } else {
            const float w = (j == 2 || j == 5 || j == 9) ? 1.0f : 2.0f;
            S.r += w * k.r; S.th += w * k.th; S.ph += w * k.ph; S.pr += w * k.pr; S.pth += w * k.pth;
            if (TIME) S.t += w * k.t;
        }
        kp = k;
        if (j == 2) {
            const float h6 = h * (1.0f / 6.0f);
            inc = DerivF{h6 * S.r, h6 * S.th, h6 * S.ph, h6 * S.pr, h6 * S.pth};
            full_r = ff_add(y.r, inc.r);
            full_th = ff_add(y.th, inc.th);
            S = k1;
            kp = k1;
        } else if (j == 5 || j == 9) {
            const float h6 = h2 * (1.0f / 6.0f);
            const FF5& b = (j == 9) ? half1 : y;
            FF5 o;
            o.r = ff_add(b.r, h6 * S.r);
            o.th = ff_add(b.th, h6 * S.th);
            o.ph = ff_add(b.ph, h6 * S.ph);
            o.pr = ff_add(b.pr, h6 * S.pr);
            o.pth = ff_add(b.pth, h6 * S.pth);
            if (TIME) o.t = ff_add(b.t, h6 * S.t);
            if (j == 5) {
                half1 = o;
                br = ff_f(o.r); bthl = o.th.lo; bpr = ff_f(o.pr); bpth = ff_f(o.pth);
                sincos_f(o.th.hi, bs, bc);
            } else {
                half2 = o;
            }
        }
    }
}

// Error-norm term of one component for a float attempt: |full - half2|
// over atol + rtol max(|full|, |half2|), after removing the part of the
// difference that float rounding of the stage values alone can produce
// (kNoise ulps of the step's magnitude, taken as the full step's increment
// plus h |k1|).  Differences at that floor carry no truncation information;
// counting them would reject sound far-field steps of components that pass
// near zero (p_r at periapsis, p_theta at a theta turning point).
// Both differences are increment-sized: full - y0 is the float increment
// inc itself and full - half2 = inc - (half2 - y0), each within two ulps of
// the increment, well under the floor.
constexpr float kNoise = 16.0f * 5.9604645e-8f;
__device__ __forceinline__ float errf(float inc, ff y2, ff y0, float hk1, float rt, float at) {
    const float d2 = ff_diff(y2, y0);
    const float sc = at + rt * fmaxf(fabsf(y0.hi + inc), fabsf(y2.hi));
    const float floor_ = kNoise * (fabsf(inc) + fabsf(hk1));
    return fmaxf(fabsf(inc - d2) - floor_, 0.0f) * rcpf(fmaxf(sc, 1e-30f));
}

// Domain check of a float attempt's result on the high words of r and theta
// (thresholds of FrameConst); NaN fails.  Results in the sliver between
// these thresholds and the FP64 bounds go to the FP64 re-run like any other
// failed float attempt.
__device__ __forceinline__ bool okf(const FrameConst& F, float r_hi, float th_hi) {
    return r_hi > F.r_okf && th_hi > F.th_okf_lo && th_hi < F.th_okf_hi;
}

// Error level below which a ray's later steps at the maximum step size
// beyond F.r_lazyf skip their estimate: the estimate would have to grow past
// 0.59 (0.9 e^-0.2 < 1) to shrink the next step, and past 1 to reject, while
// there the truncation error of a 0.2 M step is ~1e-5 of the tolerance.
constexpr float kLazyError = 1e-3f;

// Float error level above which a float attempt's rejection is final (no
// FP64 re-run): twice the tolerance.
constexpr float kClearReject = 2.0f;

// Deferred FP64 attempts of the FP32C path (StepF32C::kDefer): the wave runs
// its FP64 block once kDeferLanes lanes wait for it or kDeferIters
// iterations after it last ran.
constexpr int kDeferLanes = 24;
constexpr uint32_t kDeferIters = 16;

struct StepF32C {
    struct Ray : RayBase {
        FF5 y;            // y.t, the coordinate time, is integrated by TIME kernels only
        DerivF k1;
        ff lam;           // affine parameter, compensated sum of the accepted steps (TIME)
        float E, L, h_adapt, h;
        // sqrt(Q + a^2 E^2) >= |p_theta| anywhere on the ray (Q = Carter's
        // constant): bounds how far one step can move theta
        float qb;
        bool k1_valid;
        // 0: float attempt; 1: FP64 (step started inside r_isl); 2: FP64
        // re-run of a float attempt that failed or was rejected
        uint8_t mode;
        // 1 once an error estimate beyond r_lazyf at the maximum step size
        // came out below kLazyError; cleared when the ray leaves that regime
        uint8_t skip;
    };

    // p_t and p_phi in double live in the lane row (FP64 attempts and
    // crossings read them); float steps use E and L.
    __device__ static double pt(const Ray&) { return ld_d(lane_row(), LS_PT); }
    __device__ static double pph(const Ray&) { return ld_d(lane_row(), LS_PPH); }

    template <bool INV>
    __device__ static void start(const FrameConst& F, Ray& R, float n_r, float n_th, float n_ph,
                                 float cam_y) {
        double pt, pph, pr, pth;
        init_momentum(F, n_r, n_th, n_ph, pt, pr, pth, pph);
        R.y.r = ff_from_d(F.r_cam);
        R.y.th = ff_from_d(F.th_cam);
        R.y.ph = ff_from_d(F.ph_cam);
        R.y.pr = ff_from_d(pr);
        R.y.pth = ff_from_d(pth);
        R.y.t = ff{0.0f, 0.0f};
        R.lam = ff{0.0f, 0.0f};
        R.E = (float)(-pt);
        R.L = (float)pph;
        R.h_adapt = 0.0f;
        R.h = 0.0f;
        R.k1_valid = false;
        R.mode = 0;
        R.skip = 0;
        {
            const double E = -pt, L = pph, a2E2 = F.a * F.a * E * E;
            const double Q = pth * pth + F.c2_cam * (L * L / fmax(F.s2_cam, 1e-300) - a2E2);
            R.qb = (float)(sqrt(fmax(Q + a2E2, 0.0)) * (1.0 + 1e-6));
        }
        base_start<INV>(F, R, cam_y, pt, pph);
    }

    __device__ static double r(const Ray& R) { return ff_to_d(R.y.r); }
    __device__ static double th(const Ray& R) { return ff_to_d(R.y.th); }
    __device__ static double pr(const Ray& R) { return ff_to_d(R.y.pr); }
    __device__ static double pth(const Ray& R) { return ff_to_d(R.y.pth); }
    __device__ static double t(const Ray& R) { return ff_to_d(R.y.t); }
    __device__ static double phi(const Ray& R) { return ff_to_d(R.y.ph); }
    __device__ static double lam(const Ray& R) { return ff_to_d(R.lam); }
    // an accepted step of size R.h advances the affine parameter (two-sum)
    __device__ static void advance(Ray& R) { R.lam = ff_add(R.lam, R.h); }

    // With F.renorm, p_r is set from the constraint H = 0 after the steps a
    // wave accepts on every kRenormEvery-th attempt iteration (`tick`, the
    // same for all lanes, so the FP64 block runs for the whole wave at once
    // instead of for some lane on nearly every iteration; the drift between
    // two renormalisations is ~1e-10 E^2): in the Carter form 2 Sigma H =
    // Delta p_r^2 - P^2/Delta + p_theta^2 + (L/sin - a E sin)^2, so Delta^2
    // p_r^2 = P^2 - Delta K with K = p_theta^2 + (L/sin - a E sin)^2.  Only
    // where radial motion dominates (4 Delta^2 p_r^2 >= P^2): near a radial
    // turning point the square root would amplify the rounding of K.  FP64 on
    // the float-float state; the sign of p_r is kept.
    static constexpr unsigned kRenormEvery = 16;

    // FP64 attempts (mode 1 and 2) may be deferred by the persistent loop:
    // a lane whose next attempt is FP64 waits until enough lanes of its wave
    // need one, or a few iterations have passed, so the wave runs the FP64
    // block for several lanes at once instead of beside every float
    // iteration.  The ray's own sequence of attempts is unchanged.
    static constexpr bool kDefer = true;
    __device__ static bool pending64(const Ray& R) { return R.mode != 0; }

    // step-size choice at the start of a step (render.cu:185-191) and the
    // precision of its attempts: FP64 inside r_isl.  Idempotent until the
    // attempt runs.
    __device__ static void start_step(const FrameConst& F, Ray& R) {
        if (R.attempt == 0 && R.mode != 2) {
            const float hmin = (float)HMIN, hmax = (float)HMAX;
            const float rf = ff_f(R.y.r);
            const float scale = fminf(fmaxf(1.0f, rf / (0.5f * F.Rs)), 100.0f);
            const float h_base = fminf(fmaxf(STEP0 * scale, hmin), hmax);
            const float h = (R.h_adapt > 0.0f) ? R.h_adapt : h_base;
            R.h = fminf(fmaxf(h, hmin), hmax);
            R.mode = (ff_to_d(R.y.r) < F.r_isl) ? 1 : 0;
        }
    }

    __device__ static void post_accept(const FrameConst& F, Ray& R, bool tick) {
        if (!F.renorm || !tick) return;
        const double r = ff_to_d(R.y.r), th = ff_to_d(R.y.th), pr = ff_to_d(R.y.pr), pth = ff_to_d(R.y.pth);
        const double E = -pt(R), L = pph(R), a = F.a, a2 = a * a;
        double s, c;
        sincos_d(th, s, c);
        const double Delta = r * r - 2.0 * F.M * r + a2;
        const double P = (r * r + a2) * E - a * L;
        const double w = L / s - a * E * s;
        const double Rr = P * P - Delta * (pth * pth + w * w);
        if (Delta > 0.0 && Rr > 0.0 && 4.0 * Delta * Delta * pr * pr >= P * P)
            R.y.pr = ff_from_d(copysign(sqrt(Rr), pr) / Delta);
    }

    __device__ static FF5 to_ff5(const Y5& y) {
        return FF5{ff_from_d(y.r), ff_from_d(y.th), ff_from_d(y.ph), ff_from_d(y.pr), ff_from_d(y.pth),
                   ff_from_d(y.t)};
    }

    // Step size and (r, theta) derivatives at both ends of the step just
    // accepted, for the Hermite crossing: R.k1 holds f at the step's start
    // (float path, or the FP64 step's first stage); f at its end is
    // evaluated in FP64 and kept as the next k1.
    __device__ static void cross_derivs(const FrameConst& F, Ray& R, double& h, double& dr0,
                                        double& dth0, double& dr1, double& dth1) {
        h = (double)R.h;
        dr0 = (double)R.k1.r;
        dth0 = (double)R.k1.th;
        const Deriv k = kerr_rhs_c64(F, ff_to_d(R.y.r), ff_to_d(R.y.th), ff_to_d(R.y.pr),
                                     ff_to_d(R.y.pth), -pt(R), pph(R));
        dr1 = k.r;
        dth1 = k.th;
        R.k1 = DerivF{(float)k.r, (float)k.th, (float)k.ph, (float)k.pr, (float)k.pth};
        R.k1_valid = true;
    }

    // FP64 attempt with the Carter-form field: the step control of
    // StepF64::attempt (rk4_adaptive_step_d, bh_kerr.h:1058-1105, and the
    // retry loop of render.cu:193-205) on the FP64 image of the state.
    // Step sizes are held in float (a 6e-8 relative change of h); so is the
    // step factor.
    template <bool TIME>
    __device__ static int attempt64(const FrameConst& F, Ray& R, int& rhs64) {
        const double h = (double)R.h;
        Y5 half2;
        double k1r, k1th;
        const bool ok = step_doubling_c64<TIME>(F, R.y, -pt(R), pph(R), h, half2, k1r, k1th);
        rhs64 += 12;
        float h_try = R.h;
        float fac = 0.0f;
        if (!ok) {
            h_try *= 0.5f;
        } else {
            const double rt = F.rtol, at = F.atol;
            const uint32_t* row = lane_row();
            double e = 0.0;
            const double* hv = &half2.r;
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                const double fv = ld_d(row, LS_FULL + 2 * c);
                e = fmax(e, err1(fv, hv[c], at + rt * fmax(fabs(fv), fabs(hv[c]))));
            }
            const float le = __log2f(fmaxf(1e-12f, (float)e));
            if (e <= 1.0) {
                fac = fminf(2.0f, 0.9f * exp2f(-0.2f * le));
                R.y = to_ff5(half2);
                R.k1.r = (float)k1r;
                R.k1.th = (float)k1th;
                R.k1_valid = false;
            } else {
                h_try *= fmaxf(0.1f, 0.9f * exp2f(-0.25f * le));
            }
        }
        const float hmin = (float)HMIN, hmax = (float)HMAX;
        if (fac > 0.0f) {
            R.h_adapt = fminf(fmaxf(h_try * fac, hmin), hmax);
            R.attempt = 0;
            R.mode = 0;
            return 1;
        }
        R.h = fmaxf(h_try, hmin);
        R.attempt += 1;
        if (R.h <= hmin * 1.01f || R.attempt >= 8) return 2;
        return 0;
    }

    // STARTED: start_step has run for this attempt (the caller deferred it).
    template <bool ADAPTIVE, bool TIME, bool STARTED = false>
    __device__ static int attempt(const FrameConst& F, Ray& R, int& rhs_count, int& rhs64,
                                  int& lazy_att) {
        const float hmin = (float)HMIN, hmax = (float)HMAX;
        if (!ADAPTIVE) {
            const float rf = ff_f(R.y.r);
            const float scale = fminf(fmaxf(1.0f, rf / (0.5f * F.Rs)), 100.0f);
            const float h = STEP0 * scale;
            R.h = h;
            if (ff_to_d(R.y.r) < F.r_isl) {
                double k1r, k1th;
                R.y = to_ff5(rk4_c64<TIME>(F, ff5_to_y5(R.y), -pt(R), pph(R), (double)h, k1r, k1th));
                R.k1.r = (float)k1r;
                R.k1.th = (float)k1th;
                rhs64 += 4;
            } else {
                const DerivF k1 = kerr_rhs_f32(F, rf, R.y.th.hi, R.y.th.lo, ff_f(R.y.pr), ff_f(R.y.pth), R.E, R.L);
                rhs_count += 1;
                R.k1 = k1;
                R.y = rk4_f32<TIME>(F, R.y, k1, R.E, R.L, h, rhs_count);
            }
            const double rr = ff_to_d(R.y.r), tt = ff_to_d(R.y.th);
            const bool ok = rr > F.r_plus + 1e-9 && tt > 1e-9 && tt < NP_PI_D - 1e-9;
            return ok ? 1 : 2;
        }
        if (!STARTED) start_step(F, R);
        if (R.mode != 0) {
            R.skip = 0;
            return attempt64<TIME>(F, R, rhs64);
        }
        const float h = R.h;
        float s0, c0;
        sincos_f(R.y.th.hi, s0, c0);
        if (!R.k1_valid) {
            R.k1 = kerr_rhs_f32_sc(F, ff_f(R.y.r), s0, c0, R.y.th.lo, ff_f(R.y.pr), ff_f(R.y.pth), R.E, R.L);
            rhs_count += 1;
            R.k1_valid = true;
        }
        DerivF inc;
        ff full_r, full_th;
        FF5 half1, half2;
        const bool far = h == hmax && R.y.r.hi >= F.r_lazyf;
        bool lazy = false;
        if (far && R.skip) {
            // The skipped full step stays inside the domain: every stage is
            // beyond r - 1 >= 99 M from the hole, where |dtheta/dlambda| =
            // |p_theta| / Sigma <= qb / (r - 1)^2; its theta moves by at most
            // h times that, kept 1e-6 clear of the pole thresholds.
            const float rm = R.y.r.hi - 1.0f * F.Mf;
            const float dth = h * R.qb * rcpf(rm * rm) + 1e-6f;
            lazy = R.y.th.hi - dth > F.th_okf_lo && R.y.th.hi + dth < F.th_okf_hi;
        }
        R.skip = lazy;
        // the full step runs when any float lane of the wave estimates its
        // error (a uniform branch: a lane's own result never depends on it)
        const bool full_step = __ballot(!lazy) != 0;
        step_doubling_f32<TIME>(F, R.y, R.k1, s0, c0, R.E, R.L, h, full_step, inc, full_r, full_th, half1,
                          half2);
        rhs_count += full_step ? 10 : 7;
        lazy_att = full_step ? 0 : 1;
        if (lazy) {
            if (okf(F, half1.r.hi, half1.th.hi) && okf(F, half2.r.hi, half2.th.hi)) {
                R.y = half2;
                R.k1_valid = false;
                R.h_adapt = hmax;
                R.attempt = 0;
                return 1;
            }
            R.skip = 0;
            R.mode = 2;
            return 0;
        }
        if (okf(F, full_r.hi, full_th.hi) && okf(F, half1.r.hi, half1.th.hi) &&
            okf(F, half2.r.hi, half2.th.hi)) {
            const float rt = F.rtolf, at = F.atolf;
            const DerivF& k = R.k1;
            const float e = fmaxf(fmaxf(errf(inc.r, half2.r, R.y.r, h * k.r, rt, at),
                                        errf(inc.th, half2.th, R.y.th, h * k.th, rt, at)),
                                  fmaxf(errf(inc.ph, half2.ph, R.y.ph, h * k.ph, rt, at),
                                        fmaxf(errf(inc.pr, half2.pr, R.y.pr, h * k.pr, rt, at),
                                              errf(inc.pth, half2.pth, R.y.pth, h * k.pth, rt, at))));
            if (e <= 1.0f) {
                const float le = __log2f(fmaxf(1e-12f, e));
                const float fac = fminf(2.0f, 0.9f * exp2f(-0.2f * le));
                R.y = half2;
                R.k1_valid = false;
                R.h_adapt = fminf(fmaxf(h * fac, hmin), hmax);
                R.attempt = 0;
                R.skip = F.lazy && far && e < kLazyError;
                return 1;
            }
            if (e > kClearReject) {
                // A clear rejection: the noise floor is subtracted from every
                // difference, so the float error level is a lower bound of
                // the FP64 one and the step is rejected in both.  The retry
                // runs in float with the step law of render.cu:193-205 on
                // the float error level; no FP64 re-run is needed to decide.
                const float le = __log2f(e);
                R.h = fmaxf(h * fmaxf(0.1f, 0.9f * exp2f(-0.25f * le)), hmin);
                R.attempt += 1;
                R.skip = 0;
                if (R.h <= hmin * 1.01f || R.attempt >= 8) return 2;
                return 0;
            }
        }
        // A float attempt that leaves the domain or fails the tolerance by a
        // margin within float rounding is re-run in FP64 with the same step;
        // that run decides acceptance, the next step size and the retry
        // budget exactly as the FP64 path.
        R.mode = 2;
        return 0;
    }
};


// ---------------------------------------------------------------------------
// FP64 comparator with the FP32C path's formulas (NP_PREC_FP64_CARTER): the
// Carter-form field kerr_rhs_c64 with the polynomial sincos_d, plain stage
// sums, and the FP32C step control (step doubling with the reference's
// constants, the far-field estimate skip under a wave-uniform full step),
// all in double.  It exists to time the precision alone: FP32C against this
// path differs only in the arithmetic type.
// ---------------------------------------------------------------------------
template <bool TIME>
__device__ __forceinline__ void step_doubling_c64_lazy(const FrameConst& F, const Y5& y, const Deriv& k1,
                                                       const double E, const double L, const double h,
                                                       const bool full_step, Y5& full, Y5& half1,
                                                       Y5& half2) {
    const double h2 = 0.5 * h;
    Deriv S = k1, kp = k1;
    Y5 b = y;
#pragma unroll
    for (int j = 0; j < 10; ++j) {
        if (j < 3 && !full_step) continue;
        const double c = (j <= 1) ? 0.5 * h : (j == 2) ? h : (j == 5 || j == 9) ? h2 : 0.5 * h2;
        const double cc = (j == 6) ? 0.0 : c;
        const Deriv k = kerr_rhs_c64(F, b.r + cc * kp.r, b.th + cc * kp.th, b.pr + cc * kp.pr,
                                     b.pth + cc * kp.pth, E, L);
        if (j == 6) {
            S = k;
        } else {
            const double w = (j == 2 || j == 5 || j == 9) ? 1.0 : 2.0;
            S.r += w * k.r; S.th += w * k.th; S.ph += w * k.ph; S.pr += w * k.pr; S.pth += w * k.pth;
            if (TIME) S.t += w * k.t;
        }
        kp = k;
        if (j == 2 || j == 5 || j == 9) {
            const double h6 = (j == 2 ? h : h2) / 6.0;
            const Y5& base = (j == 9) ? half1 : y;
            Y5 o;
            o.r = base.r + h6 * S.r;
            o.th = base.th + h6 * S.th;
            o.ph = base.ph + h6 * S.ph;
            o.pr = base.pr + h6 * S.pr;
            o.pth = base.pth + h6 * S.pth;
            o.t = TIME ? base.t + h6 * S.t : 0.0;
            if (j == 2) {
                full = o;
                S = k1;
                kp = k1;
            } else if (j == 5) {
                half1 = o;
                b = o;
            } else {
                half2 = o;
            }
        }
    }
}

struct StepC64 {
    struct Ray : RayBase {
        Y5 y;
        Deriv k1;
        double pt, pph;   // read by every right-hand side
        double h_adapt, h, qb;
        double lam;       // affine parameter (TIME kernels)
        bool k1_valid, skip;
    };

    __device__ static double pt(const Ray& R) { return R.pt; }
    __device__ static double pph(const Ray& R) { return R.pph; }
    __device__ static double t(const Ray& R) { return R.y.t; }
    __device__ static double phi(const Ray& R) { return R.y.ph; }
    __device__ static double lam(const Ray& R) { return R.lam; }
    __device__ static void advance(Ray& R) { R.lam += R.h; }
    static constexpr unsigned kRenormEvery = 16;
    static constexpr bool kDefer = false;
    __device__ static void post_accept(const FrameConst&, Ray&, bool) {}

    template <bool INV>
    __device__ static void start(const FrameConst& F, Ray& R, float n_r, float n_th, float n_ph,
                                 float cam_y) {
        init_momentum(F, n_r, n_th, n_ph, R.pt, R.y.pr, R.y.pth, R.pph);
        R.y.r = F.r_cam;
        R.y.th = F.th_cam;
        R.y.ph = F.ph_cam;
        R.y.t = 0.0;
        R.lam = 0.0;
        R.h_adapt = 0.0;
        R.h = 0.0;
        R.k1_valid = false;
        R.skip = false;
        const double E = -R.pt, L = R.pph, a2E2 = F.a * F.a * E * E;
        const double Q = R.y.pth * R.y.pth + F.c2_cam * (L * L / fmax(F.s2_cam, 1e-300) - a2E2);
        R.qb = sqrt(fmax(Q + a2E2, 0.0)) * (1.0 + 1e-6);
        base_start<INV>(F, R, cam_y, R.pt, R.pph);
    }

    __device__ static double r(const Ray& R) { return R.y.r; }
    __device__ static double th(const Ray& R) { return R.y.th; }
    __device__ static double pr(const Ray& R) { return R.y.pr; }
    __device__ static double pth(const Ray& R) { return R.y.pth; }

    __device__ static void cross_derivs(const FrameConst& F, Ray& R, double& h, double& dr0,
                                        double& dth0, double& dr1, double& dth1) {
        h = R.h;
        dr0 = R.k1.r;
        dth0 = R.k1.th;
        R.k1 = kerr_rhs_c64(F, R.y.r, R.y.th, R.y.pr, R.y.pth, -R.pt, R.pph);
        R.k1_valid = true;
        dr1 = R.k1.r;
        dth1 = R.k1.th;
    }

    template <bool ADAPTIVE, bool TIME>
    __device__ static int attempt(const FrameConst& F, Ray& R, int& rhs_count, int& /*rhs64*/,
                                  int& lazy_att) {
        const double E = -R.pt, L = R.pph;
        const double Rs = F.Rs_d;
        if (!ADAPTIVE) {
            const double scale = fmin(fmax(1.0, R.y.r / (0.5 * Rs)), 100.0);
            const double h = (double)STEP0 * scale;
            double k1r, k1th;
            R.y = rk4_c64<TIME>(F, R.y, E, L, h, k1r, k1th);
            R.k1.r = k1r;
            R.k1.th = k1th;
            R.k1_valid = false;
            R.h = h;
            rhs_count += 4;
            const bool ok = R.y.r > F.r_plus + 1e-9 && R.y.th > 1e-9 && R.y.th < NP_PI_D - 1e-9;
            return ok ? 1 : 2;
        }
        if (R.attempt == 0) {
            const double scale = fmin(fmax(1.0, R.y.r / (0.5 * Rs)), 100.0);
            const double h_base = fmin(fmax((double)STEP0 * scale, HMIN), HMAX);
            const double h = (R.h_adapt > 0.0) ? R.h_adapt : h_base;
            R.h = fmin(fmax(h, HMIN), HMAX);
        }
        const double h = R.h;
        if (!R.k1_valid) {
            R.k1 = kerr_rhs_c64(F, R.y.r, R.y.th, R.y.pr, R.y.pth, E, L);
            rhs_count += 1;
            R.k1_valid = true;
        }
        const bool far = h == HMAX && R.y.r >= (double)F.r_lazyf;
        bool lazy = false;
        if (far && R.skip) {
            const double rm = R.y.r - F.M;
            const double dth = h * R.qb / (rm * rm) + 1e-6;
            lazy = R.y.th - dth > 1e-9 + 1e-6 && R.y.th + dth < NP_PI_D - 1e-9 - 1e-6;
        }
        R.skip = lazy;
        const bool full_step = __ballot(!lazy) != 0;
        Y5 full, half1, half2;
        step_doubling_c64_lazy<TIME>(F, R.y, R.k1, E, L, h, full_step, full, half1, half2);
        rhs_count += full_step ? 10 : 7;
        lazy_att = full_step ? 0 : 1;
        const double lim = NP_PI_D - 1e-9;
        const bool ok2 = half1.r > F.r_plus + 1e-12 && half1.th > 1e-9 && half1.th < lim && half2.r > F.r_plus + 1e-12 &&
                         half2.th > 1e-9 && half2.th < lim;
        double h_try = h, fac = 0.0, e = 0.0;
        if (lazy && ok2) {
            fac = 2.0;
        } else if (!lazy && ok2 && full.r > F.r_plus + 1e-12 && full.th > 1e-9 && full.th < lim) {
            const double rt = F.rtol, at = F.atol;
            e = fmax(fmax(err1(full.r, half2.r, at + rt * fmax(fabs(full.r), fabs(half2.r))),
                          err1(full.th, half2.th, at + rt * fmax(fabs(full.th), fabs(half2.th)))),
                     fmax(err1(full.ph, half2.ph, at + rt * fmax(fabs(full.ph), fabs(half2.ph))),
                          fmax(err1(full.pr, half2.pr, at + rt * fmax(fabs(full.pr), fabs(half2.pr))),
                               err1(full.pth, half2.pth, at + rt * fmax(fabs(full.pth), fabs(half2.pth))))));
            if (e <= 1.0) fac = fmin(2.0, SAFETY * pow(fmax(1e-12, e), -0.2));
            else h_try *= fmax(0.1, SAFETY * pow(fmax(1e-12, e), -0.25));
        } else {
            h_try *= 0.5;
        }
        if (fac > 0.0) {
            R.y = half2;
            R.k1_valid = false;
            R.h_adapt = fmin(fmax(h_try * fac, HMIN), HMAX);
            R.attempt = 0;
            if (!lazy) R.skip = F.lazy && far && e < (double)kLazyError;
            return 1;
        }
        R.skip = false;
        R.h = fmax(h_try, HMIN);
        R.attempt += 1;
        if (R.h <= HMIN * 1.01 || R.attempt >= 8) return 2;
        return 0;
    }
};

// ---------------------------------------------------------------------------
// Persistent renderer: work item = pixel (or a chunk of consecutive pixels),
// lane-level refill, one integrator attempt per loop iteration.
//
// Every launch renders a table of frame records in device memory (one
// record for a single tile: its row window; one per frame for a batch).  A
// wave works on one frame at a time, so the frame constants stay
// wave-uniform and are read with scalar loads from the record (they are not
// held in SGPRs across the whole kernel), refilling its lanes from that
// frame's work counter, and moves to the next frame with work left once all
// of its lanes are idle.  Waves start staggered over the frames, so the
// drain at a frame switch is rare.  A tile and a frame of a batch run the
// same kernel code.
// ---------------------------------------------------------------------------
// The frame record's address behind an empty asm: loads through it cannot
// be hoisted out of the attempt loop, so the constants the per-ray set-up and
// shading read (tetrad, camera, background) are fetched when used instead of
// occupying SGPRs for the whole frame.
__device__ __forceinline__ const FrameConst& cold(const BatchFrame* B) {
    const BatchFrame* p = B;
    asm volatile("" : "+s"(p));
    return p->F;
}

// TIME: also integrate the coordinate time and the affine parameter and
// write each sample's end t, phi and lambda to times (same indexing as the
// ray records).
template <class S, bool ADAPTIVE, bool INV, bool TIME, bool DEFER_KERNEL>
__device__ __forceinline__ void
kerr_body(const BatchFrame* __restrict__ frames, const int nframes, const int spp,
          float* __restrict__ out_base, const int accumulate,
          float* __restrict__ inv_max3, np_trace_counters* __restrict__ counters,
          np_ray_record* __restrict__ records_base, unsigned long long* __restrict__ work,
          const int chunk, np_ray_times* __restrict__ times) {
    const int wave_global = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    int fid = __builtin_amdgcn_readfirstlane(wave_global % nframes);
    typename S::Ray R;
    bool has = false, done = false;
    uint32_t* row = lane_row();
    st_u64(row, LS_PIX, 0);
    st_u64(row, LS_CEND, 0);
    row[LS_SAMPLE] = 0;
#pragma unroll
    for (int k = LS_N_RAYS; k <= LS_N_ATT; ++k) row[k] = 0;
    if (INV) {
        uint32_t* iv = inv_row();
        st_f(iv, IS_INVL + 0, 0.0f);
        st_f(iv, IS_INVL + 1, 0.0f);
        st_f(iv, IS_INVL + 2, 0.0f);
    }
    // Right-hand sides per lane (a register: it moves on every attempt),
    // flushed before it can wrap; FP64 and lazy attempts as wave totals
    // (scalar registers, from ballots).
    uint32_t c_rhs = 0;
    uint64_t w_att64 = 0, w_lazy = 0;
    uint32_t iter = 0;   // attempt iterations of this wave (scalar)
    uint32_t wait64 = 0; // iterations since the deferred FP64 block last ran (scalar)
    // kernels built with deferral (DEFER_KERNEL) defer in frames with
    // F.defer64 set; the others carry none of its bookkeeping
    constexpr bool DEFER = DEFER_KERNEL && S::kDefer && ADAPTIVE;
    if (spp <= 0) done = true;

    // Outer loop: one frame at a time; inner loop: one integrator attempt per
    // iteration, in which the frame record's address is invariant, so the
    // constants every attempt reads are loaded once per frame.
    while (true) {
        const BatchFrame* B = frames + fid;
        const FrameConst& F = B->F;
        const int W = F.width;
        const uint64_t npix = B->npix;
        if (F.max_steps <= 0) done = true;
        while (true) {
            const bool need = !has && !done;
            if (__ballot(need)) {
                const uint64_t pix = ld_u64(row, LS_PIX), chunk_end = ld_u64(row, LS_CEND);
                const bool in_chunk = need && (pix + 1 < chunk_end);
                const bool fetch = need && !in_chunk;
                const uint64_t unit = wave_fetch(fetch, work + fid);
                if (need) {
                    uint64_t p;
                    if (in_chunk) {
                        p = pix + 1;
                    } else {
                        p = unit * (uint64_t)chunk;
                        st_u64(row, LS_CEND, min(p + (uint64_t)chunk, npix));
                    }
                    if (p < npix) {
                        st_u64(row, LS_PIX, p);
                        row[LS_SAMPLE] = 0;
                        st_f(row, LS_ACC + 0, 0.0f);
                        st_f(row, LS_ACC + 1, 0.0f);
                        st_f(row, LS_ACC + 2, 0.0f);
                        const FrameConst& Fc = cold(B);
                        const uint64_t ap = B->pix0 + p;
                        const int x = (int)(ap % (uint64_t)W);
                        const int yy = (int)(ap / (uint64_t)W);
                        float n_r, n_th, n_ph, cx, cy;
                        camera_dir(Fc, x, yy, B->sample0, n_r, n_th, n_ph, cx, cy);
                        S::template start<INV>(Fc, R, n_r, n_th, n_ph, cy);
                        has = true;
                    } else {
                        done = true;
                    }
                }
            }
            if (__ballot(has) == 0) break;
            if (counters && __ballot(c_rhs > 0x40000000u)) {
                wave_add_u64(&counters->rhs_evals, c_rhs);
                c_rhs = 0;
            }
            int rhs64 = 0, lazy_att = 0;
            bool ended = false;
            ++iter;
            bool act = has;
            if constexpr (DEFER) {
                bool p64 = false;
                if (has) {
                    S::start_step(F, R);
                    p64 = S::pending64(R);
                }
                // F.defer64: lanes whose next attempt is FP64 sit out this
                // iteration unless kDeferLanes of them wait, kDeferIters
                // iterations have passed since the FP64 block last ran, no
                // lane has a float attempt to make, or the frame has no
                // pixels left for this wave (its last rays then finish at
                // full speed) -- a wave-uniform decision
                const uint64_t pend = __ballot(p64);
                if (F.defer64 && pend) {
                    ++wait64;
                    const bool run64 = __ballot(has && !p64) == 0 || __ballot(done) != 0 ||
                                       __popcll(pend) >= kDeferLanes || wait64 >= kDeferIters;
                    if (run64) wait64 = 0;
                    else act = has && !p64;
                }
            }
            if (act) {
                int rhs_count = 0;
                row[LS_ATT_RAY] += 1;
                int st;
                if constexpr (DEFER)   // start_step has run above
                    st = S::template attempt<ADAPTIVE, TIME, true>(F, R, rhs_count, rhs64, lazy_att);
                else
                    st = S::template attempt<ADAPTIVE, TIME>(F, R, rhs_count, rhs64, lazy_att);
                c_rhs += (uint32_t)(rhs_count + rhs64);
                int cls = -1;
                if (st == 1) S::post_accept(F, R, (iter % S::kRenormEvery) == 0);
                if (TIME && st == 1) S::advance(R);
                if (st == 2) cls = NP_RAY_CAPTURED;
                else if (st == 1) cls = after_accept<S, INV>(F, R, S::r(R), S::th(R), S::pr(R), S::pth(R));
                if (cls >= 0) {
                    ended = true;
                    const FrameConst& Fc = cold(B);
                    float col[3];
                    shade_final(Fc, R, cls, col);
                    const uint32_t orders = row[LS_ORDERS], att = row[LS_ATT_RAY];
                    if (INV) {
                        if (cls != NP_RAY_CAPTURED && !(orders > 0 && Fc.disk_on))
                            invariants_lane(Fc, S::r(R), S::th(R), S::pt(R), S::pr(R), S::pth(R), S::pph(R));
                        uint32_t* iv = inv_row();
#pragma unroll
                        for (int k = 0; k < 3; ++k)
                            st_f(iv, IS_INVL + k, fmaxf(ld_f(iv, IS_INVL + k), ld_f(iv, IS_INV + k)));
                    }
                    st_f(row, LS_ACC + 0, ld_f(row, LS_ACC + 0) + col[0]);
                    st_f(row, LS_ACC + 1, ld_f(row, LS_ACC + 1) + col[1]);
                    st_f(row, LS_ACC + 2, ld_f(row, LS_ACC + 2) + col[2]);
                    row[LS_N_RAYS] += 1;
                    row[LS_N_STEPS] += (uint32_t)R.step;
                    row[LS_N_CAP] += (cls == NP_RAY_CAPTURED);
                    row[LS_N_ESC] += (cls == NP_RAY_ESCAPED);
                    row[LS_N_STEPCAP] += (cls == NP_RAY_STEPCAP);
                    row[LS_N_DISK] += (orders > 0);
                    row[LS_N_ATT] += att;
                    const uint64_t pix = ld_u64(row, LS_PIX);
                    const uint32_t sample = row[LS_SAMPLE];
                    if (records_base) {
                        np_ray_record rec;
                        rec.cls = cls;
                        rec.orders = (int)orders;
                        rec.steps = R.step;
                        rec.attempts = (int)att;
                        rec.min_r = R.min_r;
                        rec.r_end = S::r(R);
                        records_base[B->rec_off + pix * (uint64_t)spp + (uint64_t)sample] = rec;
                    }
                    if (TIME) {
                        np_ray_times tm;
                        tm.t = S::t(R);
                        tm.phi = S::phi(R);
                        tm.lambda = S::lam(R);
                        times[B->rec_off + pix * (uint64_t)spp + (uint64_t)sample] = tm;
                    }
                    if ((int)sample + 1 < spp) {
                        row[LS_SAMPLE] = sample + 1;
                        const uint64_t ap = B->pix0 + pix;
                        const int x = (int)(ap % (uint64_t)W);
                        const int yy = (int)(ap / (uint64_t)W);
                        float n_r, n_th, n_ph, cx, cy;
                        camera_dir(Fc, x, yy, B->sample0 + sample + 1u, n_r, n_th, n_ph, cx, cy);
                        S::template start<INV>(Fc, R, n_r, n_th, n_ph, cy);
                    } else {
                        float* o = out_base + B->out_off + 3 * pix;
                        const float a0 = ld_f(row, LS_ACC + 0), a1 = ld_f(row, LS_ACC + 1),
                                    a2 = ld_f(row, LS_ACC + 2);
                        if (accumulate) {
                            o[0] += a0;
                            o[1] += a1;
                            o[2] += a2;
                        } else {
                            o[0] = a0;
                            o[1] = a1;
                            o[2] = a2;
                        }
                        has = false;
                    }
                }
            }
            w_att64 += (uint64_t)__popcll(__ballot(rhs64 > 0));
            w_lazy += (uint64_t)__popcll(__ballot(lazy_att != 0));
            // the lane-row totals grow only when a ray ends
            if (counters &&
                __ballot(ended && (row[LS_N_STEPS] > 0x40000000u || row[LS_N_ATT] > 0x40000000u))) {
                wave_add_u64(&counters->steps, row[LS_N_STEPS]);
                wave_add_u64(&counters->attempts, row[LS_N_ATT]);
                row[LS_N_STEPS] = 0;
                row[LS_N_ATT] = 0;
            }
        }
        // every lane is idle: move to the next frame that still has work
        int next = -1;
        for (int k = 1; k <= nframes; ++k) {
            const int f = (fid + k) % nframes;
            const unsigned long long taken =
                __hip_atomic_load(work + f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (taken * (unsigned long long)chunk < frames[f].npix) {
                next = f;
                break;
            }
        }
        if (next < 0) break;
        fid = __builtin_amdgcn_readfirstlane(next);
        done = (spp <= 0);
        st_u64(row, LS_PIX, 0);
        st_u64(row, LS_CEND, 0);
    }
    if (counters) {
        // FP64 attempts run 12 right-hand sides (4 with fixed steps)
        const uint64_t rhs_per64 = ADAPTIVE ? 12u : 4u;
        wave_add_u64(&counters->rays, row[LS_N_RAYS]);
        wave_add_u64(&counters->steps, row[LS_N_STEPS]);
        wave_add_u64(&counters->attempts, row[LS_N_ATT]);
        wave_add_u64(&counters->rhs_evals, c_rhs);
        wave_add_u64(&counters->captured, row[LS_N_CAP]);
        wave_add_u64(&counters->escaped, row[LS_N_ESC]);
        wave_add_u64(&counters->stepcap, row[LS_N_STEPCAP]);
        wave_add_u64(&counters->disk_hits, row[LS_N_DISK]);
        wave_add_uniform_u64(&counters->attempts_f64, w_att64);
        wave_add_uniform_u64(&counters->rhs_evals_f64, w_att64 * rhs_per64);
        wave_add_uniform_u64(&counters->attempts_lazy, w_lazy);
    }
    if (INV && inv_max3) {
        const uint32_t* iv = inv_row();
        atomic_max_nonneg(&inv_max3[0], ld_f(iv, IS_INVL + 0));
        atomic_max_nonneg(&inv_max3[1], ld_f(iv, IS_INVL + 1));
        atomic_max_nonneg(&inv_max3[2], ld_f(iv, IS_INVL + 2));
    }
}

#define NP_KERNEL_ARGS                                                                          \
    const BatchFrame* __restrict__ frames, const int nframes, const int spp,                    \
        float* __restrict__ out, const int accumulate, float* __restrict__ inv_max3,            \
        np_trace_counters* __restrict__ counters, np_ray_record* __restrict__ records,          \
        unsigned long long* __restrict__ work, const int chunk, np_ray_times* __restrict__ times
#define NP_KERNEL_PASS frames, nframes, spp, out, accumulate, inv_max3, counters, records, work, chunk, times

// Kernel entry points.  Variant 0 leaves the register budget to the
// compiler; variant W asks for W waves per SIMD, which caps the kernel at
// 512 / W VGPRs (rounded down to the allocation granule): FP32C W = 3, 4
// (168, 128), the FP64 comparator W = 3, the reference-formula FP64 path
// W = 2 (256, no AGPRs).
template <class S, bool ADAPTIVE, bool INV, bool DEFER = false>
__global__ void __launch_bounds__(256) kerr_render(NP_KERNEL_ARGS) {
    kerr_body<S, ADAPTIVE, INV, false, DEFER>(NP_KERNEL_PASS);
}

template <class S, bool ADAPTIVE, bool INV, int W, bool TIME = false, bool DEFER = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, W)))
kerr_render_capped(NP_KERNEL_ARGS) {
    kerr_body<S, ADAPTIVE, INV, TIME, DEFER>(NP_KERNEL_PASS);
}

#undef NP_KERNEL_ARGS
#undef NP_KERNEL_PASS

template <class S, bool A>
const void* kernel_ptr(bool inv) {
    return inv ? reinterpret_cast<const void*>(&kerr_render<S, A, true>)
               : reinterpret_cast<const void*>(&kerr_render<S, A, false>);
}

template <class S, bool A, int W>
const void* kernel_ptr_capped(bool inv) {
    return inv ? reinterpret_cast<const void*>(&kerr_render_capped<S, A, true, W>)
               : reinterpret_cast<const void*>(&kerr_render_capped<S, A, false, W>);
}

template <bool A>
const void* kernel_ptr_variant(int precision, bool inv, int variant, bool defer) {
    if constexpr (A) {
        // FP32C with deferred FP64 attempts (adaptive steps, no monitor)
        if (precision == NP_PREC_FP32C && defer && !inv) {
            if (variant == 4)
                return reinterpret_cast<const void*>(&kerr_render_capped<StepF32C, A, false, 4, false, true>);
            if (variant == 3)
                return reinterpret_cast<const void*>(&kerr_render_capped<StepF32C, A, false, 3, false, true>);
            return reinterpret_cast<const void*>(&kerr_render<StepF32C, A, false, true>);
        }
    }
    if (precision == NP_PREC_FP32C) {
        if (variant == 4) return kernel_ptr_capped<StepF32C, A, 4>(inv);
        if (variant == 3) return kernel_ptr_capped<StepF32C, A, 3>(inv);
        return kernel_ptr<StepF32C, A>(inv);
    }
    if (precision == NP_PREC_FP64_CARTER) {
        if (variant == 3) return kernel_ptr_capped<StepC64, A, 3>(inv);
        return kernel_ptr<StepC64, A>(inv);
    }
    if (variant == 2) return kernel_ptr_capped<StepF64, A, 2>(inv);
    return kernel_ptr<StepF64, A>(inv);
}

// The kernels that also integrate t and lambda (TIME): each precision's
// default variant, without the invariant monitor.
template <bool A>
const void* kernel_ptr_time(int precision) {
    if (precision == NP_PREC_FP32C)
        return reinterpret_cast<const void*>(&kerr_render_capped<StepF32C, A, false, 4, true>);
    if (precision == NP_PREC_FP64_CARTER)
        return reinterpret_cast<const void*>(&kerr_render_capped<StepC64, A, false, 3, true>);
    return reinterpret_cast<const void*>(&kerr_render_capped<StepF64, A, false, 2, true>);
}

// variant (FrameConst::variant): waves-per-SIMD cap, 0 = none.  time:
// the TIME kernel of the precision (nullptr with the invariant monitor).
// defer: the FP32C kernel that defers FP64 attempts in frames with
// FrameConst::defer64 set (adaptive, no monitor, no TIME).
const void* select_kernel(int precision, bool adaptive, bool inv, int variant, bool time, bool defer) {
    if (time) {
        if (inv) return nullptr;
        return adaptive ? kernel_ptr_time<true>(precision) : kernel_ptr_time<false>(precision);
    }
    return adaptive ? kernel_ptr_variant<true>(precision, inv, variant, defer)
                    : kernel_ptr_variant<false>(precision, inv, variant, defer);
}

__global__ void init_frame_record(const BatchFrame rec, BatchFrame* __restrict__ dst,
                                  unsigned long long* __restrict__ counter) {
    if (threadIdx.x == 0) {
        *dst = rec;
        *counter = 0ull;
    }
}

// The same for a pass whose first sample id lives in device memory: the
// record takes *next_sample, which then advances by spp (so a captured
// sequence of passes continues where the previous replay stopped).
__global__ void init_pass_record(const BatchFrame rec, BatchFrame* __restrict__ dst,
                                 unsigned long long* __restrict__ counter,
                                 uint32_t* __restrict__ next_sample, const uint32_t spp) {
    if (threadIdx.x == 0) {
        const uint32_t s0 = *next_sample;
        *dst = rec;
        dst->sample0 = s0;
        *counter = 0ull;
        *next_sample = s0 + spp;
    }
}

__global__ void set_u32(uint32_t* __restrict__ p, const uint32_t v) {
    if (threadIdx.x == 0) *p = v;
}

}  // namespace

// Dynamic LDS of a launch: the lane rows, and the monitor region when the
// invariant monitor is on.
size_t lane_lds_bytes(bool inv, int block) {
    return (size_t)block * sizeof(uint32_t) * (size_t)(kLaneSlots + (inv ? kInvSlots : 0));
}

int launch_kerr(const BatchFrame* d_frames, int nframes, int precision, bool adaptive, bool inv,
                int variant, int spp, float* d_out, int accumulate, float* d_inv,
                np_trace_counters* d_counters, np_ray_record* d_records,
                unsigned long long* d_work, hipStream_t stream, int grid, int block, int chunk,
                np_ray_times* d_times, bool defer) {
    if (block <= 0 || block > 256 || block % 64 != 0) return NP_EINVAL;
    const void* fn = select_kernel(precision, adaptive, inv, variant, d_times != nullptr, defer);
    if (!fn) return NP_EINVAL;
    void* args[] = {&d_frames, &nframes, &spp, &d_out, &accumulate, &d_inv, &d_counters, &d_records,
                    &d_work, &chunk, &d_times};
    const hipError_t e =
        hipLaunchKernel(fn, dim3(grid), dim3(block), args, lane_lds_bytes(inv, block), stream);
    return (e == hipSuccess && hipGetLastError() == hipSuccess) ? NP_OK : NP_ELAUNCH;
}

int occupancy_kerr(int precision, bool adaptive, bool inv, int variant, int block, int* blocks_per_cu,
                   bool time, bool defer) {
    const void* fn = select_kernel(precision, adaptive, inv, variant, time, defer);
    if (!fn) return NP_EINVAL;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, fn, block,
                                                                      lane_lds_bytes(inv, block));
    return e == hipSuccess ? NP_OK : NP_ELAUNCH;
}

int write_frame_record(const BatchFrame& rec, BatchFrame* d_dst, unsigned long long* d_counter,
                       hipStream_t stream) {
    hipLaunchKernelGGL(init_frame_record, dim3(1), dim3(64), 0, stream, rec, d_dst, d_counter);
    return hipGetLastError() == hipSuccess ? NP_OK : NP_ELAUNCH;
}

int write_pass_record(const BatchFrame& rec, BatchFrame* d_dst, unsigned long long* d_counter,
                      uint32_t* d_next_sample, int spp, hipStream_t stream) {
    hipLaunchKernelGGL(init_pass_record, dim3(1), dim3(64), 0, stream, rec, d_dst, d_counter, d_next_sample,
                       (uint32_t)spp);
    return hipGetLastError() == hipSuccess ? NP_OK : NP_ELAUNCH;
}

int set_device_u32(uint32_t* d_p, uint32_t v, hipStream_t stream) {
    hipLaunchKernelGGL(set_u32, dim3(1), dim3(64), 0, stream, d_p, v);
    return hipGetLastError() == hipSuccess ? NP_OK : NP_ELAUNCH;
}

}  // namespace np
