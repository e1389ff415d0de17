/*
 * kerr_oracle.c -- FP64 CPU restatement of the reference Kerr integrator,
 * cameras, Doppler factor and per-pixel renderer.
 *
 * TEST INFRASTRUCTURE ONLY (see np_oracle.h).  Built with
 * -ffp-contract=off so every multiply and add rounds separately, as the
 * reference's source text reads.  Differences from the reference at the
 * libm boundary: rsqrt/rsqrtf are computed as 1/sqrt (CUDA's are not
 * correctly rounded); sin, cos, pow, exp, tan come from glibc.
 */
#include <math.h>
#include <string.h>
#include <stdlib.h>
#include "np_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

#define BH_PI_F 3.14159265358979323846f
#define BH_PI_D 3.141592653589793238462643383279502884
#define BH_EPS_D 1e-12
#define BH_EPS_SIN_D 1e-12
#define BH_RK_SAFETY 0.9
#define BH_RK_HMIN 1e-6
#define BH_RK_HMAX 0.2
#define BH_INTEGRATION_STEP 0.01f

enum { T_ = 0, R_ = 1, TH_ = 2, PH_ = 3, PT_ = 4, PR_ = 5, PTH_ = 6, PPH_ = 7 };

/* BhKleinSum<double>::add / result (bh_common.h:451-483). */
typedef struct { double sum, cs, ccs; } klein_d;
static void klein_init(klein_d* k) { k->sum = 0.0; k->cs = 0.0; k->ccs = 0.0; }
static void klein_add(klein_d* k, double x) {
    double t = k->sum + x;
    double c;
    if (fabs(k->sum) >= fabs(x)) c = (k->sum - t) + x;
    else c = (x - t) + k->sum;
    k->sum = t;
    t = k->cs + c;
    double cc;
    if (fabs(k->cs) >= fabs(c)) cc = (k->cs - t) + c;
    else cc = (c - t) + k->cs;
    k->cs = t;
    k->ccs += cc;
}
static double klein_result(const klein_d* k) { return k->sum + k->cs + k->ccs; }

/* bh_rk4_sum (bh_common.h:485-493): k1 + 2k2 + 2k3 + k4, compensated. */
double orc_rk4_sum_d(double k1, double k2, double k3, double k4) {
    klein_d s; klein_init(&s);
    klein_add(&s, k1);
    klein_add(&s, 2.0 * k2);
    klein_add(&s, 2.0 * k3);
    klein_add(&s, k4);
    return klein_result(&s);
}

/* kerr_blocks_cov_d (bh_kerr.h:575-592). */
void orc_kerr_blocks_cov_d(double M, double a, double r, double th, double out[5]) {
    double ct = cos(th);
    double st = sin(th);
    double st2 = st * st;
    double r2 = r * r;
    double a2 = a * a;
    double Sigma = r2 + a2 * ct * ct;
    double Delta = r2 - 2.0 * M * r + a2;
    double A = (r2 + a2) * (r2 + a2) - a2 * Delta * st2;
    out[0] = -(1.0 - 2.0 * M * r / fmax(BH_EPS_D, Sigma));
    out[1] = -2.0 * M * a * r * st2 / fmax(BH_EPS_D, Sigma);
    out[2] = A * st2 / fmax(BH_EPS_D, Sigma);
    out[3] = Sigma / fmax(BH_EPS_D, Delta);
    out[4] = Sigma;
}

/* kerr_blocks_contra_d (bh_kerr.h:600-617). */
void orc_kerr_blocks_contra_d(double M, double a, double r, double th, double out[5]) {
    double ct = cos(th);
    double st = sin(th);
    double st2 = fmax(BH_EPS_SIN_D, st * st);
    double r2 = r * r;
    double a2 = a * a;
    double Sigma = fmax(BH_EPS_D, r2 + a2 * ct * ct);
    double Delta = r2 - 2.0 * M * r + a2;
    double A = (r2 + a2) * (r2 + a2) - a2 * Delta * st2;
    out[3] = fmax(BH_EPS_D, Delta) / Sigma;
    out[4] = 1.0 / Sigma;
    out[0] = -A / fmax(BH_EPS_D, Sigma * Delta);
    out[1] = -2.0 * M * a * r / fmax(BH_EPS_D, Sigma * Delta);
    out[2] = (Delta - a2 * st2) / fmax(BH_EPS_D, Sigma * Delta * st2);
}

/* kerr_blocks_contra_derivs_d (bh_kerr.h:624-672). */
void orc_kerr_blocks_contra_derivs_d(double M, double a, double r, double th, double o[10]) {
    double s = sin(th), c = cos(th);
    double s2 = (s * s > 1e-30 ? s * s : 1e-30);
    double r2 = r * r;
    double a2 = a * a;
    double B = r2 + a2;
    double Sigma = r2 + a2 * c * c;
    double Delta = r2 - 2.0 * M * r + a2;
    double A = (B * B) - a2 * Delta * s2;
    double D = Sigma * Delta;
    double D2 = D * D + 1e-48;
    double s2_th = 2.0 * s * c;

    double Sigma_r = 2.0 * r;
    double Sigma_th = -2.0 * a2 * s * c;
    double Delta_r = 2.0 * (r - M);
    double A_r = 4.0 * r * B - a2 * Delta_r * s2;
    double A_th = -a2 * Delta * s2_th;
    double D_r = Sigma_r * Delta + Sigma * Delta_r;
    double D_th = Sigma_th * Delta;

    o[0] = -(A_r * D - A * D_r) / D2;
    o[5] = -(A_th * D - A * D_th) / D2;

    double Nt = 2.0 * M * a * r;
    double Nt_r = 2.0 * M * a;
    o[1] = -(Nt_r * D - Nt * D_r) / D2;
    o[6] = -(0.0 * D - Nt * D_th) / D2;

    o[3] = (Delta_r * Sigma - Delta * Sigma_r) / (Sigma * Sigma);
    o[8] = -(Delta * Sigma_th) / (Sigma * Sigma);
    o[4] = -Sigma_r / (Sigma * Sigma);
    o[9] = -Sigma_th / (Sigma * Sigma);

    double N = Delta - a2 * s2;
    double N_r = Delta_r;
    double N_th = -a2 * s2_th;
    double Den = Delta * Sigma * s2 + 1e-48;
    double Den_r = Delta_r * Sigma * s2 + Delta * Sigma_r * s2;
    double Den_th = Delta * Sigma_th * s2 + Delta * Sigma * s2_th;
    double Den2 = Den * Den + 1e-48;
    o[2] = (N_r * Den - N * Den_r) / Den2;
    o[7] = (N_th * Den - N * Den_th) / Den2;
}

/* computeGeodesicDerivativesKerrD (bh_kerr.h:679-727). */
void orc_kerr_rhs_d(const double st[8], double M, double a, double d[8]) {
    double r = st[R_];
    double th = st[TH_];
    double disc = M * M - a * a;
    if (disc < 0.0) disc = 0.0;
    double r_plus = M + sqrt(disc);
    if (r <= r_plus + 1e-12 || r <= 1e-12) {
        memset(d, 0, 8 * sizeof(double));
        return;
    }
    double g[5], dg[10];
    orc_kerr_blocks_contra_d(M, a, r, th, g);
    orc_kerr_blocks_contra_derivs_d(M, a, r, th, dg);
    const double pt = st[PT_], pr = st[PR_], pth = st[PTH_], pph = st[PPH_];

    d[T_] = g[0] * pt + g[1] * pph;
    d[R_] = g[3] * pr;
    d[TH_] = g[4] * pth;
    d[PH_] = g[1] * pt + g[2] * pph;
    d[PT_] = 0.0;
    d[PPH_] = 0.0;

    klein_d sr; klein_init(&sr);
    klein_add(&sr, dg[0] * pt * pt);
    klein_add(&sr, 2.0 * dg[1] * pt * pph);
    klein_add(&sr, dg[2] * pph * pph);
    klein_add(&sr, dg[3] * pr * pr);
    klein_add(&sr, dg[4] * pth * pth);
    double S_r = klein_result(&sr);

    klein_d sth; klein_init(&sth);
    klein_add(&sth, dg[5] * pt * pt);
    klein_add(&sth, 2.0 * dg[6] * pt * pph);
    klein_add(&sth, dg[7] * pph * pph);
    klein_add(&sth, dg[8] * pr * pr);
    klein_add(&sth, dg[9] * pth * pth);
    double S_th = klein_result(&sth);

    d[PR_] = -0.5 * S_r;
    d[PTH_] = -0.5 * S_th;
}

/* H = 1/2 g^{mu nu} p_mu p_nu, plain sum as in render.cu:236-237. */
double orc_kerr_hamiltonian_d(const double s[8], double M, double a) {
    double g[5];
    orc_kerr_blocks_contra_d(M, a, s[R_], s[TH_], g);
    return 0.5 * (g[0] * s[PT_] * s[PT_] + 2.0 * g[1] * s[PT_] * s[PPH_] +
                  g[2] * s[PPH_] * s[PPH_] + g[3] * s[PR_] * s[PR_] +
                  g[4] * s[PTH_] * s[PTH_]);
}

/* Classic RK4 stages shared by integrateGeodesicKerrD (bh_kerr.h:734-774)
 * and rk4_step_d (bh_kerr.h:987-1032). */
static void rk4_core(const double st[8], double M, double a, double h, double out[8]) {
    double k1[8], k2[8], k3[8], k4[8], tmp[8];
    orc_kerr_rhs_d(st, M, a, k1);
    for (int i = 0; i < 8; ++i) tmp[i] = st[i] + 0.5 * h * k1[i];
    orc_kerr_rhs_d(tmp, M, a, k2);
    for (int i = 0; i < 8; ++i) tmp[i] = st[i] + 0.5 * h * k2[i];
    orc_kerr_rhs_d(tmp, M, a, k3);
    for (int i = 0; i < 8; ++i) tmp[i] = st[i] + h * k3[i];
    orc_kerr_rhs_d(tmp, M, a, k4);
    double h6 = h / 6.0;
    for (int i = 0; i < 8; ++i) out[i] = st[i] + h6 * orc_rk4_sum_d(k1[i], k2[i], k3[i], k4[i]);
}

static double r_plus_d(double M, double a) {
    double disc = M * M - a * a;
    if (disc < 0.0) disc = 0.0;
    return M + sqrt(disc);
}

int orc_kerr_rk4_fixed_d(double s[8], double M, double a, double h) {
    double out[8];
    rk4_core(s, M, a, h, out);
    memcpy(s, out, sizeof(out));
    double rp = r_plus_d(M, a);
    return (s[R_] > rp + 1e-9 && s[TH_] > 1e-9 && s[TH_] < BH_PI_D - 1e-9);
}

int orc_kerr_rk4_step_d(const double s[8], double M, double a, double h, double out[8]) {
    rk4_core(s, M, a, h, out);
    double rp = r_plus_d(M, a);
    return (out[R_] > rp + 1e-12 && out[TH_] > 1e-9 && out[TH_] < BH_PI_D - 1e-9);
}

static double err_norm(double x, double y, double scale) {
    return fabs(x - y) / fmax(scale, 1e-30);
}

/* rk4_adaptive_step_d (bh_kerr.h:1058-1105). */
double orc_kerr_adaptive_step_d(double s[8], double M, double a, double* h,
                                double rel_tol, double abs_tol) {
    double s_full[8], s_half1[8], s_half2[8];
    int ok1 = orc_kerr_rk4_step_d(s, M, a, *h, s_full);
    double h2 = 0.5 * (*h);
    int ok2 = orc_kerr_rk4_step_d(s, M, a, h2, s_half1);
    int ok3 = orc_kerr_rk4_step_d(s_half1, M, a, h2, s_half2);
    if (!(ok1 && ok2 && ok3)) {
        *h *= 0.5;
        return 0.0;
    }
    double sc_r = abs_tol + rel_tol * fmax(fabs(s_full[R_]), fabs(s_half2[R_]));
    double sc_th = abs_tol + rel_tol * fmax(fabs(s_full[TH_]), fabs(s_half2[TH_]));
    double sc_ph = abs_tol + rel_tol * fmax(fabs(s_full[PH_]), fabs(s_half2[PH_]));
    double sc_pr = abs_tol + rel_tol * fmax(fabs(s_full[PR_]), fabs(s_half2[PR_]));
    double sc_pth = abs_tol + rel_tol * fmax(fabs(s_full[PTH_]), fabs(s_half2[PTH_]));
    double sc_pph = abs_tol + rel_tol * fmax(fabs(s_full[PPH_]), fabs(s_half2[PPH_]));
    double e = fmax(
        fmax(err_norm(s_full[R_], s_half2[R_], sc_r), err_norm(s_full[TH_], s_half2[TH_], sc_th)),
        fmax(fmax(err_norm(s_full[PH_], s_half2[PH_], sc_ph), err_norm(s_full[PR_], s_half2[PR_], sc_pr)),
             fmax(err_norm(s_full[PTH_], s_half2[PTH_], sc_pth),
                  err_norm(s_full[PPH_], s_half2[PPH_], sc_pph))));
    if (e <= 1.0) {
        memcpy(s, s_half2, sizeof(s_half2));
        return fmin(2.0, BH_RK_SAFETY * pow(fmax(1e-12, e), -0.2));
    }
    double fac = fmax(0.1, BH_RK_SAFETY * pow(fmax(1e-12, e), -0.25));
    *h *= fac;
    return 0.0;
}

/* init_ray_from_camera_general_d (bh_kerr.h:781-826). */
void orc_init_static_d(double M, double a, double r_cam, double th_cam, double ph_cam,
                       double n_r, double n_th, double n_ph, double s[8]) {
    s[T_] = 0.0; s[R_] = r_cam; s[TH_] = th_cam; s[PH_] = ph_cam;
    double g[5];
    orc_kerr_blocks_cov_d(M, a, r_cam, th_cam, g);
    const double g_tt = g[0], g_tphi = g[1], g_phiphi = g[2], g_rr = g[3], g_thth = g[4];
    double ut = 1.0 / sqrt(fmax(BH_EPS_D, -g_tt));
    double er_r = 1.0 / sqrt(fmax(BH_EPS_D, g_rr));
    double eth_th = 1.0 / sqrt(fmax(BH_EPS_D, g_thth));
    double ephi_t, ephi_phi;
    if (fabs(g_tphi) > BH_EPS_D) {
        double beta = -g_tt / g_tphi;
        klein_d n2; klein_init(&n2);
        klein_add(&n2, g_tt);
        klein_add(&n2, 2.0 * g_tphi * beta);
        klein_add(&n2, g_phiphi * beta * beta);
        double norm2 = klein_result(&n2);
        double inv = 1.0 / sqrt(fmax(BH_EPS_D, fabs(norm2)));
        ephi_t = inv;
        ephi_phi = beta * inv;
    } else {
        ephi_t = 0.0;
        ephi_phi = 1.0 / sqrt(fmax(BH_EPS_D, g_phiphi));
    }
    double E_loc = 1.0;
    double pt_c = E_loc * (ut + n_ph * ephi_t);
    double pr_c = E_loc * (n_r * er_r);
    double pth_c = E_loc * (n_th * eth_th);
    double pph_c = E_loc * (n_ph * ephi_phi);
    s[PT_] = g_tt * pt_c + g_tphi * pph_c;
    s[PR_] = g_rr * pr_c;
    s[PTH_] = g_thth * pth_c;
    s[PPH_] = g_tphi * pt_c + g_phiphi * pph_c;
}

/* kerr_static_timelike (bh_kerr.h:880-884). */
int orc_static_timelike_d(double M, double a, double r, double th) {
    double g[5];
    orc_kerr_blocks_cov_d(M, a, r, th, g);
    return g[0] < 0.0;
}

/* init_ray_from_camera_zamo_d (bh_kerr.h:910-967). */
void orc_init_zamo_d(double M, double a, double r_cam, double th_cam, double ph_cam,
                     double n_r, double n_th, double n_ph, double s[8]) {
    s[T_] = 0.0; s[R_] = r_cam; s[TH_] = th_cam; s[PH_] = ph_cam;
    double ct = cos(th_cam), st = sin(th_cam);
    double r2 = r_cam * r_cam, a2 = a * a;
    double Sigma = fmax(BH_EPS_D, r2 + a2 * ct * ct);
    double Delta = r2 - 2.0 * M * r_cam + a2;
    double A = fmax(BH_EPS_D, (r2 + a2) * (r2 + a2) - a2 * Delta * st * st);
    double alpha = sqrt(fmax(BH_EPS_D, (Delta * Sigma) / A));
    double omega = (2.0 * M * a * r_cam) / A;
    double e_t_t = 1.0 / fmax(BH_EPS_D, alpha);
    double e_t_phi = omega / fmax(BH_EPS_D, alpha);
    double e_r_r = sqrt(fmax(BH_EPS_D, Delta / Sigma));
    double e_th_th = 1.0 / sqrt(Sigma);
    double e_phi_phi = sqrt(fmax(BH_EPS_D, Sigma / A)) / fmax(BH_EPS_D, fabs(st));
    double pt_c = e_t_t;
    double pr_c = n_r * e_r_r;
    double pth_c = n_th * e_th_th;
    double pph_c = e_t_phi + n_ph * e_phi_phi;
    double g[5];
    orc_kerr_blocks_cov_d(M, a, r_cam, th_cam, g);
    s[PT_] = g[0] * pt_c + g[1] * pph_c;
    s[PR_] = g[3] * pr_c;
    s[PTH_] = g[4] * pth_c;
    s[PPH_] = g[1] * pt_c + g[2] * pph_c;
}

/* doppler_g_general_d (bh_kerr.h:833-862). */
double orc_doppler_d(const double st[8], double M, double a) {
    double r = st[R_];
    double th = BH_PI_D * 0.5;
    double g[5];
    orc_kerr_blocks_cov_d(M, a, r, th, g);
    double M_safe = fmax(BH_EPS_D, M);
    double r_over_M = r / M_safe;
    double Omega = 1.0 / (a / M_safe + pow(fmax(1.0, r_over_M), 1.5)) / M_safe;
    klein_d ds; klein_init(&ds);
    klein_add(&ds, g[0]);
    klein_add(&ds, 2.0 * g[1] * Omega);
    klein_add(&ds, g[2] * Omega * Omega);
    double denom = -klein_result(&ds);
    double u_t = 1.0 / sqrt(denom > 1e-24 ? denom : 1e-24);
    double u_phi = Omega * u_t;
    double E_inf = -st[PT_];
    klein_d ks; klein_init(&ks);
    klein_add(&ks, st[PT_] * u_t);
    klein_add(&ks, st[PPH_] * u_phi);
    double k_u = -klein_result(&ks);
    double gg = E_inf / (k_u > 1e-18 ? k_u : 1e-18);
    return gg > 0.0 ? gg : 0.0;
}

/* ---------------------------------------------------------------------- */
/* Per-pixel renderer (render.cu:65-99, 124-363), FP64 geodesic path.       */
/* ---------------------------------------------------------------------- */

/* wanghash (render.cu:88-95). */
uint32_t orc_wanghash(uint32_t a) {
    a = (a ^ 61u) ^ (a >> 16);
    a *= 9u;
    a = a ^ (a >> 4);
    a *= 0x27d4eb2du;
    a = a ^ (a >> 15);
    return a;
}

/* rnd01 (render.cu:97-99). */
static float rnd01(uint32_t seed) {
    return (float)(orc_wanghash(seed) & 0x00ffffffu) * (1.0f / 16777216.0f);
}

/* Jitter keys of render.cu:145-146; a nonzero seed is folded into both
 * keys (an extension; seed 0 leaves the reference keys unchanged). */
void orc_camera_dir(const np_render_params* p, int x, int y, uint32_t sample_id,
                    float n[3], float cam[2]) {
    float aspect = (float)p->width / (float)p->height;
    float fov_y = p->fov_y_deg * BH_PI_F / 180.0f;
    float tan_half = tanf(0.5f * fov_y);
    uint32_t salt = p->seed ? orc_wanghash(p->seed) : 0u;
    uint32_t ux = (uint32_t)x, uy = (uint32_t)y;
    float jx = rnd01((ux * 73856093u ^ uy * 19349663u ^ sample_id * 83492791u) ^ salt) - 0.5f;
    float jy = rnd01((ux * 83492791u ^ uy * 2654435761u ^ sample_id * 19349663u) ^ salt) - 0.5f;
    float ndc_x = (((float)x + 0.5f + jx) / (float)p->width) * 2.0f - 1.0f;
    float ndc_y = (((float)y + 0.5f + jy) / (float)p->height) * 2.0f - 1.0f;
    float cam_x = ndc_x * aspect * tan_half;
    float cam_y = -ndc_y * tan_half;
    float cam_z = -1.0f;
    float norm = 1.0f / sqrtf(cam_x * cam_x + cam_y * cam_y + cam_z * cam_z);
    cam_x *= norm; cam_y *= norm; cam_z *= norm;
    n[0] = cam_z; n[1] = cam_y; n[2] = cam_x;
    cam[0] = cam_x; cam[1] = cam_y;
}

/* shade_background (render.cu:74-86). */
static void shade_background(float u, float v, float c[3]) {
    (void)u;
    float t = fminf(fmaxf(v, 0.f), 1.f);
    c[0] = 0.02f * (1 - t) + 0.08f * t;
    c[1] = 0.03f * (1 - t) + 0.09f * t;
    c[2] = 0.05f * (1 - t) + 0.12f * t;
    float band = expf(-50.0f * (v - 0.5f) * (v - 0.5f));
    c[0] += 0.4f * band; c[1] += 0.35f * band; c[2] += 0.3f * band;
    for (int i = 0; i < 3; ++i) c[i] = fminf(fmaxf(c[i], 0.f), 1.f);
}

static void inv_update(float inv[3], const double st[8], double M, double a,
                       double pt0, double pph0) {
    double Hc = orc_kerr_hamiltonian_d(st, M, a);
    float Habs = (float)fabs(Hc);
    float Erel = (float)(fabs(st[PT_] - pt0) / fmax(1e-12, fabs(pt0)));
    float Lrel = (float)(fabs(st[PPH_] - pph0) / fmax(1e-12, fabs(pph0)));
    inv[0] = fmaxf(inv[0], Habs);
    inv[1] = fmaxf(inv[1], Erel);
    inv[2] = fmaxf(inv[2], Lrel);
}

/* Crossing radius from the cubic Hermite interpolants of theta(lambda) and
 * r(lambda) over an accepted step of size h (an extension of this library,
 * np_crossing HERMITE): Newton on theta(s) = pi/2, s = lambda/h in [0, 1],
 * from the linear estimate s0, four iterations; the linear radius is kept
 * when an iterate leaves [0, 1] or the slope vanishes. */
double orc_hermite_crossing(double h, double r0, double th0, double dr0, double dth0, double r1,
                            double th1, double dr1, double dth1, double s0, double r_lin) {
    const double m0 = h * dth0, m1 = h * dth1, n0 = h * dr0, n1 = h * dr1;
    const double g0 = th0 - BH_PI_D * 0.5, g1 = th1 - BH_PI_D * 0.5;
    double s = s0;
    for (int it = 0; it < 4; ++it) {
        double s2 = s * s, s3 = s2 * s;
        double f = (2 * s3 - 3 * s2 + 1) * g0 + (s3 - 2 * s2 + s) * m0 + (3 * s2 - 2 * s3) * g1 +
                   (s3 - s2) * m1;
        double fp = (6 * s2 - 6 * s) * g0 + (3 * s2 - 4 * s + 1) * m0 + (6 * s - 6 * s2) * g1 +
                    (3 * s2 - 2 * s) * m1;
        if (!(fabs(fp) > 0.0)) return r_lin;
        s -= f / fp;
        if (!(s >= 0.0 && s <= 1.0)) return r_lin;
    }
    double s2 = s * s, s3 = s2 * s;
    return (2 * s3 - 3 * s2 + 1) * r0 + (s3 - 2 * s2 + s) * n0 + (3 * s2 - 2 * s3) * r1 + (s3 - s2) * n1;
}

/* Initial state of one sample (render.cu:143-169): camera direction, then
 * the static or ZAMO tetrad (ZAMO forced where the static observer is not
 * timelike). */
static void ray_start(const np_render_params* p, int x, int y, uint32_t sample_id, double state[8],
                      float cam[2]) {
    const float Rs = 2.0f * p->mass;
    const float Mf = p->mass;
    const float af = p->spin * Mf;
    const float r_cam = p->r_cam_Rs * Rs;
    const double M = (double)Mf, a = (double)af;
    float n[3];
    orc_camera_dir(p, x, y, sample_id, n, cam);
    int static_ok = orc_static_timelike_d(M, a, (double)r_cam, (double)p->theta_cam);
    if (p->camera == NP_CAMERA_ZAMO || !static_ok)
        orc_init_zamo_d(M, a, r_cam, p->theta_cam, p->phi_cam, n[0], n[1], n[2], state);
    else
        orc_init_static_d(M, a, r_cam, p->theta_cam, p->phi_cam, n[0], n[1], n[2], state);
}

void orc_ray_p0_pixels(const np_render_params* p, int npix, const int32_t* xs, const int32_t* ys,
                       int sample_offset, int spp, double* p0) {
    for (int i = 0; i < npix; ++i)
        for (int s = 0; s < spp; ++s) {
            double st[8];
            float cam[2];
            ray_start(p, xs[i], ys[i], (uint32_t)(sample_offset + s), st, cam);
            double* o = p0 + 4 * ((size_t)i * spp + s);
            o[0] = st[PT_]; o[1] = st[PR_]; o[2] = st[PTH_]; o[3] = st[PPH_];
        }
}

void orc_render_ray(const np_render_params* p, int x, int y, uint32_t sample_id,
                    orc_ray_info* info) {
    const float Rs = 2.0f * p->mass;
    const float Mf = p->mass;
    const float af = p->spin * Mf;
    const float r_cam = p->r_cam_Rs * Rs;
    const double M = (double)Mf, a = (double)af;
    float cam[2];
    double state[8];
    ray_start(p, x, y, sample_id, state, cam);

    const double pt0 = state[PT_], pph0 = state[PPH_];
    info->p0[0] = state[PT_]; info->p0[1] = state[PR_]; info->p0[2] = state[PTH_]; info->p0[3] = state[PPH_];
    info->ncross = 0;
    double min_r = state[R_];
    int captured = 0, orders = 0;
    float disk_rgb[3] = {0, 0, 0};
    double prev[8];
    memcpy(prev, state, sizeof(prev));
    const double rin = p->disk_rin, rout = p->disk_rout;
    const double r_far = (double)r_cam + 10.0 * (double)Rs;
    double h_adapt = 0.0;
    int64_t steps = 0, attempts = 0, rhs = 0;
    int reached_far = 0;
    float inv[3] = {0, 0, 0};
    int step;
    double h_acc = 0.0;   /* size of the last accepted step */
    double lam = 0.0;     /* affine parameter: sum of the accepted steps */
    for (step = 0; step < p->max_steps; ++step) {
        double scale = fmin(fmax(1.0, state[R_] / (0.5 * (double)Rs)), 100.0);
        int ok = 0;
        if (p->integrator == NP_INTEGRATOR_ADAPTIVE) {
            double h_base = (double)BH_INTEGRATION_STEP * scale;
            h_base = fmin(fmax(h_base, BH_RK_HMIN), BH_RK_HMAX);
            double h = (h_adapt > 0.0) ? h_adapt : h_base;
            h = fmin(fmax(h, BH_RK_HMIN), BH_RK_HMAX);
            for (int attempt = 0; attempt < 8; ++attempt) {
                double h_try = h;
                ++attempts; rhs += 12;
                double fac = orc_kerr_adaptive_step_d(state, M, a, &h_try, p->rtol, p->atol);
                if (fac > 0.0) {
                    ok = 1;
                    h_acc = h;
                    h_adapt = fmin(fmax(h_try * fac, BH_RK_HMIN), BH_RK_HMAX);
                    break;
                }
                h = fmax(h_try, BH_RK_HMIN);
                if (h <= BH_RK_HMIN * 1.01) break;
            }
        } else {
            ++attempts; rhs += 4;
            h_acc = (double)BH_INTEGRATION_STEP * scale;
            ok = orc_kerr_rk4_fixed_d(state, M, a, h_acc);
        }
        if (!ok) { captured = 1; break; }
        ++steps;
        lam += h_acc;
        min_r = fmin(min_r, state[R_]);
        double d_prev = prev[TH_] - BH_PI_D * 0.5;
        double d_curr = state[TH_] - BH_PI_D * 0.5;
        {
            double eps = 1e-8;
            int sign_change = (d_prev > 0 && d_curr < 0) || (d_prev < 0 && d_curr > 0);
            int near_zero = (fabs(d_prev) < eps && fabs(d_curr) >= eps) ||
                            (fabs(d_curr) < eps && fabs(d_prev) >= eps);
            if ((sign_change || near_zero) && orders < p->disk_max_orders) {
                double denom = fabs(d_prev) + fabs(d_curr) + 1e-30;
                double alpha = fabs(d_prev) / denom;
                double r_cross = prev[R_] + alpha * (state[R_] - prev[R_]);
                if (p->crossing == NP_CROSSING_HERMITE && sign_change) {
                    double d0[8], d1[8];
                    orc_kerr_rhs_d(prev, M, a, d0);
                    orc_kerr_rhs_d(state, M, a, d1);
                    r_cross = orc_hermite_crossing(h_acc, prev[R_], prev[TH_], d0[R_], d0[TH_], state[R_],
                                                   state[TH_], d1[R_], d1[TH_], alpha, r_cross);
                }
                if (info->ncross < 4) info->xr[info->ncross++] = r_cross;
                if (r_cross >= rin && r_cross <= rout) {
                    double cross[8];
                    memcpy(cross, state, sizeof(cross));
                    cross[R_] = r_cross;
                    cross[TH_] = BH_PI_D * 0.5;
                    cross[PT_] = prev[PT_] + alpha * (state[PT_] - prev[PT_]);
                    cross[PPH_] = prev[PPH_] + alpha * (state[PPH_] - prev[PPH_]);
                    float g_boost = (float)orc_doppler_d(cross, M, a);
                    /* shade_disk (render.cu:65-72) with the emissivity extension */
                    float rr = p->emiss_m_units ? (float)(r_cross / M) : (float)r_cross;
                    float j = powf(fmaxf(rr, 1.0f), -p->emiss_q);
                    float I = j * g_boost * g_boost * g_boost;
                    disk_rgb[0] += (1.0f * I) * p->disk_gain;
                    disk_rgb[1] += (0.75f * I) * p->disk_gain;
                    disk_rgb[2] += (0.45f * I) * p->disk_gain;
                    orders++;
                    if (p->check_inv) inv_update(inv, cross, M, a, pt0, pph0);
                }
            }
        }
        memcpy(prev, state, sizeof(prev));
        if (state[R_] > r_far && step > 10) { reached_far = 1; break; }
    }

    float color[3];
    if (captured) {
        if (p->bg_mode == NP_BG_POINT) {
            float r_ph = 1.5f * Rs;
            float width = fmaxf(0.02f * Rs, 0.005f * Rs + 0.02f * p->bg_sigma_rad * Rs);
            float dr = ((float)min_r - r_ph);
            float w = expf(-0.5f * (dr / width) * (dr / width));
            float I = p->bg_peak * w + 0.0f;
            color[0] = color[1] = color[2] = I;
        } else {
            color[0] = color[1] = color[2] = 0.0f;
        }
    } else if (orders > 0 && p->disk_on) {
        memcpy(color, disk_rgb, sizeof(color));
    } else {
        if (p->check_inv) inv_update(inv, state, M, a, pt0, pph0);
        if (p->bg_mode == NP_BG_POINT) {
            float r_ph = 1.5f * Rs;
            float width = fmaxf(0.02f * Rs, 0.005f * Rs + 0.02f * p->bg_sigma_rad * Rs);
            float dr = ((float)min_r - r_ph);
            float w = expf(-0.5f * (dr / width) * (dr / width));
            float I = p->bg_peak * w + 0.02f;
            color[0] = color[1] = color[2] = I;
        } else {
            shade_background(cam[0] * 0.5f + 0.5f, cam[1] * 0.5f + 0.5f, color);
        }
    }
    info->cls = captured ? NP_RAY_CAPTURED : (reached_far ? NP_RAY_ESCAPED : NP_RAY_STEPCAP);
    info->orders = orders;
    info->steps = steps;
    info->attempts = attempts;
    info->rhs_evals = rhs;
    info->min_r = min_r;
    info->t_end = state[T_];
    info->phi_end = state[PH_];
    info->lam_end = lam;
    info->r_end = state[R_];
    info->th_end = state[TH_];
    memcpy(info->inv, inv, sizeof(inv));
    memcpy(info->rgb, color, sizeof(color));
}

void orc_render_pixels(const np_render_params* p, int npix, const int32_t* xs,
                       const int32_t* ys, int sample_offset, int spp, float* rgb,
                       orc_ray_info* info, int nthreads) {
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int i = 0; i < npix; ++i) {
        float acc[3] = {0.0f, 0.0f, 0.0f};
        for (int s = 0; s < spp; ++s) {
            orc_ray_info ri;
            orc_render_ray(p, xs[i], ys[i], (uint32_t)(sample_offset + s), &ri);
            acc[0] += ri.rgb[0]; acc[1] += ri.rgb[1]; acc[2] += ri.rgb[2];
            if (info) info[(size_t)i * spp + s] = ri;
        }
        rgb[3 * (size_t)i + 0] = acc[0];
        rgb[3 * (size_t)i + 1] = acc[1];
        rgb[3 * (size_t)i + 2] = acc[2];
    }
    (void)nthreads;
}
