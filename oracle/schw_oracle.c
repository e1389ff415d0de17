/*
 * schw_oracle.c -- CPU restatement of the reference Schwarzschild equatorial
 * sweep (include/bh_schwarzschild.h, schwarzschild_blackhole.cu) in its
 * original FP32 form and as the FP64 "plain method" sweep.
 *
 * TEST INFRASTRUCTURE ONLY (see np_oracle.h).
 */
#include <math.h>
#include <string.h>
#include "np_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_F 3.14159265358979323846f
#define BH_EPS_F 1e-6f
#define BH_EPS_SIN_F 1e-6f

typedef struct { float t, r, theta, phi, pt, pr, ptheta, pphi; } gs_f;

/* computeGeodesicDerivatives (bh_schwarzschild.h:124-188), FP32. */
static void schw_rhs_f(const gs_f* s, float Rs, gs_f* d) {
    float r = s->r, theta = s->theta;
    if (r <= Rs + 0.001f || r <= 0.001f) { memset(d, 0, sizeof(*d)); return; }
    float sin_theta = sinf(theta);
    float cos_theta = cosf(theta);
    float r2 = r * r;
    float r3 = r2 * r;
    float A = fmaxf(BH_EPS_F, 1.0f - Rs / r);
    d->t = (-1.0f / A) * s->pt;
    d->r = A * s->pr;
    d->theta = s->ptheta / r2;
    float sin2 = fmaxf(BH_EPS_SIN_F, sin_theta * sin_theta);
    d->phi = s->pphi / (r2 * sin2);
    d->pt = 0.0f;
    float dgtt = Rs / (r2 * A * A);
    float dgrr = Rs / r2;
    float dgthth = -2.0f / r3;
    float dgphph = -2.0f / (r3 * sin2);
    d->pr = -0.5f * (dgtt * s->pt * s->pt + dgrr * s->pr * s->pr +
                     dgthth * s->ptheta * s->ptheta + dgphph * s->pphi * s->pphi);
    float sin_safe = (fabsf(sin_theta) < 1e-6f) ? (sin_theta >= 0.0f ? 1e-6f : -1e-6f) : sin_theta;
    d->ptheta = (cos_theta / (r2 * sin2 * sin_safe)) * s->pphi * s->pphi;
    d->pphi = 0.0f;
}

#define AXPY_F(dst, src, k, c) do { \
    (dst).t = (src).t + (c) * (k).t; (dst).r = (src).r + (c) * (k).r; \
    (dst).theta = (src).theta + (c) * (k).theta; (dst).phi = (src).phi + (c) * (k).phi; \
    (dst).pt = (src).pt + (c) * (k).pt; (dst).pr = (src).pr + (c) * (k).pr; \
    (dst).ptheta = (src).ptheta + (c) * (k).ptheta; (dst).pphi = (src).pphi + (c) * (k).pphi; } while (0)

/* integrateGeodesic (bh_schwarzschild.h:235-295), FP32. */
static int schw_rk4_f(gs_f* s, float Rs, float h) {
    gs_f k1, k2, k3, k4, tmp;
    schw_rhs_f(s, Rs, &k1);
    float hh = 0.5f * h;
    tmp.t = s->t + hh * k1.t; tmp.r = s->r + hh * k1.r; tmp.theta = s->theta + hh * k1.theta;
    tmp.phi = s->phi + hh * k1.phi; tmp.pt = s->pt + hh * k1.pt; tmp.pr = s->pr + hh * k1.pr;
    tmp.ptheta = s->ptheta + hh * k1.ptheta; tmp.pphi = s->pphi + hh * k1.pphi;
    schw_rhs_f(&tmp, Rs, &k2);
    tmp.t = s->t + hh * k2.t; tmp.r = s->r + hh * k2.r; tmp.theta = s->theta + hh * k2.theta;
    tmp.phi = s->phi + hh * k2.phi; tmp.pt = s->pt + hh * k2.pt; tmp.pr = s->pr + hh * k2.pr;
    tmp.ptheta = s->ptheta + hh * k2.ptheta; tmp.pphi = s->pphi + hh * k2.pphi;
    schw_rhs_f(&tmp, Rs, &k3);
    AXPY_F(tmp, *s, k3, h);
    schw_rhs_f(&tmp, Rs, &k4);
    float h6 = h / 6.0f;
    s->t += h6 * (k1.t + 2.0f * k2.t + 2.0f * k3.t + k4.t);
    s->r += h6 * (k1.r + 2.0f * k2.r + 2.0f * k3.r + k4.r);
    s->theta += h6 * (k1.theta + 2.0f * k2.theta + 2.0f * k3.theta + k4.theta);
    s->phi += h6 * (k1.phi + 2.0f * k2.phi + 2.0f * k3.phi + k4.phi);
    s->pt += h6 * (k1.pt + 2.0f * k2.pt + 2.0f * k3.pt + k4.pt);
    s->pr += h6 * (k1.pr + 2.0f * k2.pr + 2.0f * k3.pr + k4.pr);
    s->ptheta += h6 * (k1.ptheta + 2.0f * k2.ptheta + 2.0f * k3.ptheta + k4.ptheta);
    s->pphi += h6 * (k1.pphi + 2.0f * k2.pphi + 2.0f * k3.pphi + k4.pphi);
    return (s->r > Rs + 0.001f && s->theta > 0.001f && s->theta < PI_F - 0.001f);
}

/* schwarzschild_rmin_analytic (schwarzschild_blackhole.cu:24-37). */
int orc_schw_rmin_analytic(float Rs, float b, float* rmin_out) {
    if (Rs <= 0.0f || b <= 0.0f || rmin_out == 0) return 0;
    const double beta = (double)b / (double)Rs;
    const double beta_crit = 0.5 * 3.0 * sqrt(3.0);
    if (beta <= beta_crit) return 0;
    const double p = -beta * beta;
    const double A = 2.0 * sqrt(-p / 3.0);
    double cosarg = -(3.0 * sqrt(3.0)) / (2.0 * beta);
    if (cosarg < -1.0) cosarg = -1.0;
    if (cosarg > 1.0) cosarg = 1.0;
    const double ang = (1.0 / 3.0) * acos(cosarg);
    const double u = A * cos(ang);
    *rmin_out = (float)(u * (double)Rs);
    return isfinite(*rmin_out) && *rmin_out > Rs;
}

/* precomputeGeodesicsKernel body (schwarzschild_blackhole.cu:48-139). */
static void schw_one_f(float mass, float r_start, float b, float* defl, float* rmin, int32_t* status) {
    const float Rs = 2.0f * mass;
    gs_f s;
    s.r = r_start; s.theta = PI_F / 2.0f; s.phi = 0.0f; s.t = 0.0f;
    float energy = 1.0f;
    s.pt = -energy; s.pphi = b * energy; s.ptheta = 0.0f;
    float r = s.r;
    float A = 1.0f - Rs / r;
    float L = s.pphi;
    float pr2 = (energy * energy) / (A * A) - (L * L) / (A * r * r);
    s.pr = -sqrtf(fmaxf(0.0f, pr2));
    float min_r = s.r, total = 0.0f, prev_phi = s.phi;
    int escaped = 0;
    float prev_r = s.r, prev_pr = s.pr;
    for (int step = 0; step < 10000; step++) {
        float scale = fminf(100.0f, fmaxf(1.0f, s.r / (0.5f * Rs)));
        float h = 0.01f * scale;
        if (!schw_rk4_f(&s, Rs, h)) break;
        min_r = fminf(min_r, s.r);
        if (prev_pr < 0.0f && s.pr > 0.0f) {
            float denom = (s.pr - prev_pr);
            if (fabsf(denom) > 1e-12f) {
                float alpha = (-prev_pr) / denom;
                alpha = fminf(fmaxf(alpha, 0.0f), 1.0f);
                float r_turn = prev_r + alpha * (s.r - prev_r);
                min_r = fminf(min_r, r_turn);
            }
        }
        prev_r = s.r; prev_pr = s.pr;
        float dphi = s.phi - prev_phi;
        if (dphi > PI_F) dphi -= 2.0f * PI_F;
        if (dphi < -PI_F) dphi += 2.0f * PI_F;
        total += dphi;
        prev_phi = s.phi;
        if (s.r > r_start + 1.0f && s.pr > 0) { escaped = 1; break; }
    }
    *defl = total;
    float rmin_out = min_r;
    if (escaped) {
        float r_an = 0.0f;
        if (orc_schw_rmin_analytic(Rs, b, &r_an)) rmin_out = r_an;
    }
    *rmin = rmin_out;
    if (escaped || (s.r > r_start + 1.0f && s.pr > 0)) *status = 0;
    else if (s.r <= Rs + 0.01f) *status = 1;
    else if (fabsf(min_r - 1.5f * Rs) < 0.01f * Rs) *status = 2;
    else *status = 3;
}

void orc_schw_precompute_f32(float mass, float r_start, const float* b, int n,
                             float* defl, float* rmin, int32_t* status, int nthreads) {
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
#endif
    for (int i = 0; i < n; ++i) schw_one_f(mass, r_start, b[i], &defl[i], &rmin[i], &status[i]);
    (void)nthreads;
}

/* ---- FP64 sweep: the same right-hand side restated in double ---- */
typedef struct { double t, r, theta, phi, pt, pr, ptheta, pphi; } gs_d;

static void schw_rhs_d(const gs_d* s, double Rs, gs_d* d) {
    double r = s->r, theta = s->theta;
    if (r <= Rs + 0.001 || r <= 0.001) { memset(d, 0, sizeof(*d)); return; }
    double sin_theta = sin(theta);
    double cos_theta = cos(theta);
    double r2 = r * r;
    double r3 = r2 * r;
    double A = fmax(1e-6, 1.0 - Rs / r);
    d->t = (-1.0 / A) * s->pt;
    d->r = A * s->pr;
    d->theta = s->ptheta / r2;
    double sin2 = fmax(1e-6, sin_theta * sin_theta);
    d->phi = s->pphi / (r2 * sin2);
    d->pt = 0.0;
    double dgtt = Rs / (r2 * A * A);
    double dgrr = Rs / r2;
    double dgthth = -2.0 / r3;
    double dgphph = -2.0 / (r3 * sin2);
    d->pr = -0.5 * (dgtt * s->pt * s->pt + dgrr * s->pr * s->pr +
                    dgthth * s->ptheta * s->ptheta + dgphph * s->pphi * s->pphi);
    double sin_safe = (fabs(sin_theta) < 1e-6) ? (sin_theta >= 0.0 ? 1e-6 : -1e-6) : sin_theta;
    d->ptheta = (cos_theta / (r2 * sin2 * sin_safe)) * s->pphi * s->pphi;
    d->pphi = 0.0;
}

static int schw_rk4_d(gs_d* s, double Rs, double h) {
    gs_d k1, k2, k3, k4, tmp;
    schw_rhs_d(s, Rs, &k1);
    double hh = 0.5 * h;
    AXPY_F(tmp, *s, k1, hh);
    schw_rhs_d(&tmp, Rs, &k2);
    AXPY_F(tmp, *s, k2, hh);
    schw_rhs_d(&tmp, Rs, &k3);
    AXPY_F(tmp, *s, k3, h);
    schw_rhs_d(&tmp, Rs, &k4);
    double h6 = h / 6.0;
    s->t += h6 * (k1.t + 2.0 * k2.t + 2.0 * k3.t + k4.t);
    s->r += h6 * (k1.r + 2.0 * k2.r + 2.0 * k3.r + k4.r);
    s->theta += h6 * (k1.theta + 2.0 * k2.theta + 2.0 * k3.theta + k4.theta);
    s->phi += h6 * (k1.phi + 2.0 * k2.phi + 2.0 * k3.phi + k4.phi);
    s->pt += h6 * (k1.pt + 2.0 * k2.pt + 2.0 * k3.pt + k4.pt);
    s->pr += h6 * (k1.pr + 2.0 * k2.pr + 2.0 * k3.pr + k4.pr);
    s->ptheta += h6 * (k1.ptheta + 2.0 * k2.ptheta + 2.0 * k3.ptheta + k4.ptheta);
    s->pphi += h6 * (k1.pphi + 2.0 * k2.pphi + 2.0 * k3.pphi + k4.pphi);
    return (s->r > Rs + 0.001 && s->theta > 0.001 && s->theta < M_PI - 0.001);
}

static void schw_sweep_one(double M, double r0, double r_cap, double ps_band, int64_t max_steps,
                           double b, int32_t* cls, double* dphi_out, int64_t* steps_out) {
    const double Rs = 2.0 * M;
    gs_d s;
    s.t = 0.0; s.r = r0; s.theta = M_PI / 2.0; s.phi = 0.0;
    s.pt = -1.0; s.pphi = b; s.ptheta = 0.0;
    double A = 1.0 - Rs / r0;
    double pr2 = 1.0 / (A * A) - (b * b) / (A * r0 * r0);
    s.pr = -sqrt(fmax(0.0, pr2));
    double total = 0.0, prev_phi = s.phi;
    int status = 3;
    int64_t step;
    for (step = 0; step < max_steps; ++step) {
        double scale = fmin(100.0, fmax(1.0, s.r / (0.5 * Rs)));
        double h = 0.01 * scale;
        int ok = schw_rk4_d(&s, Rs, h);
        double dphi = s.phi - prev_phi;
        if (dphi > M_PI) dphi -= 2.0 * M_PI;
        if (dphi < -M_PI) dphi += 2.0 * M_PI;
        total += dphi;
        prev_phi = s.phi;
        if (!ok || s.r < r_cap) { status = 1; ++step; break; }
        if (s.r > r0 && s.pr > 0.0) { status = 0; ++step; break; }
    }
    if (fabs(b - 3.0 * sqrt(3.0) * M) < ps_band) status = 2;
    *cls = status;
    *dphi_out = total;
    *steps_out = step;
}

void orc_schw_sweep_d(double M, double r0, double r_capture, double ps_band,
                      int64_t max_steps, const double* b, int n, int32_t* cls,
                      double* dphi, int64_t* steps, int nthreads) {
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#endif
    for (int i = 0; i < n; ++i)
        schw_sweep_one(M, r0, r_capture, ps_band, max_steps, b[i], &cls[i], &dphi[i], &steps[i]);
    (void)nthreads;
}
