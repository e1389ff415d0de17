/*
 * np_oracle.h -- CPU oracle for the null-geodesic hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C FP64 restatement of the
 * reference algorithms, used by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg as the checker.  The product library
 * (libnullpath.so) never links, loads or calls it.
 *
 * Each function cites the reference file:line it restates.  How each one is
 * pinned (golden vectors, closed forms, invariants) is listed in the test
 * files that exercise it.
 */
#ifndef NP_ORACLE_H
#define NP_ORACLE_H

#include <stdint.h>
#include "../include/np_render.h"
#include "../include/nullpath.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Kerr metric blocks, FP64 (include/bh_kerr.h:575-672) ---- */
/* out = {g_tt, g_tphi, g_phiphi, g_rr, g_thth} */
void orc_kerr_blocks_cov_d(double M, double a, double r, double th, double out[5]);
/* out = {g^tt, g^tphi, g^phiphi, g^rr, g^thth} */
void orc_kerr_blocks_contra_d(double M, double a, double r, double th, double out[5]);
/* out = {d_r g^tt, d_r g^tphi, d_r g^phiphi, d_r g^rr, d_r g^thth,
 *        d_th g^tt, ..., d_th g^thth} */
void orc_kerr_blocks_contra_derivs_d(double M, double a, double r, double th, double out[10]);

/* State vectors are {t, r, theta, phi, p_t, p_r, p_theta, p_phi}
 * (GeodesicStateD member order, bh_kerr.h:553-556). */
void orc_kerr_rhs_d(const double s[8], double M, double a, double d[8]);
double orc_kerr_hamiltonian_d(const double s[8], double M, double a);
/* integrateGeodesicKerrD: in-place fixed step, returns continue flag. */
int orc_kerr_rk4_fixed_d(double s[8], double M, double a, double h);
/* rk4_step_d: out-of-place step, returns ok flag. */
int orc_kerr_rk4_step_d(const double s[8], double M, double a, double h, double out[8]);
/* rk4_adaptive_step_d: returns growth factor (>0 accepted, 0 rejected);
 * on acceptance s is replaced, on rejection *h is shrunk. */
double orc_kerr_adaptive_step_d(double s[8], double M, double a, double* h,
                                double rtol, double atol);

/* Cameras (bh_kerr.h:781-826, 880-884, 910-967) and Doppler (833-862). */
void orc_init_static_d(double M, double a, double r, double th, double ph,
                       double n_r, double n_th, double n_ph, double s[8]);
void orc_init_zamo_d(double M, double a, double r, double th, double ph,
                     double n_r, double n_th, double n_ph, double s[8]);
int orc_static_timelike_d(double M, double a, double r, double th);
double orc_doppler_d(const double s[8], double M, double a);

/* Compensated summation (bh_common.h:451-493). */
double orc_rk4_sum_d(double k1, double k2, double k3, double k4);

/* ---- Per-pixel renderer (render.cu:124-363, FP64 path) ---- */
typedef struct orc_ray_info {
    int32_t cls;          /* np_ray_class */
    int32_t orders;       /* disk crossings accumulated */
    int64_t steps;        /* accepted steps */
    int64_t attempts;     /* adaptive attempts (== steps in fixed mode) */
    int64_t rhs_evals;    /* Hamilton right-hand-side evaluations, reference count */
    double  min_r;
    float   inv[3];       /* |H|, dE/E, dL/L at crossings/ends (render.cu:233-244,327-345) */
    float   rgb[3];       /* this sample's colour */
    int32_t ncross, pad_; /* equatorial crossings detected (in the disk or not), at most 4 kept */
    double  xr[4];        /* their radii, as the crossing rule computes them */
    double  p0[4];        /* covariant p_t, p_r, p_theta, p_phi at the camera */
    double  t_end;        /* coordinate time t of the last accepted state (t = 0 at the camera) */
    double  phi_end;      /* its azimuth phi */
    double  lam_end;      /* affine parameter: the sum of the accepted step sizes */
    double  r_end;        /* r of the last accepted state */
    double  th_end;       /* its theta */
} orc_ray_info;

/* Camera-space unit direction of sample `sample_id` of pixel (x, y)
 * (render.cu:137-159).  n = {n_r, n_theta, n_phi}; cam = {cam_x, cam_y}. */
void orc_camera_dir(const np_render_params* p, int x, int y, uint32_t sample_id,
                    float n[3], float cam[2]);
uint32_t orc_wanghash(uint32_t a);

/* Trace one sample; returns its colour in info->rgb. */
/* Hermite crossing radius (extension, np_crossing HERMITE). */
double orc_hermite_crossing(double h, double r0, double th0, double dr0, double dth0, double r1,
                            double th1, double dr1, double dth1, double s0, double r_lin);
void orc_render_ray(const np_render_params* p, int x, int y, uint32_t sample_id,
                    orc_ray_info* info);
/* Sum of `spp` samples starting at sample_offset for each listed pixel, in
 * the reference's order (render.cu:141-363).  rgb is npix*3 floats.
 * info (optional) gets npix*spp entries, pixel-major.  nthreads <= 0 uses
 * every core. */
/* Initial covariant momenta (p_t, p_r, p_theta, p_phi) of every sample,
 * pixel-major then sample, without tracing. */
void orc_ray_p0_pixels(const np_render_params* p, int npix, const int32_t* xs, const int32_t* ys,
                       int sample_offset, int spp, double* p0);
void orc_render_pixels(const np_render_params* p, int npix, const int32_t* xs,
                       const int32_t* ys, int sample_offset, int spp, float* rgb,
                       orc_ray_info* info, int nthreads);

/* ---- Schwarzschild (include/bh_schwarzschild.h, schwarzschild_blackhole.cu) ---- */
/* Reference FP32 precompute kernel body (schwarzschild_blackhole.cu:48-139). */
void orc_schw_precompute_f32(float mass, float r_start, const float* b, int n,
                             float* defl, float* rmin, int32_t* status, int nthreads);
int orc_schw_rmin_analytic(float Rs, float b, float* rmin_out);
/* FP64 equatorial sweep: RK4 of the Schwarzschild Hamiltonian restated in
 * FP64, step law of schwarzschild_blackhole.cu:84-85, launched from r0 with
 * E=1, L=b; captured when r < r_capture, escaped when r > r0 with p_r > 0,
 * photon-sphere class when |b - 3 sqrt(3) M| < ps_band.  Outputs class,
 * unwrapped total dphi and step count. */
void orc_schw_sweep_d(double M, double r0, double r_capture, double ps_band,
                      int64_t max_steps, const double* b, int n, int32_t* cls,
                      double* dphi, int64_t* steps, int nthreads);

/* ---- Carlson symmetric integrals (geokerr_original.f:3444-3489, 3608-3762) ---- */
double orc_rf(double x, double y, double z);
double orc_rc(double x, double y);
double orc_rd(double x, double y, double z);
double orc_rj(double x, double y, double z, double p);

/* ---- GEOKERR GEOMU and its numerics (geokerr_original.f:771-1186, 3137-3441) ---- */
typedef struct orc_geomu_out {
    double muf;       /* final mu = cos(theta) */
    double iu;        /* I_u integral between u0 and uf */
    double u1, u2, u3, u4;  /* real roots of U(u) where the case has them */
    double u0_used, uf_used;  /* after the reference's clamps/warning edits */
    double i1mu, i3mu;
    double h1;        /* Eq. (21) root, ncase 6 */
    int32_t tpm;      /* mu turning points */
    int32_t ncase;    /* Table 1 case 1..8 (0 = error path) */
    int32_t tpr;      /* radial turning points (as modified) */
    int32_t warned;   /* 1 when the reference would print a WARNING and modify u0 */
} orc_geomu_out;

void orc_geomu(double u0, double uf, double mu0, double a, double l, double q2, int tpr,
               double su, double sm, orc_geomu_out* o);

/* ---- GEOR (geokerr_original.f:2672-3134), FIRSTPT on, PHT off ---- */
typedef struct orc_geor_out {
    double uf;        /* final inverse radius (-1 where the reference returns it) */
    double imu;       /* the mu integral from mu0 to muf after tpm turning points */
    double iu0;       /* the u0 piece of the u integral */
    double u1, u2, u3, u4;  /* real roots of U(u) where the case has them */
    double h1;        /* Eq. (21) root, ncase 6 */
    double u0_used;   /* after the reference's forbidden-region edit */
    int32_t ncase;    /* Table 1 case 1..8 (0 = no solution) */
    int32_t tpr;      /* radial turning points (as passed where not written) */
    int32_t warned;   /* 1 when the reference would print a WARNING and modify u0 */
    int32_t pad_;
} orc_geor_out;

double orc_calcimuasymf(double a, double mneg, double mpos, double muf, double muplus, double* rff);
double orc_calcimusymf(double a, double mneg, double mpos, double muf, double muplus, double imum,
                       double* rff);
void orc_geor(double u0, double uf_in, double mu0, double muf, double a, double l, double q2, int32_t tpm,
              int32_t tpr_in, double su, double sm, orc_geor_out* o);
void orc_geor_batch(const double* u0, const double* uf_in, const double* mu0, const double* muf,
                    const double* a, const double* l, const double* q2, const int32_t* tpm,
                    const double* su, const double* sm, int n, double* uf, double* imu, int32_t* tpr,
                    int32_t* ncase, int nthreads);

/* GEOKERR's variables shared by its GEOMU and GEOPHITIME calls along one
 * geodesic (geokerr_original.f:398-412): written when FIRSTPT is true (or on
 * every call, where the Fortran does so), read otherwise. */
typedef struct orc_gk_vars {
    int32_t ncase, pad_;
    double u1, u2, u3, u4, h1;
    double iu0, rffu0, rffu1;
    double i1mu, i3mu, rffmu1, rffmu2, rffmu3;
    double rdc, rjc, tu01, tu02, tu03, tu04, tmu1, tmu3, phimu1, phimu3;
} orc_gk_vars;

/* GEOMU with the Fortran's by-reference arguments (u0, uf, tpm, tpr) and
 * its FIRSTPT/PHT switches. */
void orc_gk_geomu(orc_gk_vars* v, double* u0, double* uf, double mu0, double a, double l,
                  double q2, int32_t* tpm, int32_t* tpr, double su, double sm, int pht,
                  int firstpt, double* muf, double* iu, int32_t* warned);
double orc_calcimuasymf_rf(double mneg, double mpos, double muf);
double orc_calcimusymf_rf(double mneg, double mpos, double muf, double muplus);

/* GEOPHITIME (geokerr_original.f:1648-2360): u and mu parts of t and phi
 * and the affine parameter between (u0, mu0) and (uf, muf). */
void orc_gk_geophitime(orc_gk_vars* v, double u0, double uf, double mu0, double muf, double a,
                       double l, double q2, int32_t tpm, int32_t tpr, double su, double sm,
                       double iu, int firstpt, double* phimu, double* tmu, double* phiu,
                       double* tu, double* lambda);

/* GEOR with the Fortran's argument contract: *u0, *uf, *tpr by reference,
 * IMU written into *imu (left as it was on the asymmetric-roots exit), the
 * shared variables in v written where FIRSTPT (or the Fortran
 * unconditionally) writes them and read otherwise; PHT forms RFFU0/RFFU1/
 * RFFMU2 for GEOPHITIME. */
void orc_gk_geor(orc_gk_vars* v, double* u0, double* uf, double mu0, double muf, double a, double l,
                 double q2, double* imu, int32_t tpm, int32_t* tpr, double su, double sm, int pht,
                 int firstpt, int32_t* warned);

/* FINDMROOTS (geokerr_original.f:607-642) and INDEP_MUF (:645-706). */
void orc_findmroots(double q2, double l, double a, double mu0, double* muminus, double* muplus);
void orc_indep_muf(double muminus, double muplus, double mu0, double muf, int32_t tpm0, int32_t tpmf,
                   double sm, int32_t k, int32_t npts, double offset, double* mun, int32_t* tpmk);

/* The GEOKERR driver's switches and per-geodesic arguments
 * (geokerr_original.f:357-360).  ALPHA is not read by the driver. */
typedef struct orc_gk_opts {
    int32_t nup, npts, kext, phit, usegeor, mufill;
    double offset;
} orc_gk_opts;
typedef struct orc_gk_args {
    double u0, uf, uout, mu0, muf, a, l, q2, beta, su, sm;
    int32_t tpm, tpr;
} orc_gk_args;

/* The whole GEOKERR driver (geokerr_original.f:357-604): the USEGEOR
 * branch (samples evenly in mu, GEOR + GEOPHITIME) and the u-sampled branch
 * with the MUFILL switch to mu near the u turning point (INDEP_MUF + GEOR).
 * Arrays hold nup + npts - 2 kext entries; entries the driver does not
 * write are NaN (doubles) and -1 (ints).  g comes back as the Fortran
 * leaves its by-reference arguments (U0, UF, MUF, TPM, TPR); *ncase,
 * *nup_out and *npts_out are NCASE, NUP and NPTS on return.  Samples the
 * Fortran would store outside its arrays are not computed. */
void orc_gk_geokerr_full(const orc_gk_opts* o, orc_gk_args* g, double* ufi, double* mufi, double* dti,
                         double* dphi, int32_t* tpmi, int32_t* tpri, double* lami, int32_t* ncase,
                         int32_t* nup_out, int32_t* npts_out);
/* n independent drivers; g is n entries (in/out), the arrays n x (nup +
 * npts - 2 kext), OpenMP. */
void orc_gk_geokerr_full_batch(const orc_gk_opts* o, orc_gk_args* g, int n, double* ufi, double* mufi,
                               double* dti, double* dphi, int32_t* tpmi, int32_t* tpri, double* lami,
                               int32_t* ncase, int32_t* nup_out, int32_t* npts_out, int nthreads);

/* GEOKERR driver (geokerr_original.f:357-604) in the configuration
 * geokerr_eval_lambda and geokerr_estimate_periods use: USEGEOR = .false.,
 * PHIT = .true., MUFILL with NPTS = KEXT = 0, ALPHA = BETA = 0.  uf_is_u0
 * reproduces geokerr_estimate_periods passing one variable as both U0 and
 * UF.  Arrays hold nup entries; returns how many leading entries the
 * driver wrote (the reference leaves the rest unset), after the driver's
 * own NUP update.  *u0 and *uf come back as the driver modified them. */
int orc_gk_geokerr(double* u0, double* uf, double uout, double mu0, double a, double l, double q2,
                   int32_t tpm, int32_t tpr, double su, double sm, int nup, double offset,
                   int uf_is_u0, double* ufi, double* mufi, double* dti, double* dphi,
                   int32_t* tpmi, int32_t* tpri, double* lami, int* nup_out);

/* geokerr_ref_api (geokerr_eval_lambda.f90). */
void orc_geokerr_eval_lambda(double u0, double mu0, double uf, double a, double lz, double q2,
                             double t0, double lam0, double dlam, int s, double* u, double* mu,
                             double* t, double* phi, double* affine, int* tpmi, int* tpri,
                             int* info);
void orc_geokerr_estimate_periods(double u0, double mu0, double a, double lz, double q2,
                                  double* tr, double* ttheta, int* info);
double orc_geokerr_phi_omega(double u0, double mu0, double a, double lz, double q2);
/* n independent eval_lambda calls (inputs SoA, outputs n x s), OpenMP. */
void orc_geokerr_eval_lambda_batch(const double* u0, const double* mu0, const double* uf,
                                   const double* a, const double* lz, const double* q2,
                                   const double* t0, double lam0, double dlam, int s, int n,
                                   double* u, double* mu, double* t, double* phi, double* affine,
                                   int* tpmi, int* tpri, int* info, int nthreads);
void orc_sncndn(double uu, double emmc, double* sn, double* cn, double* dn);
/* coefficients and roots are C99 double _Complex arrays, constant term first */
void orc_zroots(const double _Complex* a, int m, double _Complex* roots, int polish);
double orc_ellcubicreal(const int p[5], double a1, double b1, double a2, double b2, double a3,
                        double b3, double a4, double b4, double* rff, double y, double x);
double orc_ellcubiccomplex(const int p[5], double a1, double b1, double a4, double b4, double f,
                           double g, double h, double* rff, double y, double x);
double orc_ellquarticreal(const int p[5], double a1, double b1, double a2, double b2, double a3,
                          double b3, double a4, double b4, double a5, double b5, double* rff,
                          double y, double x);
double orc_ellquarticcomplex(const int p[5], double a1, double b1, double a4, double b4, double a5,
                             double b5, double f, double g, double h, double* rff, double y, double x);
double orc_elldoublecomplex(const int p[5], double f1, double g1, double h1, double f2, double g2,
                            double h2, double a5, double b5, double* rff, double y, double x);
void orc_calcimuasym(double a, double mneg, double mpos, double mu0, double muplus,
                     double* imu0, double* imum, double* rf0, double* rfc);
void orc_calcimusym(double a, double mneg, double mpos, double mu0, double muplus,
                    double* imu0, double* imum, double* rf0, double* rfc);
double orc_asech(double x);
void orc_geokerr_camera(double alpha, double beta, double mu0, double a_in, double* L,
                        double* Q2, double* a_out, double* su, double* sm);
/* Batch of GEOMU solves from screen coordinates (config: u0 fixed, tpr = 0). */
void orc_geomu_batch(const double* alpha, const double* beta, const double* a, const double* mu0,
                     const double* uf, double u0, int n, double* muf, int32_t* tpm, int32_t* ncase,
                     double* iu, int nthreads);

/* ---- Differentiable disk render, FP64 forward mode (adjoint_oracle.c) ---- */
/* Per pixel: I and dI/d(spin, theta_obs, q, gain) (dI is npix x 4). */
void orc_adj_pixels(const np_adjoint_params* p, int npix, const int32_t* xs, const int32_t* ys,
                    double* rad, double* drad, int32_t* ncross, int nthreads);
double orc_adj_isco(double a, double* d_da);
double orc_adj_disk_term(double r, double E, double L, double a, double q, double gain, double rout,
                         double ew);
void orc_adj_zamo_init(double a, double th, double r_cam, const float n[3], double out[4]);
void orc_adj_rhs(const double y[4], double E, double L, double a, double out[4]);

#ifdef __cplusplus
}
#endif

#endif /* NP_ORACLE_H */
