/*
 * nullpath.h -- C ABI of libnullpath.so, the MI355X (gfx950) null-geodesic
 * library.  Plain pointers and sizes only; device pointers are HIP device
 * memory, `stream` is a hipStream_t passed as void* (NULL = default stream).
 *
 * Every entry point returns NP_OK (0) or a negative np_status code and
 * never aborts the process.
 */
#ifndef NULLPATH_H
#define NULLPATH_H

#include <stddef.h>
#include <stdint.h>
#include "np_render.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum np_status {
    NP_OK = 0,
    NP_EINVAL = -1,    /* bad argument (sizes, pointers, parameters) */
    NP_EALLOC = -2,    /* device or pinned-host allocation failed */
    NP_ELAUNCH = -3,   /* kernel launch or execution failed */
    NP_ENODEV = -4,    /* no HIP device */
    NP_EDISABLED = -5, /* entry point switched off in this build (see its comment) */
    NP_ECAPACITY = -6  /* request beyond a fixed capacity (np_render_batch_noalloc) */
} np_status;

/* Human-readable text for a status code (static storage). */
const char* np_status_string(int code);

/* Counters a tracer launch adds to (device memory, caller-zeroed, may be
 * NULL).  Steps are accepted integration steps; attempts include rejected
 * adaptive attempts; rhs_evals counts Hamilton right-hand-side evaluations
 * actually performed.  attempts_f64 and rhs_evals_f64 count the part of
 * attempts and rhs_evals that an NP_PREC_FP32C launch ran in FP64 (inside
 * fp64_radius, or a float attempt's FP64 re-run); they stay 0 for FP64.
 * attempts_lazy counts the float attempts that ran without the full step
 * (estimate_every_step = 0: the two half steps only). */
typedef struct np_trace_counters {
    uint64_t rays;
    uint64_t steps;
    uint64_t attempts;
    uint64_t rhs_evals;
    uint64_t captured;
    uint64_t escaped;
    uint64_t stepcap;
    uint64_t disk_hits;
    uint64_t attempts_f64;
    uint64_t rhs_evals_f64;
    uint64_t attempts_lazy;
} np_trace_counters;

/* Optional per-sample diagnostic record (extension; the reference's kernel
 * keeps these in registers only). */
typedef struct np_ray_record {
    int32_t cls;       /* np_ray_class */
    int32_t orders;    /* disk crossings accumulated */
    int32_t steps;     /* accepted integration steps */
    int32_t attempts;  /* integrator attempts (rejections included) */
    double  min_r;     /* closest approach, absolute units */
    double  r_end;     /* r when the ray stopped */
} np_ray_record;

/* Optional per-sample end values of the integrated coordinates (extension):
 * the coordinate time t (dt/dlambda = g^{tt} p_t + g^{t phi} p_phi,
 * bh_kerr.h:679-727, integrated with the state from t = 0 at the camera, as
 * the reference's RK4 does), the azimuth phi, and the affine parameter
 * lambda (the sum of the accepted step sizes), all at the last accepted
 * state of the sample.  NP_PREC_FP32C accumulates t, phi and lambda in
 * float-float with two-sum updates. */
typedef struct np_ray_times {
    double t;
    double phi;
    double lambda;
} np_ray_times;

/*
 * Trace `spp` samples (sample ids sample_offset .. sample_offset+spp-1) of
 * every pixel in rows [y0, y0+rows) of the frame described by *p, and write
 * each pixel's colour sum to d_rgb (device, rows*width*3 floats, row-major,
 * RGB interleaved: the float3 tile of the reference).  Rows past p->height
 * are left untouched.
 *
 * Replaces the renderKernel launch of render.cu:118-123 / 540-550 (same
 * per-pixel sample loop and colour rules, render.cu:124-363).  d_inv_max3
 * (optional, 3 floats) receives atomic maxima of |H|, dE/E, dL/L when
 * p->check_inv is set (render.cu:233-244, 327-345).  d_records (optional,
 * rows*width*spp entries, pixel-major then sample) receives one
 * np_ray_record per traced sample.
 */
int np_render_tile(const np_render_params* p, int y0, int rows, int spp, int sample_offset,
                   float* d_rgb, float* d_inv_max3, np_trace_counters* d_counters,
                   np_ray_record* d_records, void* stream);

/*
 * np_render_tile that also integrates the coordinate time and the affine
 * parameter and writes one np_ray_times per traced sample to d_times
 * (device, rows*width*spp entries, the indexing of d_records).  It has no
 * counterpart in the reference, whose integrator advances t with the state
 * (bh_kerr.h:734-774, 987-1032) but whose renderer never reads it.  The
 * image, records and counters are those of np_render_tile; the launch runs
 * the precision's default kernel variant (p->waves_per_simd is ignored).
 * NP_EINVAL when d_times is NULL or p->check_inv is set.
 */
int np_render_tile_times(const np_render_params* p, int y0, int rows, int spp, int sample_offset,
                         float* d_rgb, np_trace_counters* d_counters, np_ray_record* d_records,
                         np_ray_times* d_times, void* stream);

/*
 * Progressive render of a whole frame into a host accumulation buffer
 * h_accum (width*height*3 floats, caller-zeroed or holding earlier passes),
 * `passes` passes of `samples_per_pass` samples starting at sample id
 * spp_done, accumulated in pass order as render.cu:523-580 does.
 * inv_max (optional, 3 floats) receives the invariant maxima.
 */
int np_render_accumulate(const np_render_params* p, int spp_done, int samples_per_pass,
                         int passes, int tile_h, float* h_accum, float* inv_max);

/*
 * One progressive pass with the accumulation on the device: adds the colour
 * sums of samples sample_offset .. sample_offset+spp-1 of every pixel to
 * d_accum (device, width*height*3 floats, row-major RGB), in the order the
 * reference's host loop adds each pass's tiles (render.cu:523-580, one
 * launch per pass instead of per tile; the same per-pixel sample loop).
 * d_inv_max3 as in np_render_tile.
 */
int np_render_pass(const np_render_params* p, int sample_offset, int spp, float* d_accum,
                   float* d_inv_max3, void* stream);

/*
 * Progressive passes captured once in a HIP graph and replayed: the graph
 * holds `passes` passes of `spp` samples each (as np_render_pass) into
 * d_accum (and d_inv_max3, optional), and its own device-side sample
 * counter: a replay (np_pass_graph_launch) adds samples next .. next +
 * passes*spp - 1 and advances the counter, so successive replays continue
 * the progressive render of render.cu:523-580 (the first replay starts at
 * first_sample; np_pass_graph_seek moves the counter, stream-ordered).  The
 * graph contains kernels only (no memset or copy nodes); it keeps its own
 * frame record and work counter, so replays of one graph must not overlap
 * each other.  Create on the device the replays run on; destroy after the
 * last replay has finished.
 */
typedef struct np_pass_graph np_pass_graph;
int np_pass_graph_create(const np_render_params* p, int spp, int passes, float* d_accum,
                         float* d_inv_max3, int first_sample, np_pass_graph** graph);
int np_pass_graph_launch(np_pass_graph* graph, void* stream);
int np_pass_graph_seek(np_pass_graph* graph, int next_sample, void* stream);
void np_pass_graph_destroy(np_pass_graph* graph);

/*
 * Launch-shape choices of the batch renderer, picked by np_autotune_batch.
 * block_size: 64, 128 or 256 threads; waves_per_simd: kernel variant as in
 * np_render_params (FP32C: 4 = 128-VGPR cap, 3 = 168-VGPR cap, 2 = compiler budget), 0 = each
 * frame's own; refill_chunk: consecutive pixels a lane takes per refill; fp64_radius: FP32C
 * radius (absolute units) of FP64 steps, 0 = each frame's own; defer_fp64: as in
 * np_render_params (1 = defer FP64 attempts, -1 = never), 0 = each frame's own.
 */
typedef struct np_tuning {
    int32_t block_size;
    int32_t waves_per_simd;
    int32_t refill_chunk;
    double  fp64_radius;
    int32_t defer_fp64;
} np_tuning;

/*
 * Render nframes whole frames of uneven resolution, spin and inclination in
 * one persistent launch (an extension: the reference renders one frame per
 * process, render.cu:523-638).  All frames share one precision and one
 * integrator; check_inv must be 0.  d_rgb holds the frames back to back
 * (frame i starts at 3 * sum_{j<i} width_j * height_j floats; colour sums
 * over spp samples); d_records (optional) likewise with spp records per
 * pixel; d_counters (optional) sums over all frames.  tuning NULL = block
 * 256, the frames' own FP32C variant (4 waves by default), chunk 1, frame radii.
 */
int np_render_batch(const np_render_params* frames, int nframes, int spp, int sample_offset,
                    float* d_rgb, np_ray_record* d_records, np_trace_counters* d_counters,
                    const np_tuning* tuning, void* stream);

/* Frames np_render_batch_noalloc accepts per call. */
#define NP_NOALLOC_MAX_FRAMES 128

/*
 * np_render_batch without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, kernels and results as np_render_batch; the
 * frame table and the per-frame work counters live in a fixed device array
 * of the library (about 51 KB per GPU, held until the process ends) instead
 * of a per-call allocation.  At most NP_NOALLOC_MAX_FRAMES frames
 * (NP_ECAPACITY beyond).  The caller still provides d_rgb and the optional
 * d_records / d_counters.  The call copies the frame table into that array
 * and waits for the copy on `stream` before launching (the host staging
 * buffer is reused), so it returns after earlier work on `stream` has
 * finished.  Every call on one GPU shares the array: calls that overlap on
 * one GPU (other streams or threads) race.
 */
int np_render_batch_noalloc(const np_render_params* frames, int nframes, int spp, int sample_offset,
                            float* d_rgb, np_ray_record* d_records, np_trace_counters* d_counters,
                            const np_tuning* tuning, void* stream);

/*
 * Time every candidate tuning (block 128/256 x chunk 1/4, and for FP32C the
 * kernel variant 4/3/2 x FP64 radius 0.25/0.5/1.5 x the largest disk_rout) on
 * the batch with each frame's resolution divided by `subsample`, and write
 * the fastest to *best.  Optional trials / trial_ms (capacity max_trials)
 * receive every candidate and its kernel time; *n_trials their count.  The
 * choice is cached per batch signature for the life of the process (a
 * cached answer reports no trials).  Synchronizes the device.
 */
int np_autotune_batch(const np_render_params* frames, int nframes, int spp, int subsample,
                      np_tuning* best, np_tuning* trials, double* trial_ms, int max_trials,
                      int* n_trials);

/* ---------------------------------------------------------------------
 * GEOKERR semi-analytic solver (geokerr_port/reference/geokerr_original.f)
 * ------------------------------------------------------------------- */

enum np_carlson_kind { NP_CARLSON_RF = 0, NP_CARLSON_RC = 1, NP_CARLSON_RD = 2, NP_CARLSON_RJ = 3 };

/*
 * Batched Carlson symmetric integrals in FP64: out[i] = R_kind(x[i], y[i]
 * [, z[i] [, p[i]]]).  RC ignores z and p, RF and RD ignore p; RJ with
 * p <= 0 returns the Cauchy principal value.  Replaces the device functions
 * carlson_rf/rc/rd/rj and elliptic_batch_kernel of
 * geokerr_port/cuda/carlson_elliptic.cu:77-351 with the algorithm of
 * geokerr_original.f:3444-3489, 3608-3762 (the CUDA file's RJ is wrong).
 */
int np_carlson_batch(int kind, const double* d_x, const double* d_y, const double* d_z,
                     const double* d_p, int64_t n, double* d_out, void* stream);

/*
 * GEOMU for n rays (geokerr_original.f:771-1186, FIRSTPT = .true.):
 * final mu = cos(theta) at inverse radius uf[i] for a photon that starts at
 * (u0[i], mu0[i]) with constants of motion L[i], Q2[i], spin a = spin[i],
 * initial signs su[i] = du/dlambda, sm[i] = dmu/dlambda, and tpr[i] radial
 * turning points between u0 and uf (NULL = all zero).  Outputs: muf, and
 * (each optional) tpm = mu turning points, ncase = Table 1 case, iu = the
 * Mino-time integral I_u.  All arrays are device arrays of n elements.
 */
int np_geomu_batch(const double* d_u0, const double* d_uf, const double* d_mu0,
                   const double* d_spin, const double* d_L, const double* d_Q2,
                   const double* d_su, const double* d_sm, const int32_t* d_tpr, int64_t n,
                   double* d_muf, int32_t* d_tpm, int32_t* d_ncase, double* d_iu, void* stream);

/*
 * GEOMU from screen coordinates: per ray (alpha, beta) at infinity, spin and
 * mu0 = cos(theta_obs) give L = -alpha sqrt(1 - mu0^2),
 * Q2 = beta^2 - (a^2 - alpha^2) mu0^2, su = 1, sm = +1 for beta >= 0
 * (INITIALIZE_CAMERA_GEOKERR, geokerr_original.f:238-320, inputs below 1e-5
 * zeroed), then GEOMU from the common start u0 to uf[i] with tpr = 0.
 */
int np_geomu_camera_batch(const double* d_alpha, const double* d_beta, const double* d_spin,
                          const double* d_mu0, const double* d_uf, double u0, int64_t n,
                          double* d_muf, int32_t* d_tpm, int32_t* d_ncase, double* d_iu,
                          void* stream);

/*
 * The same with the quartic's roots in mixed precision (the GEOKERR batch's
 * compared mode: FP32 bulk, FP64 for near-degenerate arguments): ZROOTS'
 * deflation (LAGUER, geokerr_original.f:3527-3606) runs in float, each root
 * is polished on the full polynomial with three Newton steps in double, and
 * the FP64 ZROOTS runs
 * instead where the float phase cannot tell a real root from a complex pair
 * (imaginary part between 1e-5 and 1e-2 of the modulus) or finds two real
 * roots within 1e-3 of each other.  d_fp64_roots (n flags, or NULL) marks
 * those rays.  Everything after the roots is FP64 as in
 * np_geomu_camera_batch.
 */
int np_geomu_camera_batch_mixed(const double* d_alpha, const double* d_beta, const double* d_spin,
                                const double* d_mu0, const double* d_uf, double u0, int64_t n,
                                double* d_muf, int32_t* d_tpm, int32_t* d_ncase, double* d_iu,
                                int32_t* d_fp64_roots, void* stream);

/*
 * GEOR for n rays (geokerr_original.f:2672-3134, FIRSTPT = .true., PHT =
 * .false.): the inverse radius uf at which a photon from (u0[i], mu0[i])
 * with constants of motion L[i], Q2[i], spin[i] and initial signs su[i],
 * sm[i] reaches mu = muf[i] after tpm[i] mu turning points (with muf = 0,
 * an equatorial crossing: the disk-crossing radius).  Outputs uf (-1 where the
 * reference has no solution: equatorial orbits, a muf outside the allowed
 * band, or more Mino time than remains before the horizon or infinity) and,
 * optional, imu = the Mino-time integral, tpr = radial turning points passed
 * (0 where the reference does not set it), ncase = Table 1 case (0 = no
 * solution).  uf_in (optional, NULL = -1) is the UF value the reference's
 * equal-root cubic branch reads.  Device arrays of n.
 */
int np_geor_batch(const double* d_u0, const double* d_mu0, const double* d_muf, const double* d_spin,
                  const double* d_L, const double* d_Q2, const double* d_su, const double* d_sm,
                  const int32_t* d_tpm, const double* d_uf_in, int64_t n, double* d_uf, double* d_imu,
                  int32_t* d_tpr, int32_t* d_ncase, void* stream);

/* Samples of the GEOKERR driver behind geokerr_estimate_periods
 * (geokerr_eval_lambda.f90:233). */
#define NP_GEOKERR_PERIOD_SAMPLES 4096
/* geokerr_ref_api info for a failed GPU call: 100 - np_status. */
#define NP_GEOKERR_INFO_DEVICE 100

/*
 * Device workspace (256-byte aligned) for the GEOKERR driver over n
 * geodesics of nup samples each: a per-geodesic record (roots, case, the
 * cached u0 / mu0 integrals) and the driver's sample arrays, 48 bytes per
 * sample.  Returns -1 for nup <= 0 or n < 0.
 */
int64_t np_geokerr_workspace_bytes(int32_t nup, int64_t n);

/*
 * n independent geokerr_eval_lambda calls (geokerr_eval_lambda.f90:62-198;
 * host form in geokerr_ref_api.h) on device arrays: inputs u0, mu0, uf, a,
 * lz, q2, t0, lam0, dlam have n entries; outputs u, mu, t, phi, affine are
 * n x s (row i = geodesic i), tpmi, tpri and info (optional) have n.
 * workspace: np_geokerr_workspace_bytes(s, n) bytes.  Replaces the
 * per-call Fortran (GEOKERR, geokerr_original.f:357-604) and the port
 * solver geokerr_complete_stream_cuda (geokerr_port/cuda/
 * geokerr_complete.cu:446-490).
 */
int np_geokerr_eval_lambda_batch(const double* d_u0, const double* d_mu0, const double* d_uf,
                                 const double* d_a, const double* d_lz, const double* d_q2,
                                 const double* d_t0, const double* d_lam0, const double* d_dlam,
                                 int32_t s, int64_t n, double* d_u, double* d_mu, double* d_t,
                                 double* d_phi, double* d_affine, int32_t* d_tpmi,
                                 int32_t* d_tpri, int32_t* d_info, void* d_workspace,
                                 void* stream);

/*
 * n geokerr_estimate_periods calls (geokerr_eval_lambda.f90:200-295);
 * workspace: np_geokerr_workspace_bytes(NP_GEOKERR_PERIOD_SAMPLES, n).
 */
int np_geokerr_estimate_periods_batch(const double* d_u0, const double* d_mu0, const double* d_a,
                                      const double* d_lz, const double* d_q2, int64_t n,
                                      double* d_tr, double* d_ttheta, int32_t* d_info,
                                      void* d_workspace, void* stream);

/* n geokerr_phi_omega values (geokerr_eval_lambda.f90:297-306). */
int np_geokerr_phi_omega_batch(const double* d_u0, const double* d_mu0, const double* d_a,
                               const double* d_lz, const double* d_q2, int64_t n, double* d_out,
                               void* stream);

/*
 * The whole GEOKERR driver (SUBROUTINE GEOKERR, geokerr_original.f:357-604)
 * for n geodesics: the USEGEOR branch (NUP samples evenly in mu between MU0
 * and MUF, each a GEOR + GEOPHITIME solve) and the u-sampled branch with
 * the MUFILL switch to mu as the independent variable across the u turning
 * point (FINDMROOTS :607-642, INDEP_MUF :645-706, GEOR).  The switches and
 * sizes are the Fortran's NUP, NPTS, KEXT, PHIT, USEGEOR, MUFILL, OFFSET;
 * per geodesic the Fortran's U0, UF, UOUT, MU0, MUF, A, L, Q2, BETA, TPM,
 * TPR, SU, SM (ALPHA is not read by the driver).  Each geodesic's sample
 * arrays have NUP + NPTS - 2 KEXT entries (the Fortran's DIMENSION), row
 * after row; entries the driver does not write are NaN (doubles) and -1
 * (ints), and samples the Fortran would store outside its arrays are not
 * computed.  With PHIT false DTI, DPHI and LAMBDAI are 0 (the Fortran
 * stores unset locals).  The *_out arrays (each may be NULL) receive the
 * by-reference arguments as the driver leaves them and NCASE, NUP, NPTS.
 * NP_EINVAL for NUP < 1, NPTS < 0, KEXT < 0, NUP + NPTS - 2 KEXT < 1, a
 * NULL input or sample array, n < 0.  Numerics follow the oracle's
 * restatement, including the reference's quirks stated there.
 */
typedef struct np_geokerr_opts {
    int32_t nup, npts, kext;
    int32_t phit, usegeor, mufill;   /* 0 / 1 */
    double offset;
} np_geokerr_opts;

typedef struct np_geokerr_rays {     /* device arrays, n entries each */
    const double *u0, *uf, *uout, *mu0, *muf, *a, *l, *q2, *beta, *su, *sm;
    const int32_t *tpm, *tpr;
} np_geokerr_rays;

typedef struct np_geokerr_samples {  /* device arrays */
    double *ufi, *mufi, *dti, *dphi, *lambdai;   /* n x (nup + npts - 2 kext) */
    int32_t *tpmi, *tpri;                        /* n x (nup + npts - 2 kext) */
    double *u0_out, *uf_out, *muf_out;           /* n each, or NULL */
    int32_t *tpm_out, *tpr_out, *ncase_out, *nup_out, *npts_out;   /* n each, or NULL */
} np_geokerr_samples;

/* np_geokerr_batch runs a geodesic's samples in parallel: one lane per
 * geodesic runs the solves whose results later samples read (the first
 * solve, the FIRSTPT sample, MUFILL's far-side solve, the last sample of
 * each leg), then one lane per sample solves from its leg's starting state;
 * a geodesic whose samples meet one of the driver's sticky states (GEOR
 * leaving with NCASE = 0, GEOR's equal-root cubic branch, an input edit of
 * U0 or UF) is rerun serially.  The results are those of the serial walk.
 * USEGEOR samples after one where GEOR leaves its quartic case analysis
 * (NCASE = 0, the Fortran's exit when the mu target lies beyond the u
 * range) start again from the state that sample left.  It allocates a
 * workspace of about 1.1 KB per geodesic (u-sampled) or 0.6 KB (USEGEOR),
 * stream-ordered on `stream` where the device allows.
 * NP_EALLOC when that fails. */
int np_geokerr_batch(const np_geokerr_opts* opts, const np_geokerr_rays* rays, int64_t n,
                     const np_geokerr_samples* out, void* stream);

/* The same driver, one lane walking every sample of its geodesic in the
 * Fortran's order; no workspace.  The form np_geokerr_batch is checked
 * against, and its fallback for flagged geodesics. */
int np_geokerr_batch_serial(const np_geokerr_opts* opts, const np_geokerr_rays* rays, int64_t n,
                            const np_geokerr_samples* out, void* stream);

/* ---------------------------------------------------------------------
 * Schwarzschild equatorial sweeps (schwarzschild_blackhole.cu)
 * ------------------------------------------------------------------- */

typedef struct GeodesicLookupTable GeodesicLookupTable;   /* bh_types.h */

/*
 * The reference's FP32 lookup-table kernel, per impact parameter b[i]:
 * photon from r_start with E = 1, L = b, fixed-step RK4 with
 * h = 0.01 clamp(r / (Rs/2), 1, 100), at most 10000 steps; outputs the
 * unwrapped deflection, closest approach (analytic for escaped rays) and
 * status 0 escaped / 1 captured / 2 photon sphere / 3 inconclusive.
 * Replaces the precomputeGeodesicsKernel launch of
 * schwarzschild_blackhole.cu:44-139 / 211-217.  Device arrays of n.
 */
int np_schw_precompute(float mass, float r_start, const float* d_b, int64_t n, float* d_defl,
                       float* d_rmin, int32_t* d_status, void* stream);

/*
 * Host driver of precomputeGeodesics (schwarzschild_blackhole.cu:142-266):
 * log-spaced b on [0.3, 20] b_crit (env BH_B_MIN_RS / BH_B_MAX_RS in Rs),
 * fills *table with new[] arrays (release with np_free_lookup_table or
 * delete[]).  The C++ entry point GeodesicLookupTable*
 * precomputeGeodesics(float, int) is declared in bh_schwarzschild.h.
 */
int np_precompute_geodesics(float mass, int num_rays, GeodesicLookupTable* table);
void np_free_lookup_table(GeodesicLookupTable* table);

/*
 * FP64 plain-method sweep: the Schwarzschild Hamiltonian right-hand side and
 * RK4 of bh_schwarzschild.h:124-295 in FP64, step law
 * h = 0.01 clamp(r / M, 1, 100) (schwarzschild_blackhole.cu:84-85), photon
 * launched from r0 with E = 1, L = b[i], p_r < 0 from H = 0.  A ray is
 * captured when r < r_capture (or the integrator stops at the horizon),
 * escaped when r > r0 with p_r > 0, class 2 when |b - 3 sqrt(3) M| <
 * ps_band (0 turns the band off), class 3 when max_steps run out.  Outputs
 * per ray: class, total unwrapped delta-phi; optional: steps, the turning
 * radius r_min (for escaped rays the largest root of r^3 - b^2 r + b^2 Rs,
 * NaN otherwise) and the largest |H| over the accepted steps (H = 1/2
 * g^{mu nu} p_mu p_nu, the accuracy plan's invariant monitor,
 * docs/PLAN.md:38-43); d_total_steps (optional, one u64) accumulates the
 * step count.
 */
int np_schw_sweep(double M, double r0, double r_capture, double ps_band, int64_t max_steps,
                  const double* d_b, int64_t n, int32_t* d_cls, double* d_dphi, int64_t* d_steps,
                  double* d_rmin, double* d_hmax, uint64_t* d_total_steps, void* stream);

/*
 * Differentiable disk render (no counterpart in the reference, which has
 * no gradients; it reuses the reference's camera (render.cu:143-159,
 * bh_kerr.h:897-964), Keplerian redshift (bh_kerr.h:833-862) and crossing
 * interpolation (render.cu:211-226)).  M = 1.  One sample per pixel; fixed
 * RK4 steps of size h in s, dlambda = r ds; a crossing at radius r adds
 * gain W(r) r^-q g^3 (1, 0.75, 0.45), W a C^1 smoothstep over edge_width
 * rising from the ISCO and falling to disk_rout.  Gradients are taken with
 * respect to (spin, theta_obs, emiss_q, gain).
 */
typedef struct np_adjoint_params {
    int32_t width, height;
    float fov_y_deg;
    uint32_t seed;       /* pixel jitter salt, 0 = the renderer's unsalted jitter */
    double r_cam;        /* camera radius (M) */
    double theta_obs;    /* camera polar angle (rad) */
    double spin;         /* a/M, |a| < 1 */
    double emiss_q;      /* emissivity index q */
    double gain;         /* disk gain */
    double disk_rout;    /* outer disk edge (M) */
    double edge_width;   /* smoothstep width of both edges (M) */
    double h;            /* step in s */
    int32_t n_steps;     /* steps per ray */
    int32_t ckpt_every;  /* checkpoint interval (steps) */
    int32_t max_orders;  /* crossings counted, 1..3 */
    int32_t reserved;
} np_adjoint_params;

/* 512 x 512, r_cam = 1000, theta_obs = 60 deg, a = 0.7, q = 3, gain = 1,
 * disk to 20 M with 0.5 M edges, 4096 steps of h = 0.005, checkpoint every
 * 64 steps, 3 crossings, fov 2.8 deg. */
static inline void np_adjoint_params_default(np_adjoint_params* p) {
    p->width = 512;
    p->height = 512;
    p->fov_y_deg = 2.8f;
    p->seed = 1u;
    p->r_cam = 1000.0;
    p->theta_obs = 60.0 * 3.14159265358979323846 / 180.0;
    p->spin = 0.7;
    p->emiss_q = 3.0;
    p->gain = 1.0;
    p->disk_rout = 20.0;
    p->edge_width = 0.5;
    p->h = 0.005;
    p->n_steps = 4096;
    p->ckpt_every = 64;
    p->max_orders = 3;
    p->reserved = 0;
}

/* Device workspace bytes np_adjoint_forward (FP32C) and np_adjoint_backward
 * need (checkpoints, crossings, the backward's work list, the chunk maps
 * of split rays and segment recompute), or -1 for invalid parameters.
 * About 4.1 KB per pixel at the defaults. */
int64_t np_adjoint_workspace_bytes(const np_adjoint_params* p);

/* Forward render into d_image (float RGB, width*height*3).  precision
 * NP_PREC_FP32C writes the checkpoints into d_workspace (required);
 * NP_PREC_FP64 needs no workspace and, when d_isum is not NULL, also
 * writes the per-pixel FP64 radiance sum (for finite differences). */
int np_adjoint_forward(const np_adjoint_params* p, int precision, float* d_image, void* d_workspace,
                       double* d_isum, void* stream);

/* Backward: given dLoss/dImage (float RGB), writes the four gradients
 * dLoss/d(spin, theta_obs, emiss_q, gain) into d_grad4 (double[4]).  Uses
 * the workspace of the last FP32C forward with the same parameters. */
int np_adjoint_backward(const np_adjoint_params* p, const float* d_grad_image, void* d_workspace,
                        double* d_grad4, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NULLPATH_H */
