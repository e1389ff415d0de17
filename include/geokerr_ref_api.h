/*
 * geokerr_ref_api.h -- drop-in C ABI of the reference's geokerr_ref_api
 * module (geokerr_port/reference/geokerr_eval_lambda.f90), the prototypes
 * its validation harness compiles against (geokerr_port/validation/
 * gkval_adapters.c:33-44).  Exported by libnullpath.so; each call runs the
 * GEOKERR driver on the GPU for one geodesic (M = 1, E = 1, u = 1/r,
 * mu = cos(theta)) and returns host results synchronously on the null
 * stream.  Batched device-array forms are np_geokerr_*_batch in nullpath.h.
 */
#ifndef GEOKERR_REF_API_H
#define GEOKERR_REF_API_H

#ifdef __cplusplus
extern "C" {
#endif

/*
 * geokerr_eval_lambda.f90:62-198.  GEOKERR with NUP = S samples from u0
 * toward uf (through the radial turning point when there is one),
 * resampled linearly onto lambda_k = lam0 + k dlam, k = 0..S-1:
 * u, mu, t (= t0 + dt), phi (unwrapped), affine (= lambda_k); sample 0 is
 * forced to (u0, mu0).  tpmi / tpri: first grid index at or after the first
 * mu / u turning point, -1 if none.  info: 1 if S <= 0 (outputs untouched),
 * else 0; 100 - np_status (101..104) if the GPU call failed (an extension:
 * the reference has no failure path).  Where the reference reads samples
 * its driver never wrote (a ray that ends at the horizon), only the
 * written samples are used; a ray with none gives NaN u, mu, t, phi.
 */
void geokerr_eval_lambda(double u0, double mu0, double uf, double a, double Lz, double Q2, double t0,
                         double lam0, double dlam, int S, double* u, double* mu, double* t,
                         double* phi, double* affine, int* tpmi, int* tpri, int* info);

/*
 * geokerr_eval_lambda.f90:200-295.  GEOKERR with NUP = 4096 from u0 back to
 * u0 (one variable passed as both U0 and UF, SU = 1); Ttheta = twice the
 * affine distance between the first and second mu turning points, -1 if
 * fewer; Tr likewise for u turning points, which the driver never counts
 * past one, so Tr is always -1.  info as above (0 on success).
 */
void geokerr_estimate_periods(double u0, double mu0, double a, double Lz, double Q2, double* Tr,
                              double* Ttheta, int* info);

/*
 * geokerr_eval_lambda.f90:297-306 with dphi_dlam (:41-51): dphi/dlambda
 * (Mino time) at r = max(1e-12, 1/max(1e-12, u0)), mu0, with Delta and
 * 1 - mu0^2 floored at 1e-30.  Q2 is unused, as in the reference.
 */
double geokerr_phi_omega(double u0, double mu0, double a, double Lz, double Q2);

#ifdef __cplusplus
}
#endif

#endif /* GEOKERR_REF_API_H */
