/*
 * geokerr_port.h -- drop-in C ABI of the reference's GEOKERR port-solver
 * entry points, the symbols its validation harness links and calls through
 * the "port" solver (gkval_launch_cuda_integration,
 * geokerr_port/validation/gkval_adapters.c:661-695, declared there at
 * :661-664).  Exported by libnullpath.so.
 *
 * The harness compares what these return with the Fortran
 * geokerr_eval_lambda on the same inputs.  Both symbols here run the GEOKERR
 * driver on the GPU and resample it exactly as geokerr_eval_lambda does
 * (geokerr_ref_api.h) with t0 = 0, so the port fields equal the Fortran
 * ones.  (The reference's own bodies are a linear interpolation between the
 * end states, geokerr_port/cuda/geokerr_complete.cu:422-445, and a
 * one-thread RK4 integration, geokerr_port/cuda/geokerr_stream.cu:57-156.)
 *
 * Arguments (M = 1, E = 1): u0, uf inverse radii, mu0 = cos(theta0), spin a,
 * L = Lz, Q2 = Carter constant, lam0 / dlam the lambda grid, S samples;
 * five host arrays of S doubles receive u, mu, t, phi, affine.  Synchronous
 * on the null stream; device memory is allocated per call.  Returns 0 on
 * success, -1 for S <= 0 or a null output (outputs untouched), -2 if device
 * memory could not be allocated, -3 if a launch or a kernel failed -- the
 * reference's codes.
 */
#ifndef GEOKERR_PORT_H
#define GEOKERR_PORT_H

#ifdef __cplusplus
extern "C" {
#endif

/* geokerr_port/cuda/geokerr_complete.cu:446-490 */
int geokerr_complete_stream_cuda(double u0, double uf, double mu0, double a, double L, double Q2,
                                 double lam0, double dlam, int S, double* h_u, double* h_mu,
                                 double* h_t, double* h_phi, double* h_affine);

/* geokerr_port/cuda/geokerr_stream.cu:158-193 (same contract) */
int geokerr_stream_cuda(double u0, double uf, double mu0, double a, double L, double Q2, double lam0,
                        double dlam, int S, double* h_u, double* h_mu, double* h_t, double* h_phi,
                        double* h_affine);

#ifdef __cplusplus
}
#endif

#endif /* GEOKERR_PORT_H */
