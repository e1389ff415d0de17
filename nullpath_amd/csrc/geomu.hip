// geomu.hip -- gfx950 kernels for GEOKERR's semi-analytic solver: Carlson
// RF/RC/RD/RJ batches and the GEOMU batch (mu at a given inverse radius).
//
// One ray per lane, grid-stride loop over SoA inputs (coalesced 8-byte
// loads), FP64 throughout.  The numerics (np_geokerr.hpp) follow
// geokerr_port/reference/geokerr_original.f: rf/rc/rd/rj 3444-3489 and
// 3608-3762, GEOMU 771-1186 with the reductions of 1188-1645, calcimu*
// 3137-3274, SNCNDN 3277-3340, ZROOTS/LAGUER 3343-3441 and 3527-3606.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "np_geokerr.hpp"
#include "nullpath.h"

namespace np {

namespace {

using gk::GkVars;

struct GeomuResult {
    double muf, iu;
    int tpm, ncase, fp64_roots;
};

// GEOMU with FIRSTPT = .true., PHT = .false. on a fresh geodesic; inlined
// so that no kernel here calls a function (a long leaf function's
// long-branch sequences may reuse its return-address registers).
template <bool MIXED = false>
__device__ __forceinline__ GeomuResult geomu_once(double u0, double uf, double mu0, double a, double l, double q2, int tpr,
                                         double su, double sm) {
    GkVars v{};
    int tpm = 0;
    GeomuResult r;
    gk::geomu<true, false, MIXED>(v, u0, uf, false, mu0, a, l, q2, tpm, tpr, su, sm, r.muf, r.iu);
    r.tpm = tpm;
    r.ncase = v.ncase;
    r.fp64_roots = v.fp64_roots;
    return r;
}

// Camera -> constants (INITIALIZE_CAMERA_GEOKERR, standard run).
__device__ __forceinline__ void camera_constants(double alpha, double beta, double mu0, double a_in,
                                                 double& L, double& Q2, double& a, double& su,
                                                 double& sm) {
    L = -alpha * sqrt(1.0 - mu0 * mu0);
    Q2 = beta * beta - (a_in * a_in - alpha * alpha) * mu0 * mu0;
    a = a_in;
    if (fabs(Q2) < 1e-5 * 1e-5) Q2 = 0.0;
    if (fabs(a) < 1e-5) a = 0.0;
    if (fabs(L) < 1e-5) L = 0.0;
    su = 1.0;
    sm = (mu0 < 1.0 && beta >= 0.0) ? 1.0 : -1.0;
}

// Both GEOMU kernels are capped at 256 registers (2 waves per SIMD): left
// alone the compiler takes 256 VGPRs plus AGPRs, 1 wave per SIMD, and the
// 2^20-ray batch runs 1.4x slower.
// MIXED: the quartic's roots from a float deflation polished in double
// (gk::zroots4_mixed), FP64 ZROOTS where that hands back; fp64_roots flags
// those rays.
template <bool MIXED>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
geomu_camera_kernel(const double* __restrict__ alpha, const double* __restrict__ beta,
                    const double* __restrict__ spin, const double* __restrict__ mu0,
                    const double* __restrict__ uf, const double u0, const int64_t n,
                    double* __restrict__ muf, int32_t* __restrict__ tpm, int32_t* __restrict__ ncase,
                    double* __restrict__ iu, int32_t* __restrict__ fp64_roots) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double L, Q2, a, su, sm;
        camera_constants(alpha[i], beta[i], mu0[i], spin[i], L, Q2, a, su, sm);
        const GeomuResult r = geomu_once<MIXED>(u0, uf[i], mu0[i], a, L, Q2, 0, su, sm);
        muf[i] = r.muf;
        if (tpm) tpm[i] = r.tpm;
        if (ncase) ncase[i] = r.ncase;
        if (iu) iu[i] = r.iu;
        if (MIXED && fp64_roots) fp64_roots[i] = r.fp64_roots;
    }
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
geomu_kernel(const double* __restrict__ u0, const double* __restrict__ uf,
             const double* __restrict__ mu0, const double* __restrict__ spin,
             const double* __restrict__ L, const double* __restrict__ Q2,
             const double* __restrict__ su, const double* __restrict__ sm,
             const int32_t* __restrict__ tpr, const int64_t n, double* __restrict__ muf,
             int32_t* __restrict__ tpm, int32_t* __restrict__ ncase, double* __restrict__ iu) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const GeomuResult r = geomu_once(u0[i], uf[i], mu0[i], spin[i], L[i], Q2[i], tpr ? tpr[i] : 0,
                                         su[i], sm[i]);
        muf[i] = r.muf;
        if (tpm) tpm[i] = r.tpm;
        if (ncase) ncase[i] = r.ncase;
        if (iu) iu[i] = r.iu;
    }
}

__global__ void __launch_bounds__(256)
carlson_kernel(const int kind, const double* __restrict__ x, const double* __restrict__ y,
               const double* __restrict__ z, const double* __restrict__ p, const int64_t n,
               double* __restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double v;
        switch (kind) {
            case NP_CARLSON_RF: v = gk::rf_d(x[i], y[i], z[i]); break;
            case NP_CARLSON_RC: v = gk::rc_d(x[i], y[i]); break;
            case NP_CARLSON_RD: v = gk::rd_d(x[i], y[i], z[i]); break;
            default: v = gk::rj_d(x[i], y[i], z[i], p[i]); break;
        }
        out[i] = v;
    }
}

// GEOR per ray (one lane per ray); capped at 2 waves per SIMD like the GEOMU
// kernels (same FP64 numerics, same register appetite).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
geor_kernel(const double* __restrict__ u0, const double* __restrict__ mu0, const double* __restrict__ muf,
            const double* __restrict__ spin, const double* __restrict__ L, const double* __restrict__ Q2,
            const double* __restrict__ su, const double* __restrict__ sm, const int32_t* __restrict__ tpm,
            const double* __restrict__ uf_in, const int64_t n, double* __restrict__ uf, double* __restrict__ imu,
            int32_t* __restrict__ tpr, int32_t* __restrict__ ncase) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const gk::GeorResult r = gk::geor(u0[i], uf_in ? uf_in[i] : -1.0, mu0[i], muf[i], spin[i], L[i], Q2[i],
                                          tpm[i], 0, su[i], sm[i]);
        uf[i] = r.uf;
        if (imu) imu[i] = r.imu;
        if (tpr) tpr[i] = r.tpr;
        if (ncase) ncase[i] = r.ncase;
    }
}

int grid_for(int64_t n) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int64_t want = (n + 255) / 256;
    const int64_t cap = (int64_t)cus * 16;
    return (int)(want < cap ? (want < 1 ? 1 : want) : cap);
}

}  // namespace

}  // namespace np

extern "C" {

int np_geomu_camera_batch(const double* d_alpha, const double* d_beta, const double* d_spin,
                          const double* d_mu0, const double* d_uf, double u0, int64_t n,
                          double* d_muf, int32_t* d_tpm, int32_t* d_ncase, double* d_iu,
                          void* stream) {
    if (n < 0 || (n > 0 && (!d_alpha || !d_beta || !d_spin || !d_mu0 || !d_uf || !d_muf)))
        return NP_EINVAL;
    if (n == 0) return NP_OK;
    hipLaunchKernelGGL(np::geomu_camera_kernel<false>, dim3(np::grid_for(n)), dim3(256), 0,
                       (hipStream_t)stream, d_alpha, d_beta, d_spin, d_mu0, d_uf, u0, n, d_muf,
                       d_tpm, d_ncase, d_iu, (int32_t*)nullptr);
    return hipGetLastError() == hipSuccess ? NP_OK : NP_ELAUNCH;
}

int np_geomu_camera_batch_mixed(const double* d_alpha, const double* d_beta, const double* d_spin,
                                const double* d_mu0, const double* d_uf, double u0, int64_t n,
                                double* d_muf, int32_t* d_tpm, int32_t* d_ncase, double* d_iu,
                                int32_t* d_fp64_roots, void* stream) {
    if (n < 0 || (n > 0 && (!d_alpha || !d_beta || !d_spin || !d_mu0 || !d_uf || !d_muf)))
        return NP_EINVAL;
    if (n == 0) return NP_OK;
    hipLaunchKernelGGL(np::geomu_camera_kernel<true>, dim3(np::grid_for(n)), dim3(256), 0,
                       (hipStream_t)stream, d_alpha, d_beta, d_spin, d_mu0, d_uf, u0, n, d_muf,
                       d_tpm, d_ncase, d_iu, d_fp64_roots);
    return hipGetLastError() == hipSuccess ? NP_OK : NP_ELAUNCH;
}

int np_geomu_batch(const double* d_u0, const double* d_uf, const double* d_mu0,
                   const double* d_spin, const double* d_L, const double* d_Q2,
                   const double* d_su, const double* d_sm, const int32_t* d_tpr, int64_t n,
                   double* d_muf, int32_t* d_tpm, int32_t* d_ncase, double* d_iu, void* stream) {
    if (n < 0 || (n > 0 && (!d_u0 || !d_uf || !d_mu0 || !d_spin || !d_L || !d_Q2 || !d_su ||
                            !d_sm || !d_muf)))
        return NP_EINVAL;
    if (n == 0) return NP_OK;
    hipLaunchKernelGGL(np::geomu_kernel, dim3(np::grid_for(n)), dim3(256), 0, (hipStream_t)stream,
                       d_u0, d_uf, d_mu0, d_spin, d_L, d_Q2, d_su, d_sm, d_tpr, n, d_muf, d_tpm,
                       d_ncase, d_iu);
    return hipGetLastError() == hipSuccess ? NP_OK : NP_ELAUNCH;
}

int np_geor_batch(const double* d_u0, const double* d_mu0, const double* d_muf, const double* d_spin,
                  const double* d_L, const double* d_Q2, const double* d_su, const double* d_sm,
                  const int32_t* d_tpm, const double* d_uf_in, int64_t n, double* d_uf, double* d_imu,
                  int32_t* d_tpr, int32_t* d_ncase, void* stream) {
    if (n < 0 || (n > 0 && (!d_u0 || !d_mu0 || !d_muf || !d_spin || !d_L || !d_Q2 || !d_su || !d_sm ||
                            !d_tpm || !d_uf)))
        return NP_EINVAL;
    if (n == 0) return NP_OK;
    hipLaunchKernelGGL(np::geor_kernel, dim3(np::grid_for(n)), dim3(256), 0, (hipStream_t)stream, d_u0, d_mu0,
                       d_muf, d_spin, d_L, d_Q2, d_su, d_sm, d_tpm, d_uf_in, n, d_uf, d_imu, d_tpr, d_ncase);
    return hipGetLastError() == hipSuccess ? NP_OK : NP_ELAUNCH;
}

int np_carlson_batch(int kind, const double* d_x, const double* d_y, const double* d_z,
                     const double* d_p, int64_t n, double* d_out, void* stream) {
    if (kind < NP_CARLSON_RF || kind > NP_CARLSON_RJ || n < 0) return NP_EINVAL;
    if (n > 0 && (!d_x || !d_y || !d_out)) return NP_EINVAL;
    if (n > 0 && kind != NP_CARLSON_RC && !d_z) return NP_EINVAL;
    if (n > 0 && kind == NP_CARLSON_RJ && !d_p) return NP_EINVAL;
    if (n == 0) return NP_OK;
    hipLaunchKernelGGL(np::carlson_kernel, dim3(np::grid_for(n)), dim3(256), 0, (hipStream_t)stream,
                       kind, d_x, d_y, d_z, d_p, n, d_out);
    return hipGetLastError() == hipSuccess ? NP_OK : NP_ELAUNCH;
}

}  // extern "C"
