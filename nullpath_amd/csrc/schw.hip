// schw.hip -- gfx950 Schwarzschild equatorial sweeps.
//
// * np_schw_precompute / precomputeGeodesics(): the reference's FP32
//   lookup-table kernel (schwarzschild_blackhole.cu:44-139, 142-266) with
//   the same per-ray loop, step law, turning-point refinement and status
//   rules.
// * np_schw_sweep: the FP64 "plain method" sweep: the same Hamiltonian
//   right-hand side (bh_schwarzschild.h:124-188) and RK4 (235-295) in FP64,
//   with caller-set capture radius, escape radius, photon-sphere band and
//   step budget.  One ray per lane on a persistent grid with lane-level
//   refill from a wave64 work queue: captured rays (~1e4 steps) and
//   escaping rays (~3e4 steps) interleave without idling lanes.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <cstring>
#include <vector>

#include "bh_types.h"
#include "np_device.hpp"
#include "np_kernels.hpp"
#include "np_math.hpp"
#include "nullpath.h"

namespace np {

namespace {

// ---------------- FP32 reference kernel ----------------

struct GS {
    float t, r, th, ph, pt, pr, pth, pph;
};

__device__ __forceinline__ GS schw_rhs_f(const GS& s, float Rs) {
    GS d;
    const float r = s.r, theta = s.th;
    if (r <= Rs + 0.001f || r <= 0.001f) {
        d.t = d.r = d.th = d.ph = d.pt = d.pr = d.pth = d.pph = 0.0f;
        return d;
    }
    float sin_t, cos_t;
    sincosf(theta, &sin_t, &cos_t);
    const float r2 = r * r;
    const float r3 = r2 * r;
    const float A = fmaxf(1e-6f, 1.0f - Rs / r);
    d.t = (-1.0f / A) * s.pt;
    d.r = A * s.pr;
    d.th = s.pth / r2;
    const float sin2 = fmaxf(1e-6f, sin_t * sin_t);
    d.ph = s.pph / (r2 * sin2);
    d.pt = 0.0f;
    const float dgtt = Rs / (r2 * A * A);
    const float dgrr = Rs / r2;
    const float dgthth = -2.0f / r3;
    const float dgphph = -2.0f / (r3 * sin2);
    d.pr = -0.5f * (dgtt * s.pt * s.pt + dgrr * s.pr * s.pr + dgthth * s.pth * s.pth +
                    dgphph * s.pph * s.pph);
    const float sin_safe = (fabsf(sin_t) < 1e-6f) ? (sin_t >= 0.0f ? 1e-6f : -1e-6f) : sin_t;
    d.pth = (cos_t / (r2 * sin2 * sin_safe)) * s.pph * s.pph;
    d.pph = 0.0f;
    return d;
}

__device__ __forceinline__ GS axpy(const GS& y, float c, const GS& k) {
    GS o;
    o.t = y.t + c * k.t; o.r = y.r + c * k.r; o.th = y.th + c * k.th; o.ph = y.ph + c * k.ph;
    o.pt = y.pt + c * k.pt; o.pr = y.pr + c * k.pr; o.pth = y.pth + c * k.pth; o.pph = y.pph + c * k.pph;
    return o;
}

// integrateGeodesic (bh_schwarzschild.h:235-295)
__device__ __forceinline__ bool schw_rk4_f(GS& s, float Rs, float h) {
    const GS k1 = schw_rhs_f(s, Rs);
    const GS k2 = schw_rhs_f(axpy(s, 0.5f * h, k1), Rs);
    const GS k3 = schw_rhs_f(axpy(s, 0.5f * h, k2), Rs);
    const GS k4 = schw_rhs_f(axpy(s, h, k3), Rs);
    const float h6 = h / 6.0f;
    s.t += h6 * (k1.t + 2.0f * k2.t + 2.0f * k3.t + k4.t);
    s.r += h6 * (k1.r + 2.0f * k2.r + 2.0f * k3.r + k4.r);
    s.th += h6 * (k1.th + 2.0f * k2.th + 2.0f * k3.th + k4.th);
    s.ph += h6 * (k1.ph + 2.0f * k2.ph + 2.0f * k3.ph + k4.ph);
    s.pt += h6 * (k1.pt + 2.0f * k2.pt + 2.0f * k3.pt + k4.pt);
    s.pr += h6 * (k1.pr + 2.0f * k2.pr + 2.0f * k3.pr + k4.pr);
    s.pth += h6 * (k1.pth + 2.0f * k2.pth + 2.0f * k3.pth + k4.pth);
    s.pph += h6 * (k1.pph + 2.0f * k2.pph + 2.0f * k3.pph + k4.pph);
    return s.r > Rs + 0.001f && s.th > 0.001f && s.th < BH_PI - 0.001f;
}

// schwarzschild_rmin_analytic (schwarzschild_blackhole.cu:24-37)
__device__ __host__ bool rmin_analytic(float Rs, float b, float* rmin_out) {
    if (Rs <= 0.0f || b <= 0.0f) return false;
    const double beta = (double)b / (double)Rs;
    const double beta_crit = 0.5 * 3.0 * sqrt(3.0);
    if (beta <= beta_crit) return false;
    const double A = 2.0 * sqrt(beta * beta / 3.0);
    double cosarg = -(3.0 * sqrt(3.0)) / (2.0 * beta);
    if (cosarg < -1.0) cosarg = -1.0;
    if (cosarg > 1.0) cosarg = 1.0;
    const double u = A * cos((1.0 / 3.0) * acos(cosarg));
    *rmin_out = (float)(u * (double)Rs);
    return isfinite(*rmin_out) && *rmin_out > Rs;
}

__global__ void __launch_bounds__(256)
schw_precompute_kernel(const float mass, const float r_start, const float* __restrict__ b_in,
                       float* __restrict__ defl_out, float* __restrict__ rmin_out_arr,
                       int32_t* __restrict__ status_out, const int64_t n) {
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const float Rs = 2.0f * mass;
        const float b = b_in[idx];
        GS s;
        s.r = r_start; s.th = BH_PI / 2.0f; s.ph = 0.0f; s.t = 0.0f;
        const float energy = 1.0f;
        s.pt = -energy; s.pph = b * energy; s.pth = 0.0f;
        const float A = 1.0f - Rs / s.r;
        const float pr2 = (energy * energy) / (A * A) - (s.pph * s.pph) / (A * s.r * s.r);
        s.pr = -sqrtf(fmaxf(0.0f, pr2));
        float min_r = s.r, total = 0.0f, prev_phi = s.ph, prev_r = s.r, prev_pr = s.pr;
        bool escaped = false;
        for (int step = 0; step < BH_MAX_INTEGRATION_STEPS_DEFAULT; step++) {
            const float scale = fminf(100.0f, fmaxf(1.0f, s.r / (0.5f * Rs)));
            const float h = BH_INTEGRATION_STEP_DEFAULT * scale;
            if (!schw_rk4_f(s, Rs, h)) break;
            min_r = fminf(min_r, s.r);
            if (prev_pr < 0.0f && s.pr > 0.0f) {
                const float denom = s.pr - prev_pr;
                if (fabsf(denom) > 1e-12f) {
                    float alpha = (-prev_pr) / denom;
                    alpha = fminf(fmaxf(alpha, 0.0f), 1.0f);
                    min_r = fminf(min_r, prev_r + alpha * (s.r - prev_r));
                }
            }
            prev_r = s.r;
            prev_pr = s.pr;
            float dphi = s.ph - prev_phi;
            if (dphi > BH_PI) dphi -= 2.0f * BH_PI;
            if (dphi < -BH_PI) dphi += 2.0f * BH_PI;
            total += dphi;
            prev_phi = s.ph;
            if (s.r > r_start + 1.0f && s.pr > 0) {
                escaped = true;
                break;
            }
        }
        defl_out[idx] = total;
        float rmin_o = min_r;
        if (escaped) {
            float r_an = 0.0f;
            if (rmin_analytic(Rs, b, &r_an)) rmin_o = r_an;
        }
        rmin_out_arr[idx] = rmin_o;
        int st;
        if (escaped || (s.r > r_start + 1.0f && s.pr > 0)) st = 0;
        else if (s.r <= Rs + 0.01f) st = 1;
        else if (fabsf(min_r - BH_PHOTON_SPHERE_MULTIPLIER * Rs) < 0.01f * Rs) st = 2;
        else st = 3;
        status_out[idx] = st;
    }
}

// ---------------- FP64 plain-method sweep ----------------

struct GD {
    double r, th, ph, pr, pth;   // t is never read; p_t, p_phi are constants
};

// The right-hand side of bh_schwarzschild.h:124-295 in FP64: the same
// terms, with the powers of 1/r, 1/A and 1/sin^2 formed from one reciprocal
// each instead of a division per term, and the polynomial sincos_d.
__device__ __forceinline__ GD schw_rhs_d(const GD& s, double Rs, double pt, double pph) {
    GD d;
    const double r = s.r;
    if (r <= Rs + 0.001 || r <= 0.001) {
        d.r = d.th = d.ph = d.pr = d.pth = 0.0;
        return d;
    }
    double sin_t, cos_t;
    sincos_d(s.th, sin_t, cos_t);
    const double ir = 1.0 / r;
    const double ir2 = ir * ir;
    const double ir3 = ir2 * ir;
    const double A = fmax(1e-6, 1.0 - Rs * ir);
    const double iA = 1.0 / A;
    d.r = A * s.pr;
    d.th = s.pth * ir2;
    const double sin2 = fmax(1e-6, sin_t * sin_t);
    const double is2 = 1.0 / sin2;
    d.ph = pph * ir2 * is2;
    const double dgtt = Rs * ir2 * iA * iA;
    const double dgrr = Rs * ir2;
    const double dgthth = -2.0 * ir3;
    const double dgphph = -2.0 * ir3 * is2;
    d.pr = -0.5 * (dgtt * pt * pt + dgrr * s.pr * s.pr + dgthth * s.pth * s.pth + dgphph * pph * pph);
    const double sin_safe = (fabs(sin_t) < 1e-6) ? (sin_t >= 0.0 ? 1e-6 : -1e-6) : sin_t;
    d.pth = cos_t * ir2 * is2 / sin_safe * pph * pph;
    return d;
}

// H = 1/2 g^{mu nu} p_mu p_nu of the Schwarzschild metric at the state.
__device__ __forceinline__ double schw_hamiltonian_d(const GD& s, double Rs, double pt, double pph) {
    double sin_t, cos_t;
    sincos_d(s.th, sin_t, cos_t);
    const double ir2 = 1.0 / (s.r * s.r);
    const double A = 1.0 - Rs / s.r;
    return 0.5 * (-pt * pt / A + A * s.pr * s.pr + ir2 * s.pth * s.pth + ir2 * pph * pph / (sin_t * sin_t));
}

// Turning radius of an equatorial photon with impact parameter b: the
// largest root of r^3 - b^2 r + b^2 Rs = 0 (exists for b >= b_crit =
// (3 sqrt 3 / 2) Rs), by the trigonometric form of the cubic's roots,
// r = (2b/sqrt 3) cos(arccos(-(3 sqrt 3 / 2) Rs / b) / 3), polished by two
// Newton steps; NaN below b_crit.
__device__ __forceinline__ double schw_turning_radius(double b, double Rs) {
    const double x = -2.598076211353316 * Rs / b;   // -(3 sqrt 3 / 2) Rs / b
    if (!(x >= -1.0)) return NAN;
    double r = 1.1547005383792515 * b * cos(acos(x) / 3.0);   // 2 / sqrt 3
    for (int i = 0; i < 2; ++i) {
        const double f = (r * r - b * b) * r + b * b * Rs;
        const double fp = 3.0 * r * r - b * b;
        if (fp != 0.0) r -= f / fp;
    }
    return r;
}

__device__ __forceinline__ GD axpy_d(const GD& y, double c, const GD& k) {
    GD o;
    o.r = y.r + c * k.r; o.th = y.th + c * k.th; o.ph = y.ph + c * k.ph;
    o.pr = y.pr + c * k.pr; o.pth = y.pth + c * k.pth;
    return o;
}

// DIAG: also the turning radius and the largest |H| per ray (a separate
// instantiation, so the plain sweep keeps its register budget).
template <bool DIAG>
__global__ void __launch_bounds__(256)
schw_sweep_kernel(const double M, const double r0, const double r_cap, const double ps_band,
                  const int64_t max_steps, const double* __restrict__ b_in, const int64_t n,
                  int32_t* __restrict__ cls_out, double* __restrict__ dphi_out,
                  int64_t* __restrict__ steps_out, double* __restrict__ rmin_out,
                  double* __restrict__ hmax_out, unsigned long long* __restrict__ work,
                  uint64_t* __restrict__ total_steps) {
    const double Rs = 2.0 * M;
    const double pt = -1.0;
    GD s;
    double pph = 0.0, b = 0.0, total = 0.0, prev_phi = 0.0, hmax = 0.0;
    int64_t step = 0;
    uint64_t idx = 0, lane_steps = 0;
    bool has = false, done = false;
    while (true) {
        const bool need = !has && !done;
        if (__ballot(need)) {
            const uint64_t k = wave_fetch(need, work);
            if (need) {
                if (k < (uint64_t)n) {
                    idx = k;
                    b = b_in[idx];
                    s.r = r0; s.th = NP_PI_D / 2.0; s.ph = 0.0; s.pth = 0.0;
                    pph = b;
                    const double A = 1.0 - Rs / r0;
                    const double pr2 = 1.0 / (A * A) - (b * b) / (A * r0 * r0);
                    s.pr = -sqrt(fmax(0.0, pr2));
                    total = 0.0;
                    prev_phi = 0.0;
                    hmax = 0.0;
                    step = 0;
                    has = true;
                } else {
                    done = true;
                }
            }
        }
        if (__ballot(has) == 0) break;
        if (has) {
            // one RK4 step with the reference's step law (schwarzschild_blackhole.cu:84-85)
            const double scale = fmin(100.0, fmax(1.0, s.r / (0.5 * Rs)));
            const double h = 0.01 * scale;
            const GD k1 = schw_rhs_d(s, Rs, pt, pph);
            const GD k2 = schw_rhs_d(axpy_d(s, 0.5 * h, k1), Rs, pt, pph);
            const GD k3 = schw_rhs_d(axpy_d(s, 0.5 * h, k2), Rs, pt, pph);
            const GD k4 = schw_rhs_d(axpy_d(s, h, k3), Rs, pt, pph);
            const double h6 = h / 6.0;
            s.r += h6 * (k1.r + 2.0 * k2.r + 2.0 * k3.r + k4.r);
            s.th += h6 * (k1.th + 2.0 * k2.th + 2.0 * k3.th + k4.th);
            s.ph += h6 * (k1.ph + 2.0 * k2.ph + 2.0 * k3.ph + k4.ph);
            s.pr += h6 * (k1.pr + 2.0 * k2.pr + 2.0 * k3.pr + k4.pr);
            s.pth += h6 * (k1.pth + 2.0 * k2.pth + 2.0 * k3.pth + k4.pth);
            const bool ok = s.r > Rs + 0.001 && s.th > 0.001 && s.th < NP_PI_D - 0.001;
            double dphi = s.ph - prev_phi;
            if (dphi > NP_PI_D) dphi -= 2.0 * NP_PI_D;
            if (dphi < -NP_PI_D) dphi += 2.0 * NP_PI_D;
            total += dphi;
            prev_phi = s.ph;
            step += 1;
            if constexpr (DIAG) {
                if (hmax_out && ok) hmax = fmax(hmax, fabs(schw_hamiltonian_d(s, Rs, pt, pph)));
            }
            int st = -1;
            if (!ok || s.r < r_cap) st = 1;
            else if (s.r > r0 && s.pr > 0.0) st = 0;
            else if (step >= max_steps) st = 3;
            if (st >= 0) {
                if (fabs(b - 3.0 * sqrt(3.0) * M) < ps_band) st = 2;
                cls_out[idx] = st;
                dphi_out[idx] = total;
                if (steps_out) steps_out[idx] = step;
                if constexpr (DIAG) {
                    if (rmin_out) rmin_out[idx] = st == 0 ? schw_turning_radius(b, Rs) : NAN;
                    if (hmax_out) hmax_out[idx] = hmax;
                }
                lane_steps += (uint64_t)step;
                has = false;
            }
        }
    }
    if (total_steps) wave_add_u64(total_steps, lane_steps);
}

int grid_for(int64_t n, int per_cu_cap) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int64_t want = (n + 255) / 256;
    const int64_t cap = (int64_t)cus * per_cu_cap;
    return (int)(want < cap ? (want < 1 ? 1 : want) : cap);
}

}  // namespace

}  // namespace np

extern "C" {

int np_schw_precompute(float mass, float r_start, const float* d_b, int64_t n, float* d_defl,
                       float* d_rmin, int32_t* d_status, void* stream) {
    if (n < 0 || !(mass > 0.0f) || (n > 0 && (!d_b || !d_defl || !d_rmin || !d_status)))
        return NP_EINVAL;
    if (n == 0) return NP_OK;
    hipLaunchKernelGGL(np::schw_precompute_kernel, dim3(np::grid_for(n, 32)), dim3(256), 0,
                       (hipStream_t)stream, mass, r_start, d_b, d_defl, d_rmin, d_status, n);
    return hipGetLastError() == hipSuccess ? NP_OK : NP_ELAUNCH;
}

int np_schw_sweep(double M, double r0, double r_capture, double ps_band, int64_t max_steps,
                  const double* d_b, int64_t n, int32_t* d_cls, double* d_dphi, int64_t* d_steps,
                  double* d_rmin, double* d_hmax, uint64_t* d_total_steps, void* stream) {
    if (n < 0 || !(M > 0.0) || !(r0 > 2.0 * M) || max_steps <= 0 ||
        (n > 0 && (!d_b || !d_cls || !d_dphi)))
        return NP_EINVAL;
    if (n == 0) return NP_OK;
    unsigned long long* slot = nullptr;
    int err = np::take_work_slot((hipStream_t)stream, &slot);
    if (err != NP_OK) return err;
    const bool diag = d_rmin || d_hmax;
    const void* fn = diag ? (const void*)np::schw_sweep_kernel<true> : (const void*)np::schw_sweep_kernel<false>;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0) != hipSuccess || per_cu <= 0)
        per_cu = 1;
    if (diag)
        hipLaunchKernelGGL(np::schw_sweep_kernel<true>, dim3(np::grid_for(n, per_cu)), dim3(256), 0,
                           (hipStream_t)stream, M, r0, r_capture, ps_band, max_steps, d_b, n, d_cls,
                           d_dphi, d_steps, d_rmin, d_hmax, slot, d_total_steps);
    else
        hipLaunchKernelGGL(np::schw_sweep_kernel<false>, dim3(np::grid_for(n, per_cu)), dim3(256), 0,
                           (hipStream_t)stream, M, r0, r_capture, ps_band, max_steps, d_b, n, d_cls,
                           d_dphi, d_steps, d_rmin, d_hmax, slot, d_total_steps);
    return hipGetLastError() == hipSuccess ? NP_OK : NP_ELAUNCH;
}

int np_precompute_geodesics(float mass, int num_rays, GeodesicLookupTable* table) {
    if (!table || num_rays < 2 || !(mass > 0.0f)) return NP_EINVAL;
    // impact parameters and start radius of schwarzschild_blackhole.cu:149-185
    const float Rs = SCHWARZSCHILD_MULTIPLIER * mass;
    const float bcrit_Rs = bh_bcrit_from_Rs(Rs);
    float b_min = 0.30f * bcrit_Rs;
    float b_max = 20.0f * bcrit_Rs;
    if (const char* e = std::getenv("BH_B_MIN_RS")) {
        float v = strtof(e, nullptr);
        if (v > 0.f) b_min = v * Rs;
    }
    if (const char* e = std::getenv("BH_B_MAX_RS")) {
        float v = strtof(e, nullptr);
        if (v > 0.f) b_max = v * Rs;
    }
    if (b_min >= b_max) {
        b_min = Rs;
        b_max = 50.0f * Rs;
    }
    std::vector<float> b(num_rays), defl(num_rays), rmin(num_rays), bend(num_rays);
    std::vector<int32_t> st(num_rays);
    for (int i = 0; i < num_rays; i++) {
        const float t = (float)i / (num_rays - 1);
        b[i] = b_min * powf(b_max / b_min, t);
    }
    float r_start = 1000.0f;
    if (1.2f * b_max > r_start) r_start = 1.2f * b_max;

    float *d_b = nullptr, *d_defl = nullptr, *d_rmin = nullptr;
    int32_t* d_st = nullptr;
    const size_t fb = (size_t)num_rays * sizeof(float);
    int err = NP_OK;
    if (hipMalloc(&d_b, fb) != hipSuccess || hipMalloc(&d_defl, fb) != hipSuccess ||
        hipMalloc(&d_rmin, fb) != hipSuccess || hipMalloc(&d_st, (size_t)num_rays * sizeof(int32_t)) != hipSuccess)
        err = NP_EALLOC;
    if (err == NP_OK && hipMemcpy(d_b, b.data(), fb, hipMemcpyHostToDevice) != hipSuccess) err = NP_ELAUNCH;
    if (err == NP_OK) err = np_schw_precompute(mass, r_start, d_b, num_rays, d_defl, d_rmin, d_st, nullptr);
    if (err == NP_OK && hipDeviceSynchronize() != hipSuccess) err = NP_ELAUNCH;
    if (err == NP_OK &&
        (hipMemcpy(defl.data(), d_defl, fb, hipMemcpyDeviceToHost) != hipSuccess ||
         hipMemcpy(rmin.data(), d_rmin, fb, hipMemcpyDeviceToHost) != hipSuccess ||
         hipMemcpy(st.data(), d_st, (size_t)num_rays * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess))
        err = NP_ELAUNCH;
    (void)hipFree(d_b);
    (void)hipFree(d_defl);
    (void)hipFree(d_rmin);
    (void)hipFree(d_st);
    if (err != NP_OK) return err;
    // bending = deflection - pi for escaped rays, NaN otherwise (:250-257)
    for (int i = 0; i < num_rays; ++i) bend[i] = (st[i] == 0) ? defl[i] - PI_F : NAN;
    table->mass = mass;
    table->num_entries = num_rays;
    table->impact_parameters = new float[num_rays];
    table->deflection_angles = new float[num_rays];
    table->closest_approaches = new float[num_rays];
    table->bending_angles = new float[num_rays];
    table->ray_status = new int[num_rays];
    std::memcpy(table->impact_parameters, b.data(), fb);
    std::memcpy(table->deflection_angles, defl.data(), fb);
    std::memcpy(table->closest_approaches, rmin.data(), fb);
    std::memcpy(table->bending_angles, bend.data(), fb);
    std::memcpy(table->ray_status, st.data(), (size_t)num_rays * sizeof(int));
    return NP_OK;
}

void np_free_lookup_table(GeodesicLookupTable* t) {
    if (!t) return;
    delete[] t->impact_parameters;
    delete[] t->deflection_angles;
    delete[] t->closest_approaches;
    delete[] t->bending_angles;
    delete[] t->ray_status;
    t->impact_parameters = t->deflection_angles = t->closest_approaches = t->bending_angles = nullptr;
    t->ray_status = nullptr;
}

}  // extern "C"

// The reference's C++-linkage entry point (schwarzschild_blackhole.cu:142):
// returns a new table whose arrays the caller delete[]s, or nullptr with a
// message on stderr.
GeodesicLookupTable* precomputeGeodesics(float mass, int num_rays) {
    std::printf("Precomputing %d geodesics for black hole mass %.2f...\n", num_rays, mass);
    if (num_rays < 2) {
        std::fprintf(stderr, "num_rays must be >= 2.\n");
        return nullptr;
    }
    const float Rs = SCHWARZSCHILD_MULTIPLIER * mass;
    std::printf("[ref] r_ph = %.6f (%.3f Rs), b_crit = %.6f (%.6f M)\n", bh_r_ph(Rs),
                BH_PHOTON_SPHERE_MULTIPLIER, bh_bcrit_from_Rs(Rs), bh_bcrit_from_M(mass));
    GeodesicLookupTable* table = new GeodesicLookupTable();
    const int err = np_precompute_geodesics(mass, num_rays, table);
    if (err != NP_OK) {
        std::fprintf(stderr, "precomputeGeodesics failed: %s\n", np_status_string(err));
        delete table;
        return nullptr;
    }
    std::printf("Geodesic precomputation complete!\n");
    return table;
}
