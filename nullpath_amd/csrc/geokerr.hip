// geokerr.hip -- gfx950 GEOKERR driver and the geokerr_ref_api entry points
// (geokerr_port/reference/geokerr_eval_lambda.f90, GEOKERR driver
// geokerr_original.f:357-604).  geokerr_eval_lambda's configuration
// (USEGEOR = .false., PHIT = .true., MUFILL with NPTS = KEXT = 0) runs
// through the split kernels below; np_geokerr_batch runs every
// configuration of the driver, one lane per geodesic (gk_full_*_kernel).
//
// The Fortran samples one geodesic sequentially: a first GEOMU on (u0, uf)
// fixes the case, roots and the u0/mu0 pieces; every later GEOMU/GEOPHITIME
// pair reads those and writes nothing a later sample reads.  So the work
// splits into
//   1. gk_setup_kernel   one lane per geodesic: the first GEOMU, the
//                        driver's case logic (TPR, UB, DU, KMAX) and the
//                        FIRSTPT GEOPHITIME pieces of sample 1 -> GkRecord;
//   2. gk_sample_kernel  one lane per (geodesic, sample): GEOMU + GEOPHITIME
//                        with FIRSTPT = false from the record, NUP x n lanes;
//   3. gk_resample_kernel  one 256-thread block per geodesic, the samples'
//                        lambda staged in LDS: eval_lambda's linear
//                        resampling onto the lambda grid (binary search when
//                        the samples' lambda is sorted, the reference's
//                        sequential scan otherwise) and the turning-point
//                        indices; gk_periods_kernel (one wave per geodesic)
//                        for estimate_periods' search.
// Each sample reproduces the sequential driver's value: the record holds
// exactly the state the Fortran carries from sample to sample.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "geokerr_ref_api.h"
#include "np_geokerr.hpp"
#include "nullpath.h"

namespace np {
namespace gk {
namespace {

// Per-geodesic state after the driver's first solve.
struct GkRecord {
    GkVars v;
    double u0, uf, uout, mu0, a, l, q2, su, sm, ub, du, offset;
    int32_t tpm0, tpr, kmax, nwritten;
};

struct GkInputs {
    const double *u0, *mu0, *uf, *a, *lz, *q2;
};

// SoA sample arrays of the driver, nup entries per geodesic.
struct GkSamples {
    double *ufi, *mufi, *dti, *dphi, *lami;
    int32_t *tpmi, *tpri;
};

constexpr size_t kAlign = 256;
__host__ __device__ inline size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Workspace: records, then the sample arrays.
__host__ inline size_t workspace_bytes(int32_t nup, int64_t n) {
    const size_t per = (size_t)nup * (size_t)n;
    return align_up(sizeof(GkRecord) * (size_t)n) + 5 * align_up(per * sizeof(double)) +
           2 * align_up(per * sizeof(int32_t));
}

__host__ inline GkSamples carve(void* ws, int32_t nup, int64_t n, GkRecord** rec) {
    char* p = (char*)ws;
    const size_t per = (size_t)nup * (size_t)n;
    *rec = (GkRecord*)p;
    p += align_up(sizeof(GkRecord) * (size_t)n);
    GkSamples s;
    double** d[5] = {&s.ufi, &s.mufi, &s.dti, &s.dphi, &s.lami};
    for (int k = 0; k < 5; ++k) {
        *d[k] = (double*)p;
        p += align_up(per * sizeof(double));
    }
    s.tpmi = (int32_t*)p;
    p += align_up(per * sizeof(int32_t));
    s.tpri = (int32_t*)p;
    return s;
}

// One sample of the driver (1-based k) from the record: first leg for
// k <= KMAX (GEOMU's u0 slot holds U0, tpr = 0), the leg back from the
// turning point otherwise (UF in that slot, the driver's TPR).  The whole
// GEOMU/GEOPHITIME body is inlined into each kernel: as a separate leaf
// function it outgrows the 16-bit branch range, and the long-branch
// sequences the compiler then emits reuse s[30:31], the return-address
// register, so the function returns into its own body.
template <bool FIRSTPT>
__device__ __forceinline__ void driver_sample(GkVars& v, const GkRecord& R, int k, double& un, double& mun,
                                              double& dt, double& dphi, int& tpm, int& tpri, double& lambda) {
    const bool leg0 = (k <= R.kmax);
    const double along = R.uout + ((double)k - R.offset) * R.du;
    un = leg0 ? along : 2.0 * R.ub - along;
    double ustart = leg0 ? R.u0 : R.uf;
    int tprv = leg0 ? 0 : R.tpr;
    double iu;
    tpm = R.tpm0;
    geomu<false, true>(v, ustart, un, false, R.mu0, R.a, R.l, R.q2, tpm, tprv, R.su, R.sm, mun, iu);
    const PhiTime p = geophitime<FIRSTPT>(v, ustart, un, R.mu0, mun, R.a, R.l, R.q2, tpm, tprv, R.su, R.sm, iu);
    dt = p.tmu + p.tu;
    dphi = p.phimu + p.phiu;
    lambda = p.lambda;
    tpri = leg0 ? 0 : 1;
}

// mode 0: geokerr_eval_lambda (uf given, SU = sign(uf - u0));
// mode 1: geokerr_estimate_periods (one variable as U0 and UF, SU = 1).
// The setup and sample kernels are capped at 256 registers (2 waves per
// SIMD): left alone the compiler takes 256 VGPRs plus AGPRs, 1 wave per
// SIMD, and these latency-bound FP64 chains run 4 % slower.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
gk_setup_kernel(GkInputs in, int mode, int32_t nup, int64_t n, GkRecord* __restrict__ rec) {
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g >= n) return;
    GkRecord R{};
    const double u0 = in.u0[g], mu0 = in.mu0[g], a = in.a[g], l = in.lz[g], q2 = in.q2[g];
    const bool alias = (mode == 1);
    const double uf_in = alias ? u0 : in.uf[g];
    const double su = alias ? 1.0 : (((uf_in - u0) >= 0.0) ? 1.0 : -1.0);
    const double sm = (mu0 <= 0.0) ? 1.0 : -1.0;
    const double uout = u0, offset = 0.5;
    const double uplus = 1.0 / (1.0 + sqrt(1.0 - a * a));
    double U0 = u0, UFs = uf_in;
    double& UF = alias ? U0 : UFs;
    int tpm = 0, tpr = 1;
    double muf, iu;
    if (alias) geomu<true, true>(R.v, U0, U0, true, mu0, a, l, q2, tpm, tpr, su, sm, muf, iu);
    else geomu<true, true>(R.v, U0, UF, false, mu0, a, l, q2, tpm, tpr, su, sm, muf, iu);
    const int ncase = R.v.ncase;
    if ((ncase > 2 && ncase < 7) || (ncase == 1 && su < 0.0) || (ncase == 7 && su < 0.0) ||
        (ncase == 2 && su > 0.0) || (ncase == 8 && su > 0.0))
        tpr = 0;
    double ub;
    if (ncase > 2 && ncase < 7) {
        ub = (su > 0.0) ? uplus : 0.0;
    } else if (ncase == 1 || ncase == 7) {
        ub = R.v.u2;
        if (tpr == 1 && UF == ub) UF = uout;
    } else {
        ub = R.v.u3;
        if (tpr == 1 && UF == ub) UF = uout;
    }
    double du = sgn1(ub - uout) * su * ((ub - uout) + (2.0 * tpr - 1.0) * (ub - UF));
    du = du / (double)nup;
    int kmax = 0;
    if (du != 0.0) {
        const double xk = (ub - uout) / du + offset;
        if (!(xk < (double)nup)) kmax = nup;
        else if (xk < -1.0) kmax = -1;
        else kmax = (int)xk;
    }
    R.u0 = U0; R.uf = UF; R.uout = uout; R.mu0 = mu0; R.a = a; R.l = l; R.q2 = q2;
    R.su = su; R.sm = sm; R.ub = ub; R.du = du; R.offset = offset;
    R.tpm0 = tpm; R.tpr = tpr; R.kmax = kmax;
    const bool go = (sgn1(uout - ub) != sgn1(uout - U0)) || (uout == U0);
    R.nwritten = go ? (tpr == 1 ? nup : (kmax > 0 ? kmax : 0)) : 0;
    // FIRSTPT pieces of GEOPHITIME from the driver's first sample: sample 1
    // of the first leg, or of the return leg when NUP = 1.
    if (go && (kmax >= 1 || (tpr == 1 && nup == 1))) {
        double un, mun, dt, dphi, lambda;
        int tpm1, tpri;
        driver_sample<true>(R.v, R, 1, un, mun, dt, dphi, tpm1, tpri, lambda);
    }
    rec[g] = R;
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
gk_sample_kernel(const GkRecord* __restrict__ rec, int32_t nup, int64_t n, GkSamples s) {
    const int64_t total = (int64_t)nup * n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = i / nup;
        const int k = (int)(i - g * nup) + 1;
        const GkRecord& R = rec[g];
        if (k > R.nwritten) continue;
        GkVars v = R.v;
        double un, mun, dt, dphi, lambda;
        int tpm, tpri;
        driver_sample<false>(v, R, k, un, mun, dt, dphi, tpm, tpri, lambda);
        s.ufi[i] = un;
        s.mufi[i] = mun;
        s.dti[i] = dt;
        s.dphi[i] = dphi;
        s.lami[i] = lambda;
        s.tpmi[i] = tpm;
        s.tpri[i] = tpri;
    }
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ int first_lane(uint64_t m) { return m ? __ffsll((unsigned long long)m) - 1 : -1; }

// First grid index i in [0, s) with lam0 + dlam * i >= L, -1 if none.
__device__ inline int first_grid_at_or_after(double lam0, double dlam, int s, double L, int lane) {
    for (int base = 0; base < s; base += 64) {
        const int i = base + lane;
        const bool hit = (i < s) && (lam0 + dlam * (double)i >= L);
        const uint64_t m = ballot(hit);
        if (m) return base + first_lane(m);
    }
    return -1;
}

// Samples of lambda staged in LDS per geodesic by the resampling kernel
// (32 KB); longer sample arrays are searched in place in HBM.
constexpr int kLdsLam = 4096;

// geokerr_eval_lambda's resampling and turning-point indices
// (geokerr_eval_lambda.f90:128-195) for one geodesic per 256-thread block,
// with lam the geodesic's nw driver samples of lambda (LDS or HBM).
__device__ __forceinline__ void resample_geodesic(const double* lam, const GkRecord& R, int nw,
                                                  const GkSamples& smp, size_t o, double t0, double lam0,
                                                  double dlam, int32_t s, size_t so, double* __restrict__ u,
                                                  double* __restrict__ mu, double* __restrict__ t,
                                                  double* __restrict__ phi, double* __restrict__ affine,
                                                  int32_t* tp_out) {
    const int tid = threadIdx.x, lane = tid & 63;
    const double *ufi = smp.ufi + o, *mufi = smp.mufi + o, *dti = smp.dti + o, *dph = smp.dphi + o;
    const double base = (nw > 0) ? lam[0] : 0.0;
    // sorted check: lam[j] <= lam[j+1] for all j (NaN fails it)
    int unsorted = 0;
    for (int j = tid; j + 1 < nw; j += blockDim.x)
        if (!(lam[j] <= lam[j + 1])) unsorted = 1;
    const bool sorted = !__syncthreads_or(unsorted);
    const double* lami = lam;

    auto emit = [&](int i, int j) {
        const double lamk = lam0 + dlam * (double)i;
        double uu, mm, tt, pp;
        if (nw <= 0) {
            uu = mm = tt = pp = NAN;
        } else if (lamk <= (lami[0] - base)) {
            uu = ufi[0]; mm = mufi[0]; tt = t0 + dti[0]; pp = dph[0];
        } else if (lamk >= (lami[nw - 1] - base)) {
            uu = ufi[nw - 1]; mm = mufi[nw - 1]; tt = t0 + dti[nw - 1]; pp = dph[nw - 1];
        } else {
            const int j1 = (j + 1 < nw) ? j + 1 : nw - 1;
            const double x0 = lami[j] - base, x1 = lami[j1] - base;
            const double w = (x1 <= x0 + 1.0e-30) ? 0.0 : (lamk - x0) / (x1 - x0);
            uu = (1.0 - w) * ufi[j] + w * ufi[j1];
            mm = (1.0 - w) * mufi[j] + w * mufi[j1];
            tt = t0 + ((1.0 - w) * dti[j] + w * dti[j1]);
            pp = (1.0 - w) * dph[j] + w * dph[j1];
        }
        if (i == 0) {
            uu = R.u0;
            mm = R.mu0;
        }
        u[so + i] = uu;
        mu[so + i] = mm;
        t[so + i] = tt;
        phi[so + i] = pp;
        affine[so + i] = lamk;
    };

    if (nw <= 0 || sorted) {
        // j = first index whose successor reaches the key (n-1 if none); the
        // scan's running j is the key of the largest grid value seen so far
        for (int i = tid; i < s; i += blockDim.x) {
            int j = 0;
            if (nw > 1) {
                const double lamk = lam0 + dlam * (double)((dlam >= 0.0) ? i : 0);
                const double key = lamk - 1.0e-30;
                int lo = 1, hi = nw;   // first m in [1, nw-1] with lami[m]-base >= key, nw if none
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if ((lami[mid] - base) < key) lo = mid + 1;
                    else hi = mid;
                }
                j = lo - 1;
            }
            emit(i, j);
        }
    } else if (tid == 0) {
        int j = 0;
        for (int i = 0; i < s; ++i) {
            const double lamk = lam0 + dlam * (double)i;
            while (j < nw - 1 && (lami[j + 1] - base) < lamk - 1.0e-30) ++j;
            emit(i, j);
        }
    }

    // first turning-point indices on the grid (wave 0)
    if (tid < 64) {
        int tp[2] = {-1, -1};
        const int32_t* tpa[2] = {smp.tpmi + o, smp.tpri + o};
        for (int which = 0; which < 2; ++which) {
            for (int jb = 0; jb < nw && tp[which] < 0; jb += 64) {
                const int jj = jb + lane;
                uint64_t m = ballot(jj < nw && tpa[which][jj] > 0);
                while (m && tp[which] < 0) {
                    const int jl = first_lane(m);
                    m &= m - 1;
                    const double L = lami[jb + jl] - base;
                    tp[which] = first_grid_at_or_after(lam0, dlam, s, L, lane);
                    if (tp[which] < 0 && sorted) {
                        // later samples have larger lambda: none of them lands either
                        jb = nw;
                        break;
                    }
                }
            }
        }
        if (lane == 0) {
            tp_out[0] = tp[0];
            tp_out[1] = tp[1];
        }
    }
}

__global__ void __launch_bounds__(256)
gk_resample_kernel(const GkRecord* __restrict__ rec, GkSamples smp, int32_t nup, int64_t n,
                   const double* __restrict__ t0a, const double* __restrict__ lam0a,
                   const double* __restrict__ dlama, int32_t s, double* __restrict__ u, double* __restrict__ mu,
                   double* __restrict__ t, double* __restrict__ phi, double* __restrict__ affine,
                   int32_t* __restrict__ tpmi_o, int32_t* __restrict__ tpri_o, int32_t* __restrict__ info) {
    __shared__ double s_lam[kLdsLam];
    const int64_t g = blockIdx.x;
    if (g >= n) return;
    const GkRecord& R = rec[g];
    const int nw = R.nwritten;
    const size_t o = (size_t)g * (size_t)nup;
    const size_t so = (size_t)g * (size_t)s;
    const double t0 = t0a[g], lam0 = lam0a[g], dlam = dlama[g];
    int32_t tp[2];
    if (nw <= kLdsLam) {
        for (int j = threadIdx.x; j < nw; j += blockDim.x) s_lam[j] = smp.lami[o + j];
        __syncthreads();
        resample_geodesic(s_lam, R, nw, smp, o, t0, lam0, dlam, s, so, u, mu, t, phi, affine, tp);
    } else {
        resample_geodesic(smp.lami + o, R, nw, smp, o, t0, lam0, dlam, s, so, u, mu, t, phi, affine, tp);
    }
    if (threadIdx.x == 0) {
        tpmi_o[g] = tp[0];
        tpri_o[g] = tp[1];
        if (info) info[g] = 0;
    }
}

// geokerr_estimate_periods' turning-point search (geokerr_eval_lambda.f90:
// 264-293), one wave per geodesic.
__global__ void __launch_bounds__(256)
gk_periods_kernel(const GkRecord* __restrict__ rec, GkSamples smp, int32_t nup, int64_t n,
                  double* __restrict__ tr, double* __restrict__ ttheta, int32_t* __restrict__ info) {
    const int lane = threadIdx.x & 63;
    const int64_t g = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    if (g >= n) return;
    const int nw = rec[g].nwritten;
    const size_t o = (size_t)g * (size_t)nup;
    // first index with tpmi >= 1 / >= 2 and tpri >= 1 / >= 2: with the
    // running maximum of the reference's scan these are plain first hits
    int idx[4] = {-1, -1, -1, -1};
    for (int jb = 0; jb < nw; jb += 64) {
        const int jj = jb + lane;
        const int32_t tm = (jj < nw) ? smp.tpmi[o + jj] : 0;
        const int32_t trr = (jj < nw) ? smp.tpri[o + jj] : 0;
        const bool hit[4] = {jj < nw && tm >= 1, jj < nw && tm >= 2, jj < nw && trr >= 1, jj < nw && trr >= 2};
        for (int q = 0; q < 4; ++q) {
            const uint64_t m = ballot(hit[q]);
            if (idx[q] < 0 && m) idx[q] = jb + first_lane(m);
        }
    }
    if (lane == 0) {
        const double* lami = smp.lami + o;
        const double base = (nw > 0) ? lami[0] : 0.0;
        tr[g] = (idx[2] >= 0 && idx[3] >= 0) ? 2.0 * fmax(1.0e-18, (lami[idx[3]] - base) - (lami[idx[2]] - base)) : -1.0;
        ttheta[g] = (idx[0] >= 0 && idx[1] >= 0) ? 2.0 * fmax(1.0e-18, (lami[idx[1]] - base) - (lami[idx[0]] - base))
                                                 : -1.0;
        if (info) info[g] = 0;
    }
}

__host__ __device__ inline double phi_omega(double u0, double mu0, double a, double lz) {
    const double r = fmax(1.0e-12, 1.0 / fmax(1.0e-12, u0));
    const double delta = fmax(1.0e-30, r * r - 2.0 * r + a * a);
    const double sin2 = fmax(1.0e-30, 1.0 - mu0 * mu0);
    return a * (((r * r + a * a) - a * lz) / delta) + lz / sin2 - a;
}

__global__ void __launch_bounds__(256)
gk_phi_omega_kernel(const double* __restrict__ u0, const double* __restrict__ mu0, const double* __restrict__ a,
                    const double* __restrict__ lz, int64_t n, double* __restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = phi_omega(u0[i], mu0[i], a[i], lz[i]);
}

// ------------------------ the whole GEOKERR driver ------------------------
// One lane per geodesic running the Fortran's solves in its order
// (geokerr_original.f:357-604), every by-reference argument and shared
// variable carried as the Fortran carries it.  The u-sampled branch runs as
// one loop over its steps (first leg, MUFILL's far-side solve, MUFILL
// samples, the leg back from the turning point) so that GEOMU, GEOR and
// each GEOPHITIME form are inlined once.

__device__ __forceinline__ void full_prefill(const np_geokerr_samples& s, size_t base, int total) {
    const double qnan = __builtin_nan("");
    for (int j = 0; j < total; ++j) {
        s.ufi[base + j] = qnan;
        s.mufi[base + j] = qnan;
        s.dti[base + j] = qnan;
        s.dphi[base + j] = qnan;
        s.lambdai[base + j] = qnan;
        s.tpmi[base + j] = -1;
        s.tpri[base + j] = -1;
    }
}

__device__ __forceinline__ void full_store(const np_geokerr_samples& s, size_t base, int total, int64_t idx1,
                                           double u, double mu, double dt, double dphi, int tpm, int tpr,
                                           double lam) {
    const int64_t j = idx1 - 1;
    if (j < 0 || j >= (int64_t)total) return;
    s.ufi[base + j] = u;
    s.mufi[base + j] = mu;
    s.dti[base + j] = dt;
    s.dphi[base + j] = dphi;
    s.tpmi[base + j] = tpm;
    s.tpri[base + j] = tpr;
    s.lambdai[base + j] = lam;
}

__device__ __forceinline__ void full_finish(const np_geokerr_samples& s, int64_t g, double U0, double UF, double MUF,
                                            int TPM, int TPR, int ncase, int nup, int npts) {
    if (s.u0_out) s.u0_out[g] = U0;
    if (s.uf_out) s.uf_out[g] = UF;
    if (s.muf_out) s.muf_out[g] = MUF;
    if (s.tpm_out) s.tpm_out[g] = TPM;
    if (s.tpr_out) s.tpr_out[g] = TPR;
    if (s.ncase_out) s.ncase_out[g] = ncase;
    if (s.nup_out) s.nup_out[g] = nup;
    if (s.npts_out) s.npts_out[g] = npts;
}

// USEGEOR: samples evenly in mu from MU0 to MUF (:421-467).  GEOR and
// GEOMU always form their PHT pieces: they write only RF values that
// GEOPHITIME alone reads, so PHIT = false changes no position.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
gk_full_geor_kernel(np_geokerr_opts o, np_geokerr_rays in, int64_t n, np_geokerr_samples s) {
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int total = o.nup + o.npts - 2 * o.kext;
    const size_t base = (size_t)g * (size_t)total;
    full_prefill(s, base, total);
    GkVars v{};
    const double mu0 = in.mu0[g], a = in.a[g], l = in.l[g], q2 = in.q2[g], su = in.su[g], sm = in.sm[g];
    const double beta = in.beta ? in.beta[g] : 0.0;
    double U0 = in.u0[g], UF = in.uf[g];
    const double MUF = in.muf[g];
    int TPM = in.tpm[g], TPR = in.tpr[g];
    double IU = 0.0, dt = 0.0, dph = 0.0, lam = 0.0;
    for (int k = 1; k <= o.nup; ++k) {
        const double MUN = mu0 + ((double)k - o.offset) * (MUF - mu0) / (double)o.nup;
        if (k >= 2) TPM = f_int((sgn1(mu0) * sm + 1.0) / 2.0);
        if (mu0 != 0.0 || beta != 0.0) {
            if (k == 1) geor_v<true, true>(v, U0, UF, mu0, MUN, a, l, q2, IU, TPM, TPR, su, sm);
            else geor_v<false, true>(v, U0, UF, mu0, MUN, a, l, q2, IU, TPM, TPR, su, sm);
            if (o.phit) {
                const PhiTime p = geophitime<true>(v, U0, UF, mu0, MUN, a, l, q2, TPM, TPR, su, sm, IU);
                dt = p.tmu + p.tu;
                dph = p.phimu + p.phiu;
                lam = p.lambda;
            }
        } else {
            UF = U0;
        }
        full_store(s, base, total, k, UF, MUN, dt, dph, TPM, TPR, lam);
    }
    full_finish(s, g, U0, UF, MUF, TPM, TPR, v.ncase, o.nup, o.npts);
}

// The u-sampled branch (:468-597) with the MUFILL switch (:539-575).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
gk_full_u_kernel(np_geokerr_opts o, np_geokerr_rays in, int64_t n, np_geokerr_samples s) {
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int total = o.nup + o.npts - 2 * o.kext;
    const size_t base = (size_t)g * (size_t)total;
    full_prefill(s, base, total);
    GkVars v{};
    const double uout = in.uout[g], mu0 = in.mu0[g], a = in.a[g], l = in.l[g], q2 = in.q2[g];
    const double su = in.su[g], sm = in.sm[g], offset = o.offset;
    const int nup = o.nup, kext = o.kext;
    double U0 = in.u0[g], UF = in.uf[g], MUF = in.muf[g];
    int TPM = in.tpm[g], TPR = in.tpr[g];
    const double uplus = 1.0 / (1.0 + sqrt(1.0 - a * a));
    double IU = 0.0, dt = 0.0, dph = 0.0, lam = 0.0;

    // first solve on (U0, UF): case, roots, cached u0 / mu0 pieces
    geomu<true, true>(v, U0, UF, false, mu0, a, l, q2, TPM, TPR, su, sm, MUF, IU);
    const int nc = v.ncase;
    if ((nc > 2 && nc < 7) || (nc == 1 && su < 0.0) || (nc == 7 && su < 0.0) || (nc == 2 && su > 0.0) ||
        (nc == 8 && su > 0.0))
        TPR = 0;
    double ub;
    if (nc > 2 && nc < 7) {
        ub = (su > 0.0) ? uplus : 0.0;
    } else if (nc == 1 || nc == 7) {
        ub = v.u2;
        if (TPR == 1 && UF == ub) UF = uout;
    } else {
        ub = v.u3;
        if (TPR == 1 && UF == ub) UF = uout;
    }
    double du = sgn1(ub - uout) * su * ((ub - uout) + (2.0 * TPR - 1.0) * (ub - UF));
    du = du / (double)nup;
    int kmax = f_int((ub - uout) / du + offset);
    if (kmax > nup) kmax = nup;
    if (du == 0.0) kmax = 0;
    if (!(sgn1(uout - ub) != sgn1(uout - U0) || uout == U0)) {
        full_finish(s, g, U0, UF, MUF, TPM, TPR, v.ncase, 0, o.npts);
        return;
    }
    const bool back = (TPR == 1);
    const bool fill = back && o.mufill;
    const int npts2 = fill ? o.npts - 2 * kext : o.npts;
    const int64_t nfirst = (kmax == 0) ? 0 : (kmax > 1 ? kmax : 1);
    int64_t kfirst = (int64_t)kmax + 1 + kext;
    if (kfirst < 1 - (int64_t)npts2) kfirst = 1 - (int64_t)npts2;
    const int64_t nret = back ? ((int64_t)nup - kfirst + 1 > 0 ? (int64_t)nup - kfirst + 1 : 0) : 0;
    const int64_t nfill = fill ? 1 + (int64_t)o.npts : 0;
    const int64_t nsteps = nfirst + nfill + nret;

    int TPR1 = 0, TPM1 = TPM, tpmstart = 0;
    double UN = 0.0, MUN = 0.0, muminus = 0.0, muplus = 0.0, mustart = mu0;
    for (int64_t it = 0; it < nsteps; ++it) {
        // phase 0: first leg; 1: MUFILL far side; 2: MUFILL samples; 3: back
        int phase;
        int64_t k;
        if (it < nfirst) { phase = 0; k = it + 1; }
        else if (it < nfirst + nfill) { k = it - nfirst; phase = (k == 0) ? 1 : 2; }
        else { phase = 3; k = kfirst + (it - nfirst - nfill); }
        if (phase == 2) {
            indep_muf(muminus, muplus, mustart, MUF, tpmstart, TPM1, sm, (int)k, o.npts, offset, MUN, TPM);
            geor_v<false, true>(v, U0, UN, mu0, MUN, a, l, q2, IU, TPM, TPR, su, sm);
        } else {
            const double kk = (phase == 1) ? (double)((int64_t)kmax + kext + 1) : (double)k;
            UN = (phase == 0) ? uout + (kk - offset) * du : 2.0 * ub - (uout + (kk - offset) * du);
            double us = (phase == 0) ? U0 : UF;
            int tpmv = (phase == 1) ? TPM1 : TPM;
            int tprv = (phase == 0) ? TPR1 : TPR;
            geomu<false, true>(v, us, UN, false, mu0, a, l, q2, tpmv, tprv, su, sm, MUN, IU);
            if (phase == 0) U0 = us; else UF = us;
            if (phase == 1) TPM1 = tpmv; else TPM = tpmv;
            if (phase == 0) TPR1 = tprv; else TPR = tprv;
        }
        if (phase == 1) {
            MUF = MUN;
            findmroots(q2, l, a, mu0, muminus, muplus);
            const int64_t jb = (int64_t)kmax - kext;   // the Fortran reads MUFI(KMAX-KEXT)
            if (jb >= 1 && jb <= (int64_t)total) {
                mustart = s.mufi[base + (size_t)(jb - 1)];
                tpmstart = s.tpmi[base + (size_t)(jb - 1)];
            }
            continue;
        }
        if (o.phit) {
            const bool fp = (phase == 0 && k == 1) || (phase == 3 && nup == 1);
            const double us = (phase == 3) ? UF : U0;
            const int tprv = (phase == 0) ? TPR1 : TPR;
            const PhiTime p = fp ? geophitime<true>(v, us, UN, mu0, MUN, a, l, q2, TPM, tprv, su, sm, IU)
                                 : geophitime<false>(v, us, UN, mu0, MUN, a, l, q2, TPM, tprv, su, sm, IU);
            dt = p.tmu + p.tu;
            dph = p.phimu + p.phiu;
            lam = p.lambda;
        }
        const int64_t idx = (phase == 0) ? k : (phase == 2) ? k + kmax - kext : k + npts2;
        const int tpri = (phase == 0) ? 0 : (phase == 2) ? TPR : 1;
        full_store(s, base, total, idx, UN, MUN, dt, dph, TPM, tpri, lam);
    }
    full_finish(s, g, U0, UF, MUF, TPM, TPR, v.ncase, back ? nup + npts2 : nup, npts2);
}

int cus() {
    int dev = 0, c = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev);
    return c;
}

int blocks_for(int64_t work, int per_block, int cap_per_cu) {
    const int64_t want = (work + per_block - 1) / per_block;
    const int64_t cap = (int64_t)cus() * cap_per_cu;
    return (int)(want < 1 ? 1 : (want < cap ? want : cap));
}

int run_driver(const GkInputs& in, int mode, int32_t nup, int64_t n, void* ws, hipStream_t st, GkRecord** rec,
               GkSamples* smp) {
    *smp = carve(ws, nup, n, rec);
    hipLaunchKernelGGL(gk_setup_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, in, mode, nup, n, *rec);
    if (hipGetLastError() != hipSuccess) return NP_ELAUNCH;
    hipLaunchKernelGGL(gk_sample_kernel, dim3(blocks_for((int64_t)nup * n, 256, 64)), dim3(256), 0, st, *rec, nup,
                       n, *smp);
    return hipGetLastError() == hipSuccess ? NP_OK : NP_ELAUNCH;
}

bool aligned_ws(const void* ws) { return ws && ((uintptr_t)ws % kAlign) == 0; }

// Switch for the driver kernels (setup/sample/resample/periods): when false,
// np_geokerr_eval_lambda_batch / np_geokerr_estimate_periods_batch return
// NP_EDISABLED after validating their arguments and launch nothing.
constexpr bool kDriverEnabled = true;

}  // namespace
}  // namespace gk
}  // namespace np

extern "C" {

int64_t np_geokerr_workspace_bytes(int32_t nup, int64_t n) {
    if (nup <= 0 || n < 0) return -1;
    return (int64_t)np::gk::workspace_bytes(nup, n);
}

int np_geokerr_eval_lambda_batch(const double* d_u0, const double* d_mu0, const double* d_uf, const double* d_a,
                                 const double* d_lz, const double* d_q2, const double* d_t0,
                                 const double* d_lam0, const double* d_dlam, int32_t s, int64_t n, double* d_u,
                                 double* d_mu, double* d_t, double* d_phi, double* d_affine, int32_t* d_tpmi,
                                 int32_t* d_tpri, int32_t* d_info, void* d_workspace, void* stream) {
    using namespace np::gk;
    if (s <= 0 || n < 0) return NP_EINVAL;
    if (n == 0) return NP_OK;
    if (!d_u0 || !d_mu0 || !d_uf || !d_a || !d_lz || !d_q2 || !d_t0 || !d_lam0 || !d_dlam || !d_u || !d_mu ||
        !d_t || !d_phi || !d_affine || !d_tpmi || !d_tpri || !aligned_ws(d_workspace))
        return NP_EINVAL;
    if (!kDriverEnabled) return NP_EDISABLED;
    hipStream_t st = (hipStream_t)stream;
    GkInputs in{d_u0, d_mu0, d_uf, d_a, d_lz, d_q2};
    GkRecord* rec;
    GkSamples smp;
    int rc = run_driver(in, 0, s, n, d_workspace, st, &rec, &smp);
    if (rc != NP_OK) return rc;
    hipLaunchKernelGGL(gk_resample_kernel, dim3((unsigned)n), dim3(256), 0, st, rec, smp, s, n, d_t0,
                       d_lam0, d_dlam, s, d_u, d_mu, d_t, d_phi, d_affine, d_tpmi, d_tpri, d_info);
    return hipGetLastError() == hipSuccess ? NP_OK : NP_ELAUNCH;
}

int np_geokerr_estimate_periods_batch(const double* d_u0, const double* d_mu0, const double* d_a,
                                      const double* d_lz, const double* d_q2, int64_t n, double* d_tr,
                                      double* d_ttheta, int32_t* d_info, void* d_workspace, void* stream) {
    using namespace np::gk;
    if (n < 0) return NP_EINVAL;
    if (n == 0) return NP_OK;
    if (!d_u0 || !d_mu0 || !d_a || !d_lz || !d_q2 || !d_tr || !d_ttheta || !aligned_ws(d_workspace))
        return NP_EINVAL;
    if (!kDriverEnabled) return NP_EDISABLED;
    hipStream_t st = (hipStream_t)stream;
    GkInputs in{d_u0, d_mu0, d_u0, d_a, d_lz, d_q2};
    GkRecord* rec;
    GkSamples smp;
    int rc = run_driver(in, 1, NP_GEOKERR_PERIOD_SAMPLES, n, d_workspace, st, &rec, &smp);
    if (rc != NP_OK) return rc;
    hipLaunchKernelGGL(gk_periods_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, rec, smp,
                       NP_GEOKERR_PERIOD_SAMPLES, n, d_tr, d_ttheta, d_info);
    return hipGetLastError() == hipSuccess ? NP_OK : NP_ELAUNCH;
}

int np_geokerr_phi_omega_batch(const double* d_u0, const double* d_mu0, const double* d_a, const double* d_lz,
                               const double* d_q2, int64_t n, double* d_out, void* stream) {
    using namespace np::gk;
    (void)d_q2;
    if (n < 0) return NP_EINVAL;
    if (n == 0) return NP_OK;
    if (!d_u0 || !d_mu0 || !d_a || !d_lz || !d_out) return NP_EINVAL;
    hipLaunchKernelGGL(gk_phi_omega_kernel, dim3(blocks_for(n, 256, 16)), dim3(256), 0, (hipStream_t)stream, d_u0,
                       d_mu0, d_a, d_lz, n, d_out);
    return hipGetLastError() == hipSuccess ? NP_OK : NP_ELAUNCH;
}

int np_geokerr_batch(const np_geokerr_opts* opts, const np_geokerr_rays* rays, int64_t n,
                     const np_geokerr_samples* out, void* stream) {
    using namespace np::gk;
    if (!opts || !rays || !out || n < 0) return NP_EINVAL;
    const np_geokerr_opts o = *opts;
    if (o.nup < 1 || o.npts < 0 || o.kext < 0) return NP_EINVAL;
    const int64_t total = (int64_t)o.nup + o.npts - 2 * (int64_t)o.kext;
    if (total < 1 || total > INT32_MAX) return NP_EINVAL;
    if (n == 0) return NP_OK;
    const np_geokerr_rays r = *rays;
    const np_geokerr_samples s = *out;
    if (!r.u0 || !r.uf || !r.uout || !r.mu0 || !r.muf || !r.a || !r.l || !r.q2 || !r.su || !r.sm || !r.tpm ||
        !r.tpr)
        return NP_EINVAL;
    if (!s.ufi || !s.mufi || !s.dti || !s.dphi || !s.lambdai || !s.tpmi || !s.tpri) return NP_EINVAL;
    if (n > INT64_MAX / total || (n + 255) / 256 > INT32_MAX) return NP_EINVAL;
    const dim3 grid((unsigned)((n + 255) / 256));
    if (o.usegeor)
        hipLaunchKernelGGL(gk_full_geor_kernel, grid, dim3(256), 0, (hipStream_t)stream, o, r, n, s);
    else
        hipLaunchKernelGGL(gk_full_u_kernel, grid, dim3(256), 0, (hipStream_t)stream, o, r, n, s);
    return hipGetLastError() == hipSuccess ? NP_OK : NP_ELAUNCH;
}

// ---- geokerr_ref_api drop-ins (host arrays, one geodesic, synchronous) ----

namespace {

// Device scratch for one call: inputs, outputs and workspace in one block.
struct OneCall {
    char* base = nullptr;
    ~OneCall() {
        if (base) (void)hipFree(base);
    }
};

int gk_status(hipError_t e) { return e == hipSuccess ? NP_OK : NP_ELAUNCH; }

// One geokerr_eval_lambda on the GPU for host arrays (S > 0): args = {u0,
// mu0, uf, a, Lz, Q2, t0, lam0, dlam}; hosts = {u, mu, t, phi, affine}
// (null entries are not copied back); ints_out = {tpmi, tpri, info}, left
// as passed unless the call succeeds.  Returns an np status.
int eval_lambda_one(const double args[9], int S, double* const hosts[5], int32_t ints_out[3]) {
    using namespace np::gk;
    double in[9];
    for (int k = 0; k < 9; ++k) in[k] = args[k];
    const size_t ws = align_up(workspace_bytes(S, 1));
    const size_t outd = align_up(5 * (size_t)S * sizeof(double));
    const size_t total = ws + align_up(sizeof(in)) + outd + align_up(3 * sizeof(int32_t));
    OneCall m;
    int rc = (hipMalloc((void**)&m.base, total) == hipSuccess) ? NP_OK : NP_EALLOC;
    double* d_in = (double*)(m.base + ws);
    double* d_out = (double*)(m.base + ws + align_up(sizeof(in)));
    int32_t* d_ints = (int32_t*)(m.base + ws + align_up(sizeof(in)) + outd);
    if (rc == NP_OK) rc = gk_status(hipMemcpy(d_in, in, sizeof(in), hipMemcpyHostToDevice));
    if (rc == NP_OK)
        rc = np_geokerr_eval_lambda_batch(d_in + 0, d_in + 1, d_in + 2, d_in + 3, d_in + 4, d_in + 5, d_in + 6,
                                          d_in + 7, d_in + 8, S, 1, d_out, d_out + S, d_out + 2 * (size_t)S,
                                          d_out + 3 * (size_t)S, d_out + 4 * (size_t)S, d_ints, d_ints + 1,
                                          d_ints + 2, m.base, nullptr);
    if (rc == NP_OK) rc = gk_status(hipStreamSynchronize(nullptr));
    for (int k = 0; k < 5 && rc == NP_OK; ++k)
        if (hosts[k]) rc = gk_status(hipMemcpy(hosts[k], d_out + k * (size_t)S, S * sizeof(double), hipMemcpyDeviceToHost));
    int32_t ints[3];
    if (rc == NP_OK) rc = gk_status(hipMemcpy(ints, d_ints, sizeof(ints), hipMemcpyDeviceToHost));
    if (rc == NP_OK)
        for (int k = 0; k < 3; ++k) ints_out[k] = ints[k];
    return rc;
}

// The port-solver entry points: geokerr_eval_lambda with t0 = 0, return
// codes 0 / -1 (bad arguments) / -2 (device allocation) / -3 (launch or
// kernel failure).
int port_stream(double u0, double uf, double mu0, double a, double L, double Q2, double lam0, double dlam, int S,
                double* h_u, double* h_mu, double* h_t, double* h_phi, double* h_affine) {
    if (S <= 0 || !h_u || !h_mu || !h_t || !h_phi || !h_affine) return -1;
    if (!np::gk::kDriverEnabled) return -3;
    const double in[9] = {u0, mu0, uf, a, L, Q2, 0.0, lam0, dlam};
    double* const hosts[5] = {h_u, h_mu, h_t, h_phi, h_affine};
    int32_t ints[3] = {-1, -1, 0};
    const int rc = eval_lambda_one(in, S, hosts, ints);
    return rc == NP_OK ? 0 : rc == NP_EALLOC ? -2 : -3;
}

}  // namespace

void geokerr_eval_lambda(double u0, double mu0, double uf, double a, double Lz, double Q2, double t0, double lam0,
                         double dlam, int S, double* u, double* mu, double* t, double* phi, double* affine,
                         int* tpmi, int* tpri, int* info) {
    using namespace np::gk;
    if (info) *info = 0;
    if (S <= 0) {
        if (info) *info = 1;
        return;
    }
    if (!kDriverEnabled) {
        if (info) *info = NP_GEOKERR_INFO_DEVICE - NP_EDISABLED;
        return;
    }
    const double in[9] = {u0, mu0, uf, a, Lz, Q2, t0, lam0, dlam};
    double* const hosts[5] = {u, mu, t, phi, affine};
    int32_t ints[3] = {-1, -1, 0};
    const int rc = eval_lambda_one(in, S, hosts, ints);
    if (tpmi) *tpmi = ints[0];
    if (tpri) *tpri = ints[1];
    if (info) *info = (rc == NP_OK) ? ints[2] : NP_GEOKERR_INFO_DEVICE - rc;
}

int geokerr_complete_stream_cuda(double u0, double uf, double mu0, double a, double L, double Q2, double lam0,
                                 double dlam, int S, double* h_u, double* h_mu, double* h_t, double* h_phi,
                                 double* h_affine) {
    return port_stream(u0, uf, mu0, a, L, Q2, lam0, dlam, S, h_u, h_mu, h_t, h_phi, h_affine);
}

int geokerr_stream_cuda(double u0, double uf, double mu0, double a, double L, double Q2, double lam0, double dlam,
                        int S, double* h_u, double* h_mu, double* h_t, double* h_phi, double* h_affine) {
    return port_stream(u0, uf, mu0, a, L, Q2, lam0, dlam, S, h_u, h_mu, h_t, h_phi, h_affine);
}

void geokerr_estimate_periods(double u0, double mu0, double a, double Lz, double Q2, double* Tr, double* Ttheta,
                              int* info) {
    using namespace np::gk;
    if (!kDriverEnabled) {
        if (Tr) *Tr = -1.0;
        if (Ttheta) *Ttheta = -1.0;
        if (info) *info = NP_GEOKERR_INFO_DEVICE - NP_EDISABLED;
        return;
    }
    const double in[5] = {u0, mu0, a, Lz, Q2};
    const size_t ws = align_up(workspace_bytes(NP_GEOKERR_PERIOD_SAMPLES, 1));
    const size_t total = ws + align_up(sizeof(in)) + align_up(2 * sizeof(double) + sizeof(int32_t));
    OneCall m;
    int rc = (hipMalloc((void**)&m.base, total) == hipSuccess) ? NP_OK : NP_EALLOC;
    double* d_in = (double*)(m.base + ws);
    double* d_out = (double*)(m.base + ws + align_up(sizeof(in)));
    int32_t* d_info = (int32_t*)(d_out + 2);
    if (rc == NP_OK) rc = gk_status(hipMemcpy(d_in, in, sizeof(in), hipMemcpyHostToDevice));
    if (rc == NP_OK)
        rc = np_geokerr_estimate_periods_batch(d_in + 0, d_in + 1, d_in + 2, d_in + 3, d_in + 4, 1, d_out, d_out + 1,
                                               d_info, m.base, nullptr);
    if (rc == NP_OK) rc = gk_status(hipStreamSynchronize(nullptr));
    double outv[2] = {-1.0, -1.0};
    int32_t inf = 0;
    if (rc == NP_OK) rc = gk_status(hipMemcpy(outv, d_out, sizeof(outv), hipMemcpyDeviceToHost));
    if (rc == NP_OK) rc = gk_status(hipMemcpy(&inf, d_info, sizeof(inf), hipMemcpyDeviceToHost));
    if (Tr) *Tr = outv[0];
    if (Ttheta) *Ttheta = outv[1];
    if (info) *info = (rc == NP_OK) ? inf : NP_GEOKERR_INFO_DEVICE - rc;
}

double geokerr_phi_omega(double u0, double mu0, double a, double Lz, double Q2) {
    (void)Q2;
    return np::gk::phi_omega(u0, mu0, a, Lz);
}

}  // extern "C"
