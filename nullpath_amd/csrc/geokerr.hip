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
// np_geokerr_batch runs the Fortran's solves (geokerr_original.f:357-604)
// with every by-reference argument and shared variable carried as the
// Fortran carries it.  The steps of one geodesic are written once, as
// u_step (the u-sampled branch) and g_step (USEGEOR), and run in two forms:
//
// * serial (gk_full_u_kernel / gk_full_geor_kernel): one lane walks every
//   step of its geodesic (np_geokerr_batch_serial);
// * split (np_geokerr_batch): gk_fu_setup / gk_fg_setup, one lane per
//   geodesic, run the steps whose results later steps read -- the first
//   solve, the FIRSTPT sample, the last sample of the first leg, MUFILL's
//   far-side solve, the last MUFILL sample and the last sample overall --
//   and store the state each leg starts from; gk_fu_step / gk_fg_step then
//   run every other step in a lane of its own from its leg's state.
//
// The split gives the serial results because a step with FIRSTPT off writes
// only values that it, or its GEOPHITIME, recomputes before reading them:
// each step of a leg reads the state the leg started from.  The exceptions
// are the Fortran's sticky states, which a step detects and reports:
//   - a solve that changes NCASE (GEOR leaving with NCASE = 0 when imu lies
//     beyond the u range or muf in the forbidden mu band; later quartic
//     solves then skip the case analysis);
//   - GEOR's equal-root cubic branch, which reads the previous sample's UF;
//   - the forbidden-region retry editing U0 or UF.
// A geodesic with such a step is flagged, and gk_full_*_kernel<true> then
// reruns the flagged geodesics serially over the same outputs.

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

__device__ __forceinline__ void gap_store(const np_geokerr_samples& s, size_t at) {
    const double qnan = __builtin_nan("");
    s.ufi[at] = qnan;
    s.mufi[at] = qnan;
    s.dti[at] = qnan;
    s.dphi[at] = qnan;
    s.lambdai[at] = qnan;
    s.tpmi[at] = -1;
    s.tpri[at] = -1;
}

// One driver sample: its 1-based output index and values.
struct GkSample {
    int64_t idx;
    double u, mu, dt, dphi, lam;
    int32_t tpm, tpr;
};

__device__ __forceinline__ void full_store(const np_geokerr_samples& s, size_t base, int total, const GkSample& x) {
    const int64_t j = x.idx - 1;
    if (j < 0 || j >= (int64_t)total) return;
    s.ufi[base + j] = x.u;
    s.mufi[base + j] = x.mu;
    s.dti[base + j] = x.dt;
    s.dphi[base + j] = x.dphi;
    s.tpmi[base + j] = x.tpm;
    s.tpri[base + j] = x.tpr;
    s.lambdai[base + j] = x.lam;
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

// GEOR's equal-root cubic branch (u2 = u3 set exactly, rows 1/2): the only
// GEOR branch that reads the UF it is given.
__device__ __forceinline__ bool equal_root_cubic(const GkVars& v) {
    return (v.ncase == 1 || v.ncase == 2) && v.u2 == v.u3;
}

// ---- USEGEOR: samples evenly in mu from MU0 to MUF (:421-467) ----
// GEOR and GEOMU always form their PHT pieces: they write only RF values
// that GEOPHITIME alone reads, so PHIT = false changes no position.

struct GConst {
    double mu0, a, l, q2, su, sm, beta, MUF;
};

struct GState {
    GkVars v;
    double U0, UF, IU, dt, dph, lam;
    int32_t TPM, TPR;
};

struct FgRecord {
    GConst c;
    GState s1;   // after sample 1
    GState se;   // after the first sample that left the quartic case analysis
};

__device__ __forceinline__ void g_init(const np_geokerr_rays& in, int64_t g, GConst& c, GState& st) {
    c.mu0 = in.mu0[g]; c.a = in.a[g]; c.l = in.l[g]; c.q2 = in.q2[g];
    c.su = in.su[g]; c.sm = in.sm[g];
    c.beta = in.beta ? in.beta[g] : 0.0;
    c.MUF = in.muf[g];
    st.v = GkVars{};
    st.U0 = in.u0[g]; st.UF = in.uf[g];
    st.IU = 0.0; st.dt = 0.0; st.dph = 0.0; st.lam = 0.0;
    st.TPM = in.tpm[g]; st.TPR = in.tpr[g];
}

// What a USEGEOR sample did to state later samples read.
enum GEffect : int {
    kNone = 0,      // nothing: later samples start from the same state
    kExit = 1,      // GEOR left the quartic case analysis (NCASE = 0)
    kChained = 2,   // the equal-root cubic branch, an edit of U0, a cubic case change
};

// Sample k (1-based).  FIRST: k may be 1 (GEOR with FIRSTPT on); without
// it the caller guarantees k >= 2.  FP: GEOPHITIME with FIRSTPT on, as the
// Fortran calls it in this branch; off, it reads the pieces sample 1 formed
// (the u0 integrals of rows 1, 2, 7, 8 and the mu0 / turning-point pieces
// of calcphitmu), which depend only on the geodesic and the case: the
// caller guarantees sample 1 ran in the case this sample is in.  Returns
// the sample's GEffect (k >= 2).  ROW6 = false: the caller guarantees the
// geodesic is not in Table 1 row 6 (its code is left out).
template <bool FIRST, bool FP, bool ROW6 = true>
__device__ __forceinline__ int g_step(const np_geokerr_opts& o, const GConst& c, GState& st, int k, GkSample& out) {
    const double MUN = c.mu0 + ((double)k - o.offset) * (c.MUF - c.mu0) / (double)o.nup;
    if (k >= 2) st.TPM = f_int((sgn1(c.mu0) * c.sm + 1.0) / 2.0);
    int effect = kNone;
    if (c.mu0 != 0.0 || c.beta != 0.0) {
        const double u0_in = st.U0;
        const int nc_in = st.v.ncase;
        if (FIRST && k == 1)
            geor_v<true, true>(st.v, st.U0, st.UF, c.mu0, MUN, c.a, c.l, c.q2, st.IU, st.TPM, st.TPR, c.su, c.sm);
        else
            geor_v<false, true, ROW6>(st.v, st.U0, st.UF, c.mu0, MUN, c.a, c.l, c.q2, st.IU, st.TPM, st.TPR, c.su,
                                      c.sm);
        if (k >= 2) {
            const double ee = -c.a * c.a * c.q2;   // GEOR's test for the quartic case
            if (st.U0 != u0_in || equal_root_cubic(st.v)) effect = kChained;
            else if (st.v.ncase != nc_in) effect = (ee != 0.0 && st.v.ncase == 0) ? kExit : kChained;
        }
        if (o.phit) {
            const PhiTime p = geophitime<FP, ROW6>(st.v, st.U0, st.UF, c.mu0, MUN, c.a, c.l, c.q2, st.TPM, st.TPR,
                                                   c.su, c.sm, st.IU);
            st.dt = p.tmu + p.tu;
            st.dph = p.phimu + p.phiu;
            st.lam = p.lambda;
        }
    } else {
        st.UF = st.U0;
    }
    out = GkSample{k, st.UF, MUN, st.dt, st.dph, st.lam, st.TPM, st.TPR};
    return effect;
}

__device__ __forceinline__ void g_finish(const np_geokerr_opts& o, const np_geokerr_samples& s, int64_t g,
                                         const GConst& c, const GState& st) {
    full_finish(s, g, st.U0, st.UF, c.MUF, st.TPM, st.TPR, st.v.ncase, o.nup, o.npts);
}

// Serial walk; ONLY_FLAGGED: the geodesics the split form flagged.
template <bool ONLY_FLAGGED>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
gk_full_geor_kernel(np_geokerr_opts o, np_geokerr_rays in, int64_t n, np_geokerr_samples s,
                    const int32_t* __restrict__ flags) {
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g >= n) return;
    if (ONLY_FLAGGED && !flags[g]) return;
    const int total = o.nup + o.npts - 2 * o.kext;
    const size_t base = (size_t)g * (size_t)total;
    full_prefill(s, base, total);
    GConst c;
    GState st;
    g_init(in, g, c, st);
    GkSample x;
    for (int k = 1; k <= o.nup; ++k) {
        (void)g_step<true, true>(o, c, st, k, x);
        full_store(s, base, total, x);
    }
    g_finish(o, s, g, c, st);
}

// Per geodesic: sample 1 (FIRSTPT on) and the state after it; the final
// state from sample NUP run from there (right unless an earlier sample
// changes the state, which the later kernels then handle).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
gk_fg_setup_kernel(np_geokerr_opts o, np_geokerr_rays in, int64_t n, np_geokerr_samples s, FgRecord* __restrict__ rec,
                   int32_t* __restrict__ chained, int32_t* __restrict__ first_exit) {
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int total = o.nup + o.npts - 2 * o.kext;
    const size_t base = (size_t)g * (size_t)total;
    GConst c;
    GState st;
    g_init(in, g, c, st);
    GkSample x;
    (void)g_step<true, true>(o, c, st, 1, x);
    full_store(s, base, total, x);
    rec[g].c = c;
    rec[g].s1 = st;
    if (o.nup >= 2) (void)g_step<false, false>(o, c, st, o.nup, x);   // its lane stores the sample
    g_finish(o, s, g, c, st);
    chained[g] = 0;
    first_exit[g] = INT32_MAX;
}

// Lane (g, slot): samples k = slot + 1, slot + 1 + width, ... (k >= 2) of
// geodesic g from the state after sample 1.  A sample that leaves the
// quartic case analysis records itself in first_exit[g] (the lowest wins);
// a chained state marks the geodesic for the serial walk.  ROW6 = false
// runs the geodesics outside Table 1 row 6 (whose sextic root finders
// would set this kernel's register budget) and writes the never-written
// output slots beyond NUP; ROW6 = true runs the row-6 geodesics only.
template <bool ROW6>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
gk_fg_step_kernel(np_geokerr_opts o, int64_t n, int64_t width, np_geokerr_samples s,
                  const FgRecord* __restrict__ rec, int32_t* __restrict__ chained, int32_t* __restrict__ first_exit) {
    const int total = o.nup + o.npts - 2 * o.kext;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n * width;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = i / width;
        const int64_t slot = i - g * width;
        const size_t base = (size_t)g * (size_t)total;
        if (!ROW6)
            for (int64_t j = slot; j < total; j += width)
                if (j >= o.nup) gap_store(s, base + (size_t)j);
        if (slot == 0 || (rec[g].s1.v.ncase == 6) != ROW6) continue;
        const GConst& c = rec[g].c;
        for (int64_t k = slot + 1; k <= o.nup; k += width) {
            GState st = rec[g].s1;
            GkSample x;
            const int effect = g_step<false, false, ROW6>(o, c, st, (int)k, x);
            full_store(s, base, total, x);
            if (effect == kExit) atomicMin(&first_exit[g], (int)k);
            else if (effect == kChained) chained[g] = 1;
        }
    }
}

// Per geodesic whose samples left the quartic case analysis: the state
// after the first such sample e (from the state after sample 1, as its lane
// ran it), and the final state from there.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
gk_fg_exit_kernel(np_geokerr_opts o, int64_t n, np_geokerr_samples s, FgRecord* __restrict__ rec,
                  const int32_t* __restrict__ chained, const int32_t* __restrict__ first_exit) {
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int e = first_exit[g];
    if (e > o.nup || chained[g]) return;
    const GConst c = rec[g].c;
    GState st = rec[g].s1;
    GkSample x;
    (void)g_step<false, false>(o, c, st, e, x);
    rec[g].se = st;
    if (e < o.nup) (void)g_step<false, false>(o, c, st, o.nup, x);
    g_finish(o, s, g, c, st);
}

// Lane (g, slot): the samples after the first exit e, from the state after e.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
gk_fg_restep_kernel(np_geokerr_opts o, int64_t n, int64_t width, np_geokerr_samples s,
                    const FgRecord* __restrict__ rec, int32_t* __restrict__ chained,
                    const int32_t* __restrict__ first_exit) {
    const int total = o.nup + o.npts - 2 * o.kext;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n * width;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = i / width;
        const int64_t slot = i - g * width;
        const int e = first_exit[g];
        if (e > o.nup || chained[g]) continue;
        const size_t base = (size_t)g * (size_t)total;
        const GConst& c = rec[g].c;
        for (int64_t k = slot + 1; k <= o.nup; k += width) {
            if (k <= e) continue;
            GState st = rec[g].se;   // NCASE = 0: no case's u integrals run
            GkSample x;
            const int effect = g_step<false, false, false>(o, c, st, (int)k, x);
            full_store(s, base, total, x);
            if (effect != kNone) chained[g] = 1;
        }
    }
}

// ---- the u-sampled branch (:468-597) with the MUFILL switch (:539-575) ----

// Per geodesic, fixed by the first solve.
struct UConst {
    double uout, mu0, a, l, q2, su, sm, ub, du;
    int64_t nfirst, nfill, nret, kfirst, nsteps;
    int32_t kmax, npts2, go, back, fill, pad;
};

// The serial loop's variables.
struct UState {
    GkVars v;
    double U0, UF, MUF, UN, MUN, IU, dt, dph, lam, muminus, muplus, mustart;
    int32_t TPM, TPR, TPR1, TPM1, tpmstart, pad;
};

struct FuRecord {
    UConst c;
    UState s0;   // after the first leg's first sample (the FIRSTPT GEOPHITIME)
    UState s1;   // after MUFILL's far-side solve
    UState s2;   // where the leg back starts: after the last MUFILL sample, or the last first-leg sample
};

// The first solve on (U0, UF) -- case, roots, cached u0 / mu0 pieces -- and
// the driver's sample counts (:468-537).
__device__ __forceinline__ void u_first_solve(const np_geokerr_opts& o, const np_geokerr_rays& in, int64_t g,
                                              UConst& c, UState& st) {
#pragma clang fp contract(off)   // DU as the Fortran rounds it (see u_step)
    c = UConst{};
    st = UState{};
    c.uout = in.uout[g]; c.mu0 = in.mu0[g]; c.a = in.a[g]; c.l = in.l[g]; c.q2 = in.q2[g];
    c.su = in.su[g]; c.sm = in.sm[g];
    st.U0 = in.u0[g]; st.UF = in.uf[g]; st.MUF = in.muf[g];
    st.TPM = in.tpm[g]; st.TPR = in.tpr[g];
    const int nup = o.nup, kext = o.kext;
    const double uplus = 1.0 / (1.0 + sqrt(1.0 - c.a * c.a));
    geomu<true, true>(st.v, st.U0, st.UF, false, c.mu0, c.a, c.l, c.q2, st.TPM, st.TPR, c.su, c.sm, st.MUF, st.IU);
    const int nc = st.v.ncase;
    if ((nc > 2 && nc < 7) || (nc == 1 && c.su < 0.0) || (nc == 7 && c.su < 0.0) || (nc == 2 && c.su > 0.0) ||
        (nc == 8 && c.su > 0.0))
        st.TPR = 0;
    double ub;
    if (nc > 2 && nc < 7) {
        ub = (c.su > 0.0) ? uplus : 0.0;
    } else if (nc == 1 || nc == 7) {
        ub = st.v.u2;
        if (st.TPR == 1 && st.UF == ub) st.UF = c.uout;
    } else {
        ub = st.v.u3;
        if (st.TPR == 1 && st.UF == ub) st.UF = c.uout;
    }
    double du = sgn1(ub - c.uout) * c.su * ((ub - c.uout) + (2.0 * st.TPR - 1.0) * (ub - st.UF));
    du = du / (double)nup;
    int kmax = f_int((ub - c.uout) / du + o.offset);
    if (kmax > nup) kmax = nup;
    if (du == 0.0) kmax = 0;
    c.ub = ub;
    c.du = du;
    c.kmax = kmax;
    c.go = (sgn1(c.uout - ub) != sgn1(c.uout - st.U0) || c.uout == st.U0) ? 1 : 0;
    if (!c.go) return;
    c.back = (st.TPR == 1) ? 1 : 0;
    c.fill = (c.back && o.mufill) ? 1 : 0;
    c.npts2 = c.fill ? o.npts - 2 * kext : o.npts;
    c.nfirst = (kmax == 0) ? 0 : (kmax > 1 ? kmax : 1);
    c.kfirst = (int64_t)kmax + 1 + kext;
    if (c.kfirst < 1 - (int64_t)c.npts2) c.kfirst = 1 - (int64_t)c.npts2;
    c.nret = c.back ? ((int64_t)nup - c.kfirst + 1 > 0 ? (int64_t)nup - c.kfirst + 1 : 0) : 0;
    c.nfill = c.fill ? 1 + (int64_t)o.npts : 0;
    c.nsteps = c.nfirst + c.nfill + c.nret;
    st.TPR1 = 0;
    st.TPM1 = st.TPM;
    st.tpmstart = 0;
    st.mustart = c.mu0;
}

// Step `it` of the u-sampled loop.  Phase 0: the first leg; 1: MUFILL's
// far-side solve; 2: MUFILL samples; 3: the leg back from the turning
// point.  Returns the phase; for phases 0, 2 and 3 `out` holds the sample.
// The far-side solve reads MUFI/TPMI(KMAX-KEXT) as they stand at that
// point: jb_in says whether that index lies inside the arrays, jb_mu and
// jb_tpm are its values.  ALLOW_FP: the step may be one with FIRSTPT on in
// its GEOPHITIME (the first sample, or the leg back when NUP = 1).  KIND:
// kUStep or kMuStep when the caller guarantees a GEOMU step (phases 0, 1,
// 3) or a GEOR step (phase 2), so that only that solver is inlined.
enum UStepKind : int { kAnyStep = 0, kUStep = 1, kMuStep = 2 };

template <bool ALLOW_FP, int KIND = kAnyStep, bool ROW6 = true>
__device__ __forceinline__ int u_step(const np_geokerr_opts& o, const UConst& c, UState& st, int64_t it, bool jb_in,
                                      double jb_mu, int jb_tpm, GkSample& out, bool& sticky) {
    // The sample positions UN = UOUT + (K - OFFSET) DU are rounded as the
    // Fortran rounds them (no fused multiply-add): a sample placed on the u
    // turning point enters the integrals through sqrt(u2 - u), where one ulp
    // of UN moves phi by ~1e-8, so every kernel running this step has to
    // place it identically.
#pragma clang fp contract(off)
    int phase;
    int64_t k;
    if (it < c.nfirst) { phase = 0; k = it + 1; }
    else if (it < c.nfirst + c.nfill) { k = it - c.nfirst; phase = (k == 0) ? 1 : 2; }
    else { phase = 3; k = c.kfirst + (it - c.nfirst - c.nfill); }
    const int nc_in = st.v.ncase;
    bool mu_step;
    if constexpr (KIND == kMuStep) mu_step = true;
    else if constexpr (KIND == kUStep) mu_step = false;
    else mu_step = (phase == 2);
    if (mu_step) {
        const double u0_in = st.U0;
        indep_muf(st.muminus, st.muplus, st.mustart, st.MUF, st.tpmstart, st.TPM1, c.sm, (int)k, o.npts, o.offset,
                  st.MUN, st.TPM);
        geor_v<false, true, ROW6>(st.v, st.U0, st.UN, c.mu0, st.MUN, c.a, c.l, c.q2, st.IU, st.TPM, st.TPR, c.su,
                                  c.sm);
        if (st.U0 != u0_in || equal_root_cubic(st.v)) sticky = true;
    } else {
        const double kk = (phase == 1) ? (double)((int64_t)c.kmax + o.kext + 1) : (double)k;
        st.UN = (phase == 0) ? c.uout + (kk - o.offset) * c.du : 2.0 * c.ub - (c.uout + (kk - o.offset) * c.du);
        double us = (phase == 0) ? st.U0 : st.UF;
        const double us_in = us;
        int tpmv = (phase == 1) ? st.TPM1 : st.TPM;
        int tprv = (phase == 0) ? st.TPR1 : st.TPR;
        geomu<false, true, false, ROW6>(st.v, us, st.UN, false, c.mu0, c.a, c.l, c.q2, tpmv, tprv, c.su, c.sm, st.MUN,
                                        st.IU);
        if (us != us_in) sticky = true;
        if (phase == 0) st.U0 = us; else st.UF = us;
        if (phase == 1) st.TPM1 = tpmv; else st.TPM = tpmv;
        if (phase == 0) st.TPR1 = tprv; else st.TPR = tprv;
    }
    if (st.v.ncase != nc_in) sticky = true;
    if (phase == 1) {
        st.MUF = st.MUN;
        findmroots(c.q2, c.l, c.a, c.mu0, st.muminus, st.muplus);
        if (jb_in) {
            st.mustart = jb_mu;
            st.tpmstart = jb_tpm;
        }
        return phase;
    }
    if (o.phit) {
        const bool fp = ALLOW_FP && ((phase == 0 && k == 1) || (phase == 3 && o.nup == 1));
        const double us = (phase == 3) ? st.UF : st.U0;
        const int tprv = (phase == 0) ? st.TPR1 : st.TPR;
        PhiTime p;
        if (ALLOW_FP && fp)
            p = geophitime<true, ROW6>(st.v, us, st.UN, c.mu0, st.MUN, c.a, c.l, c.q2, st.TPM, tprv, c.su, c.sm,
                                       st.IU);
        else
            p = geophitime<false, ROW6>(st.v, us, st.UN, c.mu0, st.MUN, c.a, c.l, c.q2, st.TPM, tprv, c.su, c.sm,
                                        st.IU);
        st.dt = p.tmu + p.tu;
        st.dph = p.phimu + p.phiu;
        st.lam = p.lambda;
    }
    out.idx = (phase == 0) ? k : (phase == 2) ? k + c.kmax - o.kext : k + c.npts2;
    out.u = st.UN;
    out.mu = st.MUN;
    out.dt = st.dt;
    out.dphi = st.dph;
    out.lam = st.lam;
    out.tpm = st.TPM;
    out.tpr = (phase == 0) ? 0 : (phase == 2) ? st.TPR : 1;
    return phase;
}

__device__ __forceinline__ void u_finish(const np_geokerr_opts& o, const np_geokerr_samples& s, int64_t g,
                                         const UConst& c, const UState& st) {
    if (!c.go) full_finish(s, g, st.U0, st.UF, st.MUF, st.TPM, st.TPR, st.v.ncase, 0, o.npts);
    else full_finish(s, g, st.U0, st.UF, st.MUF, st.TPM, st.TPR, st.v.ncase, c.back ? o.nup + c.npts2 : o.nup,
                     c.npts2);
}

// Output index ranges of the MUFILL samples and of the leg back.
__device__ __forceinline__ bool in_fill_range(const np_geokerr_opts& o, const UConst& c, int64_t idx) {
    return c.fill && o.npts >= 1 && idx >= 1 + (int64_t)c.kmax - o.kext &&
           idx <= (int64_t)o.npts + c.kmax - o.kext;
}
__device__ __forceinline__ bool in_back_range(const np_geokerr_opts& o, const UConst& c, int64_t idx) {
    return c.nret >= 1 && idx >= c.kfirst + c.npts2 && idx <= (int64_t)o.nup + c.npts2;
}

// Whether a sample of `phase` at idx is the last write there in the serial
// order (later phases overwrite earlier ones; within a phase the indices
// differ).
__device__ __forceinline__ bool u_last_writer(const np_geokerr_opts& o, const UConst& c, int phase, int64_t idx) {
    if (phase == 3) return true;
    if (in_back_range(o, c, idx)) return false;
    return phase == 2 || !in_fill_range(o, c, idx);
}

// Output slot idx (1-based) written by no step.
__device__ __forceinline__ bool u_gap(const np_geokerr_opts& o, const UConst& c, int64_t idx) {
    if (!c.go) return true;
    return !(idx <= c.nfirst || in_fill_range(o, c, idx) || in_back_range(o, c, idx));
}

// Serial walk; ONLY_FLAGGED: the geodesics the split form flagged.
template <bool ONLY_FLAGGED>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
gk_full_u_kernel(np_geokerr_opts o, np_geokerr_rays in, int64_t n, np_geokerr_samples s,
                 const int32_t* __restrict__ flags) {
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g >= n) return;
    if (ONLY_FLAGGED && !flags[g]) return;
    const int total = o.nup + o.npts - 2 * o.kext;
    const size_t base = (size_t)g * (size_t)total;
    full_prefill(s, base, total);
    UConst c;
    UState st;
    u_first_solve(o, in, g, c, st);
    bool sticky = false;
    GkSample x;
    for (int64_t it = 0; c.go && it < c.nsteps; ++it) {
        // the far-side solve reads MUFI/TPMI(KMAX-KEXT) as the loop left them
        const int64_t jb = (int64_t)c.kmax - o.kext;
        const bool jb_in = (it == c.nfirst && c.fill) && jb >= 1 && jb <= (int64_t)total;
        const double jb_mu = jb_in ? s.mufi[base + (size_t)(jb - 1)] : 0.0;
        const int jb_tpm = jb_in ? s.tpmi[base + (size_t)(jb - 1)] : 0;
        if (u_step<true>(o, c, st, it, jb_in, jb_mu, jb_tpm, x, sticky) != 1) full_store(s, base, total, x);
    }
    u_finish(o, s, g, c, st);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
gk_fu_setup_kernel(np_geokerr_opts o, np_geokerr_rays in, int64_t n, np_geokerr_samples s, FuRecord* __restrict__ rec,
                   int32_t* __restrict__ flags) {
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int total = o.nup + o.npts - 2 * o.kext;
    const size_t base = (size_t)g * (size_t)total;
    UConst c;
    UState st;
    u_first_solve(o, in, g, c, st);
    bool sticky = false;
    GkSample x;
    rec[g].c = c;
    if (!c.go) {
        u_finish(o, s, g, c, st);
        flags[g] = 0;
        return;
    }
    // the first leg: sample 1 here (FIRSTPT), samples 2.. in their lanes
    double mu_1 = 0.0;
    int tpm_1 = 0;
    if (c.nfirst >= 1) {
        u_step<true>(o, c, st, 0, false, 0.0, 0, x, sticky);
        if (u_last_writer(o, c, 0, x.idx)) full_store(s, base, total, x);
        mu_1 = x.mu;
        tpm_1 = x.tpm;
    }
    rec[g].s0 = st;
    UState last = st;   // after the first leg's last sample
    if (c.nfirst >= 2) u_step<false>(o, c, last, c.nfirst - 1, false, 0.0, 0, x, sticky);
    UState back = last;
    if (c.fill) {
        // MUFI/TPMI(KMAX-KEXT) when the far-side solve runs: written by the
        // first leg's sample KMAX-KEXT, else still the prefilled NaN / -1
        const int64_t jb = (int64_t)c.kmax - o.kext;
        const bool jb_in = jb >= 1 && jb <= (int64_t)total;
        double jb_mu = __builtin_nan("");
        int jb_tpm = -1;
        if (jb_in && jb <= c.nfirst) {
            if (jb == 1) {
                jb_mu = mu_1;
                jb_tpm = tpm_1;
            } else {
                UState t = rec[g].s0;
                u_step<false>(o, c, t, jb - 1, false, 0.0, 0, x, sticky);
                jb_mu = x.mu;
                jb_tpm = x.tpm;
            }
        }
        UState far = last;
        u_step<true>(o, c, far, c.nfirst, jb_in, jb_mu, jb_tpm, x, sticky);
        rec[g].s1 = far;
        back = far;
        if (o.npts >= 1) u_step<false>(o, c, back, c.nfirst + o.npts, false, 0.0, 0, x, sticky);
    }
    rec[g].s2 = back;
    UState fin = back;
    if (c.nret >= 1 && o.nup == 1) {
        // NUP = 1: every sample of the leg back has FIRSTPT on in its
        // GEOPHITIME, so the leg runs here, in order
        for (int64_t it = c.nfirst + c.nfill; it < c.nsteps; ++it) {
            u_step<true>(o, c, fin, it, false, 0.0, 0, x, sticky);
            full_store(s, base, total, x);
        }
    } else if (c.nret >= 1) {
        // the last step overall, for the final state (its lane stores it)
        u_step<true>(o, c, fin, c.nsteps - 1, false, 0.0, 0, x, sticky);
    }
    u_finish(o, s, g, c, fin);
    flags[g] = sticky ? 1 : 0;
}

// MU = false: lane (g, slot) runs steps slot, slot + width, ... of geodesic
// g on the first leg and the leg back (GEOMU) that the setup did not run,
// each from the state its leg started from, and writes the output slots no
// step writes (width >= the row length).  MU = true: lane (g, slot) runs
// MUFILL samples slot + 1, slot + 1 + width, ... (GEOR) from the state
// after the far-side solve.  ROW6 = false runs the geodesics outside Table 1
// row 6 (and writes the gaps); ROW6 = true the row-6 geodesics only.
template <bool MU, bool ROW6>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
gk_fu_step_kernel(np_geokerr_opts o, int64_t n, int64_t width, np_geokerr_samples s,
                  const FuRecord* __restrict__ rec, int32_t* __restrict__ flags) {
    const int total = o.nup + o.npts - 2 * o.kext;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n * width;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = i / width;
        const int64_t slot = i - g * width;
        const size_t base = (size_t)g * (size_t)total;
        const UConst& c = rec[g].c;
        if (!MU && !ROW6)
            for (int64_t j = slot; j < total; j += width)
                if (u_gap(o, c, j + 1)) gap_store(s, base + (size_t)j);
        if (!c.go || (rec[g].s0.v.ncase == 6) != ROW6) continue;
        bool sticky = false;
        if (MU) {
            for (int64_t k = slot + 1; c.fill && k <= o.npts; k += width) {
                UState st = rec[g].s1;
                GkSample x;
                u_step<false, kMuStep, ROW6>(o, c, st, c.nfirst + k, false, 0.0, 0, x, sticky);
                if (u_last_writer(o, c, 2, x.idx)) full_store(s, base, total, x);
            }
        } else {
            for (int64_t it = slot; it < c.nsteps; it += width) {
                const int phase = (it < c.nfirst) ? 0 : (it < c.nfirst + c.nfill) ? ((it == c.nfirst) ? 1 : 2) : 3;
                if (phase == 1 || phase == 2 || (phase == 0 && it == 0) || (phase == 3 && o.nup == 1)) continue;
                UState st = (phase == 0) ? rec[g].s0 : rec[g].s2;
                GkSample x;
                u_step<false, kUStep, ROW6>(o, c, st, it, false, 0.0, 0, x, sticky);
                if (u_last_writer(o, c, phase, x.idx)) full_store(s, base, total, x);
            }
        }
        if (sticky) flags[g] = 1;
    }
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

namespace {

// Argument checks shared by both forms of the driver.
int driver_args(const np_geokerr_opts* opts, const np_geokerr_rays* rays, int64_t n, const np_geokerr_samples* out) {
    if (!opts || !rays || !out || n < 0) return NP_EINVAL;
    const np_geokerr_opts o = *opts;
    if (o.nup < 1 || o.npts < 0 || o.kext < 0) return NP_EINVAL;
    const int64_t total = (int64_t)o.nup + o.npts - 2 * (int64_t)o.kext;
    if (total < 1 || total > INT32_MAX) return NP_EINVAL;
    if (n == 0) return NP_OK;
    const np_geokerr_rays& r = *rays;
    const np_geokerr_samples& s = *out;
    if (!r.u0 || !r.uf || !r.uout || !r.mu0 || !r.muf || !r.a || !r.l || !r.q2 || !r.su || !r.sm || !r.tpm ||
        !r.tpr)
        return NP_EINVAL;
    if (!s.ufi || !s.mufi || !s.dti || !s.dphi || !s.lambdai || !s.tpmi || !s.tpri) return NP_EINVAL;
    // the split form's lanes: n x (NUP + NPTS + 2)
    const int64_t width = (int64_t)o.nup + o.npts + 2;
    if (n > INT64_MAX / width || (n + 255) / 256 > INT32_MAX) return NP_EINVAL;
    return NP_OK;
}

void launch_serial(const np_geokerr_opts& o, const np_geokerr_rays& r, int64_t n, const np_geokerr_samples& s,
                   hipStream_t st) {
    using namespace np::gk;
    const dim3 grid((unsigned)((n + 255) / 256));
    const int32_t* none = nullptr;
    if (o.usegeor) hipLaunchKernelGGL(gk_full_geor_kernel<false>, grid, dim3(256), 0, st, o, r, n, s, none);
    else hipLaunchKernelGGL(gk_full_u_kernel<false>, grid, dim3(256), 0, st, o, r, n, s, none);
}

// Workspace of the split form: per geodesic two flags, then its record.
size_t split_workspace_bytes(const np_geokerr_opts& o, int64_t n) {
    using namespace np::gk;
    const size_t rec = o.usegeor ? sizeof(FgRecord) : sizeof(FuRecord);
    return 2 * align_up(sizeof(int32_t) * (size_t)n) + rec * (size_t)n;
}

// The split form: setup (one lane per geodesic), steps (one lane per
// geodesic and step slot), for USEGEOR the restart after a sample that left
// the quartic case analysis, then the serial walk over flagged geodesics.
int launch_split(const np_geokerr_opts& o, const np_geokerr_rays& r, int64_t n, const np_geokerr_samples& s,
                 void* ws, hipStream_t st) {
    using namespace np::gk;
    const dim3 grid((unsigned)((n + 255) / 256));
    const int64_t total = (int64_t)o.nup + o.npts - 2 * (int64_t)o.kext;
    int32_t* flags = (int32_t*)ws;
    int32_t* first_exit = (int32_t*)((char*)ws + align_up(sizeof(int32_t) * (size_t)n));
    void* rec = (char*)ws + 2 * align_up(sizeof(int32_t) * (size_t)n);
    if (o.usegeor) {
        const int64_t width = total > o.nup ? total : (int64_t)o.nup;
        const dim3 lanes(blocks_for(n * width, 256, 32));
        FgRecord* R = (FgRecord*)rec;
        hipLaunchKernelGGL(gk_fg_setup_kernel, grid, dim3(256), 0, st, o, r, n, s, R, flags, first_exit);
        hipLaunchKernelGGL(gk_fg_step_kernel<false>, lanes, dim3(256), 0, st, o, n, width, s, (const FgRecord*)R,
                           flags, first_exit);
        hipLaunchKernelGGL(gk_fg_step_kernel<true>, lanes, dim3(256), 0, st, o, n, width, s, (const FgRecord*)R,
                           flags, first_exit);
        hipLaunchKernelGGL(gk_fg_exit_kernel, grid, dim3(256), 0, st, o, n, s, R, (const int32_t*)flags,
                           (const int32_t*)first_exit);
        hipLaunchKernelGGL(gk_fg_restep_kernel, lanes, dim3(256), 0, st, o, n, width, s, (const FgRecord*)R, flags,
                           (const int32_t*)first_exit);
        hipLaunchKernelGGL(gk_full_geor_kernel<true>, grid, dim3(256), 0, st, o, r, n, s, (const int32_t*)flags);
    } else {
        const int64_t width = (int64_t)o.nup + o.npts + 2;
        FuRecord* R = (FuRecord*)rec;
        hipLaunchKernelGGL(gk_fu_setup_kernel, grid, dim3(256), 0, st, o, r, n, s, R, flags);
        const dim3 lanes(blocks_for(n * width, 256, 32));
        hipLaunchKernelGGL((gk_fu_step_kernel<false, false>), lanes, dim3(256), 0, st, o, n, width, s,
                           (const FuRecord*)R, flags);
        hipLaunchKernelGGL((gk_fu_step_kernel<false, true>), lanes, dim3(256), 0, st, o, n, width, s,
                           (const FuRecord*)R, flags);
        if (o.mufill && o.npts > 0) {
            const dim3 fill(blocks_for(n * o.npts, 256, 32));
            hipLaunchKernelGGL((gk_fu_step_kernel<true, false>), fill, dim3(256), 0, st, o, n, (int64_t)o.npts, s,
                               (const FuRecord*)R, flags);
            hipLaunchKernelGGL((gk_fu_step_kernel<true, true>), fill, dim3(256), 0, st, o, n, (int64_t)o.npts, s,
                               (const FuRecord*)R, flags);
        }
        hipLaunchKernelGGL(gk_full_u_kernel<true>, grid, dim3(256), 0, st, o, r, n, s, (const int32_t*)flags);
    }
    return hipGetLastError() == hipSuccess ? NP_OK : NP_ELAUNCH;
}

}  // namespace

int np_geokerr_batch(const np_geokerr_opts* opts, const np_geokerr_rays* rays, int64_t n,
                     const np_geokerr_samples* out, void* stream) {
    using namespace np::gk;
    int err = driver_args(opts, rays, n, out);
    if (err != NP_OK || n == 0) return err;
    const np_geokerr_opts o = *opts;
    hipStream_t st = (hipStream_t)stream;
    const size_t bytes = split_workspace_bytes(o, n);
    void* ws = nullptr;
    bool stream_ordered = true;
    if (hipMallocAsync(&ws, bytes, st) != hipSuccess) {
        (void)hipGetLastError();
        stream_ordered = false;
        if (hipMalloc(&ws, bytes) != hipSuccess) return NP_EALLOC;
    }
    err = launch_split(o, *rays, n, *out, ws, st);
    if (stream_ordered) {
        if (hipFreeAsync(ws, st) != hipSuccess && err == NP_OK) err = NP_ELAUNCH;
    } else {
        // without stream-ordered allocation the records must outlive the kernels
        if (hipStreamSynchronize(st) != hipSuccess && err == NP_OK) err = NP_ELAUNCH;
        (void)hipFree(ws);
    }
    return err;
}

int np_geokerr_batch_serial(const np_geokerr_opts* opts, const np_geokerr_rays* rays, int64_t n,
                            const np_geokerr_samples* out, void* stream) {
    int err = driver_args(opts, rays, n, out);
    if (err != NP_OK || n == 0) return err;
    launch_serial(*opts, *rays, n, *out, (hipStream_t)stream);
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
