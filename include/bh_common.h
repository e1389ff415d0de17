/*
 * bh_common.h -- constants, ray classification and compensated summation
 * shared with code written against the reference's include/bh_common.h.
 *
 * Same names and values as bh_common.h:50-253 (BH_* constants),
 * :295-331 (RayStatus) and :451-493 (BhKleinSum, bh_rk4_sum), so host code
 * that includes the reference header compiles unchanged against this one.
 * Functions are host/device when compiled by hipcc and plain inline
 * otherwise.  The GPU error macro aborts with the HIP error string.
 */
#ifndef BH_COMMON_H
#define BH_COMMON_H

#include <cmath>
#include <cstdio>
#include <cstdlib>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define BH_HD __host__ __device__
#else
#define BH_HD
#endif

constexpr float BH_PI = 3.14159265358979323846f;
constexpr double BH_PI_D = 3.141592653589793238462643383279502884;
constexpr float BH_EPS = 1e-6f;
constexpr double BH_EPS_D = 1e-12;
constexpr float BH_EPS_SIN = 1e-6f;
constexpr double BH_EPS_SIN_D = 1e-12;
constexpr float BH_SCHWARZSCHILD_MULTIPLIER = 2.0f;
constexpr float BH_PHOTON_SPHERE_MULTIPLIER = 1.5f;
constexpr int BH_MAX_INTEGRATION_STEPS_DEFAULT = 10000;
constexpr float BH_INTEGRATION_STEP_DEFAULT = 0.01f;
constexpr double BH_RK_REL_TOL = 1e-10;
constexpr double BH_RK_ABS_TOL = 1e-12;
constexpr double BH_RK_SAFETY = 0.9;
constexpr double BH_RK_HMIN = 1e-6;
constexpr double BH_RK_HMAX = 0.2;

#ifndef SCHWARZSCHILD_MULTIPLIER
#define SCHWARZSCHILD_MULTIPLIER BH_SCHWARZSCHILD_MULTIPLIER
#endif
#ifndef PHOTON_SPHERE_MULTIPLIER
#define PHOTON_SPHERE_MULTIPLIER BH_PHOTON_SPHERE_MULTIPLIER
#endif
#ifndef MAX_INTEGRATION_STEPS
#define MAX_INTEGRATION_STEPS BH_MAX_INTEGRATION_STEPS_DEFAULT
#endif
#ifndef INTEGRATION_STEP_SIZE
#define INTEGRATION_STEP_SIZE BH_INTEGRATION_STEP_DEFAULT
#endif
#ifndef PI_F
#define PI_F BH_PI
#endif

BH_HD inline float bh_r_s(float M) { return BH_SCHWARZSCHILD_MULTIPLIER * M; }
BH_HD inline float bh_r_ph(float Rs) { return BH_PHOTON_SPHERE_MULTIPLIER * Rs; }
BH_HD inline float bh_bcrit_from_M(float M) { return 3.0f * sqrtf(3.0f) * M; }
BH_HD inline float bh_bcrit_from_Rs(float Rs) { return 0.5f * 3.0f * sqrtf(3.0f) * Rs; }
BH_HD inline float bh_to_M_units(float x, float M) { return x / fmaxf(M, BH_EPS); }
BH_HD inline float bh_to_Rs_units(float x, float Rs) { return x / fmaxf(Rs, BH_EPS); }

/* Ray classification; the integer values are part of the interface. */
enum class RayStatus : int {
    Escaped = 0,
    Captured = 1,
    PhotonSphere = 2,
    Inconclusive = 3
};

/* Abort with the HIP error text when a runtime call fails. */
#if defined(__HIPCC__)
#define BH_HIP_CHECK(expr)                                                                   \
    do {                                                                                     \
        hipError_t _err = (expr);                                                            \
        if (_err != hipSuccess) {                                                            \
            std::fprintf(stderr, "[HIP] %s failed at %s:%d: %s\n", #expr, __FILE__, __LINE__, \
                         hipGetErrorString(_err));                                           \
            std::abort();                                                                    \
        }                                                                                    \
    } while (0)
#endif

BH_HD inline float bh_abs(float x) { return fabsf(x); }
BH_HD inline double bh_abs(double x) { return fabs(x); }
BH_HD inline float bh_clamp(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
BH_HD inline float bh_clamp01(float x) { return bh_clamp(x, 0.0f, 1.0f); }
BH_HD inline float bh_safe_sin2(float theta) {
    float s = sinf(theta);
    return fmaxf(s * s, BH_EPS_SIN);
}
BH_HD inline float bh_safe_div(float num, float den, float eps = BH_EPS) {
    float d = (fabsf(den) < eps) ? copysignf(eps, den) : den;
    return num / d;
}

/* Second-order compensated (Kahan-Babuska-Klein) summation. */
template <typename Real>
struct BhKleinSum {
    Real sum;
    Real cs;
    Real ccs;

    BH_HD inline BhKleinSum() : sum(Real(0)), cs(Real(0)), ccs(Real(0)) {}

    BH_HD inline void add(Real x) {
        Real t = sum + x;
        Real c = (bh_abs(sum) >= bh_abs(x)) ? ((sum - t) + x) : ((x - t) + sum);
        sum = t;
        Real t2 = cs + c;
        Real cc = (bh_abs(cs) >= bh_abs(c)) ? ((cs - t2) + c) : ((c - t2) + cs);
        cs = t2;
        ccs += cc;
    }

    BH_HD inline Real result() const { return sum + cs + ccs; }
};

/* k1 + 2 k2 + 2 k3 + k4 through the compensated sum, in that order. */
template <typename Real>
BH_HD inline Real bh_rk4_sum(Real k1, Real k2, Real k3, Real k4) {
    BhKleinSum<Real> s;
    s.add(k1);
    s.add(Real(2) * k2);
    s.add(Real(2) * k3);
    s.add(k4);
    return s.result();
}

#endif /* BH_COMMON_H */
