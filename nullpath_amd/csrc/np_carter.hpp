// np_carter.hpp -- the Kerr geodesic right-hand side in separable (Carter)
// form with the radial force taken on the constraint surface H = 0, and its
// hand-written vector-Jacobian product.  Geometric units with M = 1.
//
//   2 Sigma H = Delta p_r^2 - P^2/Delta + p_th^2 + sin^2 w^2,
//   P = (r^2 + a^2) E - a L,  w = L / sin^2 - a E,  Q = p_th^2 + sin^2 w^2,
//   X = 2 E (r^2 - a^2) - 2 a L (r - 1),  Delta = r^2 - 2 r + a^2,
//   dr/dl   = Delta p_r / Sigma
//   dth/dl  = p_th / Sigma
//   dpr/dl  = ((r - 1) Q / Delta - P X / Delta^2) / Sigma
//   dpth/dl = cos sin w (L / sin^2 + a E) / Sigma.
// (phi and t do not feed back into the motion and are left out.)
#pragma once

#include <cmath>

#include "np_math.hpp"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define NP_HD __host__ __device__
#else
#define NP_HD
#endif

namespace np {

template <typename Real>
struct CState {
    Real r, th, pr, pth;
};

NP_HD inline void carter_sincos(float x, float& s, float& c) { sincos_f(x, s, c); }
NP_HD inline void carter_sincos(double x, double& s, double& c) {
    s = sin(x);
    c = cos(x);
}
// Reciprocals of the field: in float on the device the hardware reciprocal
// (v_rcp_f32, 1 ulp) instead of the IEEE division sequence, on the host
// and in double the division.
NP_HD inline float carter_rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
NP_HD inline double carter_rcp(double x) { return 1.0 / x; }
NP_HD inline float carter_fmax(float a, float b) { return fmaxf(a, b); }
NP_HD inline double carter_fmax(double a, double b) { return fmax(a, b); }

// Right-hand side.  rp, rm are the horizon radii 1 +- sqrt(1 - a^2).
template <typename Real>
NP_HD inline CState<Real> carter_rhs(const CState<Real>& y, Real E, Real L, Real a,
                                                   Real rp, Real rm) {
    Real s, c;
    carter_sincos(y.th, s, c);
    const Real r = y.r, a2 = a * a, r2 = r * r;
    const Real s2 = carter_fmax(s * s, Real(1e-12));
    const Real Sigma = r2 + a2 * (c * c);
    const Real Delta = (r - rp) * (r - rm);
    const Real iS = carter_rcp(Sigma), iD = carter_rcp(Delta), is2 = carter_rcp(s2);
    const Real P = (r2 + a2) * E - a * L;
    const Real w = L * is2 - a * E;
    const Real Q = y.pth * y.pth + s2 * (w * w);
    const Real X = Real(2) * E * (r2 - a2) - Real(2) * a * L * (r - Real(1));
    CState<Real> d;
    d.r = Delta * y.pr * iS;
    d.th = y.pth * iS;
    d.pr = ((r - Real(1)) * Q * iD - P * X * (iD * iD)) * iS;
    d.pth = c * s * w * (L * is2 + a * E) * iS;
    return d;
}

// Vector-Jacobian product: given v (weights of dr, dth, dpr, dpth), returns
// u = (df/dy)^T v and accumulates v . df/dE, v . df/dL, v . df/da.
template <typename Real>
NP_HD inline CState<Real> carter_vjp(const CState<Real>& y, Real E, Real L, Real a,
                                                   Real rp, Real rm, const CState<Real>& v,
                                                   Real& gE, Real& gL, Real& ga) {
    Real s, c;
    carter_sincos(y.th, s, c);
    const Real r = y.r, a2 = a * a, r2 = r * r, rm1 = r - Real(1);
    const Real s2 = carter_fmax(s * s, Real(1e-12));
    const Real c2 = c * c;
    const Real Sigma = r2 + a2 * c2;
    const Real Delta = (r - rp) * (r - rm);
    const Real iS = carter_rcp(Sigma), iD = carter_rcp(Delta), is2 = carter_rcp(s2);
    const Real iD2 = iD * iD;
    const Real P = (r2 + a2) * E - a * L;
    const Real w = L * is2 - a * E;
    const Real Z = L * is2 + a * E;
    const Real Q = y.pth * y.pth + s2 * (w * w);
    const Real X = Real(2) * E * (r2 - a2) - Real(2) * a * L * rm1;
    const Real cs = c * s;
    // g_i: numerators of f_i = g_i / Sigma
    const Real g_r = Delta * y.pr;
    const Real g_th = y.pth;
    const Real g_pr = rm1 * Q * iD - P * X * iD2;
    const Real g_pth = cs * w * Z;
    const Real G = v.r * g_r + v.th * g_th + v.pr * g_pr + v.pth * g_pth;
    // partials of Sigma, Delta
    const Real Sig_r = Real(2) * r, Sig_th = Real(-2) * a2 * cs, Sig_a = Real(2) * a * c2;
    const Real Del_r = Real(2) * rm1, Del_a = Real(2) * a;
    const Real iD_r = -iD2 * Del_r, iD_a = -iD2 * Del_a;
    // partials of P, X, Q
    const Real P_r = Real(2) * r * E, P_E = r2 + a2, P_L = -a, P_a = Real(2) * a * E - L;
    const Real X_r = Real(4) * E * r - Real(2) * a * L, X_E = Real(2) * (r2 - a2), X_L = Real(-2) * a * rm1;
    const Real X_a = Real(-4) * E * a - Real(2) * L * rm1;
    const Real Q_th = Real(-2) * cs * w * Z;   // d(sin^2 w^2)/dth
    const Real Q_pth = Real(2) * y.pth, Q_E = Real(-2) * a * s2 * w, Q_L = Real(2) * w;
    const Real Q_a = Real(-2) * E * s2 * w;
    // d g_pr
    const Real gpr_r = Q * iD + rm1 * Q * iD_r - (P_r * X + P * X_r) * iD2 - P * X * Real(2) * iD * iD_r;
    const Real gpr_th = rm1 * iD * Q_th;
    const Real gpr_pth = rm1 * iD * Q_pth;
    const Real gpr_E = rm1 * iD * Q_E - (P_E * X + P * X_E) * iD2;
    const Real gpr_L = rm1 * iD * Q_L - (P_L * X + P * X_L) * iD2;
    const Real gpr_a = rm1 * Q * iD_a + rm1 * iD * Q_a - (P_a * X + P * X_a) * iD2 - P * X * Real(2) * iD * iD_a;
    // d g_pth = d(cs w Z)
    const Real gpth_th = (c2 - s * s) * w * Z - Real(2) * L * c2 * is2 * (Z + w);
    const Real gpth_E = a * cs * (w - Z);
    const Real gpth_L = cs * is2 * (Z + w);
    const Real gpth_a = E * cs * (w - Z);
    // assemble: dF = iS dG - G iS^2 dSigma
    const Real GiS2 = G * iS * iS;
    CState<Real> u;
    u.r = iS * (v.r * Del_r * y.pr + v.pr * gpr_r) - GiS2 * Sig_r;
    u.th = iS * (v.pr * gpr_th + v.pth * gpth_th) - GiS2 * Sig_th;
    u.pr = iS * (v.r * Delta);
    u.pth = iS * (v.th + v.pr * gpr_pth);
    gE += iS * (v.pr * gpr_E + v.pth * gpth_E);
    gL += iS * (v.pr * gpr_L + v.pth * gpth_L);
    ga += iS * (v.r * Del_a * y.pr + v.pr * gpr_a + v.pth * gpth_a) - GiS2 * Sig_a;
    return u;
}

// The same flow in the parameter s with dlambda = r ds: F = r f.  A fixed
// step in s is a geometric step in r far from the hole (dr/ds ~ -r), so a
// fixed-step integration spans r = 1000 to the horizon in a few hundred
// steps while keeping small steps near the photon orbit.
template <typename Real>
NP_HD inline CState<Real> carter_rhs_r(const CState<Real>& y, Real E, Real L, Real a, Real rp, Real rm) {
    const CState<Real> f = carter_rhs(y, E, L, a, rp, rm);
    return CState<Real>{y.r * f.r, y.r * f.th, y.r * f.pr, y.r * f.pth};
}

// VJP of F = r f: (dF/dy)^T v = r (df/dy)^T v + e_r (f . v), and the
// E, L, a partials are those of f weighted by r v.
template <typename Real>
NP_HD inline CState<Real> carter_vjp_r(const CState<Real>& y, Real E, Real L, Real a, Real rp, Real rm,
                                       const CState<Real>& v, Real& gE, Real& gL, Real& ga) {
    const CState<Real> f = carter_rhs(y, E, L, a, rp, rm);
    const CState<Real> rv{y.r * v.r, y.r * v.th, y.r * v.pr, y.r * v.pth};
    CState<Real> u = carter_vjp(y, E, L, a, rp, rm, rv, gE, gL, ga);
    u.r += f.r * v.r + f.th * v.th + f.pr * v.pr + f.pth * v.pth;
    return u;
}

}  // namespace np
