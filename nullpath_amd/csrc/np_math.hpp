// np_math.hpp -- host/device FP64 elementary functions of this library.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#ifndef NP_HD
#define NP_HD __host__ __device__
#endif
#else
#ifndef NP_HD
#define NP_HD
#endif
#endif

namespace np {

// sin and cos of a double angle of moderate size (|x| < 1e5; the tracers
// call it with polar angles in (0, pi)): Cody-Waite reduction by pi/2 with
// a three-part split of pi/2, then the Taylor series of sin and cos on
// [-pi/4, pi/4] through x^17 and x^18 (truncation below 1e-19 relative),
// evaluated by Horner's rule.  Error about 1 ulp; no large-argument path,
// so the register footprint stays small.
NP_HD inline void sincos_d(double x, double& s, double& c) {
    const double k = std::rint(x * 0.6366197723675814);                  // 2/pi
    // pi/2 = hi + mid + lo, hi and mid with 33 significant bits, so k*hi
    // and k*mid are exact for |k| < 2^20
    const double r = ((x - k * 1.5707963267341256e+00) - k * 6.077100506303966e-11) -
                     k * 2.0222662487959506e-21;
    const double r2 = r * r;
    // sin r = r + r r^2 P(r^2), P = -1/3! + r^2/5! - ... + r^14/17!
    double ps = 1.0 / 355687428096000.0;            //  1/17!
    ps = ps * r2 - 1.0 / 1307674368000.0;           // -1/15!
    ps = ps * r2 + 1.0 / 6227020800.0;              //  1/13!
    ps = ps * r2 - 1.0 / 39916800.0;                // -1/11!
    ps = ps * r2 + 1.0 / 362880.0;                  //  1/9!
    ps = ps * r2 - 1.0 / 5040.0;                    // -1/7!
    ps = ps * r2 + 1.0 / 120.0;                     //  1/5!
    ps = ps * r2 - 1.0 / 6.0;                       // -1/3!
    const double sr = r + r * (r2 * ps);
    // cos r = 1 - r^2 C(r^2), C = 1/2! - r^2/4! + ... + r^16/18!
    double pc = 1.0 / 6402373705728000.0;           //  1/18!
    pc = pc * r2 - 1.0 / 20922789888000.0;          // -1/16!
    pc = pc * r2 + 1.0 / 87178291200.0;             //  1/14!
    pc = pc * r2 - 1.0 / 479001600.0;               // -1/12!
    pc = pc * r2 + 1.0 / 3628800.0;                 //  1/10!
    pc = pc * r2 - 1.0 / 40320.0;                   // -1/8!
    pc = pc * r2 + 1.0 / 720.0;                     //  1/6!
    pc = pc * r2 - 1.0 / 24.0;                      // -1/4!
    pc = pc * r2 + 0.5;                             //  1/2!
    const double cr = 1.0 - r2 * pc;
    const int q = (int)k & 3;
    s = (q == 0) ? sr : (q == 1) ? cr : (q == 2) ? -sr : -cr;
    c = (q == 0) ? cr : (q == 1) ? -sr : (q == 2) ? -cr : sr;
}

// sin and cos of a float angle of moderate size (|x| < 1e3; the Carter
// field calls it with polar angles in (0, pi)): the same Cody-Waite
// reduction by pi/2, here with a three-part float split of pi/2 applied by
// fused multiply-adds, then the Taylor series of sin through x^9 and of cos
// through x^10 on [-pi/4, pi/4] (truncation below 3e-9 relative).  Like
// sincos_d it has no large-argument path, which keeps the register
// footprint of the float field small.
NP_HD inline void sincos_f(float x, float& s, float& c) {
    const float k = rintf(x * 0.63661977236758134f);                     // 2/pi
    float r = fmaf(-k, 1.5707963705062866f, x);                           // float(pi/2)
    r = fmaf(-k, -4.371138828673793e-08f, r);                             // pi/2 - float(pi/2)
    r = fmaf(-k, -1.7151245100058819e-15f, r);                             // the rest, rounded
    const float r2 = r * r;
    float ps = 1.0f / 362880.0f;                    //  1/9!
    ps = ps * r2 - 1.0f / 5040.0f;                  // -1/7!
    ps = ps * r2 + 1.0f / 120.0f;                   //  1/5!
    ps = ps * r2 - 1.0f / 6.0f;                     // -1/3!
    const float sr = r + r * (r2 * ps);
    float pc = 1.0f / 3628800.0f;                   //  1/10!
    pc = pc * r2 - 1.0f / 40320.0f;                 // -1/8!
    pc = pc * r2 + 1.0f / 720.0f;                   //  1/6!
    pc = pc * r2 - 1.0f / 24.0f;                    // -1/4!
    pc = pc * r2 + 0.5f;                            //  1/2!
    const float cr = 1.0f - r2 * pc;
    const int q = (int)k & 3;
    s = (q == 0) ? sr : (q == 1) ? cr : (q == 2) ? -sr : -cr;
    c = (q == 0) ? cr : (q == 1) ? -sr : (q == 2) ? -cr : sr;
}

}  // namespace np
