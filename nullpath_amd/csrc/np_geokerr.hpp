// np_geokerr.hpp -- device numerics of GEOKERR's semi-analytic solver, FP64:
// Carlson RF/RC/RD/RJ, the Carlson reductions of Table 1 with every branch,
// ZROOTS/LAGUER, SNCNDN, GEOMU with GEOKERR's between-call cache, and
// GEOPHITIME.  Shared by geomu.hip (GEOMU batches) and geokerr.hip (the
// GEOKERR driver and geokerr_ref_api).
//
// Source: geokerr_port/reference/geokerr_original.f
//   rf/rc/rd/rj            3444-3489, 3608-3645, 3648-3694, 3697-3762
//   ellcubicreal/complex   1188-1262 / 1264-1340
//   ellquarticreal/complex 1362-1442 / 1444-1529, elldoublecomplex 1531-1645
//   GEOMU                  771-1186, calcimu* 3137-3274, SNCNDN 3277-3340
//   ZROOTS/LAGUER          3343-3441, 3527-3606
//   GEOPHITIME             1648-2360, calcphitmu* 2361-2507, ellphitmu 2508-2592,
//                          phifnkerr/tfnkerr 2593-2671
//
// The Fortran's INT() is a truncating cast, SIGN(1,x) is copysign(1,x) and
// (-1)**n goes by parity.  FIRSTPT and PHT are template parameters: the
// per-sample path (FIRSTPT = false) carries no root finder for the quartic
// cases.  Where the Fortran reads a variable it never set, the value used
// is the one stated at that spot (the same choice as the CPU restatement).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Every FP64 product and sum below is rounded on its own (no fused
// multiply-add), as in the Fortran and the oracle.  GEOKERR's results hinge
// on comparisons of sample positions against turning points and on
// cancelling sums (eq. 48 near a = 1, ellphitmu near the pole); with fusion
// left to the compiler, two kernels inlining this same source rounded them
// differently and a sample on a turning point moved by 1e-8 or became NaN.
// Fast contraction is restored at the end of this header.
#pragma clang fp contract(off)

namespace np {
namespace gk {

constexpr int kCarlsonMaxIter = 100;  // a converged argument needs ~30; stops NaN inputs
constexpr double kPi = 3.141592653589793;
constexpr double kHalfPi = 1.5707963267948966;

__device__ __forceinline__ double sgn1(double x) { return copysign(1.0, x); }
// (-1)**n by parity; unsigned so that n + 1 at INT_MAX wraps as in the
// reference's 32-bit integer arithmetic instead of being undefined.
__device__ __forceinline__ double neg1pow(unsigned n) { return (n & 1u) ? -1.0 : 1.0; }
__device__ __forceinline__ int add_i32(int a, int b) { return (int)((unsigned)a + (unsigned)b); }
// Fortran INT() as the reference's x86-64 build executes it (cvttsd2si):
// truncation toward zero, INT_MIN for NaN and for values outside int.  A
// plain C++ cast of those is undefined.
__device__ __forceinline__ int f_int(double x) {
    return (x > -2147483649.0 && x < 2147483648.0) ? (int)x : (int)0x80000000u;
}

// ------------------------------ Carlson forms ------------------------------

__device__ inline double rf_d(double x, double y, double z) {
    const double ERRTOL = 0.0025, THIRD = 1.0 / 3.0, A1 = 1.0 / 24.0, C2 = 0.1, C3 = 3.0 / 44.0,
                 C4 = 1.0 / 14.0;
    double xt = x, yt = y, zt = z, ave = 0.0, dx = 0.0, dy = 0.0, dz = 0.0;
    for (int it = 0; it < kCarlsonMaxIter; ++it) {
        const double sx = sqrt(xt), sy = sqrt(yt), sz = sqrt(zt);
        const double lam = sx * (sy + sz) + sy * sz;
        xt = 0.25 * (xt + lam);
        yt = 0.25 * (yt + lam);
        zt = 0.25 * (zt + lam);
        ave = THIRD * (xt + yt + zt);
        if (ave == 0.0) {
            dx = dy = dz = 0.0;
        } else {
            dx = (ave - xt) / ave;
            dy = (ave - yt) / ave;
            dz = (ave - zt) / ave;
        }
        if (!(fmax(fmax(fabs(dx), fabs(dy)), fabs(dz)) > ERRTOL)) break;
    }
    const double e2 = dx * dy - dz * dz;
    const double e3 = dx * dy * dz;
    return (1.0 + (A1 * e2 - C2 - C3 * e3) * e2 + C4 * e3) / sqrt(ave);
}

__device__ inline double rc_d(double x, double y) {
    const double ERRTOL = 0.0012, THIRD = 1.0 / 3.0, A1 = 0.3, C2 = 1.0 / 7.0, C3 = 0.375,
                 C4 = 9.0 / 22.0;
    double xt, yt, w, ave = 0.0, s = 0.0;
    if (y > 0.0) {
        xt = x; yt = y; w = 1.0;
    } else {
        xt = x - y; yt = -y; w = sqrt(x) / sqrt(xt);
    }
    for (int it = 0; it < kCarlsonMaxIter; ++it) {
        const double lam = 2.0 * sqrt(xt) * sqrt(yt) + yt;
        xt = 0.25 * (xt + lam);
        yt = 0.25 * (yt + lam);
        ave = THIRD * (xt + yt + yt);
        s = (yt - ave) / ave;
        if (!(fabs(s) > ERRTOL)) break;
    }
    return w * (1.0 + s * s * (A1 + s * (C2 + s * (C3 + s * C4)))) / sqrt(ave);
}

__device__ inline double rd_d(double x, double y, double z) {
    const double ERRTOL = 0.0015, A1 = 3.0 / 14.0, C2 = 1.0 / 6.0, C3 = 9.0 / 22.0, C4 = 3.0 / 26.0,
                 C5 = 0.25 * C3, C6 = 1.5 * C4;
    double xt = x, yt = y, zt = z, sum = 0.0, fac = 1.0, ave = 0.0, dx = 0.0, dy = 0.0, dz = 0.0;
    for (int it = 0; it < kCarlsonMaxIter; ++it) {
        const double sx = sqrt(xt), sy = sqrt(yt), sz = sqrt(zt);
        const double lam = sx * (sy + sz) + sy * sz;
        sum = sum + fac / (sz * (zt + lam));
        fac = 0.25 * fac;
        xt = 0.25 * (xt + lam);
        yt = 0.25 * (yt + lam);
        zt = 0.25 * (zt + lam);
        ave = 0.2 * (xt + yt + 3.0 * zt);
        dx = (ave - xt) / ave;
        dy = (ave - yt) / ave;
        dz = (ave - zt) / ave;
        if (!(fmax(fmax(fabs(dx), fabs(dy)), fabs(dz)) > ERRTOL)) break;
    }
    const double ea = dx * dy, eb = dz * dz, ec = ea - eb, ed = ea - 6.0 * eb, ee = ed + ec + ec;
    return 3.0 * sum + fac * (1.0 + ed * (-A1 + C5 * ed - C6 * dz * ee) +
                              dz * (C2 * ee + dz * (-C3 * ec + dz * C4 * ea))) /
                           (ave * sqrt(ave));
}

__device__ inline double rj_d(double x, double y, double z, double p) {
    const double ERRTOL = 0.0015, A1 = 3.0 / 14.0, C2 = 1.0 / 3.0, C3 = 3.0 / 22.0, C4 = 3.0 / 26.0,
                 C5 = 0.75 * C3, C6 = 1.5 * C4, C7 = 0.5 * C2, C8 = C3 + C3;
    double sum = 0.0, fac = 1.0, xt, yt, zt, pt, a = 0.0, b = 0.0, rcx = 0.0;
    double ave = 0.0, dx = 0.0, dy = 0.0, dz = 0.0, dp = 0.0;
    if (p > 0.0) {
        xt = x; yt = y; zt = z; pt = p;
    } else {
        xt = fmin(fmin(x, y), z);
        zt = fmax(fmax(x, y), z);
        yt = x + y + z - xt - zt;
        a = 1.0 / (yt - p);
        b = a * (zt - yt) * (yt - xt);
        pt = yt + b;
        const double rho = xt * zt / yt;
        const double tau = p * pt / yt;
        rcx = rc_d(rho, tau);
    }
    for (int it = 0; it < kCarlsonMaxIter; ++it) {
        const double sx = sqrt(xt), sy = sqrt(yt), sz = sqrt(zt);
        const double lam = sx * (sy + sz) + sy * sz;
        double alpha = pt * (sx + sy + sz) + sx * sy * sz;
        alpha = alpha * alpha;
        const double beta = pt * (pt + lam) * (pt + lam);
        sum = sum + fac * rc_d(alpha, beta);
        fac = 0.25 * fac;
        xt = 0.25 * (xt + lam);
        yt = 0.25 * (yt + lam);
        zt = 0.25 * (zt + lam);
        pt = 0.25 * (pt + lam);
        ave = 0.2 * (xt + yt + zt + pt + pt);
        dx = (ave - xt) / ave;
        dy = (ave - yt) / ave;
        dz = (ave - zt) / ave;
        dp = (ave - pt) / ave;
        if (!(fmax(fmax(fabs(dx), fabs(dy)), fmax(fabs(dz), fabs(dp))) > ERRTOL)) break;
    }
    const double ea = dx * (dy + dz) + dy * dz;
    const double eb = dx * dy * dz;
    const double ec = dp * dp;
    const double ed = ea - 3.0 * ec;
    const double ee = eb + 2.0 * dp * (ea - ec);
    double rj = 3.0 * sum + fac * (1.0 + ed * (-A1 + C5 * ed - C6 * ee) + eb * (C7 + dp * (-C8 + dp * C4)) +
                                   dp * ea * (C2 - dp * C3) - C2 * dp * ec) /
                                (ave * sqrt(ave));
    if (p <= 0.0) rj = a * (b * rj + 3.0 * (rcx - rf_d(xt, yt, zt)));
    return rj;
}

// ------------------------- Carlson reductions -------------------------------
// int_y^x prod (a_i + b_i t)^(p_i / 2) dt.  P is the exponent of the last
// factor: 0 computes the RF piece into rff; -2 and -4 read rff.  The -4
// forms drop a third-kind term proportional to the linear coefficient of
// the polynomial under the root, which U(u) never has.

template <int P>
__device__ inline double ellcubicreal(double a1, double b1, double a2, double b2, double a3, double b3,
                                      double a4, double b4, double& rff, double y, double x) {
    const double X1 = sqrt(a1 + b1 * x), X2 = sqrt(a2 + b2 * x), X3 = sqrt(a3 + b3 * x);
    const double Y1 = sqrt(a1 + b1 * y), Y2 = sqrt(a2 + b2 * y), Y3 = sqrt(a3 + b3 * y);
    const double U1c = (X1 * Y2 * Y3 + Y1 * X2 * X3) / (x - y);
    const double U12 = U1c * U1c;
    const double t2 = (X2 * Y1 * Y3 + Y2 * X1 * X3) / (x - y);
    const double t3 = (X3 * Y1 * Y2 + Y3 * X1 * X2) / (x - y);
    const double U22 = t2 * t2, U32 = t3 * t3;
    if constexpr (P == 0) {
        rff = rf_d(U32, U22, U12);
        return 2.0 * rff;
    } else {
        const double d12 = a1 * b2 - a2 * b1, d13 = a1 * b3 - a3 * b1, d14 = a1 * b4 - a4 * b1;
        const double d24 = a2 * b4 - a4 * b2, d34 = a3 * b4 - a4 * b3;
        const double X4 = sqrt(a4 + b4 * x), Y4 = sqrt(a4 + b4 * y);
        const double I1c = 2.0 * rff;
        if constexpr (P == -2) {
            const double W22 = U12 - b4 * d12 * d13 / d14;
            const double q = X4 * Y4 / X1 / Y1;
            const double Q22 = q * q * W22;
            const double P22 = Q22 + b4 * d24 * d34 / d14;
            const double I3c = 2.0 * rc_d(P22, Q22) - 2.0 * b1 * d12 * d13 / 3.0 / d14 * rj_d(U32, U22, U12, W22);
            return (b4 * I3c - b1 * I1c) / d14;
        } else {
            const double r12 = a1 / b1 - a2 / b2, r13 = a1 / b1 - a3 / b3;
            const double r24i = b2 * b4 / (a2 * b4 - a4 * b2), r34i = b3 * b4 / (a3 * b4 - a4 * b3);
            const double I2c = 2.0 / 3.0 * d12 * d13 * rd_d(U32, U22, U12) + 2.0 * X1 * Y1 / U1c;
            const double K2c = b2 * b3 * I2c - 2.0 * b4 * (X1 * X2 * X3 / (X4 * X4) - Y1 * Y2 * Y3 / (Y4 * Y4));
            return 0.5 * b4 / d14 / d24 / d34 * K2c +
                   (b1 / d14) * (b1 / d14) * (1.0 - 0.5 * r12 * r13 * r24i * r34i) * I1c;
        }
    }
}

template <int P>
__device__ inline double ellcubiccomplex(double a1, double b1, double a4, double b4, double f, double g,
                                         double h, double& rff, double y, double x) {
    const double X1 = sqrt(a1 + b1 * x), Y1 = sqrt(a1 + b1 * y);
    const double beta1 = g * b1 - 2.0 * h * a1;
    const double a11 = sqrt(2.0 * f * b1 * b1 - 2.0 * g * a1 * b1 + 2.0 * h * a1 * a1);
    const double xi = sqrt(f + g * x + h * x * x);
    const double eta = sqrt(f + g * y + h * y * y);
    const double s = xi + eta;
    const double mm = (X1 + Y1) * sqrt(s * s - h * (x - y) * (x - y)) / (x - y);
    const double M2 = mm * mm;
    const double Lp2 = M2 - beta1 + sqrt(2.0 * h) * a11;
    const double Lm2 = M2 - beta1 - sqrt(2.0 * h) * a11;
    if constexpr (P == 0) {
        rff = rf_d(M2, Lm2, Lp2);
        return 4.0 * rff;
    } else {
        const double X4 = sqrt(a4 + b4 * x), Y4 = sqrt(a4 + b4 * y);
        const double d14 = a1 * b4 - a4 * b1;
        const double c44 = sqrt(2.0 * f * b4 * b4 - 2.0 * g * a4 * b4 + 2.0 * h * a4 * a4);
        const double a142 = 2.0 * f * b1 * b4 - g * (a1 * b4 + a4 * b1) + 2.0 * h * a1 * a4;
        const double I1c = 4.0 * rff;
        const double U = (X1 * eta + Y1 * xi) / (x - y);
        if constexpr (P == -2) {
            const double U2 = U * U;
            const double Wp2 = M2 - b1 * (a142 + a11 * c44) / d14;
            const double W2 = U2 - a11 * a11 * b4 / 2.0 / d14;
            const double qq = X4 * Y4 / X1 / Y1;
            const double Q2 = qq * qq * W2;
            const double P2 = Q2 + c44 * c44 * b4 / 2.0 / d14;
            const double I3c = (2.0 * a11 / 3.0 / c44) *
                                   ((-4.0 * b1 / d14) * (a142 + a11 * c44) * rj_d(M2, Lm2, Lp2, Wp2) - 6.0 * rff +
                                    3.0 * rc_d(U2, W2)) +
                               2.0 * rc_d(P2, Q2);
            return (b4 * I3c - b1 * I1c) / d14;
        } else {
            const double rho = sqrt(2.0 * h) * a11 - beta1;
            const double r24Xr34 = 0.5 * c44 * c44 / h / (b4 * b4);
            const double r12Xr13 = 0.5 * a11 * a11 / h / (b1 * b1);
            const double N2c = 2.0 / 3.0 * sqrt(2.0 * h) / a11 * (4.0 * rho * rd_d(M2, Lm2, Lp2) - 6.0 * rff + 3.0 / U) +
                               2.0 / X1 / Y1 / U;
            const double K2c = 0.5 * a11 * a11 * N2c - 2.0 * d14 * (xi / X1 / (X4 * X4) - eta / Y1 / (Y4 * Y4));
            return 0.5 / d14 / (h * b4 * r24Xr34) * K2c +
                   (b1 / d14) * (b1 / d14) * (1.0 - 0.5 * r12Xr13 / r24Xr34) * I1c;
        }
    }
}

template <int P>
__device__ inline double ellquarticreal(double a1, double b1, double a2, double b2, double a3, double b3,
                                        double a4, double b4, double a5, double b5, double& rff, double y,
                                        double x) {
    const double X1 = sqrt(a1 + b1 * x), X2 = sqrt(a2 + b2 * x), X3 = sqrt(a3 + b3 * x), X4 = sqrt(a4 + b4 * x);
    const double Y1 = sqrt(a1 + b1 * y), Y2 = sqrt(a2 + b2 * y), Y3 = sqrt(a3 + b3 * y), Y4 = sqrt(a4 + b4 * y);
    const double u1 = (X1 * X2 * Y3 * Y4 + Y1 * Y2 * X3 * X4) / (x - y);
    const double u2 = (X1 * X3 * Y2 * Y4 + Y1 * Y3 * X2 * X4) / (x - y);
    const double u3 = (X1 * X4 * Y2 * Y3 + Y1 * Y4 * X2 * X3) / (x - y);
    const double U122 = u1 * u1, U132 = u2 * u2, U142 = u3 * u3;
    if constexpr (P == 0) {
        rff = rf_d(U122, U132, U142);
        return 2.0 * rff;
    } else {
        const double d12 = a1 * b2 - a2 * b1, d13 = a1 * b3 - a3 * b1, d14 = a1 * b4 - a4 * b1;
        const double d15 = a1 * b5 - a5 * b1;
        const double d25 = a2 * b5 - a5 * b2, d35 = a3 * b5 - a5 * b3, d45 = a4 * b5 - a5 * b4;
        const double I1 = 2.0 * rff;
        const double X52 = a5 + b5 * x, Y52 = a5 + b5 * y;
        if constexpr (P == -2) {
            const double W22 = U122 - d13 * d14 * d25 / d15;
            const double Q22 = X52 * Y52 / ((X1 * Y1) * (X1 * Y1)) * W22;
            const double P22 = Q22 + d25 * d35 * d45 / d15;
            const double I3 = 2.0 * d12 * d13 * d14 / 3.0 / d15 * rj_d(U122, U132, U142, W22) + 2.0 * rc_d(P22, Q22);
            return (b5 * I3 - b1 * I1) / d15;
        } else {
            const double d24 = a2 * b4 - a4 * b2, d34 = a3 * b4 - a4 * b3;
            const double I2 = 2.0 / 3.0 * d12 * d13 * rd_d(U122, U132, U142) + 2.0 * X1 * Y1 / X4 / Y4 / sqrt(U142);
            const double r12 = a1 / b1 - a2 / b2, r13 = a1 / b1 - a3 / b3;
            const double r25i = b2 * b5 / (a2 * b5 - a5 * b2), r35i = b3 * b5 / (a3 * b5 - a5 * b3);
            const double A111m1m2 = X1 * X2 * X3 / X4 / X52 - Y1 * Y2 * Y3 / Y4 / Y52;
            return 0.5 * (b5 * b5) * d24 * d34 / d15 / d25 / d35 / d45 * I2 +
                   (b1 / d15) * (b1 / d15) * (1.0 - 0.5 * r12 * r13 * r25i * r35i) * I1 -
                   (b5 * b5) / d15 / d25 / d35 * A111m1m2;
        }
    }
}

template <int P>
__device__ inline double ellquarticcomplex(double a1, double b1, double a4, double b4, double a5, double b5,
                                           double f, double g, double h, double& rff, double y, double x) {
    const double X1 = sqrt(a1 + b1 * x), X4 = sqrt(a4 + b4 * x);
    const double Y1 = sqrt(a1 + b1 * y), Y4 = sqrt(a4 + b4 * y);
    const double a11 = sqrt(2.0 * f * b1 * b1 - 2.0 * g * a1 * b1 + 2.0 * h * a1 * a1);
    const double c44 = sqrt(2.0 * f * b4 * b4 - 2.0 * g * a4 * b4 + 2.0 * h * a4 * a4);
    const double a142 = 2.0 * f * b1 * b4 - g * (a1 * b4 + a4 * b1) + 2.0 * h * a1 * a4;
    const double xi = sqrt(f + g * x + h * x * x);
    const double eta = sqrt(f + g * y + h * y * y);
    const double s = xi + eta;
    const double mm = (X1 * Y4 + Y1 * X4) * sqrt(s * s - h * (x - y) * (x - y)) / (x - y);
    const double M2 = mm * mm;
    const double Lp2 = M2 + a142 + a11 * c44;
    const double Lm2 = fmax(M2 + a142 - a11 * c44, 0.0);
    if constexpr (P == 0) {
        rff = rf_d(M2, Lm2, Lp2);
        return 4.0 * rff;
    } else {
        const double I1 = 4.0 * rff;
        const double d14 = a1 * b4 - a4 * b1, d15 = a1 * b5 - a5 * b1, d45 = a4 * b5 - a5 * b4;
        const double X52 = a5 + b5 * x, Y52 = a5 + b5 * y;
        const double c552 = 2.0 * f * b5 * b5 - 2.0 * g * a5 * b5 + 2.0 * h * a5 * a5;
        const double U = (X1 * X4 * eta + Y1 * Y4 * xi) / (x - y);
        if constexpr (P == -2) {
            const double c55 = sqrt(c552);
            const double a152 = 2.0 * f * b1 * b5 - g * (a1 * b5 + a5 * b1) + 2.0 * h * a5 * a1;
            const double U2 = U * U;
            const double Wp2 = M2 + d14 * (a152 + a11 * c55) / d15;
            const double W2 = U2 - a11 * a11 * d45 / 2.0 / d15;
            const double Q2 = X52 * Y52 / ((X1 * Y1) * (X1 * Y1)) * W2;
            const double P2 = Q2 + c55 * c55 * d45 / 2.0 / d15;
            const double I3 = (2.0 * a11 / 3.0 / c55) *
                                  ((4.0 * d14 / d15) * (a152 + a11 * c55) * rj_d(M2, Lm2, Lp2, Wp2) - 6.0 * rff +
                                   3.0 * rc_d(U2, W2)) +
                              2.0 * rc_d(P2, Q2);
            return (b5 * I3 - b1 * I1) / d15;
        } else {
            const double I2 = 2.0 * a11 / 3.0 / c44 * (4.0 * (a142 + a11 * c44) * rd_d(M2, Lm2, Lp2) - 6.0 * rff) +
                              2.0 * a11 / c44 / U + 2.0 * X1 * Y1 / X4 / Y4 / U;
            const double A111m1m2 = X1 * xi / X4 / X52 - Y1 * eta / Y4 / Y52;
            return (b5 * b5) / (2.0 * d15 * d45) * (c44 * c44) / c552 * I2 +
                   (b1 * b1) / (d15 * d15) * (1.0 - 0.5 * (b5 * b5) * (a11 * a11) / (b1 * b1) / c552) * I1 -
                   2.0 * (b5 * b5) / d15 / c552 * A111m1m2;
        }
    }
}

template <int P>
__device__ inline double elldoublecomplex(double f1, double g1, double h1, double f2, double g2, double h2,
                                          double a5, double b5, double& rff, double y, double x) {
    const double xi1 = sqrt(f1 + g1 * x + h1 * (x * x));
    const double xi2 = sqrt(f2 + g2 * x + h2 * (x * x));
    const double eta1 = sqrt(f1 + g1 * y + h1 * y * y);
    const double eta2 = sqrt(f2 + g2 * y + h2 * y * y);
    const double theta1 = 2.0 * f1 + g1 * (x + y) + 2.0 * h1 * x * y;
    const double theta2 = 2.0 * f2 + g2 * (x + y) + 2.0 * h2 * x * y;
    const double zeta1 = sqrt(2.0 * xi1 * eta1 + theta1);
    const double zeta2 = sqrt(2.0 * xi2 * eta2 + theta2);
    const double M = zeta1 * zeta2 / (x - y);
    const double M2 = M * M;
    const double delta122 = 2.0 * f1 * h2 + 2.0 * f2 * h1 - g1 * g2;
    const double delta112 = 4.0 * f1 * h1 - g1 * g1;
    const double delta222 = 4.0 * f2 * h2 - g2 * g2;
    const double Delta = sqrt(delta122 * delta122 - delta112 * delta222);
    const double Deltap = delta122 + Delta, Deltam = delta122 - Delta;
    const double Lp2 = M2 + Deltap, Lm2 = M2 + Deltam;
    if constexpr (P == 0) {
        rff = rf_d(M2, Lm2, Lp2);
        return 4.0 * rff;
    } else {
        const double U = (xi1 * eta2 + eta1 * xi2) / (x - y);
        const double U2 = U * U;
        const double alpha15 = 2.0 * f1 * b5 - g1 * a5, alpha25 = 2.0 * f2 * b5 - g2 * a5;
        const double beta15 = g1 * b5 - 2.0 * h1 * a5, beta25 = g2 * b5 - 2.0 * h2 * a5;
        const double gamma1 = 0.5 * (alpha15 * b5 - beta15 * a5), gamma2 = 0.5 * (alpha25 * b5 - beta25 * a5);
        const double Lambda = delta112 * gamma2 / gamma1;
        const double Omega2 = M2 + Lambda;
        const double psi = 0.5 * (alpha15 * beta25 - alpha25 * beta15);
        const double psi2 = psi * psi;
        const double xi5 = a5 + b5 * x, eta5 = a5 + b5 * y;
        const double Am111m1 = 1.0 / xi1 * xi2 - 1.0 / eta1 * eta2;
        const double A1111m4 = xi1 * xi2 / (xi5 * xi5) - eta1 * eta2 / (eta5 * eta5);
        const double XX = xi5 * eta5 * (theta1 * 0.5 * Am111m1 - xi5 * eta5 * A1111m4) / ((x - y) * (x - y));
        const double S = 0.5 * (M2 + delta122) - U2;
        const double S2 = S * S;
        const double mu = gamma1 * xi5 * eta5 / xi1 / eta1;
        const double T = mu * S + 2.0 * gamma1 * gamma2;
        const double T2 = T * T;
        const double V2 = (mu * mu) * (S2 + Lambda * U2);
        const double b2 = (Omega2 * Omega2) * (S2 / U2 + Lambda);
        const double a2 = b2 + (Lambda * Lambda) * psi2 / gamma1 / gamma2;
        const double H = delta112 * psi * (rj_d(M2, Lm2, Lp2, Omega2) / 3.0 + 0.5 * rc_d(a2, b2)) / (gamma1 * gamma1) -
                         XX * rc_d(T2, V2);
        if constexpr (P == -2) {
            return -2.0 * (b5 * H + beta15 * rff / gamma1);
        } else {
            const double A1111m2 = xi1 * xi2 / xi5 - eta1 * eta2 / eta5;
            const double xi1p = 0.5 * (g1 + 2.0 * h1 * x) / xi1;
            const double eta1p = 0.5 * (g1 + 2.0 * h1 * y) / eta1;
            const double B = xi1p * xi2 - eta1p * eta2;
            const double G = 2.0 / 3.0 * Delta * Deltap * rd_d(M2, Lm2, Lp2) + 0.5 * Delta / U +
                             (delta122 * theta1 - delta112 * theta2) / 4.0 / xi1 / eta1 / U;
            const double Sigma = G - Deltap * rff + B;
            return b5 * (beta15 / gamma1 + beta25 / gamma2) * H + (beta15 * beta15) * rff / (gamma1 * gamma1) +
                   (b5 * b5) * (Sigma - b5 * A1111m2) / gamma1 / gamma2;
        }
    }
}

// --------------------------------- complex ---------------------------------
struct cplx {
    double re, im;
};
__device__ __forceinline__ cplx mk(double r, double i = 0.0) { return cplx{r, i}; }
__device__ __forceinline__ cplx operator+(cplx a, cplx b) { return mk(a.re + b.re, a.im + b.im); }
__device__ __forceinline__ cplx operator-(cplx a, cplx b) { return mk(a.re - b.re, a.im - b.im); }
__device__ __forceinline__ cplx operator*(cplx a, cplx b) {
    return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
__device__ __forceinline__ cplx operator*(double s, cplx a) { return mk(s * a.re, s * a.im); }
__device__ __forceinline__ double cabs_d(cplx a) { return hypot(a.re, a.im); }
// Smith's division
__device__ __forceinline__ cplx operator/(cplx a, cplx b) {
    if (fabs(b.re) >= fabs(b.im)) {
        const double r = b.im / b.re, d = b.re + b.im * r;
        return mk((a.re + a.im * r) / d, (a.im - a.re * r) / d);
    }
    const double r = b.re / b.im, d = b.re * r + b.im;
    return mk((a.re * r + a.im) / d, (a.im * r - a.re) / d);
}
__device__ __forceinline__ bool ceq(cplx a, cplx b) { return a.re == b.re && a.im == b.im; }
// principal square root
__device__ __forceinline__ cplx csqrt_d(cplx z) {
    if (z.re == 0.0 && z.im == 0.0) return mk(0.0, z.im);
    const double m = cabs_d(z);
    double t = sqrt(0.5 * (m + fabs(z.re)));
    if (z.re >= 0.0) return mk(t, z.im / (2.0 * t));
    return mk(fabs(z.im) / (2.0 * t), copysign(t, z.im));
}

// LAGUER (geokerr_original.f:3527-3606) for a polynomial of degree M;
// a[0..M], constant term first.  The degree is a template parameter so every
// coefficient index is a constant after unrolling: the coefficient and root
// arrays of the callers stay in registers instead of private (scratch)
// memory, with the same operations in the same order as the runtime-degree
// loop of the Fortran.
template <int M>
__device__ __forceinline__ void laguer(const cplx* a, cplx& x) {
    const double frac[8] = {0.5, 0.25, 0.75, 0.13, 0.38, 0.62, 0.88, 1.0};
    const int MT = 10, MAXIT = 80;
    const double EPSS = 1e-15;
    for (int iter = 1; iter <= MAXIT; ++iter) {
        cplx b = a[M];
        double err = cabs_d(b);
        cplx d = mk(0.0), f = mk(0.0);
        const double abx = cabs_d(x);
#pragma unroll
        for (int j = M - 1; j >= 0; --j) {
            f = x * f + d;
            d = x * d + b;
            b = x * b + a[j];
            err = cabs_d(b) + abx * err;
        }
        err = EPSS * err;
        if (cabs_d(b) <= err) return;
        const cplx g = d / b;
        const cplx g2 = g * g;
        const cplx h = g2 - 2.0 * (f / b);
        const cplx sq = csqrt_d((double)(M - 1) * ((double)M * h - g2));
        cplx gp = g + sq;
        const cplx gm = g - sq;
        const double abp = cabs_d(gp), abm = cabs_d(gm);
        if (abp < abm) gp = gm;
        cplx dx;
        if (fmax(abp, abm) > 0.0) {
            dx = mk((double)M) / gp;
        } else {
            // CEXP(CMPLX(LOG(1+ABX), ITER)) in default (single) precision
            const float lr = (float)log(1.0 + abx), th = (float)iter;
            const float e = expf(lr);
            dx = mk((double)(e * cosf(th)), (double)(e * sinf(th)));
        }
        const cplx x1 = x - dx;
        if (ceq(x, x1)) return;
        if (iter % MT != 0) x = x1;
        else x = x - frac[iter / MT - 1] * dx;
    }
}

// ZROOTS' deflation (geokerr_original.f:3343-3441) from degree J down to 1,
// one level per template instance (J and the loop below are constants).
template <int M, int J>
__device__ __forceinline__ void zroots_deflate(cplx* ad, cplx* roots, cplx& xstart, int& ncc) {
    if constexpr (J >= 1) {
        const double EPS = 1e-6;
        cplx x = xstart;
        if (J < M && roots[J].im != 0.0 && ncc == 0) {
            // the reference builds (Re r, -Re r), not the conjugate
            xstart = mk(roots[J].re, -roots[J].re);
            ncc = 1;
        } else {
            xstart = mk(0.0);
            ncc = 0;
        }
        laguer<J>(ad, x);
        if (fabs(x.im) <= 2.0 * EPS * EPS * fabs(x.re)) x = mk(x.re);
        roots[J - 1] = x;
        cplx b = ad[J];
#pragma unroll
        for (int jj = J - 1; jj >= 0; --jj) {
            const cplx c = ad[jj];
            ad[jj] = b;
            b = x * b + c;
        }
        zroots_deflate<M, J - 1>(ad, roots, xstart, ncc);
    }
}

// Straight insertion by real part (ZROOTS' PIKSRT-style final loop), with
// the break of the scan as a predicate so every index is a constant: a root
// moves up past every larger real part until one at or below it (a NaN
// real part never stops the scan, as in the loop with the break).
template <typename C, int M>
__device__ __forceinline__ void insertion_sort_re(C* roots) {
#pragma unroll
    for (int j = 1; j < M; ++j) {
        const C x = roots[j];
        bool moving = true;
#pragma unroll
        for (int i = j - 1; i >= 0; --i) {
            if (moving) {
                if (roots[i].re <= x.re) {
                    moving = false;
                    roots[i + 1] = x;
                } else {
                    roots[i + 1] = roots[i];
                }
            }
        }
        if (moving) roots[0] = x;
    }
}

// ZROOTS (geokerr_original.f:3343-3441) for degree M; roots sorted by real
// part.
template <int M>
__device__ __forceinline__ void zroots(const cplx* a, cplx* roots) {
    cplx ad[M + 1];
#pragma unroll
    for (int j = 0; j <= M; ++j) ad[j] = a[j];
    cplx xstart = mk(0.0);
    int ncc = 0;
    zroots_deflate<M, M>(ad, roots, xstart, ncc);
#pragma unroll
    for (int j = 0; j < M; ++j) laguer<M>(a, roots[j]);
    insertion_sort_re<cplx, M>(roots);
}

// ---------------- mixed-precision roots of the quartic U(u) ----------------
// The deflation half of ZROOTS in float, the polish on the full polynomial
// in double (Newton steps), for GEOMU's quartic U(u) = 1 + cc u^2 +
// dd u^3 + ee u^4.  The float phase decides which roots are real; it hands
// the polynomial back to the FP64 ZROOTS (returns false) when any root's
// imaginary part lies between 1e-5 and 1e-2 of its modulus or two real
// roots lie within 1e-3 of each other (relative) -- the near-degenerate
// arguments where a float deflation could pick another case.
struct cplxf {
    float re, im;
};
__device__ __forceinline__ cplxf mkf(float r, float i = 0.0f) { return cplxf{r, i}; }
__device__ __forceinline__ cplxf operator+(cplxf a, cplxf b) { return mkf(a.re + b.re, a.im + b.im); }
__device__ __forceinline__ cplxf operator-(cplxf a, cplxf b) { return mkf(a.re - b.re, a.im - b.im); }
__device__ __forceinline__ cplxf operator*(cplxf a, cplxf b) {
    return mkf(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
__device__ __forceinline__ cplxf operator*(float s, cplxf a) { return mkf(s * a.re, s * a.im); }
// The float phase only seeds the FP64 polish, so on the device it uses the
// hardware reciprocal and square root (v_rcp_f32, v_sqrt_f32, ~1 ulp)
// instead of the IEEE division and sqrtf/hypotf sequences; a host build
// (the CPU tests) uses the IEEE operations.  |z| as sqrt(re^2 + im^2): the
// quartic's coefficients and roots here stay far from float's range limits.
__device__ __forceinline__ float rcp_f(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
__device__ __forceinline__ float sqrt_f(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(x);
#else
    return std::sqrt(x);
#endif
}
__device__ __forceinline__ float cabs_f(cplxf a) { return sqrt_f(a.re * a.re + a.im * a.im); }
__device__ __forceinline__ cplxf operator/(cplxf a, cplxf b) {
    if (fabsf(b.re) >= fabsf(b.im)) {
        const float r = b.im * rcp_f(b.re), id = rcp_f(b.re + b.im * r);
        return mkf((a.re + a.im * r) * id, (a.im - a.re * r) * id);
    }
    const float r = b.re * rcp_f(b.im), id = rcp_f(b.re * r + b.im);
    return mkf((a.re * r + a.im) * id, (a.im * r - a.re) * id);
}
__device__ __forceinline__ cplxf csqrt_f(cplxf z) {
    if (z.re == 0.0f && z.im == 0.0f) return mkf(0.0f, z.im);
    const float m = cabs_f(z);
    const float t = sqrt_f(0.5f * (m + fabsf(z.re)));
    const float h = 0.5f * rcp_f(t);
    if (z.re >= 0.0f) return mkf(t, z.im * h);
    return mkf(fabsf(z.im) * h, copysignf(t, z.im));
}

// LAGUER's iteration in float (same step control; float round-off bound),
// degree M a template parameter as in laguer.
template <int M>
__device__ __forceinline__ void laguer_f(const cplxf* a, cplxf& x) {
    const float frac[8] = {0.5f, 0.25f, 0.75f, 0.13f, 0.38f, 0.62f, 0.88f, 1.0f};
    const int MT = 10, MAXIT = 80;
    const float EPSS = 1e-7f;
    for (int iter = 1; iter <= MAXIT; ++iter) {
        cplxf b = a[M];
        float err = cabs_f(b);
        cplxf d = mkf(0.0f), f = mkf(0.0f);
        const float abx = cabs_f(x);
#pragma unroll
        for (int j = M - 1; j >= 0; --j) {
            f = x * f + d;
            d = x * d + b;
            b = x * b + a[j];
            err = cabs_f(b) + abx * err;
        }
        err = EPSS * err;
        if (cabs_f(b) <= err) return;
        const cplxf g = d / b;
        const cplxf g2 = g * g;
        const cplxf h = g2 - 2.0f * (f / b);
        const cplxf sq = csqrt_f((float)(M - 1) * ((float)M * h - g2));
        cplxf gp = g + sq;
        const cplxf gm = g - sq;
        const float abp = cabs_f(gp), abm = cabs_f(gm);
        if (abp < abm) gp = gm;
        cplxf dx;
        if (fmaxf(abp, abm) > 0.0f) {
            dx = mkf((float)M) / gp;
        } else {
            const float e = expf(logf(1.0f + abx));
            dx = mkf(e * cosf((float)iter), e * sinf((float)iter));
        }
        const cplxf x1 = x - dx;
        if (x.re == x1.re && x.im == x1.im) return;
        if (iter % MT != 0) x = x1;
        else x = x - frac[iter / MT - 1] * dx;
    }
}

// Float seeds of the four roots of U(u) = 1 + cc u^2 + dd u^3 + ee u^4 in
// closed form, with no data-dependent iteration (so the lanes of a wave stay
// together): with x = 1/u, x^4 U(1/x) = x^4 + cc x^2 + dd x + ee is a
// depressed quartic.  Ferrari's method takes the largest root m of the
// resolvent cubic m^3 + cc m^2 + (cc^2/4 - ee) m - dd^2/8 = 0 (Viete's
// trigonometric form when it has three real roots, Cardano's otherwise) and
// splits the quartic into x^2 - s x + (cc/2 + m + dd/(2s)) and x^2 + s x +
// (cc/2 + m - dd/(2s)), s = sqrt(2m), whose constant terms multiply to ee.
// False when m is not clearly positive (dd near 0: the split degenerates).
__device__ __forceinline__ bool quartic_u_seeds_f(float cc, float dd, float ee, cplxf* rf) {
    const float A = cc, B = 0.25f * cc * cc - ee, Cc = -0.125f * dd * dd;
    const float A3 = A * (1.0f / 3.0f);
    // depressed resolvent t^3 + P t + Q = 0, m = t - A/3
    const float P = B - A * A3;
    const float Q = (2.0f / 27.0f) * A * A * A - A3 * B + Cc;
    const float disc = 0.25f * Q * Q + P * P * P * (1.0f / 27.0f);
    float t;
    if (disc <= 0.0f) {
        const float sp = sqrt_f(fmaxf(-P * (1.0f / 3.0f), 0.0f));
        const float c3 = (sp > 0.0f) ? fminf(fmaxf(-0.5f * Q * rcp_f(sp * sp * sp), -1.0f), 1.0f) : 0.0f;
        t = 2.0f * sp * cosf(acosf(c3) * (1.0f / 3.0f));
    } else {
        const float w = cbrtf(-0.5f * Q - copysignf(sqrt_f(disc), Q));
        t = (w != 0.0f) ? w - P * rcp_f(3.0f * w) : 0.0f;
    }
    const float m = t - A3;
    if (!(m > 1e-4f * (fabsf(cc) + sqrt_f(fabsf(ee))))) return false;
    const float sq = sqrt_f(2.0f * m);
    const float hh = 0.5f * cc + m, kk = dd * rcp_f(2.0f * sq);
    // the two constant terms multiply to ee: the one of smaller magnitude
    // (where hh -+ kk cancels, e.g. the root u ~ -dd/ee of a tiny ee) from
    // ee over the other
    float c1 = hh + kk, c2 = hh - kk;
    if (fabsf(c1) >= fabsf(c2)) c2 = ee * rcp_f(c1);
    else c1 = ee * rcp_f(c2);
#pragma unroll
    for (int qd = 0; qd < 2; ++qd) {
        // x^2 + b x + c
        const float b = (qd == 0) ? -sq : sq, c = (qd == 0) ? c1 : c2;
        const float D = b * b - 4.0f * c;
        cplxf x1, x2;
        if (D >= 0.0f) {
            const float r1 = -0.5f * (b + copysignf(sqrt_f(D), b));
            x1 = mkf(r1);
            x2 = mkf(c * rcp_f(r1));
        } else {
            const float im = 0.5f * sqrt_f(-D);
            x1 = mkf(-0.5f * b, im);
            x2 = mkf(-0.5f * b, -im);
        }
        // u = 1 / x
        const float i1 = rcp_f(x1.re * x1.re + x1.im * x1.im), i2 = rcp_f(x2.re * x2.re + x2.im * x2.im);
        rf[2 * qd] = mkf(x1.re * i1, -x1.im * i1);
        rf[2 * qd + 1] = mkf(x2.re * i2, -x2.im * i2);
    }
    return true;
}

// The quartic's roots for GEOMU's compared mode: float closed-form seeds
// (quartic_u_seeds_f), classified real or complex in float, then three
// Newton steps in double on the full polynomial; false (the caller runs the
// FP64 ZROOTS) where the float phase cannot decide: a seed that is not
// finite, an imaginary part between 1e-5 and 1e-2 of the modulus, or two
// real seeds within 1e-3 of each other.
__device__ inline bool zroots4_mixed(const cplx* a, cplx* roots) {
    cplxf rf[4];
    if (!quartic_u_seeds_f((float)a[2].re, (float)a[3].re, (float)a[4].re, rf)) return false;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float ax = cabs_f(rf[j]);
        if (!(ax > 0.0f) || !isfinite(ax)) return false;
        const float rim = fabsf(rf[j].im) / ax;
        if (rim <= 1e-5f) rf[j] = mkf(rf[j].re);
        else if (rim < 1e-2f) return false;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = i + 1; j < 4; ++j)
            if (rf[i].im == 0.0f && rf[j].im == 0.0f &&
                fabsf(rf[i].re - rf[j].re) <= 1e-3f * fmaxf(fabsf(rf[i].re), fabsf(rf[j].re)))
                return false;
    // three Newton steps in double on the full polynomial: from a float
    // root of a simple (separated) root they reach double precision; a real
    // root stays on the real axis
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        cplx x = mk((double)rf[j].re, (double)rf[j].im);
        for (int it = 0; it < 3; ++it) {
            cplx p = a[4], dp = mk(0.0);
#pragma unroll
            for (int k = 3; k >= 0; --k) {
                dp = x * dp + p;
                p = x * p + a[k];
            }
            if (dp.re == 0.0 && dp.im == 0.0) break;
            x = x - p / dp;
        }
        roots[j] = x;
    }
    insertion_sort_re<cplx, 4>(roots);
    return true;
}

// SNCNDN (geokerr_original.f:3277-3340).  Its two arrays of the descending
// sequence, indexed by the iteration count, live on the device in LDS, one
// column per thread (every kernel that reaches this runs blocks of at most
// 256 threads): dynamically indexed, they would otherwise sit in private
// (scratch) memory.  Same values, same order of operations.
__device__ inline void sncndn(double uu, double emmc, double& sn, double& cn, double& dn) {
    const double CA = 3e-8;
    double emc = emmc, u = uu, a, b, c = 0.0, d = 0.0;
#if defined(__HIP_DEVICE_COMPILE__)
    __shared__ double s_em[13 * 256], s_en[13 * 256];
    double* const em_col = s_em + threadIdx.x;
    double* const en_col = s_en + threadIdx.x;
#define NP_EM(i) em_col[(i) * 256]
#define NP_EN(i) en_col[(i) * 256]
#else
    double em[13], en[13];
#define NP_EM(i) em[i]
#define NP_EN(i) en[i]
#endif
    int l = 0;
    if (emc != 0.0) {
        const bool bo = (emc < 0.0);
        if (bo) {
            d = 1.0 - emc;
            emc = -emc / d;
            d = sqrt(d);
            u = d * u;
        }
        a = 1.0;
        dn = 1.0;
        for (int i = 1; i <= 13; ++i) {
            l = i;
            NP_EM(i - 1) = a;
            emc = sqrt(emc);
            NP_EN(i - 1) = emc;
            c = 0.5 * (a + emc);
            if (fabs(a - emc) <= CA * a) break;
            emc = a * emc;
            a = c;
        }
        u = c * u;
        sincos(u, &sn, &cn);
        if (sn != 0.0) {
            a = cn / sn;
            c = a * c;
            for (int ii = l; ii >= 1; --ii) {
                b = NP_EM(ii - 1);
                a = c * a;
                c = dn * c;
                dn = (NP_EN(ii - 1) + a) / (b + a);
                a = c / b;
            }
            a = 1.0 / sqrt(c * c + 1.0);
            sn = (sn < 0.0) ? -a : a;
            cn = c * sn;
        }
        if (bo) {
            a = dn;
            dn = cn;
            cn = a;
            sn = sn / d;
        }
    } else {
        cn = 1.0 / cosh(u);
        dn = cn;
        sn = tanh(u);
    }
#undef NP_EM
#undef NP_EN
}

// ------------------------------- GEOMU --------------------------------------

// GEOKERR's variables shared by its GEOMU and GEOPHITIME calls along one
// geodesic (geokerr_original.f:398-412).
struct GkVars {
    int32_t ncase, fp64_roots;   // fp64_roots: set when the mixed GEOMU fell back to FP64 ZROOTS
    double u1, u2, u3, u4, h1;
    double iu0, rffu0, rffu1;
    double i1mu, i3mu, rffmu1, rffmu2, rffmu3;
    double rdc, rjc, tu01, tu02, tu03, tu04, tmu1, tmu3, phimu1, phimu3;
};

// calcimusym / calcimuasym (geokerr_original.f:3206-3240 / 3137-3172): the
// mu0 and turning-point integrals and their RF pieces.  The symmetric
// tan(phi0) < 0 test is phi0 > pi/2 compared with the double nearest pi/2,
// so the branch does not hinge on the last bit of a device tan() there; the
// asymmetric tan(asin(s0)) < 0 test never holds for s0 in [0, 1] with a
// correctly rounded asin and is omitted.
__device__ inline void calcimu(bool sym, double a, double mneg, double mpos, double mu0, double muplus,
                               double& i1mu, double& i3mu, double& rf0, double& rfc) {
    const double fac = sym ? fabs(a) * sqrt(mpos - mneg) : fabs(a) * muplus;
    const double m1 = sym ? -mneg / (mpos - mneg) : mneg / mpos;
    double s0, phi0 = 0.0;
    if (sym) {
        phi0 = acos(mu0 / muplus);
        s0 = sin(phi0);
    } else {
        s0 = sqrt((mpos - mu0 * mu0) / (mpos - mneg));
    }
    const double ak = sqrt(1.0 - m1);
    const double q0 = (1.0 - s0 * ak) * (1.0 + s0 * ak);
    rf0 = rf_d(1.0 - s0 * s0, q0, 1.0);
    rfc = rf_d(0.0, m1, 1.0);
    const double imum = 2.0 * rfc;
    double imu0 = s0 * rf0;
    if (sym && phi0 > kHalfPi) imu0 = imum - imu0;
    i1mu = imu0 / fac;
    i3mu = sym ? imum / fac : imum / fac / 2.0;
}

// RF term of calcimusymf / calcimuasymf (geokerr_original.f:3244-3274 /
// 3175-3204), the only output GEOMU keeps (RFFMU2).
__device__ inline double calcimuf_rf(bool sym, double mneg, double mpos, double muf, double muplus) {
    const double m1 = sym ? -mneg / (mpos - mneg) : mneg / mpos;
    const double sf = sym ? sin(acos(muf / muplus)) : sqrt((mpos - muf * muf) / (mpos - mneg));
    const double ak = sqrt(1.0 - m1);
    const double qf = (1.0 - sf * ak) * (1.0 + sf * ak);
    return rf_d(1.0 - sf * sf, qf, 1.0);
}

// The coefficients of U(u) = 1 + cc u^2 + dd u^3 + ee u^4 and the pieces
// beside them, shared by GEOMU and GEOR (the turning points derive from them).
struct UCoefs {
    double uplus, l2, ql2, cc, dd, ee;
};
__device__ __forceinline__ UCoefs u_coefs(double a, double l, double q2) {
    UCoefs k;
    k.uplus = 1.0 / (1.0 + sqrt(1.0 - a * a));
    k.l2 = l * l;
    k.ql2 = q2 + k.l2;
    k.cc = a * a - q2 - k.l2;
    k.dd = 2.0 * ((a - l) * (a - l) + q2);
    k.ee = -a * a * q2;
    return k;
}

// GEOMU (geokerr_original.f:771-1186).  u0, uf, tpm and tpr are the
// Fortran's by-reference arguments; `alias` reproduces a caller passing one
// variable as both u0 and uf (geokerr_estimate_periods).  With FIRSTPT =
// false the quartic roots, ncase, the u0 piece of rows 1/2/7/8 and the mu0
// integrals come from v.  ROW6 = false leaves Table 1 row 6 (no real root of
// U, ncase 6) out of the generated code: the caller guarantees it is not
// reached (its sextic root finders set the register budget of a kernel).
template <bool FIRSTPT, bool PHT, bool MIXED = false, bool ROW6 = true>
__device__ __forceinline__ void geomu(GkVars& v, double& u0, double& uf, bool alias, const double mu0, const double a,
                             const double l, const double q2, int& tpm, int& tpr, const double su,
                             const double sm, double& muf, double& iu) {
    const double one = 1.0, two = 2.0, third = 0.3333333333333333;
    const UCoefs kc = u_coefs(a, l, q2);
    const double uplus = kc.uplus, l2 = kc.l2, ql2 = kc.ql2, cc = kc.cc, dd = kc.dd, ee = kc.ee;
    double iu1 = 0.0;
    int ncase = v.ncase;
    double u1 = v.u1, u2 = v.u2, u3 = v.u3, u4 = v.u4;
    auto set_uf = [&](double x) { uf = x; if (alias) u0 = x; };
    auto set_u0 = [&](double x) { u0 = x; if (alias) uf = x; };
    iu = 0.0;
    muf = 0.0;

    if (ee == 0.0 && dd != 0.0) {
        // The cubic's roots are formed again on every call (not only with
        // FIRSTPT); a sample the driver places on the turning point u2 moves
        // mu by ~1e-8 per ulp of u2.
        const double qq = cc * cc / dd / dd / 9.0;
        const double rr = (two * (cc * cc * cc) / (dd * dd * dd) + 27.0 / dd) / 54.0;
        const double dis = rr * rr - qq * qq * qq;
        if (dis < -1e-16) {
            const double theta = acos(rr / pow(qq, 1.5));
            u1 = -two * sqrt(qq) * cos(theta / 3.0) - cc / dd / 3.0;
            u2 = -two * sqrt(qq) * cos((theta - two * kPi) / 3.0) - cc / dd / 3.0;
            u3 = -two * sqrt(qq) * cos((theta + two * kPi) / 3.0) - cc / dd / 3.0;
            for (int guard = 0; guard < 3; ++guard) {
                if (u0 <= u2) {
                    if (uf > u2) set_uf(u2);
                    ncase = 1;
                    if (FIRSTPT && u0 != u2)
                        v.iu0 = ellcubicreal<0>(-u1, one, u2, -one, u3, -one, 0.0, 0.0, v.rffu0, u0, u2);
                    else if (u0 == u2)
                        v.iu0 = 0.0;
                    if (uf != u2) iu1 = ellcubicreal<0>(-u1, one, u2, -one, u3, -one, 0.0, 0.0, v.rffu1, uf, u2);
                    iu = su * (v.iu0 - neg1pow(tpr) * iu1) / sqrt(dd);
                    break;
                } else if (u0 >= u3) {
                    if (uf < u3) set_uf(u3);
                    ncase = 2;
                    if (FIRSTPT && u0 != u3)
                        v.iu0 = -ellcubicreal<0>(-u1, one, -u2, one, -u3, one, 0.0, 0.0, v.rffu0, u3, u0);
                    else if (u0 == u3)
                        v.iu0 = 0.0;
                    if (uf != u3) iu1 = -ellcubicreal<0>(-u1, one, -u2, one, -u3, one, 0.0, 0.0, v.rffu1, u3, uf);
                    iu = su * (v.iu0 - neg1pow(tpr) * iu1) / sqrt(dd);
                    break;
                }
                set_u0((su == 1.0) ? u3 : u2);   // forbidden region: the reference's input edit
            }
        } else if (fabs(dis) < 1e-16) {
            u1 = -two * sqrt(qq) - cc / dd / 3.0;
            u2 = -two * sqrt(qq) * cos((two * kPi) / 3.0) - cc / dd / 3.0;
            u3 = u2;
            tpr = 0;
            if (u0 <= u2) {
                ncase = 1;
                if (uf > u2) {
                    set_uf(u2);
                    iu = su * 1e300;
                } else {
                    const double farg = (sqrt(uf - u1) + sqrt(u2 - u1)) / fabs(sqrt(uf - u1) - sqrt(u2 - u1));
                    const double sarg = (sqrt(u0 - u1) + sqrt(u2 - u1)) / fabs(sqrt(u0 - u1) - sqrt(u2 - u1));
                    iu = su * (log(farg) - log(sarg)) / sqrt((u2 - u1) * dd);
                }
            } else if (u0 >= u2) {
                ncase = 2;
                if (uf < u2) {
                    set_uf(u2);
                    iu = su * 1e300;
                } else {
                    const double farg = (sqrt(uf - u1) + sqrt(u2 - u1)) / fabs(sqrt(uf - u1) - sqrt(u2 - u1));
                    const double sarg = (sqrt(u0 - u1) + sqrt(u2 - u1)) / fabs(sqrt(u0 - u1) - sqrt(u2 - u1));
                    iu = -su * (log(farg) - log(sarg)) / sqrt((u2 - u1) * dd);
                }
            }
        } else {
            ncase = 3;
            const double aa = -sgn1(rr) * pow(fabs(rr) + sqrt(dis), third);
            const double bb = (aa != 0.0) ? qq / aa : 0.0;
            u1 = (aa + bb) - cc / dd / 3.0;
            const double f = -one / dd / u1;
            const double g = f / u1;
            if (uf > u0) iu = su * ellcubiccomplex<0>(-u1, one, 0.0, 0.0, f, g, one, v.rffu0, u0, uf) / sqrt(dd);
            else if (uf < u0) iu = -su * ellcubiccomplex<0>(-u1, one, 0.0, 0.0, f, g, one, v.rffu0, uf, u0) / sqrt(dd);
        }
        if (q2 == 0.0) {
            const double s1 = sgn1(mu0);
            if (l2 < a * a && mu0 != 0.0) {
                const double muplus = s1 * sqrt(one - l2 / a / a);
                const double x = mu0 / muplus;
                const double asech = (x > 0.0) ? log((1.0 + sqrt(1.0 - x * x)) / x) : 0.0;
                muf = muplus / cosh(fabs(a * muplus) * iu - s1 * sm * asech);
            }
        } else if (a == 0.0) {
            const double a1 = sm;
            double muplus = sqrt(q2 / ql2);
            if (mu0 > muplus) muplus = mu0;
            v.i1mu = acos(mu0 / muplus) / sqrt(ql2);
            v.i3mu = kPi / sqrt(ql2);
            const double i1mu = v.i1mu, i3mu = v.i3mu;
            if (sm == 1.0) tpm = add_i32(f_int((iu - i1mu) / i3mu), f_int((1 + sgn1((iu - i1mu) / i3mu)) / 2.0));
            else tpm = f_int((iu + i1mu) / i3mu);
            const double a2 = sm * neg1pow(tpm);
            const double a3 = two * f_int((two * (double)tpm + 3.0 - sm) / 4.0) - one;
            muf = -muplus * cos(sqrt(ql2) * (iu - a1 * i1mu - a3 * i3mu) / a2);
        }
    } else if (ee == 0.0 && dd == 0.0) {
        ncase = 4;
        tpr = 0;
        iu = su * (uf - u0);
    } else {
        if constexpr (FIRSTPT) {
            cplx c[5], root[4];
            c[0] = mk(one); c[1] = mk(0.0); c[2] = mk(cc); c[3] = mk(dd); c[4] = mk(ee);
            if constexpr (MIXED) {
                if (!zroots4_mixed(c, root)) {
                    zroots<4>(c, root);
                    v.fp64_roots = 1;
                }
            } else {
                zroots<4>(c, root);
            }
            int nreal = 0;
            for (int i = 0; i < 4; ++i) nreal += (root[i].im == 0.0);
            if (nreal == 2) {
                ncase = 5;
                u1 = root[0].re;
                u4 = (root[1].im == 0.0) ? root[1].re : root[3].re;
            } else if (nreal == 0) {
                ncase = 6;
            } else {
                u1 = root[0].re; u2 = root[1].re; u3 = root[2].re; u4 = root[3].re;
                for (int guard = 0; guard < 3; ++guard) {
                    if (u2 > uplus && u3 > uplus) { ncase = 5; break; }
                    else if (u0 <= u2) { ncase = 7; break; }
                    else if (u0 >= u3) { ncase = 8; break; }
                    set_u0((su == 1.0) ? 0.0 : uplus);
                }
            }
        }
        if (ncase == 5) {
            const double qs = sgn1(q2);
            const double f = -qs * one / fabs(ee) / u1 / u4;
            const double g = (u4 + u1) / u1 / u4 * f;
            if (uf > u0)
                iu = su * ellquarticcomplex<0>(-u1, one, qs * u4, -qs * one, 0.0, 0.0, f, g, one, v.rffu0, u0, uf) /
                     sqrt(fabs(ee));
            else if (uf < u0)
                iu = -su * ellquarticcomplex<0>(-u1, one, qs * u4, -qs * one, 0.0, 0.0, f, g, one, v.rffu0, uf, u0) /
                     sqrt(fabs(ee));
        } else if (ncase == 6) {
          if constexpr (ROW6) {
            // the sextic for h1 (Eq. 21) is outside the Fortran's FIRSTPT test
            const double se = sqrt(ee);
            cplx co[7], hroots[6];
            co[0] = mk(one); co[1] = mk(-cc / se); co[2] = mk(-one);
            co[3] = mk(se * (two * cc / ee - (dd / ee) * (dd / ee)));
            co[4] = mk(-one); co[5] = mk(-cc / se); co[6] = mk(one);
            zroots<6>(co, hroots);
            double h1 = 0.0;
            for (int i = 0; i < 6 && h1 == 0.0; ++i)
                if (hroots[i].im == 0.0) h1 = hroots[i].re;
            v.h1 = h1;
            const double h2 = one / h1;
            const double g1 = dd / ee / (h2 - h1);
            const double g2 = -g1;
            const double f1 = one / sqrt(ee);
            const double f2 = f1;
            if (uf > u0) iu = su * elldoublecomplex<0>(f1, g1, h1, f2, g2, h2, 0.0, 0.0, v.rffu0, u0, uf) / sqrt(fabs(ee));
            else if (uf < u0) iu = -su * elldoublecomplex<0>(f1, g1, h1, f2, g2, h2, 0.0, 0.0, v.rffu0, uf, u0) / sqrt(fabs(ee));
          }
        } else {
            if (fabs(u3 - u2) > 1e-12) {
                if (ncase == 7) {
                    if (uf > u2) set_uf(u2);
                    if (FIRSTPT && u0 != u2)
                        v.iu0 = ellquarticreal<0>(-u1, one, u2, -one, u3, -one, u4, -one, 0.0, 0.0, v.rffu0, u0, u2);
                    else if (u0 == u2)
                        v.iu0 = 0.0;
                    if (uf != u2)
                        iu1 = ellquarticreal<0>(-u1, one, u2, -one, u3, -one, u4, -one, 0.0, 0.0, v.rffu1, uf, u2);
                    iu = su * (v.iu0 - neg1pow(tpr) * iu1) / sqrt(fabs(ee));
                } else if (ncase == 8) {
                    if (uf < u3) set_uf(u3);
                    if (FIRSTPT && u0 != u3)
                        v.iu0 = -ellquarticreal<0>(-u1, one, -u2, one, -u3, one, u4, -one, 0.0, 0.0, v.rffu0, u3, u0);
                    else if (u0 == u3)
                        v.iu0 = 0.0;
                    if (uf != u3)
                        iu1 = -ellquarticreal<0>(-u1, one, -u2, one, -u3, one, u4, -one, 0.0, 0.0, v.rffu1, u3, uf);
                    iu = su * (v.iu0 - neg1pow(tpr) * iu1) / sqrt(fabs(ee));
                }
            } else {
                tpr = 0;
                if (ncase == 7) {
                    if (uf > u2) set_uf(u2);
                    v.iu0 = ellquarticreal<0>(-u1, one, u2, -one, u3, -one, u4, -one, 0.0, 0.0, v.rffu0, u0, uf);
                    iu = su * v.iu0 / sqrt(fabs(ee));
                } else if (ncase == 8) {
                    if (uf < u3) set_uf(u3);
                    v.iu0 = -ellquarticreal<0>(-u1, one, -u2, one, -u3, one, u4, -one, 0.0, 0.0, v.rffu0, u0, uf);
                    iu = su * v.iu0 / sqrt(fabs(ee));
                } else {
                    ncase = 0;
                }
            }
        }
        // roots of M(mu) (Eq. 12) and the Jacobi inversion (Eqs. 30-42)
        const double amq = a * a - ql2;
        const double yy = -0.5 * (amq + sgn1(amq) * sqrt(amq * amq + 4.0 * q2 * a * a));
        double mneg, mpos;
        if (amq < 0.0) { mneg = -yy / a / a; mpos = q2 / yy; }
        else { mneg = q2 / yy; mpos = -yy / a / a; }
        if (mpos > 1.0) mpos = 1.0;
        double muplus = sqrt(mpos), muminus;
        const bool sym = mneg < 0.0;
        if (sym) {
            if (muplus < mu0) { muplus = mu0; mpos = muplus * muplus; }
            muminus = -muplus;
        } else {
            muplus = sgn1(mu0) * muplus;
            muminus = sgn1(mu0) * sqrt(mneg);
            if (fabs(muminus) > fabs(mu0)) { muminus = mu0; mneg = muminus * muminus; }
            if (fabs(mu0) > fabs(muplus)) { muplus = mu0; mpos = muplus * muplus; }
        }
        if constexpr (FIRSTPT) calcimu(sym, a, mneg, mpos, mu0, muplus, v.i1mu, v.i3mu, v.rffmu1, v.rffmu3);
        const double i1mu = v.i1mu, i3mu = v.i3mu;
        const double a1 = sm;
        if (sm == 1.0) {
            tpm = add_i32(f_int((iu - i1mu) / i3mu),
                          sym ? f_int((1 + sgn1((iu - i1mu) / i3mu)) / 2.0)
                              : f_int(fabs((1 + sgn1((iu - i1mu) / i3mu)) / 2.0)));
        } else {
            tpm = f_int((iu + i1mu) / i3mu);
        }
        const double a2 = sm * neg1pow(tpm);
        const double a3 = two * f_int((two * (double)tpm + 3.0 - sm) / 4.0) - one;
        const double iarg = fabs(a) / a2 * (iu - a1 * i1mu - a3 * i3mu);
        double sn, cn, dn;
        if (sym) {
            sncndn(sqrt(mpos - mneg) * iarg, -mneg / (mpos - mneg), sn, cn, dn);
            muf = muminus * cn;
        } else {
            sncndn(fabs(muplus) * iarg, mneg / mpos, sn, cn, dn);
            muf = muminus / dn;
        }
        if constexpr (PHT) v.rffmu2 = calcimuf_rf(sym, mneg, mpos, muf, muplus);
    }
    v.ncase = ncase;
    v.u1 = u1; v.u2 = u2; v.u3 = u3; v.u4 = u4;
}

// -------------------------------- GEOR --------------------------------------

// calcimusymf / calcimuasymf (geokerr_original.f:3244-3274 / 3175-3204):
// the mu integral from muf to the turning point muplus from its RF piece
// rff (calcimuf_rf); i3mu is calcimusym's half-period integral (the
// symmetric tan(phif) < 0 test is phif > pi/2, as in calcimu).
__device__ __forceinline__ double calcimuf_from_rf(bool sym, double a, double mneg, double mpos, double muf,
                                                   double muplus, double i3mu, double rff) {
    const double fac = sym ? fabs(a) * sqrt(mpos - mneg) : fabs(a) * muplus;
    if (!sym) return sqrt((mpos - muf * muf) / (mpos - mneg)) * rff / fac;
    const double phif = acos(muf / muplus);
    double imuf = sin(phif) * rff;
    if (phif > kHalfPi) imuf = i3mu * fac - imuf;
    return imuf / fac;
}

struct GeorResult {
    double uf, imu;
    int tpr, ncase;
};

// GEOR (geokerr_original.f:2672-3134) with the Fortran's argument contract:
// u0, uf and tpr by reference, IMU written into imu (the caller's IU; left
// as it was on the asymmetric-roots exit), the variables shared along a
// geodesic in v written where FIRSTPT (or the Fortran unconditionally)
// writes them and read otherwise; PHT forms the RF pieces GEOPHITIME takes
// (RFFU0 in rows 3/5/6, RFFU1 in rows 1/2/7/8, RFFMU2 always).  The
// equal-root cubic branch compares against the uf passed in; row 2 with
// FIRSTPT off sets IU0 = 0 whenever u0 != u3 (:2795-2799); rows 3, 5, 6
// and the q2 = 0 / forbidden-mu cases return uf = -1, ncase = 0.  The
// forbidden-region retry runs at most 3 times (the Fortran's is unbounded).
// ROW6 as in geomu.
template <bool FIRSTPT, bool PHT, bool ROW6 = true>
__device__ __forceinline__ void geor_v(GkVars& v, double& u0, double& uf, const double mu0, const double muf,
                                       const double a, const double l, const double q2, double& imu,
                                       const int tpm, int& tpr, const double su, const double sm) {
    const double one = 1.0, two = 2.0, half = 0.5, third = 0.3333333333333333, three = 3.0;
    const UCoefs kc = u_coefs(a, l, q2);
    const double uplus = kc.uplus, l2 = kc.l2, ql2 = kc.ql2, cc = kc.cc, dd = kc.dd, ee = kc.ee;
    const unsigned utpm = (unsigned)tpm;
    double dummy = 0.0, sn, cn, dn;

    // the mu integral in the special cases (:2749-2770)
    if (q2 == 0.0) {
        if (l2 >= a * a || mu0 == 0.0) {
            imu = 0.0;
            uf = -1.0;
            v.ncase = 0;
            return;
        }
        const double s1 = sgn1(mu0);
        const double muplus = s1 * sqrt(one - l2 / a / a);
        const double x0 = mu0 / muplus, xf = muf / muplus;
        const double as0 = (x0 > 0.0) ? log((one + sqrt(one - x0 * x0)) / x0) : 0.0;
        const double asf = (xf > 0.0) ? log((one + sqrt(one - xf * xf)) / xf) : 0.0;
        v.i1mu = one / fabs(a * muplus) * as0;
        imu = s1 * sm * v.i1mu + s1 * sm * neg1pow(utpm + 1u) * (one / fabs(a * muplus) * asf);
    } else if (a == 0.0) {
        double muplus = sqrt(q2 / ql2);
        if (mu0 > muplus) muplus = mu0;
        v.i1mu = (half * kPi - asin(mu0 / muplus)) / sqrt(ql2);
        v.i3mu = kPi / sqrt(ql2);
        const double a3 = two * f_int((two * (double)tpm + 3.0 - sm) / 4.0) - one;
        const double i2mu = (half * kPi + asin(muf / muplus)) / sqrt(ql2);
        imu = sm * v.i1mu + sm * neg1pow(utpm) * i2mu + a3 * v.i3mu;
    }

    if (ee == 0.0 && dd != 0.0) {
        const double qq = cc * cc / dd / dd / 9.0;
        const double rr = (two * (cc * cc * cc) / (dd * dd * dd) + 27.0 / dd) / 54.0;
        const double dis = rr * rr - qq * qq * qq;
        if (dis < -1e-16) {
            const double theta = acos(rr / pow(qq, 1.5));
            const double u1 = -two * sqrt(qq) * cos(theta / 3.0) - cc / dd / 3.0;
            const double u2 = -two * sqrt(qq) * cos((theta - two * kPi) / 3.0) - cc / dd / 3.0;
            const double u3 = -two * sqrt(qq) * cos((theta + two * kPi) / 3.0) - cc / dd / 3.0;
            v.u1 = u1; v.u2 = u2; v.u3 = u3;
            for (int guard = 0; guard < 3; ++guard) {
                if (u0 <= u2) {
                    // Table 2 row 1
                    v.ncase = 1;
                    if (FIRSTPT && u0 != u2)
                        v.iu0 = ellcubicreal<0>(-u1, one, u2, -one, u3, -one, 0.0, 0.0, v.rffu0, u0, u2);
                    else if (u0 == u2)
                        v.iu0 = 0.0;
                    const double m1 = (u3 - u2) / (u3 - u1);
                    const double jarg = sqrt((u3 - u1) * dd) * half * (imu - su * v.iu0 / sqrt(dd));
                    sncndn(jarg, m1, sn, cn, dn);
                    uf = u1 + (u2 - u1) * (cn * cn / (dn * dn));
                    tpr = f_int((sgn1(su * (imu - su * v.iu0 / sqrt(dd))) + one) / two);
                    if (PHT) (void)ellcubicreal<0>(-u1, one, u2, -one, u3, -one, 0.0, 0.0, v.rffu1, uf, u2);
                    break;
                } else if (u0 >= u3) {
                    // Table 2 row 2
                    v.ncase = 2;
                    if (FIRSTPT && u0 != u3)
                        v.iu0 = ellcubicreal<0>(-u1, one, -u2, one, -u3, one, 0.0, 0.0, v.rffu0, u3, u0);
                    else if (u0 != u3)
                        v.iu0 = 0.0;
                    const double m1 = (u3 - u2) / (u3 - u1);
                    const double jarg = sqrt((u3 - u1) * dd) * half * (imu + su * v.iu0 / sqrt(dd));
                    sncndn(jarg, m1, sn, cn, dn);
                    uf = u1 + (u3 - u1) * (dn * dn / (cn * cn));
                    tpr = f_int((-sgn1(su * (imu + v.iu0 / sqrt(dd))) + one) / two);
                    if (PHT) (void)ellcubicreal<0>(-u1, one, -u2, one, -u3, one, 0.0, 0.0, v.rffu1, u3, uf);
                    break;
                }
                u0 = (su == 1.0) ? u3 : u2;   // forbidden region: the reference's input edit
            }
        } else if (fabs(dis) < 1e-16) {
            // equal roots u2 = u3 (:2832-2860)
            tpr = 0;
            const double u1 = -two * sqrt(qq) - cc / dd / 3.0;
            const double u2 = -two * sqrt(qq) * cos((two * kPi) / 3.0) - cc / dd / 3.0;
            v.u1 = u1; v.u2 = u2; v.u3 = u2;
            if (u0 <= u2) {
                v.ncase = 1;
                if (uf > u2) {
                    uf = u2;
                } else {
                    const double sarg = sqrt((u0 - u1) / (u2 - u1));
                    const double th = tanh(sqrt((u2 - u1) * dd) * half * su * imu + half * log((one + sarg) / (one - sarg)));
                    uf = u1 + (u2 - u1) * th * th;
                }
            } else {
                v.ncase = 2;
                if (uf < u2) {
                    uf = u2;
                } else {
                    const double sarg = sqrt((u2 - u1) / (u0 - u1));
                    const double th = tanh(-sqrt((u2 - u1) * dd) * half * su * imu + half * log((one + sarg) / (one - sarg)));
                    uf = u1 + (u2 - u1) / (th * th);
                }
            }
        } else {
            // one real root: Table 2 row 3 (:2861-2895)
            v.ncase = 3;
            tpr = 0;
            const double aa = -sgn1(rr) * pow(fabs(rr) + sqrt(dis), third);
            const double bb = (aa != 0.0) ? qq / aa : 0.0;
            const double u1 = (aa + bb) - cc / dd / 3.0;
            v.u1 = u1;
            const double f = -one / dd / u1;
            const double g = f / u1;
            const double iut = (su > 0.0)
                ? ellcubiccomplex<0>(-u1, one, 0.0, 0.0, f, g, one, dummy, u0, uplus) / sqrt(dd)
                : ellcubiccomplex<0>(-u1, one, 0.0, 0.0, f, g, one, dummy, 0.0, u0) / sqrt(dd);
            if (imu > iut) {
                uf = -1.0;
                v.ncase = 0;
                return;
            }
            if (FIRSTPT) v.iu0 = su * ellcubiccomplex<0>(-u1, one, 0.0, 0.0, f, g, one, dummy, u1, u0) / sqrt(dd);
            const double c3 = (a != 0.0 || l != 0.0) ? (a + l) / (a - l) : -one;
            const double c2 = sqrt(u1 * (three * u1 + c3));
            const double c1 = sqrt(c2 * dd);
            const double m1 = half + (6.0 * u1 + c3) / (8.0 * c2);
            sncndn(c1 * (imu + v.iu0), m1, sn, cn, dn);
            uf = (c2 + u1 - (c2 - u1) * cn) / (one + cn);
            if (PHT) (void)ellcubiccomplex<0>(-u1, one, 0.0, 0.0, f, g, one, v.rffu0, u0, uf);
        }
    } else if (ee == 0.0 && dd == 0.0) {
        v.ncase = 4;
        uf = -1.0;
        tpr = 0;
    } else {
        // roots of M(mu) and the mu integral (:2901-2945)
        const double amq = a * a - ql2;
        const double yy = -0.5 * (amq + sgn1(amq) * sqrt(amq * amq + 4.0 * q2 * a * a));
        double mneg, mpos;
        if (amq < 0.0) { mneg = -yy / a / a; mpos = q2 / yy; }
        else { mneg = q2 / yy; mpos = -yy / a / a; }
        double muplus = sqrt(mpos);
        const double a2 = sm * neg1pow(utpm + 1u);
        const double a3 = two * f_int((2.0 * (double)tpm - sm + 1.0) / 4.0);
        const bool sym = mneg < 0.0;
        if (sym) {
            if (mu0 > muplus) { muplus = mu0; mpos = muplus * muplus; }
        } else {
            if (fabs(muf) < sqrt(mneg)) {
                uf = -1.0;
                v.ncase = 0;
                tpr = 0;
                return;
            }
            if (sgn1(mu0) == -1.0) muplus = -muplus;
            if (fabs(muplus) < fabs(mu0)) { muplus = mu0; mpos = muplus * muplus; }
            mneg = fmin(mu0 * mu0, mneg);
        }
        if constexpr (FIRSTPT) calcimu(sym, a, mneg, mpos, mu0, muplus, v.i1mu, v.i3mu, v.rffmu1, v.rffmu3);
        v.rffmu2 = calcimuf_rf(sym, mneg, mpos, muf, muplus);
        const double i2mu = calcimuf_from_rf(sym, a, mneg, mpos, muf, muplus, v.i3mu, v.rffmu2);
        imu = sm * v.i1mu + a2 * i2mu + a3 * v.i3mu;

        // roots of U(u) and the case (:2947-2993)
        if constexpr (FIRSTPT) {
            cplx c[5], root[4];
            c[0] = mk(one); c[1] = mk(0.0); c[2] = mk(cc); c[3] = mk(dd); c[4] = mk(ee);
            zroots<4>(c, root);
            int nreal = 0;
            for (int i = 0; i < 4; ++i) nreal += (root[i].im == 0.0);
            if (nreal == 2) {
                v.ncase = 5;
                tpr = 0;
                v.u1 = root[0].re;
                v.u4 = (root[1].im == 0.0) ? root[1].re : root[3].re;
            } else if (nreal == 0) {
                v.ncase = 6;
                tpr = 0;
            } else {
                v.u1 = root[0].re; v.u2 = root[1].re; v.u3 = root[2].re; v.u4 = root[3].re;
                for (int guard = 0; guard < 3; ++guard) {
                    if (v.u2 > uplus && v.u3 > uplus) { v.ncase = 5; break; }
                    else if (u0 <= v.u2) { v.ncase = 7; break; }
                    else if (u0 >= v.u3) { v.ncase = 8; break; }
                    u0 = (su == 1.0) ? v.u3 : v.u2;
                }
            }
        }
        const double u1 = v.u1, u2 = v.u2, u3 = v.u3, u4 = v.u4;
        const double sqe = sqrt(fabs(ee));
        if (v.ncase == 5) {
            // Table 2 row 5 (:2995-3027)
            const double qs = sgn1(q2);
            const double f = -qs * one / fabs(ee) / u1 / u4;
            const double g = (u4 + u1) / u1 / u4 * f;
            const double iut = (su > 0.0)
                ? ellquarticcomplex<0>(-u1, one, qs * u4, -qs * one, 0.0, 0.0, f, g, one, dummy, u0, uplus) / sqe
                : ellquarticcomplex<0>(-u1, one, qs * u4, -qs * one, 0.0, 0.0, f, g, one, dummy, 0.0, u0) / sqe;
            if (imu > iut) {
                uf = -1.0;
                v.ncase = 0;
                tpr = 0;
                return;
            }
            const double ub = (qs == one) ? u1 : u4;
            const double m = -g * half;
            const double n2 = f - g * g / 4.0;
            const double c4 = sqrt((m - u4) * (m - u4) + n2);
            const double c5 = sqrt((m - u1) * (m - u1) + n2);
            const double c1 = sqrt(fabs(ee * c4 * c5));
            if (FIRSTPT)
                v.iu0 = su * ellquarticcomplex<0>(-u1, one, qs * u4, -qs * one, 0.0, 0.0, f, g, one, dummy, ub, u0) / sqe;
            const double m1 = qs * ((c4 + qs * c5) * (c4 + qs * c5) - (u4 - u1) * (u4 - u1)) / (4.0 * c4 * c5);
            sncndn(c1 * (imu + v.iu0), m1, sn, cn, dn);
            uf = (u4 * c5 + qs * u1 * c4 - (qs * u4 * c5 - u1 * c4) * cn) / ((c4 - qs * c5) * cn + (qs * c4 + c5));
            if (PHT) (void)ellquarticcomplex<0>(-u1, one, qs * u4, -qs * one, 0.0, 0.0, f, g, one, v.rffu0, u0, uf);
        } else if (v.ncase == 6) {
          if constexpr (ROW6) {
            // Table 2 row 6 (:3028-3095): h1 and mn2 from two sextics
            const double se = sqrt(ee);
            cplx co[7], hr[6];
            co[0] = mk(one); co[1] = mk(-cc / se); co[2] = mk(-one);
            co[3] = mk(se * (two * cc / ee - (dd / ee) * (dd / ee)));
            co[4] = mk(-one); co[5] = mk(-cc / se); co[6] = mk(one);
            zroots<6>(co, hr);
            double h1 = 0.0;
            for (int i = 0; i < 6 && h1 == 0.0; ++i)
                if (hr[i].im == 0.0) h1 = hr[i].re;
            if (h1 == 0.0) h1 = __builtin_nan("");
            v.h1 = h1;
            const double h2 = one / h1;
            const double g1 = dd / ee / (h2 - h1);
            const double g2 = -g1;
            const double f1 = one / sqrt(ee);
            const double f2 = f1;
            const double iut = (su > 0.0)
                ? elldoublecomplex<0>(f1, g1, h1, f2, g2, h2, 0.0, 0.0, dummy, u0, uplus) / sqe
                : elldoublecomplex<0>(f1, g1, h1, f2, g2, h2, 0.0, 0.0, dummy, 0.0, u0) / sqe;
            if (imu > iut) {
                uf = -1.0;
                v.ncase = 0;
                tpr = 0;
                return;
            }
            co[0] = mk(pow(ee, -three));
            co[1] = mk(-cc / pow(ee, three));
            co[2] = mk(-pow(ee, -two));
            co[3] = mk(-pow(ee, -two) * (pow(dd, two) / ee - two * cc));
            co[4] = mk(-one / ee);
            co[5] = mk(-cc / ee);
            co[6] = mk(one);
            zroots<6>(co, hr);
            double mn2 = 0.0;
            for (int i = 0; i < 6 && mn2 == 0.0; ++i)
                if (hr[i].im == 0.0) mn2 = hr[i].re;
            if (mn2 == 0.0) mn2 = __builtin_nan("");
            double pp = dd / (two * (ee * ee) * (mn2 * mn2 - one / ee));
            double m = -half * dd / ee - pp;
            if (m < pp) { const double t = pp; pp = m; m = t; }
            const double pr2 = one / (mn2 * ee);
            const double n = sqrt(mn2 - m * m);
            const double r = sqrt(pr2 - pp * pp);
            const double c4 = sqrt((m - pp) * (m - pp) + (n + r) * (n + r));
            const double c5 = sqrt((m - pp) * (m - pp) + (n - r) * (n - r));
            const double c1 = (c4 + c5) * half * sqe;
            const double c2 = sqrt((4.0 * n * n - (c4 - c5) * (c4 - c5)) / ((c4 + c5) * (c4 + c5) - 4.0 * n * n));
            const double c3 = m + c2 * n;
            if (FIRSTPT)
                v.iu0 = sgn1(u0 - c3) * elldoublecomplex<0>(f1, g1, h1, f2, g2, h2, 0.0, 0.0, dummy, c3, u0) / sqe;
            const double m1 = ((c4 - c5) / (c4 + c5)) * ((c4 - c5) / (c4 + c5));
            sncndn(c1 * (su * imu + v.iu0), m1, sn, cn, dn);
            const double sc = sn / cn;
            uf = c3 + (n * (one + c2 * c2) * sc) / (one - c2 * sc);
            if (PHT) (void)elldoublecomplex<0>(f1, g1, h1, f2, g2, h2, 0.0, 0.0, v.rffu0, u0, uf);
          }
        } else if (v.ncase == 7) {
            // Table 2 row 7 (:3098-3108)
            if (FIRSTPT && u0 != u2)
                v.iu0 = ellquarticreal<0>(-u1, one, u2, -one, u3, -one, u4, -one, 0.0, 0.0, v.rffu0, u0, u2);
            const double jarg = sqrt(fabs(ee) * (u3 - u1) * (u4 - u2)) * half * (imu - su * v.iu0 / sqrt(-ee));
            sncndn(jarg, (u4 - u1) * (u3 - u2) / ((u4 - u2) * (u3 - u1)), sn, cn, dn);
            uf = ((u2 - u1) * u3 * sn * sn - u2 * (u3 - u1)) / ((u2 - u1) * sn * sn - (u3 - u1));
            tpr = f_int((sgn1(su * (imu - su * v.iu0 / sqe)) + one) / two);
            if (PHT) (void)ellquarticreal<0>(-u1, one, u2, -one, u3, -one, u4, -one, 0.0, 0.0, v.rffu1, uf, u2);
        } else if (v.ncase == 8) {
            // Table 2 row 8 (:3109-3118)
            if (FIRSTPT && u0 != u3)
                v.iu0 = ellquarticreal<0>(-u1, one, -u2, one, -u3, one, u4, -one, 0.0, 0.0, v.rffu0, u3, u0);
            const double jarg = sqrt(fabs(ee) * (u3 - u1) * (u4 - u2)) * half * (imu + su * v.iu0 / sqrt(-ee));
            sncndn(jarg, (u4 - u1) * (u3 - u2) / ((u4 - u2) * (u3 - u1)), sn, cn, dn);
            uf = ((u4 - u3) * u2 * sn * sn - u3 * (u4 - u2)) / ((u4 - u3) * sn * sn - (u4 - u2));
            tpr = f_int((-sgn1(su * (imu + su * v.iu0 / sqrt(-ee))) + one) / two);
            if (PHT) (void)ellquarticreal<0>(-u1, one, -u2, one, -u3, one, u4, -one, 0.0, 0.0, v.rffu1, u3, uf);
        }
    }
}

// GEOR on a fresh geodesic (FIRSTPT on, PHT off): uf_in is the caller's UF,
// which only the equal-root cubic branch reads; tpr stays tpr_in where the
// Fortran does not write TPR.
__device__ __forceinline__ GeorResult geor(double u0, const double uf_in, const double mu0, const double muf,
                                           const double a, const double l, const double q2, const int tpm,
                                           const int tpr_in, const double su, const double sm) {
    GkVars v{};
    GeorResult R{uf_in, 0.0, tpr_in, 0};
    geor_v<true, false>(v, u0, R.uf, mu0, muf, a, l, q2, R.imu, tpm, R.tpr, su, sm);
    R.ncase = v.ncase;
    return R;
}

// FINDMROOTS (geokerr_original.f:607-642): the mu turning points (with
// a = 0 the Fortran divides by zero; the IEEE results are kept).
__device__ __forceinline__ void findmroots(double q2, double l, double a, double mu0, double& muminus,
                                           double& muplus) {
    const double ql2 = q2 + l * l;
    const double amq = a * a - ql2;
    const double yy = -0.5 * (amq + sgn1(amq) * sqrt(amq * amq + 4.0 * q2 * a * a));
    double mneg, mpos;
    if (amq < 0.0) { mneg = -yy / a / a; mpos = q2 / yy; }
    else { mneg = q2 / yy; mpos = -yy / a / a; }
    if (mpos > 1.0) mpos = 1.0;
    if (mneg > 0.0) {
        muplus = sgn1(mu0) * sqrt(mpos);
        muminus = sgn1(mu0) * sqrt(mneg);
        if (fabs(muminus) > fabs(mu0)) muminus = mu0;
        if (fabs(mu0) > fabs(muplus)) muplus = mu0;
    } else {
        muplus = sqrt(mpos);
        if (muplus < mu0) muplus = mu0;
        muminus = -muplus;
    }
}

__device__ __forceinline__ int m1pow_i(int n) { return (n & 1) ? -1 : 1; }

// INDEP_MUF (geokerr_original.f:645-706): the k-th of npts mu values spaced
// evenly in path length sum |dmu| from mu0 (after tpm0 turning points) to
// muf (after tpmf), and the turning points before it.  The Fortran's
// default-real constants in its integer expressions are kept as float.
__device__ __forceinline__ void indep_muf(double muminus, double muplus, double mu0, double muf, int tpm0,
                                          int tpmf, double sm, int k, int npts, double offset, double& mun,
                                          int& tpmk_out) {
    int tpmk;
    if (muf <= muplus && muf >= muminus) {
        const int dtpm = tpmf - tpm0;
        const int a1 = f_int(sm * (double)m1pow_i(tpm0));
        const int a2 = f_int(sm * (double)m1pow_i(tpmf));
        const int a3 = 2 * (int)floorf((2.0f * (float)dtpm + 3.0f - (float)a1) / 4.0f) - 1;
        const double dmu = a1 * (muplus - mu0) + a2 * (muf - muminus) + a3 * (muplus - muminus);
        const double w1 = (double)(((float)a1 + 1.0f) / 2.0f), w2 = (double)((1.0f - (float)a1) / 2.0f);
        const double mu1 = w1 * muplus + w2 * muminus;
        const double mu2 = w2 * muplus + w1 * muminus;
        const double dmu1 = fabs(mu1 - mu0);
        const double dmu2 = fabs(mu2 - mu1);
        const double dmuk = ((double)k - offset) * dmu / (double)(float)npts;
        int dmukdmu1 = f_int(dmuk / dmu1);
        if (dmukdmu1 < 0) dmukdmu1 = 1;
        const int tp2 = f_int((dmuk - dmu1) / dmu2);
        tpmk = (dmukdmu1 < 1 ? dmukdmu1 : 1) + (tp2 > 0 ? tp2 : 0);
        const int a1k = f_int(sm * (double)m1pow_i(tpm0));
        const int a2k = f_int(sm * (double)m1pow_i(tpmk + tpm0));
        const int a3k = (int)(2.0f * (float)(int)((2.0f * (float)tpmk + 3.0f - (float)a1k) / 4.0f) - 1.0f);
        mun = muminus + (double)(1.0f / (float)a2k) * (dmuk - a1k * (muplus - mu0) - a3k * (muplus - muminus));
    } else {
        mun = 0.0;
        tpmk = 0;
    }
    tpmk_out = tpmk + tpm0;
}

// ------------------------------ GEOPHITIME ----------------------------------

// Eq. (47): t_u from the four u integrals.
__device__ __forceinline__ double eq47(double a, double l, double uplus, double umi, double t1, double t2,
                                       double t3, double t4) {
    return (umi - 1.0 / uplus) * t1 + (umi * umi - 1.0 / (uplus * uplus)) * t2 -
           (2.0 * a * (a - l) + a * a / uplus + 1.0 / (uplus * uplus * uplus)) * t3 +
           (2.0 * a * (a - l) + a * a * umi + umi * umi * umi) * t4;
}

// Eq. (48) for one u0 / uf piece.
__device__ __forceinline__ double eq48(double a, double l, double uplus, double umi, double p1, double p2) {
    return -(l / uplus + 2.0 * (a - l)) * p1 + (l * umi + 2.0 * (a - l)) * p2;
}

// phifnkerr / tfnkerr (geokerr_original.f:2593-2628 / 2629-2671).
__device__ inline double phifnkerr(double u, double u1, double u2, double l, double a) {
    const double up = 1.0 / (1.0 + sqrt(1.0 - a * a));
    const double umi = 1.0 - sqrt(1.0 - a * a);
    const double s = sqrt(u - u1);
    const double su21 = sqrt(u2 - u1);
    const double sup = sqrt(1.0 - u1 / up);
    const double sum = sqrt(1.0 - u1 * umi);
    return (l + 2.0 * a * u2 - 2.0 * l * u2) / (su21 * (u2 * umi - 1.0) * (u2 / up - 1.0)) *
               log((su21 + s) / (su21 - s)) +
           (2.0 * a + l * (-2.0 + umi)) * sqrt(umi) * log((sum + s * sqrt(umi)) / (sum - s * sqrt(umi))) /
               (sum * (u2 * umi - 1.0) * (umi - 1.0 / up)) +
           (2.0 * a + l * (-2.0 + 1.0 / up)) * sqrt(1.0 / up) *
               log((sup + sqrt(1.0 / up) * s) / (sup - sqrt(1.0 / up) * s)) / (1.0 / up - umi) * sup *
               (u2 / up - 1.0);
}

__device__ inline double tfnkerr(double u0, double uf, double u1, double u2, double l, double a) {
    const double up = 1.0 / (1.0 + sqrt(1.0 - a * a));
    const double umi = 1.0 - sqrt(1.0 - a * a);
    const double sf = sqrt(uf - u1), ss = sqrt(u0 - u1);
    const double su21 = sqrt(u2 - u1), su1 = sqrt(-u1);
    const double sup = sqrt(1.0 - u1 / up), sum = sqrt(1.0 - u1 * umi);
    const double rup = sqrt(1.0 / up), rum = sqrt(umi);
    return sf / (uf * u1 * u2) - ss / (u0 * u1 * u2) -
           ((-u2 - 2.0 * u1 * (1.0 + u2 * umi + u2 / up)) * log((sf + su1) * (su1 - ss) / ((-sf + su1) * (ss + su1)))) /
               (2.0 * pow(-u1, 1.5) * (u2 * u2)) +
           ((1.0 - 2.0 * a * l * (u2 * u2 * u2) + a * a * (u2 * u2) * (1.0 + 2.0 * u2)) *
            log((su21 + sf) * (su21 - ss) / ((su21 - sf) * (su21 + ss)))) /
               ((u2 * u2) * su21 * (u2 * umi - 1.0) * (u2 / up - 1.0)) +
           ((-2.0 * a * l + a * a * (2.0 + umi) + umi * umi * umi) * rum *
            log((sum + sf * rum) * (sum - ss * rum) / ((sum - sf * rum) * (sum + ss * rum)))) /
               (sum * (u2 * umi - 1.0) * (umi - 1.0 / up)) +
           ((-2.0 * a * l + a * a * (2.0 + 1.0 / up) + 1.0 / (up * up * up)) * rup *
            log((sup + sf * rup) * (sup - ss * rup) / ((sup - sf * rup) * (sup + ss * rup)))) /
               ((1.0 / up - umi) * sup * (u2 / up - 1.0));
}

// calcphitmusym / calcphitmuasym with ellphitmu (geokerr_original.f:
// 2361-2592).  The mu0 and turning-point pieces (cached in v) are computed
// only with FIRSTPT; the muf pieces on every call.  When a root of M(mu)
// lies near mu = 1, n = mpos / (1 - mpos) reaches ~1e5 and rf + n RJ / 3
// cancels to ~1e-6 of its terms: one of the sums whose rounding the
// header-wide rule (no fused multiply-add) keeps the same in every kernel.
template <bool FIRSTPT>
__device__ __forceinline__ void calcphitmu(GkVars& v, bool sym, double a, double mneg, double mpos, double mu0,
                                  double muf, double muplus, double& phimuf, double& tmuf) {
    double fac, m1, n, fac2;
    if (sym) {
        fac = fabs(a) * sqrt(mpos - mneg);
        m1 = -mneg / (mpos - mneg);
    } else {
        fac = fabs(a) * muplus;
        m1 = mneg / mpos;
    }
    if (mpos != 1.0) {
        n = sym ? mpos / (1.0 - mpos) : (mpos - mneg) / (1.0 - mpos);
        fac2 = 1.0 - mpos;
    } else {
        n = -1.0;
        fac2 = 1.0;
    }
    const double phif = sym ? acos(muf / muplus) : asin(sqrt((mpos - muf * muf) / (mpos - mneg)));
    // ellphitmu with its n taken as -n (Numerical Recipes sign)
    const double nn = -n;
    const double ak = sqrt(1.0 - m1);
    const double rf0 = v.rffmu1, rff = v.rffmu2, rfc = v.rffmu3;
    double ellec = 0.0, ellpic = 0.0;
    if constexpr (FIRSTPT) {
        const double phi0 = sym ? acos(mu0 / muplus) : asin(sqrt((mpos - mu0 * mu0) / (mpos - mneg)));
        const double s0 = sin(phi0);
        const double q0 = (1.0 - s0 * ak) * (1.0 + s0 * ak);
        const double rd0 = rd_d(1.0 - s0 * s0, q0, 1.0);
        v.rdc = rd_d(0.0, m1, 1.0);
        double elle0 = s0 * (rf0 - ((s0 * ak) * (s0 * ak)) * rd0 / 3.0);
        ellec = 2.0 * (rfc - ak * ak * v.rdc / 3.0);
        double ellpi0 = 0.0;
        if (nn != 1.0) {
            const double rj0 = rj_d(1.0 - s0 * s0, q0, 1.0, 1.0 - nn * s0 * s0);
            v.rjc = rj_d(0.0, m1, 1.0, 1.0 - nn);
            ellpi0 = s0 * (rf0 + nn * s0 * s0 * rj0 / 3.0);
            ellpic = 2.0 * (rfc + nn * v.rjc / 3.0);
        } else {
            // pole-crossing orbit: the Fortran leaves RJC unset; every later
            // use is multiplied by L = 0
            v.rjc = 0.0;
        }
        if (phi0 > kHalfPi) {
            ellpi0 = ellpic - ellpi0;
            elle0 = ellec - elle0;
        }
        const double half = sym ? 1.0 : 2.0;
        v.tmu1 = fac * elle0;
        v.tmu3 = fac * ellec / half;
        v.phimu1 = ellpi0 / fac2 / fac;
        v.phimu3 = ellpic / fac2 / fac / half;
    }
    const double sf = sin(phif);
    const double qf = (1.0 - sf * ak) * (1.0 + sf * ak);
    const double rdf = rd_d(1.0 - sf * sf, qf, 1.0);
    double ellef = sf * (rff - ((sf * ak) * (sf * ak)) * rdf / 3.0);
    double ellpif = 0.0;
    if (nn != 1.0) {
        const double rjf = rj_d(1.0 - sf * sf, qf, 1.0, 1.0 - nn * sf * sf);
        ellpif = sf * (rff + nn * sf * sf * rjf / 3.0);
    }
    if (phif > kHalfPi) {
        ellpic = 2.0 * (rfc + nn * v.rjc / 3.0);
        ellec = 2.0 * (rfc - ak * ak * v.rdc / 3.0);
        ellpif = ellpic - ellpif;
        ellef = ellec - ellef;
    }
    tmuf = fac * ellef;
    phimuf = ellpif / fac2 / fac;
}

struct PhiTime {
    double phimu, tmu, phiu, tu, lambda;
};

// GEOPHITIME (geokerr_original.f:1648-2360).  Near a = 1 the two
// horizon-pole integrals tu3 and tu4 nearly cancel in eq. 48 and are scaled
// by 1 / (2 sqrt(1 - a^2)).  Always inlined: as a
// separate leaf function its body outgrows the 16-bit branch range and the
// compiler's long-branch sequences then overwrite s[30:31], the
// return-address register (seen in the generated gfx950 code).
// ROW6 as in geomu.
template <bool FIRSTPT, bool ROW6 = true>
__device__ __forceinline__ PhiTime geophitime(GkVars& v, const double u0, const double uf, const double mu0,
                                     const double muf, const double a, const double l, const double q2,
                                     const int tpm, const int tpr, const double su, const double sm,
                                     const double iu) {
    const double one = 1.0, two = 2.0;
    const double uplus = one / (one + sqrt(one - a * a));
    const double umi = one - sqrt(one - a * a);
    const double ur = -one / (two * sqrt(one - a * a));
    const double qs = sgn1(q2);
    const double dd = two * ((a - l) * (a - l) + q2);
    const double ee = -a * a * q2;
    const double l2 = l * l;
    const double ql2 = q2 + l2;
    const double a1 = sm;
    const double a2 = sm * neg1pow(tpm);
    const double a3 = two * f_int((two * (double)tpm + 3.0 - sm) / 4.0) - one;
    const double sgn_tpr = neg1pow(tpr);
    const int ncase = v.ncase;
    const double u1 = v.u1, u2 = v.u2, u3 = v.u3, u4 = v.u4;
    double tu = 0.0, phiu = 0.0, phimu = 0.0, tmu = 0.0, lambdau = 0.0;

    if (ncase == 0) {
    } else if (ncase < 3) {
        if (fabs(u3 - u2) < 1e-12) {
            // equal roots; the Fortran leaves lambdau unset here (zero)
            tu = su * tfnkerr(u0, uf, u1, u2, l, a) / sqrt(dd);
            phiu = su * (phifnkerr(uf, u1, u2, l, a) - phifnkerr(u0, u1, u2, l, a)) / sqrt(dd);
            if (u0 >= u3) {
                phiu = -phiu;
                tu = -tu;
            }
        } else if (u0 <= u2 || u0 >= u3) {
            // Table 1 rows 1 (u0 <= u2) and 2 (u0 >= u3)
            const bool row1 = (u0 <= u2);
            const double s = row1 ? one : -one;   // factor signs and integral direction
            const double ub = row1 ? u2 : u3;
            if (FIRSTPT && u0 != ub) {
                const double y = row1 ? u0 : u3, x = row1 ? u2 : u0;
                v.tu03 = s * ellcubicreal<-2>(-u1, one, s * u2, -s, s * u3, -s, one, -one / uplus, v.rffu0, y, x);
                v.tu04 = s * ellcubicreal<-2>(-u1, one, s * u2, -s, s * u3, -s, one, -umi, v.rffu0, y, x);
                v.tu02 = s * ellcubicreal<-2>(-u1, one, s * u2, -s, s * u3, -s, 0.0, one, v.rffu0, y, x);
                v.tu01 = s * ellcubicreal<-4>(-u1, one, s * u2, -s, s * u3, -s, 0.0, one, v.rffu0, y, x);
            } else if (u0 == ub) {
                v.tu01 = v.tu02 = v.tu03 = v.tu04 = 0.0;
            }
            double tu11 = 0.0, tu12 = 0.0, phiu11 = 0.0, phiu12 = 0.0;
            if (row1 ? (fabs(uf - u2) > 1e-16) : (uf != u3)) {
                const double y = row1 ? uf : u3, x = row1 ? u2 : uf;
                phiu11 = s * ellcubicreal<-2>(-u1, one, s * u2, -s, s * u3, -s, one, -one / uplus, v.rffu1, y, x);
                phiu12 = s * ellcubicreal<-2>(-u1, one, s * u2, -s, s * u3, -s, one, -umi, v.rffu1, y, x);
                tu12 = s * ellcubicreal<-2>(-u1, one, s * u2, -s, s * u3, -s, 0.0, one, v.rffu1, y, x);
                tu11 = s * ellcubicreal<-4>(-u1, one, s * u2, -s, s * u3, -s, 0.0, one, v.rffu1, y, x);
            }
            const double tu0 = eq47(a, l, uplus, umi, v.tu01, v.tu02, v.tu03, v.tu04);
            const double tuf = eq47(a, l, uplus, umi, tu11, tu12, phiu11, phiu12);
            tu = su * (tu0 - sgn_tpr * tuf) / sqrt(dd) * ur;
            const double phiu0 = eq48(a, l, uplus, umi, v.tu03, v.tu04);
            const double phiu1 = eq48(a, l, uplus, umi, phiu11, phiu12);
            phiu = su * (phiu0 - sgn_tpr * phiu1) / sqrt(dd) * ur;
            lambdau = su * (v.tu01 - sgn_tpr * tu11) / sqrt(dd);
        }
    } else if (ncase == 3) {
        // Table 1 row 3
        const double f = -one / dd / u1, g = f / u1, h = one;
        if (u0 != uf) {
            const double y = (u0 < uf) ? u0 : uf, x = (u0 < uf) ? uf : u0, sg = (u0 < uf) ? one : -one;
            const double tu1 = sg * ellcubiccomplex<-4>(-u1, one, 0.0, one, f, g, h, v.rffu0, y, x);
            const double tu2 = sg * ellcubiccomplex<-2>(-u1, one, 0.0, one, f, g, h, v.rffu0, y, x);
            const double tu3 = sg * ellcubiccomplex<-2>(-u1, one, one, -one / uplus, f, g, h, v.rffu0, y, x);
            const double tu4 = sg * ellcubiccomplex<-2>(-u1, one, one, -umi, f, g, h, v.rffu0, y, x);
            tu = su * ur / sqrt(dd) * eq47(a, l, uplus, umi, tu1, tu2, tu3, tu4);
            phiu = su * ur / sqrt(dd) * ((l / uplus + 2.0 * (a - l)) * (-tu3) - (l * umi + 2.0 * (a - l)) * (-tu4));
            lambdau = su * tu1 / sqrt(dd);
        }
    }

    if (q2 == 0.0) {
        // mu part for q2 = 0
        const double s1 = sgn1(mu0);
        const double b1 = s1 * sm;
        const double b2 = s1 * sm * neg1pow((unsigned)tpm + 1u);
        if (fabs(l) < fabs(a)) {
            const double muplus = s1 * sqrt(one - l2 / a / a);
            const double sq = sqrt(one - muplus * muplus);
            const double t0h = tan(asin(mu0 / muplus) / two), tfh = tan(asin(muf / muplus) / two);
            const double pm1 = sgn1(muplus * l / a) * (atan((muplus - t0h) / sq) + atan((muplus + t0h) / sq));
            const double pm2 = sgn1(muplus * l / a) * (atan((muplus - tfh) / sq) + atan((muplus + tfh) / sq));
            phimu = b1 * pm1 + b2 * pm2;
            tmu = fabs(muplus * a) * (b2 * sqrt(1.0 - (muf * muf) / (muplus * muplus)) +
                                      b1 * sqrt(1.0 - (mu0 * mu0) / (muplus * muplus)));
        }
    } else if (a == 0.0) {
        // mu part for a = 0
        const double muplus = sqrt(q2 / ql2);
        const double mp2 = muplus * muplus;
        double vfm = (muf - mp2) / (one - muf) / muplus;
        double vfp = -(muf + mp2) / (one + muf) / muplus;
        if (fabs(vfm) > one) vfm = sgn1(vfm) * one;
        if (fabs(vfp) > one) vfp = sgn1(vfp) * one;
        double vsm = ((mu0 - mp2) == 0.0) ? -one : (mu0 - mp2) / (one - mu0) / muplus;
        double vsp = ((mu0 + mp2) == 0.0) ? -one : -(mu0 + mp2) / (one + mu0) / muplus;
        if (fabs(vsp) > one) vsp = sgn1(vsp) * one;
        if (fabs(vsm) > one) vsm = sgn1(vsm) * one;
        const double pm1 = kPi - asin(vsm) + asin(vsp);
        const double pm2 = asin(vfm) - asin(vfp) + kPi;
        const double pm3 = two * kPi;
        phimu = -l * iu + sgn1(l) * 0.5 * (a1 * pm1 + a2 * pm2 + a3 * pm3);
    }

    if (ncase == 4) {
        // q2 = 0, l = a: U(u) = 1
        phiu = su * ur * (l * log((uf / uplus - one) / (u0 / uplus - one)) -
                          l * umi * umi * log((uf * umi - one) / (u0 * umi - one)));
        tu = su * ur *
             ((umi - one / uplus) * (one / u0 - one / uf) + (umi * umi - one / (uplus * uplus)) * log(uf / u0) +
              (a * a / uplus + one / (uplus * uplus * uplus)) * uplus * log((uf / uplus - one) / (u0 / uplus - one)) -
              (a * a * umi + umi * umi * umi) * umi * log((uf * umi - one) / (u0 * umi - one)));
        lambdau = su * (one / u0 - one / uf);
    } else if (ncase == 5 || ncase == 6) {
        // Table 1 rows 5 and 6 (row 6 leaves lambdau unset when u0 = uf: zero)
        if (u0 != uf) {
            const double y = (u0 < uf) ? u0 : uf, x = (u0 < uf) ? uf : u0, sg = (u0 < uf) ? one : -one;
            double tu1, tu2, tu3, tu4;
            if (ncase == 5) {
                const double f = -qs * one / fabs(ee) / u1 / u4;
                const double g = (u4 + u1) / u1 / u4 * f;
                tu1 = sg * ellquarticcomplex<-4>(-u1, one, u4 * qs, -one * qs, 0.0, one, f, g, one, v.rffu0, y, x);
                tu2 = sg * ellquarticcomplex<-2>(-u1, one, u4 * qs, -one * qs, 0.0, one, f, g, one, v.rffu0, y, x);
                tu3 = sg * ellquarticcomplex<-2>(-u1, one, u4 * qs, -one * qs, one, -one / uplus, f, g, one, v.rffu0, y, x);
                tu4 = sg * ellquarticcomplex<-2>(-u1, one, u4 * qs, -one * qs, one, -umi, f, g, one, v.rffu0, y, x);
            } else if constexpr (ROW6) {
                const double h1 = v.h1, h2 = one / h1;
                const double g1 = dd / ee / (h2 - h1), g2 = -g1;
                const double f1 = one / sqrt(ee), f2 = f1;
                tu1 = sg * elldoublecomplex<-4>(f1, g1, h1, f2, g2, h2, 0.0, one, v.rffu0, y, x);
                tu2 = sg * elldoublecomplex<-2>(f1, g1, h1, f2, g2, h2, 0.0, one, v.rffu0, y, x);
                tu3 = sg * elldoublecomplex<-2>(f1, g1, h1, f2, g2, h2, one, -one / uplus, v.rffu0, y, x);
                tu4 = sg * elldoublecomplex<-2>(f1, g1, h1, f2, g2, h2, one, -umi, v.rffu0, y, x);
            } else {
                tu1 = tu2 = tu3 = tu4 = 0.0;
            }
            tu = su * ur / sqrt(-qs * ee) * eq47(a, l, uplus, umi, tu1, tu2, tu3, tu4);
            phiu = su * ur / sqrt(-qs * ee) *
                   ((l / uplus + 2.0 * (a - l)) * (-tu3) - (l * umi + 2.0 * (a - l)) * (-tu4));
            lambdau = (ncase == 5) ? su * tu1 / sqrt(fabs(ee)) : su * tu1 / sqrt(-qs * ee);
        }
    } else if (ncase > 6) {
        if (fabs(u3 - u2) < 1e-12) {
            // Equal-root quartic: the Fortran combines an unset phiu0 with the
            // wrong pieces and, in row 8 ascending, calls with p(5) unset;
            // here the rows 5/6 pattern (phiu0 = -tu3, phiu1 = -tu4), p(5) = -2.
            const bool row7 = (u0 < u2), row8 = (u0 > u3);
            if (u0 != uf && (row7 || row8)) {
                const double y = (u0 < uf) ? u0 : uf, x = (u0 < uf) ? uf : u0, sg = (u0 < uf) ? one : -one;
                const double s2 = row7 ? one : -one;
                const double tu3 = sg * ellquarticreal<-2>(-u1, one, s2 * u2, -s2, s2 * u3, -s2, u4, -one, one, -one / uplus, v.rffu0, y, x);
                const double tu4 = sg * ellquarticreal<-2>(-u1, one, s2 * u2, -s2, s2 * u3, -s2, u4, -one, one, -umi, v.rffu0, y, x);
                const double tu2 = sg * ellquarticreal<-2>(-u1, one, s2 * u2, -s2, s2 * u3, -s2, u4, -one, 0.0, one, v.rffu0, y, x);
                const double tu1 = sg * ellquarticreal<-4>(-u1, one, s2 * u2, -s2, s2 * u3, -s2, u4, -one, 0.0, one, v.rffu0, y, x);
                tu = su * ur / sqrt(-qs * ee) * eq47(a, l, uplus, umi, tu1, tu2, tu3, tu4);
                phiu = su * ur / sqrt(-qs * ee) *
                       ((l / uplus + 2.0 * (a - l)) * (-tu3) - (l * umi + 2.0 * (a - l)) * (-tu4));
                lambdau = tu1 / sqrt(-qs * ee);
            }
        } else if (u0 <= u2 || u0 >= u3) {
            // Table 1 rows 7 (u0 <= u2) and 8 (u0 >= u3)
            const bool row7 = (u0 <= u2);
            const double s = row7 ? one : -one;
            const double ub = row7 ? u2 : u3;
            if (FIRSTPT && u0 != ub) {
                const double y = row7 ? u0 : u3, x = row7 ? u2 : u0;
                v.tu03 = s * ellquarticreal<-2>(-u1, one, s * u2, -s, s * u3, -s, u4, -one, one, -one / uplus, v.rffu0, y, x);
                v.tu04 = s * ellquarticreal<-2>(-u1, one, s * u2, -s, s * u3, -s, u4, -one, one, -umi, v.rffu0, y, x);
                v.tu02 = s * ellquarticreal<-2>(-u1, one, s * u2, -s, s * u3, -s, u4, -one, 0.0, one, v.rffu0, y, x);
                v.tu01 = s * ellquarticreal<-4>(-u1, one, s * u2, -s, s * u3, -s, u4, -one, 0.0, one, v.rffu0, y, x);
            } else if (u0 == ub) {
                v.tu01 = v.tu02 = v.tu03 = v.tu04 = 0.0;
            }
            double tu11 = 0.0, tu12 = 0.0, phiu11 = 0.0, phiu12 = 0.0;
            if (row7 ? (fabs(uf - u2) > 1e-12) : (uf != u3)) {
                const double y = row7 ? uf : u3, x = row7 ? u2 : uf;
                phiu11 = s * ellquarticreal<-2>(-u1, one, s * u2, -s, s * u3, -s, u4, -one, one, -one / uplus, v.rffu1, y, x);
                phiu12 = s * ellquarticreal<-2>(-u1, one, s * u2, -s, s * u3, -s, u4, -one, one, -umi, v.rffu1, y, x);
                tu12 = s * ellquarticreal<-2>(-u1, one, s * u2, -s, s * u3, -s, u4, -one, 0.0, one, v.rffu1, y, x);
                tu11 = s * ellquarticreal<-4>(-u1, one, s * u2, -s, s * u3, -s, u4, -one, 0.0, one, v.rffu1, y, x);
            }
            const double tu0 = eq47(a, l, uplus, umi, v.tu01, v.tu02, v.tu03, v.tu04);
            const double tuf = eq47(a, l, uplus, umi, tu11, tu12, phiu11, phiu12);
            tu = su * (tu0 - sgn_tpr * tuf) / sqrt(fabs(ee)) * ur;
            const double phiu0 = eq48(a, l, uplus, umi, v.tu03, v.tu04);
            const double phiu1 = eq48(a, l, uplus, umi, phiu11, phiu12);
            phiu = su * (phiu0 - sgn_tpr * phiu1) / sqrt(fabs(ee)) * ur;
            lambdau = su * (v.tu01 - sgn_tpr * tu11) / sqrt(fabs(ee));
        }
    }

    if (ncase > 4) {
        // mu part from the roots of M(mu)
        const double amq = a * a - ql2;
        const double yy = -0.5 * (amq + sgn1(amq) * sqrt(amq * amq + 4.0 * q2 * a * a));
        double mneg, mpos;
        if (amq < 0.0) { mneg = -yy / a / a; mpos = q2 / yy; }
        else { mneg = q2 / yy; mpos = -yy / a / a; }
        if (mpos > 1.0) mpos = 1.0;
        double muplus = sqrt(mpos);
        const double c1 = sm;
        const double c2 = sm * neg1pow((unsigned)tpm + 1u);
        const double c3 = 2.0 * f_int((2.0 * (double)tpm - sm + 1.0) / 4.0);
        double phimu2, tmu2;
        if (mneg < 0.0) {
            if (muplus < mu0) { muplus = mu0; mpos = muplus * muplus; }
            calcphitmu<FIRSTPT>(v, true, a, mneg, mpos, mu0, muf, muplus, phimu2, tmu2);
            tmu = a * a * mneg * iu + (c1 * v.tmu1 + c2 * tmu2 + c3 * v.tmu3);
        } else {
            if (sgn1(mu0) == -1.0) muplus = -muplus;
            if (fabs(muplus) < fabs(mu0)) { muplus = mu0; mpos = muplus * muplus; }
            if (fabs(mneg) > mu0 * mu0) mneg = mu0 * mu0;
            calcphitmu<FIRSTPT>(v, false, a, mneg, mpos, mu0, muf, muplus, phimu2, tmu2);
            tmu = c1 * v.tmu1 + c2 * tmu2 + c3 * v.tmu3;
        }
        phimu = -l * iu + l * (c1 * v.phimu1 + c2 * phimu2 + c3 * v.phimu3);
    }
    PhiTime r;
    r.lambda = lambdau + tmu;
    if (l == 0.0) phiu = phiu - sgn1(a) * kPi * tpm;
    r.phimu = phimu;
    r.tmu = tmu;
    r.phiu = phiu;
    r.tu = tu;
    return r;
}

}  // namespace gk
}  // namespace np

#pragma clang fp contract(fast)
