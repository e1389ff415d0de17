/*
 * geomu_oracle.c -- FP64 restatement of GEOKERR's GEOMU (mu at a given
 * inverse radius) and the numerics it calls, from
 * geokerr_port/reference/geokerr_original.f:
 *   GEOMU                       771-1186
 *   ellcubicreal / complex      1188-1262 / 1264-1340
 *   asech                       1342-1360
 *   ellquarticreal / complex    1362-1442 / 1444-1529
 *   elldoublecomplex            1531-1645
 *   calcimuasym(f), calcimusym(f) 3137-3274
 *   SNCNDN                      3277-3340
 *   ZROOTS, LAGUER              3343-3441, 3527-3606
 * Fortran INT() is truncation toward zero, SIGN(1,x) is copysign(1,x),
 * (-1)**n is +-1 by parity.  Every branch of the Carlson reductions is
 * restated (GEOMU itself only reaches the RF-only branches).
 *
 * TEST INFRASTRUCTURE ONLY (see np_oracle.h).
 */
#include <complex.h>
#include <math.h>
#include <string.h>
#include "np_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

static double sgn1(double x) { return copysign(1.0, x); }
/* (-1)**n by parity, unsigned so that n + 1 at INT_MAX wraps like the
 * reference's 32-bit integers. */
static double neg1pow(unsigned n) { return (n & 1u) ? -1.0 : 1.0; }
static int add_i32(int a, int b) { return (int)((unsigned)a + (unsigned)b); }
/* Fortran INT() as the reference's x86-64 build executes it (cvttsd2si):
 * truncation toward zero, INT_MIN for NaN and for values outside int (a
 * plain C cast of those is undefined). */
static int f_int(double x) { return (x > -2147483649.0 && x < 2147483648.0) ? (int)x : (int)0x80000000u; }

/* ---------------- Carlson reductions (RF, RC, RD, RJ callers) ---------------- */

/* ellcubicreal (geokerr_original.f:1188-1262), Carlson (1989). */
double orc_ellcubicreal(const int p[5], double a1, double b1, double a2, double b2, double a3,
                        double b3, double a4, double b4, double* rff, double y, double x) {
    double d12 = a1 * b2 - a2 * b1, d13 = a1 * b3 - a3 * b1, d14 = a1 * b4 - a4 * b1;
    double d24 = a2 * b4 - a4 * b2, d34 = a3 * b4 - a4 * b3;
    double X1 = sqrt(a1 + b1 * x), X2 = sqrt(a2 + b2 * x), X3 = sqrt(a3 + b3 * x), X4 = sqrt(a4 + b4 * x);
    double Y1 = sqrt(a1 + b1 * y), Y2 = sqrt(a2 + b2 * y), Y3 = sqrt(a3 + b3 * y), Y4 = sqrt(a4 + b4 * y);
    double U1c = (X1 * Y2 * Y3 + Y1 * X2 * X3) / (x - y);
    double U12 = U1c * U1c;
    double t2 = (X2 * Y1 * Y3 + Y2 * X1 * X3) / (x - y);
    double U22 = t2 * t2;
    double t3 = (X3 * Y1 * Y2 + Y3 * X1 * X2) / (x - y);
    double U32 = t3 * t3;
    double W22 = U12 - b4 * d12 * d13 / d14;
    double q = X4 * Y4 / X1 / Y1;
    double Q22 = q * q * W22;
    double P22 = Q22 + b4 * d24 * d34 / d14;
    if (p[3] == 0) {
        *rff = orc_rf(U32, U22, U12);
        return 2.0 * (*rff);
    }
    double I1c = 2.0 * (*rff);
    if (p[3] == -2) {
        double I3c = 2.0 * orc_rc(P22, Q22) - 2.0 * b1 * d12 * d13 / 3.0 / d14 * orc_rj(U32, U22, U12, W22);
        return (b4 * I3c - b1 * I1c) / d14;
    }
    double r12 = a1 / b1 - a2 / b2, r13 = a1 / b1 - a3 / b3;
    double r24i = b2 * b4 / (a2 * b4 - a4 * b2), r34i = b3 * b4 / (a3 * b4 - a4 * b3);
    double I2c = 2.0 / 3.0 * d12 * d13 * orc_rd(U32, U22, U12) + 2.0 * X1 * Y1 / U1c;
    double K2c = b2 * b3 * I2c - 2.0 * b4 * (X1 * X2 * X3 / (X4 * X4) - Y1 * Y2 * Y3 / (Y4 * Y4));
    return 0.5 * b4 / d14 / d24 / d34 * K2c + (b1 / d14) * (b1 / d14) * (1.0 - 0.5 * r12 * r13 * r24i * r34i) * I1c;
}

/* ellcubiccomplex (geokerr_original.f:1264-1340), Carlson (1991). */
double orc_ellcubiccomplex(const int p[5], double a1, double b1, double a4, double b4, double f,
                           double g, double h, double* rff, double y, double x) {
    double X1 = sqrt(a1 + b1 * x), X4 = sqrt(a4 + b4 * x);
    double Y1 = sqrt(a1 + b1 * y), Y4 = sqrt(a4 + b4 * y);
    double d14 = a1 * b4 - a4 * b1;
    double beta1 = g * b1 - 2.0 * h * a1;
    double a11 = sqrt(2.0 * f * b1 * b1 - 2.0 * g * a1 * b1 + 2.0 * h * a1 * a1);
    double c44 = sqrt(2.0 * f * b4 * b4 - 2.0 * g * a4 * b4 + 2.0 * h * a4 * a4);
    double a142 = 2.0 * f * b1 * b4 - g * (a1 * b4 + a4 * b1) + 2.0 * h * a1 * a4;
    double xi = sqrt(f + g * x + h * x * x);
    double eta = sqrt(f + g * y + h * y * y);
    double s = xi + eta;
    double mm = (X1 + Y1) * sqrt(s * s - h * (x - y) * (x - y)) / (x - y);
    double M2 = mm * mm;
    double Lp2 = M2 - beta1 + sqrt(2.0 * h) * a11;
    double Lm2 = M2 - beta1 - sqrt(2.0 * h) * a11;
    if (p[3] == 0) {
        *rff = orc_rf(M2, Lm2, Lp2);
        return 4.0 * (*rff);
    }
    double I1c = 4.0 * (*rff);
    double U = (X1 * eta + Y1 * xi) / (x - y);
    double U2 = U * U;
    double Wp2 = M2 - b1 * (a142 + a11 * c44) / d14;
    double W2 = U2 - a11 * a11 * b4 / 2.0 / d14;
    double qq = X4 * Y4 / X1 / Y1;
    double Q2 = qq * qq * W2;
    double P2 = Q2 + c44 * c44 * b4 / 2.0 / d14;
    double rho = sqrt(2.0 * h) * a11 - beta1;
    if (p[3] == -2) {
        double I3c = (2.0 * a11 / 3.0 / c44) * ((-4.0 * b1 / d14) * (a142 + a11 * c44) * orc_rj(M2, Lm2, Lp2, Wp2) -
                                               6.0 * (*rff) + 3.0 * orc_rc(U2, W2)) +
                     2.0 * orc_rc(P2, Q2);
        return (b4 * I3c - b1 * I1c) / d14;
    }
    double r24Xr34 = 0.5 * c44 * c44 / h / (b4 * b4);
    double r12Xr13 = 0.5 * a11 * a11 / h / (b1 * b1);
    double N2c = 2.0 / 3.0 * sqrt(2.0 * h) / a11 * (4.0 * rho * orc_rd(M2, Lm2, Lp2) - 6.0 * (*rff) + 3.0 / U) +
                 2.0 / X1 / Y1 / U;
    double K2c = 0.5 * a11 * a11 * N2c - 2.0 * d14 * (xi / X1 / (X4 * X4) - eta / Y1 / (Y4 * Y4));
    return 0.5 / d14 / (h * b4 * r24Xr34) * K2c + (b1 / d14) * (b1 / d14) * (1.0 - 0.5 * r12Xr13 / r24Xr34) * I1c;
}

/* asech (geokerr_original.f:1342-1360). */
double orc_asech(double x) {
    if (x > 0.0) return log((1.0 + sqrt(1.0 - x * x)) / x);
    return 0.0;
}

/* ellquarticreal (geokerr_original.f:1362-1442), Carlson (1988). */
double orc_ellquarticreal(const int p[5], double a1, double b1, double a2, double b2, double a3,
                          double b3, double a4, double b4, double a5, double b5, double* rff,
                          double y, double x) {
    double d12 = a1 * b2 - a2 * b1, d13 = a1 * b3 - a3 * b1, d14 = a1 * b4 - a4 * b1;
    double d24 = a2 * b4 - a4 * b2, d34 = a3 * b4 - a4 * b3, d15 = a1 * b5 - a5 * b1;
    double d25 = a2 * b5 - a5 * b2, d35 = a3 * b5 - a5 * b3, d45 = a4 * b5 - a5 * b4;
    double X1 = sqrt(a1 + b1 * x), X2 = sqrt(a2 + b2 * x), X3 = sqrt(a3 + b3 * x), X4 = sqrt(a4 + b4 * x);
    double Y1 = sqrt(a1 + b1 * y), Y2 = sqrt(a2 + b2 * y), Y3 = sqrt(a3 + b3 * y), Y4 = sqrt(a4 + b4 * y);
    double u1 = (X1 * X2 * Y3 * Y4 + Y1 * Y2 * X3 * X4) / (x - y);
    double u2 = (X1 * X3 * Y2 * Y4 + Y1 * Y3 * X2 * X4) / (x - y);
    double u3 = (X1 * X4 * Y2 * Y3 + Y1 * Y4 * X2 * X3) / (x - y);
    double U122 = u1 * u1, U132 = u2 * u2, U142 = u3 * u3;
    if (p[4] == 0) {
        *rff = orc_rf(U122, U132, U142);
        return 2.0 * (*rff);
    }
    double I1 = 2.0 * (*rff);
    double X52 = a5 + b5 * x, Y52 = a5 + b5 * y;
    double W22 = U122 - d13 * d14 * d25 / d15;
    double Q22 = X52 * Y52 / ((X1 * Y1) * (X1 * Y1)) * W22;
    double P22 = Q22 + d25 * d35 * d45 / d15;
    if (p[4] == -2) {
        double I3 = 2.0 * d12 * d13 * d14 / 3.0 / d15 * orc_rj(U122, U132, U142, W22) + 2.0 * orc_rc(P22, Q22);
        return (b5 * I3 - b1 * I1) / d15;
    }
    double I2 = 2.0 / 3.0 * d12 * d13 * orc_rd(U122, U132, U142) + 2.0 * X1 * Y1 / X4 / Y4 / sqrt(U142);
    double r12 = a1 / b1 - a2 / b2, r13 = a1 / b1 - a3 / b3;
    double r25i = b2 * b5 / (a2 * b5 - a5 * b2), r35i = b3 * b5 / (a3 * b5 - a5 * b3);
    double A111m1m2 = X1 * X2 * X3 / X4 / X52 - Y1 * Y2 * Y3 / Y4 / Y52;
    return 0.5 * (b5 * b5) * d24 * d34 / d15 / d25 / d35 / d45 * I2 +
           (b1 / d15) * (b1 / d15) * (1.0 - 0.5 * r12 * r13 * r25i * r35i) * I1 -
           (b5 * b5) / d15 / d25 / d35 * A111m1m2;
}

/* ellquarticcomplex (geokerr_original.f:1444-1529), Carlson (1991). */
double orc_ellquarticcomplex(const int p[5], double a1, double b1, double a4, double b4, double a5,
                             double b5, double f, double g, double h, double* rff, double y, double x) {
    double X1 = sqrt(a1 + b1 * x), X4 = sqrt(a4 + b4 * x);
    double Y1 = sqrt(a1 + b1 * y), Y4 = sqrt(a4 + b4 * y);
    double a11 = sqrt(2.0 * f * b1 * b1 - 2.0 * g * a1 * b1 + 2.0 * h * a1 * a1);
    double c44 = sqrt(2.0 * f * b4 * b4 - 2.0 * g * a4 * b4 + 2.0 * h * a4 * a4);
    double a142 = 2.0 * f * b1 * b4 - g * (a1 * b4 + a4 * b1) + 2.0 * h * a1 * a4;
    double xi = sqrt(f + g * x + h * x * x);
    double eta = sqrt(f + g * y + h * y * y);
    double s = xi + eta;
    double mm = (X1 * Y4 + Y1 * X4) * sqrt(s * s - h * (x - y) * (x - y)) / (x - y);
    double M2 = mm * mm;
    double Lp2 = M2 + a142 + a11 * c44;
    double Lm2 = fmax(M2 + a142 - a11 * c44, 0.0);
    if (p[4] == 0) {
        *rff = orc_rf(M2, Lm2, Lp2);
        return 4.0 * (*rff);
    }
    double I1 = 4.0 * (*rff);
    double d14 = a1 * b4 - a4 * b1, d15 = a1 * b5 - a5 * b1, d45 = a4 * b5 - a5 * b4;
    double X52 = a5 + b5 * x, Y52 = a5 + b5 * y;
    double c552 = 2.0 * f * b5 * b5 - 2.0 * g * a5 * b5 + 2.0 * h * a5 * a5;
    double c55 = sqrt(c552);
    double a152 = 2.0 * f * b1 * b5 - g * (a1 * b5 + a5 * b1) + 2.0 * h * a5 * a1;
    double U = (X1 * X4 * eta + Y1 * Y4 * xi) / (x - y);
    double U2 = U * U;
    double Wp2 = M2 + d14 * (a152 + a11 * c55) / d15;
    double W2 = U2 - a11 * a11 * d45 / 2.0 / d15;
    double Q2 = X52 * Y52 / ((X1 * Y1) * (X1 * Y1)) * W2;
    double P2 = Q2 + c55 * c55 * d45 / 2.0 / d15;
    if (p[4] == -2) {
        double I3 = (2.0 * a11 / 3.0 / c55) * ((4.0 * d14 / d15) * (a152 + a11 * c55) * orc_rj(M2, Lm2, Lp2, Wp2) -
                                              6.0 * (*rff) + 3.0 * orc_rc(U2, W2)) +
                    2.0 * orc_rc(P2, Q2);
        return (b5 * I3 - b1 * I1) / d15;
    }
    double I2 = 2.0 * a11 / 3.0 / c44 * (4.0 * (a142 + a11 * c44) * orc_rd(M2, Lm2, Lp2) - 6.0 * (*rff)) +
                2.0 * a11 / c44 / U + 2.0 * X1 * Y1 / X4 / Y4 / U;
    double A111m1m2 = X1 * xi / X4 / X52 - Y1 * eta / Y4 / Y52;
    return (b5 * b5) / (2.0 * d15 * d45) * (c44 * c44) / c552 * I2 +
           (b1 * b1) / (d15 * d15) * (1.0 - 0.5 * (b5 * b5) * (a11 * a11) / (b1 * b1) / c552) * I1 -
           2.0 * (b5 * b5) / d15 / c552 * A111m1m2;
}

/* elldoublecomplex (geokerr_original.f:1531-1645), Carlson (1992). */
double orc_elldoublecomplex(const int p[5], double f1, double g1, double h1, double f2, double g2,
                            double h2, double a5, double b5, double* rff, double y, double x) {
    double xi1 = sqrt(f1 + g1 * x + h1 * (x * x));
    double xi2 = sqrt(f2 + g2 * x + h2 * (x * x));
    double eta1 = sqrt(f1 + g1 * y + h1 * y * y);
    double eta2 = sqrt(f2 + g2 * y + h2 * y * y);
    double theta1 = 2.0 * f1 + g1 * (x + y) + 2.0 * h1 * x * y;
    double theta2 = 2.0 * f2 + g2 * (x + y) + 2.0 * h2 * x * y;
    double zeta1 = sqrt(2.0 * xi1 * eta1 + theta1);
    double zeta2 = sqrt(2.0 * xi2 * eta2 + theta2);
    double M = zeta1 * zeta2 / (x - y);
    double M2 = M * M;
    double delta122 = 2.0 * f1 * h2 + 2.0 * f2 * h1 - g1 * g2;
    double delta112 = 4.0 * f1 * h1 - g1 * g1;
    double delta222 = 4.0 * f2 * h2 - g2 * g2;
    double Delta = sqrt(delta122 * delta122 - delta112 * delta222);
    double Deltap = delta122 + Delta, Deltam = delta122 - Delta;
    double Lp2 = M2 + Deltap, Lm2 = M2 + Deltam;
    if (p[4] == 0) {
        *rff = orc_rf(M2, Lm2, Lp2);
        return 4.0 * (*rff);
    }
    double U = (xi1 * eta2 + eta1 * xi2) / (x - y);
    double U2 = U * U;
    double alpha15 = 2.0 * f1 * b5 - g1 * a5, alpha25 = 2.0 * f2 * b5 - g2 * a5;
    double beta15 = g1 * b5 - 2.0 * h1 * a5, beta25 = g2 * b5 - 2.0 * h2 * a5;
    double gamma1 = 0.5 * (alpha15 * b5 - beta15 * a5), gamma2 = 0.5 * (alpha25 * b5 - beta25 * a5);
    double Lambda = delta112 * gamma2 / gamma1;
    double Omega2 = M2 + Lambda;
    double psi = 0.5 * (alpha15 * beta25 - alpha25 * beta15);
    double psi2 = psi * psi;
    double xi5 = a5 + b5 * x, eta5 = a5 + b5 * y;
    double Am111m1 = 1.0 / xi1 * xi2 - 1.0 / eta1 * eta2;
    double A1111m4 = xi1 * xi2 / (xi5 * xi5) - eta1 * eta2 / (eta5 * eta5);
    double XX = xi5 * eta5 * (theta1 * 0.5 * Am111m1 - xi5 * eta5 * A1111m4) / ((x - y) * (x - y));
    double S = 0.5 * (M2 + delta122) - U2;
    double S2 = S * S;
    double mu = gamma1 * xi5 * eta5 / xi1 / eta1;
    double T = mu * S + 2.0 * gamma1 * gamma2;
    double T2 = T * T;
    double V2 = (mu * mu) * (S2 + Lambda * U2);
    double b2 = (Omega2 * Omega2) * (S2 / U2 + Lambda);
    double a2 = b2 + (Lambda * Lambda) * psi2 / gamma1 / gamma2;
    double H = delta112 * psi * (orc_rj(M2, Lm2, Lp2, Omega2) / 3.0 + 0.5 * orc_rc(a2, b2)) / (gamma1 * gamma1) -
               XX * orc_rc(T2, V2);
    if (p[4] == -2) return -2.0 * (b5 * H + beta15 * (*rff) / gamma1);
    double A1111m2 = xi1 * xi2 / xi5 - eta1 * eta2 / eta5;
    double xi1p = 0.5 * (g1 + 2.0 * h1 * x) / xi1;
    double eta1p = 0.5 * (g1 + 2.0 * h1 * y) / eta1;
    double B = xi1p * xi2 - eta1p * eta2;
    double G = 2.0 / 3.0 * Delta * Deltap * orc_rd(M2, Lm2, Lp2) + 0.5 * Delta / U +
               (delta122 * theta1 - delta112 * theta2) / 4.0 / xi1 / eta1 / U;
    double Sigma = G - Deltap * (*rff) + B;
    return b5 * (beta15 / gamma1 + beta25 / gamma2) * H + (beta15 * beta15) * (*rff) / (gamma1 * gamma1) +
           (b5 * b5) * (Sigma - b5 * A1111m2) / gamma1 / gamma2;
}

/* ---------------- mu integrals ---------------- */

/* calcimuasym (geokerr_original.f:3137-3172). */
void orc_calcimuasym(double a, double mneg, double mpos, double mu0, double muplus,
                        double* imu0, double* imum, double* rf0, double* rfc) {
    double fac = fabs(a) * muplus;
    double m1 = mneg / mpos;
    double phi0 = asin(sqrt((mpos - mu0 * mu0) / (mpos - mneg)));
    double s0 = sqrt((mpos - mu0 * mu0) / (mpos - mneg));
    double ak = sqrt(1.0 - m1);
    double q0 = (1.0 - s0 * ak) * (1.0 + s0 * ak);
    *rf0 = orc_rf(1.0 - s0 * s0, q0, 1.0);
    *rfc = orc_rf(0.0, m1, 1.0);
    *imum = 2.0 * (*rfc);
    *imu0 = s0 * (*rf0);
    if (tan(phi0) < 0.0) *imu0 = *imum - *imu0;
    *imu0 = *imu0 / fac;
    *imum = *imum / fac / 2.0;
}

/* calcimusym (geokerr_original.f:3202-3240). */
void orc_calcimusym(double a, double mneg, double mpos, double mu0, double muplus,
                       double* imu0, double* imum, double* rf0, double* rfc) {
    double fac = fabs(a) * sqrt(mpos - mneg);
    double m1 = -mneg / (mpos - mneg);
    double phi0 = acos(mu0 / muplus);
    double s0 = sin(phi0);
    double ak = sqrt(1.0 - m1);
    double q0 = (1.0 - s0 * ak) * (1.0 + s0 * ak);
    *rf0 = orc_rf(1.0 - s0 * s0, q0, 1.0);
    *rfc = orc_rf(0.0, m1, 1.0);
    *imum = 2.0 * (*rfc);
    *imu0 = s0 * (*rf0);
    if (tan(phi0) < 0.0) *imu0 = *imum - *imu0;
    *imu0 = *imu0 / fac;
    *imum = *imum / fac;
}

/* The RF term of calcimuasymf (geokerr_original.f:3175-3204): the only
 * output GEOMU keeps (RFFMU2) for GEOPHITIME. */
double orc_calcimuasymf_rf(double mneg, double mpos, double muf) {
    double m1 = mneg / mpos;
    double sf = sqrt((mpos - muf * muf) / (mpos - mneg));
    double ak = sqrt(1.0 - m1);
    double qf = (1.0 - sf * ak) * (1.0 + sf * ak);
    return orc_rf(1.0 - sf * sf, qf, 1.0);
}

/* The RF term of calcimusymf (geokerr_original.f:3244-3274). */
double orc_calcimusymf_rf(double mneg, double mpos, double muf, double muplus) {
    double m1 = -mneg / (mpos - mneg);
    double sf = sin(acos(muf / muplus));
    double ak = sqrt(1.0 - m1);
    double qf = (1.0 - sf * ak) * (1.0 + sf * ak);
    return orc_rf(1.0 - sf * sf, qf, 1.0);
}

/* SNCNDN (geokerr_original.f:3277-3340), Numerical Recipes. */
void orc_sncndn(double uu, double emmc, double* sn, double* cn, double* dn) {
    const double CA = 3e-8;
    double emc = emmc, u = uu, a, b, c = 0.0, d = 0.0, em[13], en[13];
    int l = 0, bo;
    if (emc != 0.0) {
        bo = (emc < 0.0);
        if (bo) {
            d = 1.0 - emc;
            emc = -emc / d;
            d = sqrt(d);
            u = d * u;
        }
        a = 1.0;
        *dn = 1.0;
        for (int i = 1; i <= 13; ++i) {
            l = i;
            em[i - 1] = a;
            emc = sqrt(emc);
            en[i - 1] = emc;
            c = 0.5 * (a + emc);
            if (fabs(a - emc) <= CA * a) break;
            emc = a * emc;
            a = c;
        }
        u = c * u;
        *sn = sin(u);
        *cn = cos(u);
        if (*sn != 0.0) {
            a = *cn / *sn;
            c = a * c;
            for (int ii = l; ii >= 1; --ii) {
                b = em[ii - 1];
                a = c * a;
                c = (*dn) * c;
                *dn = (en[ii - 1] + a) / (b + a);
                a = c / b;
            }
            a = 1.0 / sqrt(c * c + 1.0);
            *sn = (*sn < 0.0) ? -a : a;
            *cn = c * (*sn);
        }
        if (bo) {
            a = *dn;
            *dn = *cn;
            *cn = a;
            *sn = *sn / d;
        }
    } else {
        *cn = 1.0 / cosh(u);
        *dn = *cn;
        *sn = tanh(u);
    }
}

/* LAGUER (geokerr_original.f:3527-3606).  a[0..m] with a[0] the constant
 * term. */
static void laguer(const double complex* a, int m, double complex* x, int* its) {
    static const double frac[8] = {0.5, 0.25, 0.75, 0.13, 0.38, 0.62, 0.88, 1.0};
    const int MT = 10, MAXIT = 80;
    const double EPSS = 1e-15;
    for (int iter = 1; iter <= MAXIT; ++iter) {
        *its = iter;
        double complex b = a[m];
        double err = cabs(b);
        double complex d = 0.0, f = 0.0;
        double abx = cabs(*x);
        for (int j = m - 1; j >= 0; --j) {
            f = (*x) * f + d;
            d = (*x) * d + b;
            b = (*x) * b + a[j];
            err = cabs(b) + abx * err;
        }
        err = EPSS * err;
        double complex dx;
        if (cabs(b) <= err) return;
        {
            double complex g = d / b;
            double complex g2 = g * g;
            double complex h = g2 - 2.0 * f / b;
            double complex sq = csqrt((double)(m - 1) * ((double)m * h - g2));
            double complex gp = g + sq, gm = g - sq;
            double abp = cabs(gp), abm = cabs(gm);
            if (abp < abm) gp = gm;
            if (fmax(abp, abm) > 0.0) {
                dx = (double)m / gp;
            } else {
                /* CEXP(CMPLX(LOG(1+ABX), ITER)): default-kind (single) complex */
                float complex z = (float)log(1.0 + abx) + (float)iter * I;
                dx = (double complex)cexpf(z);
            }
        }
        double complex x1 = (*x) - dx;
        if (*x == x1) return;
        if (iter % MT != 0) *x = x1;
        else *x = (*x) - dx * frac[iter / MT - 1];
    }
}

/* ZROOTS (geokerr_original.f:3343-3441).  a[0..m] (constant term first);
 * roots sorted by real part. */
void orc_zroots(const double complex* a, int m, double complex* roots, int polish) {
    const double EPS = 1e-6;
    double complex ad[8];
    for (int j = 0; j <= m; ++j) ad[j] = a[j];
    double complex xstart = 0.0;
    int ncc = 0, its;
    for (int j = m; j >= 1; --j) {
        double complex x = xstart;
        if (j < m) {
            if (cimag(roots[j]) != 0.0 && ncc == 0) {
                /* the reference builds (Re r, -Re r), not the conjugate */
                xstart = creal(roots[j]) - creal(roots[j]) * I;
                ncc = 1;
            } else {
                xstart = 0.0;
                ncc = 0;
            }
        } else {
            xstart = 0.0;
            ncc = 0;
        }
        laguer(ad, j, &x, &its);
        if (fabs(cimag(x)) <= 2.0 * EPS * EPS * fabs(creal(x))) x = creal(x);
        roots[j - 1] = x;
        double complex b = ad[j];
        for (int jj = j - 1; jj >= 0; --jj) {
            double complex c = ad[jj];
            ad[jj] = b;
            b = x * b + c;
        }
    }
    if (polish)
        for (int j = 0; j < m; ++j) laguer(a, m, &roots[j], &its);
    for (int j = 1; j < m; ++j) {
        double complex x = roots[j];
        int i;
        for (i = j - 1; i >= 0; --i) {
            if (creal(roots[i]) <= creal(x)) break;
            roots[i + 1] = roots[i];
        }
        roots[i + 1] = x;
    }
}

/* ---------------- GEOMU (geokerr_original.f:771-1186) ----------------
 *
 * Arguments the Fortran passes by reference are pointers here: u0 and uf
 * (clamped to the turning point, or u0 moved out of a forbidden region with
 * a WARNING), tpm (left untouched when q2 = 0) and tpr (zeroed for equal
 * roots).  GEOKERR's variables shared between calls live in *v: with
 * firstpt = 0 the quartic roots, ncase, h1, the u0 piece iu0/rffu0 of the
 * Table-1 rows 1, 2, 7, 8 and the mu0 pieces i1mu/i3mu/rffmu1/rffmu3 are
 * read from it instead of recomputed.  pht = 1 also computes rffmu2
 * (calcimu*f) for GEOPHITIME. */
void orc_gk_geomu(orc_gk_vars* v, double* u0p, double* ufp, double mu0, double a, double l,
                  double q2, int32_t* tpmp, int32_t* tprp, double su, double sm, int pht,
                  int firstpt, double* mufp, double* iup, int32_t* warnedp) {
    const double one = 1.0, two = 2.0, third = 0.3333333333333333;
    double pi = acos(-one);
    double uplus = one / (one + sqrt(one - a * a));
    double l2 = l * l;
    double ql2 = q2 + l2;
    double cc = a * a - q2 - l2;
    double dd = two * ((a - l) * (a - l) + q2);
    double ee = -a * a * q2;
    double u0 = *u0p, uf = *ufp;
    /* geokerr_estimate_periods hands GEOKERR one variable as both U0 and UF,
     * so GEOMU's first call sees u0 and uf aliased: every edit of one is an
     * edit of the other. */
    const int alias = (u0p == ufp);
#define SET_UF(x) do { uf = (x); if (alias) u0 = uf; } while (0)
#define SET_U0(x) do { u0 = (x); if (alias) uf = u0; } while (0)
    int tpm = *tpmp, tpr = *tprp;
    double iu = 0.0, iu1 = 0.0, muf = 0.0;
    int ncase = v->ncase, warned = 0, p[5];
    double u1 = v->u1, u2 = v->u2, u3 = v->u3, u4 = v->u4;

    if (ee == 0.0 && dd != 0.0) {
        p[0] = -1; p[1] = -1; p[2] = -1; p[3] = 0; p[4] = 0;
        double qq = cc * cc / dd / dd / 9.0;
        double rr = (two * (cc * cc * cc) / (dd * dd * dd) + 27.0 / dd) / 54.0;
        double dis = rr * rr - qq * qq * qq;
        if (dis < -1e-16) {
            double theta = acos(rr / pow(qq, 1.5));
            u1 = -two * sqrt(qq) * cos(theta / 3.0) - cc / dd / 3.0;
            u2 = -two * sqrt(qq) * cos((theta - two * pi) / 3.0) - cc / dd / 3.0;
            u3 = -two * sqrt(qq) * cos((theta + two * pi) / 3.0) - cc / dd / 3.0;
            iu1 = 0.0;
            /* The Fortran retries the case test after its forbidden-region
             * edit without bound; where the edit does not leave the region
             * (Q2 < 0) it cycles forever.  At most 3 tries here, as in the
             * device code; the case then stays as it was. */
            for (int guard = 0; guard < 3; ++guard) {
                if (u0 <= u2) {
                    if (uf > u2) SET_UF(u2);
                    ncase = 1;
                    if (firstpt && u0 != u2)
                        v->iu0 = orc_ellcubicreal(p, -u1, one, u2, -one, u3, -one, 0.0, 0.0, &v->rffu0, u0, u2);
                    else if (u0 == u2)
                        v->iu0 = 0.0;
                    if (uf != u2) iu1 = orc_ellcubicreal(p, -u1, one, u2, -one, u3, -one, 0.0, 0.0, &v->rffu1, uf, u2);
                    iu = su * (v->iu0 - neg1pow(tpr) * iu1) / sqrt(dd);
                    break;
                } else if (u0 >= u3) {
                    if (uf < u3) SET_UF(u3);
                    ncase = 2;
                    if (firstpt && u0 != u3)
                        v->iu0 = -orc_ellcubicreal(p, -u1, one, -u2, one, -u3, one, 0.0, 0.0, &v->rffu0, u3, u0);
                    else if (u0 == u3)
                        v->iu0 = 0.0;
                    if (uf != u3) iu1 = -orc_ellcubicreal(p, -u1, one, -u2, one, -u3, one, 0.0, 0.0, &v->rffu1, u3, uf);
                    iu = su * (v->iu0 - neg1pow(tpr) * iu1) / sqrt(dd);
                    break;
                } else {
                    /* forbidden region: the reference prints a warning and moves u0 */
                    warned = 1;
                    SET_U0((su == 1.0) ? u3 : u2);
                }
            }
        } else if (fabs(dis) < 1e-16) {
            u1 = -two * sqrt(qq) - cc / dd / 3.0;
            u2 = -two * sqrt(qq) * cos((two * pi) / 3.0) - cc / dd / 3.0;
            u3 = u2;
            tpr = 0;
            if (u0 <= u2) {
                ncase = 1;
                if (uf > u2) {
                    SET_UF(u2);
                    iu = su * 1e300;
                } else {
                    double farg = (sqrt(uf - u1) + sqrt(u2 - u1)) / fabs(sqrt(uf - u1) - sqrt(u2 - u1));
                    double sarg = (sqrt(u0 - u1) + sqrt(u2 - u1)) / fabs(sqrt(u0 - u1) - sqrt(u2 - u1));
                    iu = su * (log(farg) - log(sarg)) / sqrt((u2 - u1) * dd);
                }
            } else if (u0 >= u2) {
                ncase = 2;
                if (uf < u2) {
                    SET_UF(u2);
                    iu = su * 1e300;
                } else {
                    double farg = (sqrt(uf - u1) + sqrt(u2 - u1)) / fabs(sqrt(uf - u1) - sqrt(u2 - u1));
                    double sarg = (sqrt(u0 - u1) + sqrt(u2 - u1)) / fabs(sqrt(u0 - u1) - sqrt(u2 - u1));
                    iu = -su * (log(farg) - log(sarg)) / sqrt((u2 - u1) * dd);
                }
            }
        } else {
            ncase = 3;
            double aa = -sgn1(rr) * pow(fabs(rr) + sqrt(dis), third);
            double bb = (aa != 0.0) ? qq / aa : 0.0;
            u1 = (aa + bb) - cc / dd / 3.0;
            double f = -one / dd / u1;
            double g = f / u1;
            if (uf > u0) iu = su * orc_ellcubiccomplex(p, -u1, one, 0.0, 0.0, f, g, one, &v->rffu0, u0, uf) / sqrt(dd);
            else if (uf < u0) iu = -su * orc_ellcubiccomplex(p, -u1, one, 0.0, 0.0, f, g, one, &v->rffu0, uf, u0) / sqrt(dd);
            else iu = 0.0;
        }
        if (q2 == 0.0) {
            double s1 = sgn1(mu0);
            if (l2 < a * a && mu0 != 0.0) {
                double muplus = s1 * sqrt(one - l2 / a / a);
                muf = muplus / cosh(fabs(a * muplus) * iu - s1 * sm * orc_asech(mu0 / muplus));
            } else {
                muf = 0.0;
            }
        } else if (a == 0.0) {
            double a1 = sm;
            double muplus = sqrt(q2 / ql2);
            if (mu0 > muplus) muplus = mu0;
            v->i1mu = acos(mu0 / muplus) / sqrt(ql2);
            v->i3mu = pi / sqrt(ql2);
            double i1mu = v->i1mu, i3mu = v->i3mu;
            if (sm == 1.0) tpm = add_i32(f_int((iu - i1mu) / i3mu), f_int((1 + sgn1((iu - i1mu) / i3mu)) / 2.0));
            else tpm = f_int((iu + i1mu) / i3mu);
            double a2 = sm * neg1pow(tpm);
            double a3 = two * f_int((two * (double)tpm + 3.0 - sm) / 4.0) - one;
            muf = -muplus * cos(sqrt(ql2) * (iu - a1 * i1mu - a3 * i3mu) / a2);
        }
    } else if (ee == 0.0 && dd == 0.0) {
        ncase = 4;
        tpr = 0;
        iu = su * (uf - u0);
        muf = 0.0;
    } else {
        p[0] = -1; p[1] = -1; p[2] = -1; p[3] = -1; p[4] = 0;
        if (firstpt) {
            double complex c[5], root[4];
            c[0] = one; c[1] = 0.0; c[2] = cc; c[3] = dd; c[4] = ee;
            orc_zroots(c, 4, root, 1);
            int nreal = 0;
            for (int i = 0; i < 4; ++i)
                if (cimag(root[i]) == 0.0) nreal++;
            if (nreal == 2) {
                ncase = 5;
                u1 = creal(root[0]);
                u4 = (cimag(root[1]) == 0.0) ? creal(root[1]) : creal(root[3]);
            } else if (nreal == 0) {
                ncase = 6;
            } else {
                u1 = creal(root[0]); u2 = creal(root[1]); u3 = creal(root[2]); u4 = creal(root[3]);
                for (int guard = 0; guard < 3; ++guard) {
                    if (u2 > uplus && u3 > uplus) { ncase = 5; break; }
                    else if (u0 <= u2) { ncase = 7; break; }
                    else if (u0 >= u3) { ncase = 8; break; }
                    warned = 1;
                    SET_U0((su == 1.0) ? 0.0 : uplus);
                }
            }
        }
        if (ncase == 5) {
            double qs = sgn1(q2);
            double f = -qs * one / fabs(ee) / u1 / u4;
            double g = (u4 + u1) / u1 / u4 * f;
            if (uf > u0)
                iu = su * orc_ellquarticcomplex(p, -u1, one, qs * u4, -qs * one, 0.0, 0.0, f, g, one, &v->rffu0, u0, uf) / sqrt(fabs(ee));
            else if (uf < u0)
                iu = -su * orc_ellquarticcomplex(p, -u1, one, qs * u4, -qs * one, 0.0, 0.0, f, g, one, &v->rffu0, uf, u0) / sqrt(fabs(ee));
            else
                iu = 0.0;
        } else if (ncase == 6) {
            /* The sextic for h1 (Eq. 21) sits outside the Fortran's FIRSTPT
             * test, so it is solved on every call. */
            double complex coefs[7], hroots[6];
            double se = sqrt(ee);
            coefs[0] = one; coefs[1] = -cc / se; coefs[2] = -one;
            coefs[3] = se * (two * cc / ee - (dd / ee) * (dd / ee));
            coefs[4] = -one; coefs[5] = -cc / se; coefs[6] = one;
            orc_zroots(coefs, 6, hroots, 1);
            double h1 = 0.0;
            /* first real root of the sextic (the Fortran loop has no bound) */
            for (int i = 0; i < 6 && h1 == 0.0; ++i)
                if (cimag(hroots[i]) == 0.0) h1 = creal(hroots[i]);
            v->h1 = h1;
            double h2 = one / h1;
            double g1 = dd / ee / (h2 - h1);
            double g2 = -g1;
            double f1 = one / sqrt(ee);
            double f2 = f1;
            if (uf > u0) iu = su * orc_elldoublecomplex(p, f1, g1, h1, f2, g2, h2, 0.0, 0.0, &v->rffu0, u0, uf) / sqrt(fabs(ee));
            else if (uf < u0) iu = -su * orc_elldoublecomplex(p, f1, g1, h1, f2, g2, h2, 0.0, 0.0, &v->rffu0, uf, u0) / sqrt(fabs(ee));
            else iu = 0.0;
        } else {
            if (fabs(u3 - u2) > 1e-12) {
                iu1 = 0.0;
                if (ncase == 7) {
                    if (uf > u2) SET_UF(u2);
                    if (firstpt && u0 != u2)
                        v->iu0 = orc_ellquarticreal(p, -u1, one, u2, -one, u3, -one, u4, -one, 0.0, 0.0, &v->rffu0, u0, u2);
                    else if (u0 == u2)
                        v->iu0 = 0.0;
                    if (uf != u2) iu1 = orc_ellquarticreal(p, -u1, one, u2, -one, u3, -one, u4, -one, 0.0, 0.0, &v->rffu1, uf, u2);
                    iu = su * (v->iu0 - neg1pow(tpr) * iu1) / sqrt(fabs(ee));
                } else if (ncase == 8) {
                    if (uf < u3) SET_UF(u3);
                    if (firstpt && u0 != u3)
                        v->iu0 = -orc_ellquarticreal(p, -u1, one, -u2, one, -u3, one, u4, -one, 0.0, 0.0, &v->rffu0, u3, u0);
                    else if (u0 == u3)
                        v->iu0 = 0.0;
                    if (uf != u3) iu1 = -orc_ellquarticreal(p, -u1, one, -u2, one, -u3, one, u4, -one, 0.0, 0.0, &v->rffu1, u3, uf);
                    iu = su * (v->iu0 - neg1pow(tpr) * iu1) / sqrt(fabs(ee));
                }
            } else {
                tpr = 0;
                if (ncase == 7) {
                    if (uf > u2) SET_UF(u2);
                    v->iu0 = orc_ellquarticreal(p, -u1, one, u2, -one, u3, -one, u4, -one, 0.0, 0.0, &v->rffu0, u0, uf);
                    iu = su * v->iu0 / sqrt(fabs(ee));
                } else if (ncase == 8) {
                    if (uf < u3) SET_UF(u3);
                    v->iu0 = -orc_ellquarticreal(p, -u1, one, -u2, one, -u3, one, u4, -one, 0.0, 0.0, &v->rffu0, u0, uf);
                    iu = su * v->iu0 / sqrt(fabs(ee));
                } else {
                    ncase = 0;
                }
            }
        }
        /* roots of M(mu) (Eq. 12) */
        double amq = a * a - ql2;
        double yy = -0.5 * (amq + sgn1(amq) * sqrt(amq * amq + 4.0 * q2 * a * a));
        double mneg, mpos;
        if (amq < 0.0) { mneg = -yy / a / a; mpos = q2 / yy; }
        else { mneg = q2 / yy; mpos = -yy / a / a; }
        if (mpos > 1.0) mpos = 1.0;
        double muplus = sqrt(mpos);
        if (mneg < 0.0) {
            if (muplus < mu0) { muplus = mu0; mpos = muplus * muplus; }
            double muminus = -muplus;
            if (firstpt) orc_calcimusym(a, mneg, mpos, mu0, muplus, &v->i1mu, &v->i3mu, &v->rffmu1, &v->rffmu3);
            double i1mu = v->i1mu, i3mu = v->i3mu;
            double a1 = sm;
            if (sm == 1.0) tpm = add_i32(f_int((iu - i1mu) / i3mu), f_int((1 + sgn1((iu - i1mu) / i3mu)) / 2.0));
            else tpm = f_int((iu + i1mu) / i3mu);
            double a2 = sm * neg1pow(tpm);
            double a3 = two * f_int((two * (double)tpm + 3.0 - sm) / 4.0) - one;
            double iarg = fabs(a) / a2 * (iu - a1 * i1mu - a3 * i3mu);
            double uarg = sqrt(mpos - mneg) * iarg;
            double m1 = -mneg / (mpos - mneg);
            double sn, cn, dn;
            orc_sncndn(uarg, m1, &sn, &cn, &dn);
            muf = muminus * cn;
            if (pht) v->rffmu2 = orc_calcimusymf_rf(mneg, mpos, muf, muplus);
        } else {
            muplus = sgn1(mu0) * muplus;
            double muminus = sgn1(mu0) * sqrt(mneg);
            if (fabs(muminus) > fabs(mu0)) { muminus = mu0; mneg = muminus * muminus; }
            if (fabs(mu0) > fabs(muplus)) { muplus = mu0; mpos = muplus * muplus; }
            if (firstpt) orc_calcimuasym(a, mneg, mpos, mu0, muplus, &v->i1mu, &v->i3mu, &v->rffmu1, &v->rffmu3);
            double i1mu = v->i1mu, i3mu = v->i3mu;
            double a1 = sm;
            if (sm == 1.0) tpm = add_i32(f_int((iu - i1mu) / i3mu), f_int(fabs((1 + sgn1((iu - i1mu) / i3mu)) / 2.0)));
            else tpm = f_int((iu + i1mu) / i3mu);
            double a2 = sm * neg1pow(tpm);
            double a3 = two * f_int((two * (double)tpm + 3.0 - sm) / 4.0) - one;
            double iarg = fabs(a) / a2 * (iu - a1 * i1mu - a3 * i3mu);
            double uarg = fabs(muplus) * iarg;
            double m1 = mneg / mpos;
            double sn, cn, dn;
            orc_sncndn(uarg, m1, &sn, &cn, &dn);
            muf = muminus / dn;
            if (pht) v->rffmu2 = orc_calcimuasymf_rf(mneg, mpos, muf);
        }
    }
#undef SET_UF
#undef SET_U0
    v->ncase = ncase;
    v->u1 = u1; v->u2 = u2; v->u3 = u3; v->u4 = u4;
    *u0p = u0;
    *ufp = uf;
    *tpmp = tpm;
    *tprp = tpr;
    *mufp = muf;
    *iup = iu;
    if (warnedp) *warnedp = warned;
}

void orc_geomu(double u0, double uf, double mu0, double a, double l, double q2, int tpr,
               double su, double sm, orc_geomu_out* o) {
    orc_gk_vars v;
    memset(&v, 0, sizeof v);
    int32_t tpm = 0, tpr32 = tpr, warned = 0;
    double muf = 0.0, iu = 0.0;
    orc_gk_geomu(&v, &u0, &uf, mu0, a, l, q2, &tpm, &tpr32, su, sm, 0, 1, &muf, &iu, &warned);
    o->muf = muf;
    o->iu = iu;
    o->tpm = tpm;
    o->ncase = v.ncase;
    o->tpr = tpr32;
    o->u1 = v.u1; o->u2 = v.u2; o->u3 = v.u3; o->u4 = v.u4;
    o->u0_used = u0;
    o->uf_used = uf;
    o->i1mu = v.i1mu;
    o->i3mu = v.i3mu;
    o->h1 = v.h1;
    o->warned = warned;
}

/* Camera -> constants of motion (INITIALIZE_CAMERA_GEOKERR, standard run,
 * geokerr_original.f:~270-320): L = -alpha sqrt(1-mu0^2),
 * Q2 = beta^2 - (a^2 - alpha^2) mu0^2, SU = 1, SM = +1 when mu0 < 1 and
 * beta >= 0; inputs below 1e-5 (1e-10 for Q2) are zeroed. */
void orc_geokerr_camera(double alpha, double beta, double mu0, double a_in, double* L,
                        double* Q2, double* a_out, double* su, double* sm) {
    double l = -alpha * sqrt(1.0 - mu0 * mu0);
    double q2 = beta * beta - (a_in * a_in - alpha * alpha) * mu0 * mu0;
    double a = a_in;
    if (fabs(q2) < 1e-5 * 1e-5) q2 = 0.0;
    if (fabs(a) < 1e-5) a = 0.0;
    if (fabs(l) < 1e-5) l = 0.0;
    *L = l;
    *Q2 = q2;
    *a_out = a;
    *su = 1.0;
    *sm = (mu0 < 1.0 && beta >= 0.0) ? 1.0 : -1.0;
}

void orc_geomu_batch(const double* alpha, const double* beta, const double* a, const double* mu0,
                     const double* uf, double u0, int n, double* muf, int32_t* tpm, int32_t* ncase,
                     double* iu, int nthreads) {
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
#endif
    for (int i = 0; i < n; ++i) {
        double L, Q2, aa, su, sm;
        orc_geokerr_camera(alpha[i], beta[i], mu0[i], a[i], &L, &Q2, &aa, &su, &sm);
        orc_geomu_out o;
        orc_geomu(u0, uf[i], mu0[i], aa, L, Q2, 0, su, sm, &o);
        muf[i] = o.muf;
        tpm[i] = o.tpm;
        ncase[i] = o.ncase;
        iu[i] = o.iu;
    }
    (void)nthreads;
}
