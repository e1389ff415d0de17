/*
 * carlson_oracle.c -- FP64 restatement of GEOKERR's Carlson symmetric
 * elliptic integrals RF, RC, RD, RJ (duplication algorithm with a Taylor
 * tail, Carlson 1979/1995 as coded in Numerical Recipes):
 * geokerr_port/reference/geokerr_original.f:3444-3489 (rf), 3608-3645 (rc),
 * 3648-3694 (rd), 3697-3762 (rj).  Same ERRTOL values and series
 * coefficients as the Fortran; RJ with p <= 0 uses the Cauchy principal
 * value transform through RC, as the Fortran does.
 *
 * TEST INFRASTRUCTURE ONLY (see np_oracle.h).
 */
#include <math.h>
#include "np_oracle.h"

double orc_rf(double x, double y, double z) {
    const double ERRTOL = 0.0025, THIRD = 1.0 / 3.0, A1 = 1.0 / 24.0, C2 = 0.1,
                 C3 = 3.0 / 44.0, C4 = 1.0 / 14.0;
    double xt = x, yt = y, zt = z, ave, dx, dy, dz;
    do {
        double sx = sqrt(xt), sy = sqrt(yt), sz = sqrt(zt);
        double lam = sx * (sy + sz) + sy * sz;
        xt = 0.25 * (xt + lam);
        yt = 0.25 * (yt + lam);
        zt = 0.25 * (zt + lam);
        ave = THIRD * (xt + yt + zt);
        if (ave == 0.0) { dx = dy = dz = 0.0; }
        else { dx = (ave - xt) / ave; dy = (ave - yt) / ave; dz = (ave - zt) / ave; }
    } while (fmax(fmax(fabs(dx), fabs(dy)), fabs(dz)) > ERRTOL);
    double e2 = dx * dy - dz * dz;
    double e3 = dx * dy * dz;
    return (1.0 + (A1 * e2 - C2 - C3 * e3) * e2 + C4 * e3) / sqrt(ave);
}

double orc_rc(double x, double y) {
    const double ERRTOL = 0.0012, THIRD = 1.0 / 3.0, A1 = 0.3, C2 = 1.0 / 7.0,
                 C3 = 0.375, C4 = 9.0 / 22.0;
    double xt, yt, w, ave, s;
    if (y > 0.0) { xt = x; yt = y; w = 1.0; }
    else { xt = x - y; yt = -y; w = sqrt(x) / sqrt(xt); }
    do {
        double lam = 2.0 * sqrt(xt) * sqrt(yt) + yt;
        xt = 0.25 * (xt + lam);
        yt = 0.25 * (yt + lam);
        ave = THIRD * (xt + yt + yt);
        s = (yt - ave) / ave;
    } while (fabs(s) > ERRTOL);
    return w * (1.0 + s * s * (A1 + s * (C2 + s * (C3 + s * C4)))) / sqrt(ave);
}

double orc_rd(double x, double y, double z) {
    const double ERRTOL = 0.0015, A1 = 3.0 / 14.0, C2 = 1.0 / 6.0, C3 = 9.0 / 22.0,
                 C4 = 3.0 / 26.0, C5 = 0.25 * C3, C6 = 1.5 * C4;
    double xt = x, yt = y, zt = z, sum = 0.0, fac = 1.0, ave, dx, dy, dz;
    do {
        double sx = sqrt(xt), sy = sqrt(yt), sz = sqrt(zt);
        double lam = sx * (sy + sz) + sy * sz;
        sum = sum + fac / (sz * (zt + lam));
        fac = 0.25 * fac;
        xt = 0.25 * (xt + lam);
        yt = 0.25 * (yt + lam);
        zt = 0.25 * (zt + lam);
        ave = 0.2 * (xt + yt + 3.0 * zt);
        dx = (ave - xt) / ave; dy = (ave - yt) / ave; dz = (ave - zt) / ave;
    } while (fmax(fmax(fabs(dx), fabs(dy)), fabs(dz)) > ERRTOL);
    double ea = dx * dy, eb = dz * dz, ec = ea - eb, ed = ea - 6.0 * eb, ee = ed + ec + ec;
    return 3.0 * sum + fac * (1.0 + ed * (-A1 + C5 * ed - C6 * dz * ee) +
                              dz * (C2 * ee + dz * (-C3 * ec + dz * C4 * ea))) /
                           (ave * sqrt(ave));
}

double orc_rj(double x, double y, double z, double p) {
    const double ERRTOL = 0.0015, A1 = 3.0 / 14.0, C2 = 1.0 / 3.0, C3 = 3.0 / 22.0,
                 C4 = 3.0 / 26.0, C5 = 0.75 * C3, C6 = 1.5 * C4, C7 = 0.5 * C2, C8 = C3 + C3;
    double sum = 0.0, fac = 1.0, xt, yt, zt, pt, a = 0.0, b = 0.0, rcx = 0.0;
    double ave, dx, dy, dz, dp;
    if (p > 0.0) { xt = x; yt = y; zt = z; pt = p; }
    else {
        xt = fmin(fmin(x, y), z);
        zt = fmax(fmax(x, y), z);
        yt = x + y + z - xt - zt;
        a = 1.0 / (yt - p);
        b = a * (zt - yt) * (yt - xt);
        pt = yt + b;
        double rho = xt * zt / yt;
        double tau = p * pt / yt;
        rcx = orc_rc(rho, tau);
    }
    do {
        double sx = sqrt(xt), sy = sqrt(yt), sz = sqrt(zt);
        double lam = sx * (sy + sz) + sy * sz;
        double alpha = pt * (sx + sy + sz) + sx * sy * sz;
        alpha = alpha * alpha;
        double beta = pt * (pt + lam) * (pt + lam);
        sum = sum + fac * orc_rc(alpha, beta);
        fac = 0.25 * fac;
        xt = 0.25 * (xt + lam);
        yt = 0.25 * (yt + lam);
        zt = 0.25 * (zt + lam);
        pt = 0.25 * (pt + lam);
        ave = 0.2 * (xt + yt + zt + pt + pt);
        dx = (ave - xt) / ave; dy = (ave - yt) / ave; dz = (ave - zt) / ave; dp = (ave - pt) / ave;
    } while (fmax(fmax(fabs(dx), fabs(dy)), fmax(fabs(dz), fabs(dp))) > ERRTOL);
    double ea = dx * (dy + dz) + dy * dz;
    double eb = dx * dy * dz;
    double ec = dp * dp;
    double ed = ea - 3.0 * ec;
    double ee = eb + 2.0 * dp * (ea - ec);
    double rj = 3.0 * sum + fac * (1.0 + ed * (-A1 + C5 * ed - C6 * ee) + eb * (C7 + dp * (-C8 + dp * C4)) +
                                   dp * ea * (C2 - dp * C3) - C2 * dp * ec) /
                                (ave * sqrt(ave));
    if (p <= 0.0) rj = a * (b * rj + 3.0 * (rcx - orc_rf(xt, yt, zt)));
    return rj;
}
