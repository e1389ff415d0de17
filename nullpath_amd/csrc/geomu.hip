// geomu.hip -- gfx950 kernels for GEOKERR's semi-analytic solver: Carlson
// RF/RC/RD/RJ batches and the GEOMU batch (mu at a given inverse radius).
//
// One ray per lane, grid-stride loop over SoA inputs (coalesced 8-byte
// loads), FP64 throughout.  The numerics follow
// geokerr_port/reference/geokerr_original.f: rf/rc/rd/rj 3444-3489 and
// 3608-3762 (Carlson duplication with the Fortran's ERRTOL and Taylor
// tails), GEOMU 771-1186 with the RF reductions of 1188-1645, calcimu*
// 3137-3274, SNCNDN 3277-3340, ZROOTS/LAGUER 3343-3441 and 3527-3606.
// The reference's unbounded duplication loops get an iteration cap (a
// converged argument never needs more than ~30 steps; the cap only stops a
// NaN input from spinning).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "np_device.hpp"
#include "nullpath.h"

namespace np {

namespace {

constexpr int kCarlsonMaxIter = 100;

__device__ double rf_d(double x, double y, double z) {
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

__device__ double rc_d(double x, double y) {
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

__device__ double rd_d(double x, double y, double z) {
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

__device__ double rj_d(double x, double y, double z, double p) {
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

// ------------------------------- complex ----------------------------------
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

// LAGUER (geokerr_original.f:3527-3606); a[0..m], constant term first.
__device__ void laguer(const cplx* a, int m, cplx& x) {
    const double frac[8] = {0.5, 0.25, 0.75, 0.13, 0.38, 0.62, 0.88, 1.0};
    const int MT = 10, MAXIT = 80;
    const double EPSS = 1e-15;
    for (int iter = 1; iter <= MAXIT; ++iter) {
        cplx b = a[m];
        double err = cabs_d(b);
        cplx d = mk(0.0), f = mk(0.0);
        const double abx = cabs_d(x);
        for (int j = m - 1; j >= 0; --j) {
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
        const cplx sq = csqrt_d((double)(m - 1) * ((double)m * h - g2));
        cplx gp = g + sq;
        const cplx gm = g - sq;
        const double abp = cabs_d(gp), abm = cabs_d(gm);
        if (abp < abm) gp = gm;
        cplx dx;
        if (fmax(abp, abm) > 0.0) {
            dx = mk((double)m) / gp;
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

// ZROOTS (geokerr_original.f:3343-3441); roots sorted by real part.
__device__ void zroots(const cplx* a, int m, cplx* roots) {
    const double EPS = 1e-6;
    cplx ad[8];
    for (int j = 0; j <= m; ++j) ad[j] = a[j];
    cplx xstart = mk(0.0);
    int ncc = 0;
    for (int j = m; j >= 1; --j) {
        cplx x = xstart;
        if (j < m && roots[j].im != 0.0 && ncc == 0) {
            // the reference builds (Re r, -Re r), not the conjugate
            xstart = mk(roots[j].re, -roots[j].re);
            ncc = 1;
        } else {
            xstart = mk(0.0);
            ncc = 0;
        }
        laguer(ad, j, x);
        if (fabs(x.im) <= 2.0 * EPS * EPS * fabs(x.re)) x = mk(x.re);
        roots[j - 1] = x;
        cplx b = ad[j];
        for (int jj = j - 1; jj >= 0; --jj) {
            const cplx c = ad[jj];
            ad[jj] = b;
            b = x * b + c;
        }
    }
    for (int j = 0; j < m; ++j) laguer(a, m, roots[j]);
    for (int j = 1; j < m; ++j) {
        const cplx x = roots[j];
        int i;
        for (i = j - 1; i >= 0; --i) {
            if (roots[i].re <= x.re) break;
            roots[i + 1] = roots[i];
        }
        roots[i + 1] = x;
    }
}

// SNCNDN (geokerr_original.f:3277-3340).
__device__ void sncndn(double uu, double emmc, double& sn, double& cn, double& dn) {
    const double CA = 3e-8;
    double emc = emmc, u = uu, a, b, c = 0.0, d = 0.0, em[13], en[13];
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
            em[i - 1] = a;
            emc = sqrt(emc);
            en[i - 1] = emc;
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
                b = em[ii - 1];
                a = c * a;
                c = dn * c;
                dn = (en[ii - 1] + a) / (b + a);
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
}

// RF-only Carlson reductions used by GEOMU (p(4) = 0 or p(5) = 0 branches of
// geokerr_original.f:1188-1645).
__device__ double ellcubicreal_rf(double a1, double b1, double a2, double b2, double a3, double b3,
                                  double y, double x) {
    const double X1 = sqrt(a1 + b1 * x), X2 = sqrt(a2 + b2 * x), X3 = sqrt(a3 + b3 * x);
    const double Y1 = sqrt(a1 + b1 * y), Y2 = sqrt(a2 + b2 * y), Y3 = sqrt(a3 + b3 * y);
    const double U1c = (X1 * Y2 * Y3 + Y1 * X2 * X3) / (x - y);
    const double t2 = (X2 * Y1 * Y3 + Y2 * X1 * X3) / (x - y);
    const double t3 = (X3 * Y1 * Y2 + Y3 * X1 * X2) / (x - y);
    return 2.0 * rf_d(t3 * t3, t2 * t2, U1c * U1c);
}

__device__ double ellcubiccomplex_rf(double a1, double b1, double f, double g, double h, double y,
                                     double x) {
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
    return 4.0 * rf_d(M2, Lm2, Lp2);
}

__device__ double ellquarticreal_rf(double a1, double b1, double a2, double b2, double a3, double b3,
                                    double a4, double b4, double y, double x) {
    const double X1 = sqrt(a1 + b1 * x), X2 = sqrt(a2 + b2 * x), X3 = sqrt(a3 + b3 * x), X4 = sqrt(a4 + b4 * x);
    const double Y1 = sqrt(a1 + b1 * y), Y2 = sqrt(a2 + b2 * y), Y3 = sqrt(a3 + b3 * y), Y4 = sqrt(a4 + b4 * y);
    const double u1 = (X1 * X2 * Y3 * Y4 + Y1 * Y2 * X3 * X4) / (x - y);
    const double u2 = (X1 * X3 * Y2 * Y4 + Y1 * Y3 * X2 * X4) / (x - y);
    const double u3 = (X1 * X4 * Y2 * Y3 + Y1 * Y4 * X2 * X3) / (x - y);
    return 2.0 * rf_d(u1 * u1, u2 * u2, u3 * u3);
}

__device__ double ellquarticcomplex_rf(double a1, double b1, double a4, double b4, double f, double g,
                                       double h, double y, double x) {
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
    return 4.0 * rf_d(M2, Lm2, Lp2);
}

__device__ double elldoublecomplex_rf(double f1, double g1, double h1, double f2, double g2, double h2,
                                      double y, double x) {
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
    return 4.0 * rf_d(M2, M2 + (delta122 - Delta), M2 + (delta122 + Delta));
}

__device__ __forceinline__ double sgn1(double x) { return copysign(1.0, x); }
__device__ __forceinline__ double neg1pow(int n) { return (n % 2 == 0) ? 1.0 : -1.0; }

struct GeomuResult {
    double muf, iu;
    int tpm, ncase;
};

// GEOMU (geokerr_original.f:771-1186) with FIRSTPT = .true., PHT = .false.
__device__ GeomuResult geomu(double u0, double uf, const double mu0, const double a, const double l,
                             const double q2, int tpr, const double su, const double sm) {
    const double one = 1.0, two = 2.0, third = 0.3333333333333333;
    const double pi = acos(-one);
    const double uplus = one / (one + sqrt(one - a * a));
    const double l2 = l * l;
    const double ql2 = q2 + l2;
    const double cc = a * a - q2 - l2;
    const double dd = two * ((a - l) * (a - l) + q2);
    const double ee = -a * a * q2;
    double iu = 0.0, iu0 = 0.0, iu1 = 0.0, muf = 0.0;
    double u1 = 0.0, u2 = 0.0, u3 = 0.0, u4 = 0.0;
    int tpm = 0, ncase = 0;

    if (ee == 0.0 && dd != 0.0) {
        const double qq = cc * cc / dd / dd / 9.0;
        const double rr = (two * (cc * cc * cc) / (dd * dd * dd) + 27.0 / dd) / 54.0;
        const double dis = rr * rr - qq * qq * qq;
        if (dis < -1e-16) {
            const double theta = acos(rr / pow(qq, 1.5));
            u1 = -two * sqrt(qq) * cos(theta / 3.0) - cc / dd / 3.0;
            u2 = -two * sqrt(qq) * cos((theta - two * pi) / 3.0) - cc / dd / 3.0;
            u3 = -two * sqrt(qq) * cos((theta + two * pi) / 3.0) - cc / dd / 3.0;
            for (int guard = 0; guard < 3; ++guard) {
                if (u0 <= u2) {
                    if (uf > u2) uf = u2;
                    ncase = 1;
                    iu0 = (u0 != u2) ? ellcubicreal_rf(-u1, one, u2, -one, u3, -one, u0, u2) : 0.0;
                    if (uf != u2) iu1 = ellcubicreal_rf(-u1, one, u2, -one, u3, -one, uf, u2);
                    iu = su * (iu0 - neg1pow(tpr) * iu1) / sqrt(dd);
                    break;
                } else if (u0 >= u3) {
                    if (uf < u3) uf = u3;
                    ncase = 2;
                    iu0 = (u0 != u3) ? -ellcubicreal_rf(-u1, one, -u2, one, -u3, one, u3, u0) : 0.0;
                    if (uf != u3) iu1 = -ellcubicreal_rf(-u1, one, -u2, one, -u3, one, u3, uf);
                    iu = su * (iu0 - neg1pow(tpr) * iu1) / sqrt(dd);
                    break;
                }
                u0 = (su == 1.0) ? u3 : u2;   // forbidden region: the reference's input edit
            }
        } else if (fabs(dis) < 1e-16) {
            u1 = -two * sqrt(qq) - cc / dd / 3.0;
            u2 = -two * sqrt(qq) * cos((two * pi) / 3.0) - cc / dd / 3.0;
            u3 = u2;
            tpr = 0;
            if (u0 <= u2) {
                ncase = 1;
                if (uf > u2) {
                    uf = u2;
                    iu = su * 1e300;
                } else {
                    const double farg = (sqrt(uf - u1) + sqrt(u2 - u1)) / fabs(sqrt(uf - u1) - sqrt(u2 - u1));
                    const double sarg = (sqrt(u0 - u1) + sqrt(u2 - u1)) / fabs(sqrt(u0 - u1) - sqrt(u2 - u1));
                    iu = su * (log(farg) - log(sarg)) / sqrt((u2 - u1) * dd);
                }
            } else if (u0 >= u2) {
                ncase = 2;
                if (uf < u2) {
                    uf = u2;
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
            if (uf > u0) iu = su * ellcubiccomplex_rf(-u1, one, f, g, one, u0, uf) / sqrt(dd);
            else if (uf < u0) iu = -su * ellcubiccomplex_rf(-u1, one, f, g, one, uf, u0) / sqrt(dd);
            else iu = 0.0;
        }
        if (q2 == 0.0) {
            const double s1 = sgn1(mu0);
            if (l2 < a * a && mu0 != 0.0) {
                const double muplus = s1 * sqrt(one - l2 / a / a);
                const double x = mu0 / muplus;
                const double asech = (x > 0.0) ? log((1.0 + sqrt(1.0 - x * x)) / x) : 0.0;
                muf = muplus / cosh(fabs(a * muplus) * iu - s1 * sm * asech);
            } else {
                muf = 0.0;
            }
        } else if (a == 0.0) {
            const double a1 = sm;
            double muplus = sqrt(q2 / ql2);
            if (mu0 > muplus) muplus = mu0;
            const double i1mu = acos(mu0 / muplus) / sqrt(ql2);
            const double i3mu = pi / sqrt(ql2);
            if (sm == 1.0) tpm = (int)((iu - i1mu) / i3mu) + (int)((1 + sgn1((iu - i1mu) / i3mu)) / 2.0);
            else tpm = (int)((iu + i1mu) / i3mu);
            const double a2 = sm * neg1pow(tpm);
            const double a3 = two * (int)((two * (double)tpm + 3.0 - sm) / 4.0) - one;
            muf = -muplus * cos(sqrt(ql2) * (iu - a1 * i1mu - a3 * i3mu) / a2);
        }
    } else if (ee == 0.0 && dd == 0.0) {
        ncase = 4;
        tpr = 0;
        iu = su * (uf - u0);
        muf = 0.0;
    } else {
        cplx c[5], root[4];
        c[0] = mk(one); c[1] = mk(0.0); c[2] = mk(cc); c[3] = mk(dd); c[4] = mk(ee);
        zroots(c, 4, root);
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
                u0 = (su == 1.0) ? 0.0 : uplus;
            }
        }
        if (ncase == 5) {
            const double qs = sgn1(q2);
            const double f = -qs * one / fabs(ee) / u1 / u4;
            const double g = (u4 + u1) / u1 / u4 * f;
            if (uf > u0) iu = su * ellquarticcomplex_rf(-u1, one, qs * u4, -qs * one, f, g, one, u0, uf) / sqrt(fabs(ee));
            else if (uf < u0) iu = -su * ellquarticcomplex_rf(-u1, one, qs * u4, -qs * one, f, g, one, uf, u0) / sqrt(fabs(ee));
            else iu = 0.0;
        } else if (ncase == 6) {
            const double se = sqrt(ee);
            cplx co[7], hroots[6];
            co[0] = mk(one); co[1] = mk(-cc / se); co[2] = mk(-one);
            co[3] = mk(se * (two * cc / ee - (dd / ee) * (dd / ee)));
            co[4] = mk(-one); co[5] = mk(-cc / se); co[6] = mk(one);
            zroots(co, 6, hroots);
            double h1 = 0.0;
            for (int i = 0; i < 6 && h1 == 0.0; ++i)
                if (hroots[i].im == 0.0) h1 = hroots[i].re;
            const double h2 = one / h1;
            const double g1 = dd / ee / (h2 - h1);
            const double g2 = -g1;
            const double f1 = one / sqrt(ee);
            const double f2 = f1;
            if (uf > u0) iu = su * elldoublecomplex_rf(f1, g1, h1, f2, g2, h2, u0, uf) / sqrt(fabs(ee));
            else if (uf < u0) iu = -su * elldoublecomplex_rf(f1, g1, h1, f2, g2, h2, uf, u0) / sqrt(fabs(ee));
            else iu = 0.0;
        } else {
            if (fabs(u3 - u2) > 1e-12) {
                if (ncase == 7) {
                    if (uf > u2) uf = u2;
                    iu0 = (u0 != u2) ? ellquarticreal_rf(-u1, one, u2, -one, u3, -one, u4, -one, u0, u2) : 0.0;
                    if (uf != u2) iu1 = ellquarticreal_rf(-u1, one, u2, -one, u3, -one, u4, -one, uf, u2);
                    iu = su * (iu0 - neg1pow(tpr) * iu1) / sqrt(fabs(ee));
                } else if (ncase == 8) {
                    if (uf < u3) uf = u3;
                    iu0 = (u0 != u3) ? -ellquarticreal_rf(-u1, one, -u2, one, -u3, one, u4, -one, u3, u0) : 0.0;
                    if (uf != u3) iu1 = -ellquarticreal_rf(-u1, one, -u2, one, -u3, one, u4, -one, u3, uf);
                    iu = su * (iu0 - neg1pow(tpr) * iu1) / sqrt(fabs(ee));
                }
            } else {
                tpr = 0;
                if (ncase == 7) {
                    if (uf > u2) uf = u2;
                    iu0 = ellquarticreal_rf(-u1, one, u2, -one, u3, -one, u4, -one, u0, uf);
                    iu = su * iu0 / sqrt(fabs(ee));
                } else if (ncase == 8) {
                    if (uf < u3) uf = u3;
                    iu0 = -ellquarticreal_rf(-u1, one, -u2, one, -u3, one, u4, -one, u0, uf);
                    iu = su * iu0 / sqrt(fabs(ee));
                } else {
                    ncase = 0;
                }
            }
        }
        // roots of M(mu) (Eq. 12) and Jacobi inversion (Eqs. 30-42)
        const double amq = a * a - ql2;
        const double yy = -0.5 * (amq + sgn1(amq) * sqrt(amq * amq + 4.0 * q2 * a * a));
        double mneg, mpos;
        if (amq < 0.0) { mneg = -yy / a / a; mpos = q2 / yy; }
        else { mneg = q2 / yy; mpos = -yy / a / a; }
        if (mpos > 1.0) mpos = 1.0;
        double muplus = sqrt(mpos);
        double i1mu, i3mu, uarg, m1, muminus;
        bool sym = mneg < 0.0;
        if (sym) {
            if (muplus < mu0) { muplus = mu0; mpos = muplus * muplus; }
            muminus = -muplus;
            // calcimusym (geokerr_original.f:3202-3240)
            const double fac = fabs(a) * sqrt(mpos - mneg);
            const double mm1 = -mneg / (mpos - mneg);
            const double phi0 = acos(mu0 / muplus);
            // tan(phi0) < 0 for phi0 in [0, pi] is phi0 > pi/2; compared
            // against the double nearest pi/2 so that the branch does not
            // hinge on the last bit of a device tan() at pi/2.
            const double s0 = sin(phi0);
            const double ak = sqrt(1.0 - mm1);
            const double q0 = (1.0 - s0 * ak) * (1.0 + s0 * ak);
            const double rf0 = rf_d(1.0 - s0 * s0, q0, 1.0);
            const double rfc = rf_d(0.0, mm1, 1.0);
            double imum = 2.0 * rfc, imu0 = s0 * rf0;
            if (phi0 > 1.5707963267948966) imu0 = imum - imu0;
            i1mu = imu0 / fac;
            i3mu = imum / fac;
        } else {
            muplus = sgn1(mu0) * muplus;
            muminus = sgn1(mu0) * sqrt(mneg);
            if (fabs(muminus) > fabs(mu0)) { muminus = mu0; mneg = muminus * muminus; }
            if (fabs(mu0) > fabs(muplus)) { muplus = mu0; mpos = muplus * muplus; }
            // calcimuasym (geokerr_original.f:3137-3172)
            const double fac = fabs(a) * muplus;
            const double mm1 = mneg / mpos;
            // the reference's tan(asin(s0)) < 0 test (geokerr_original.f:3165)
            // never holds for s0 in [0, 1] with a correctly rounded asin
            // (asin(1) rounds below pi/2), so the branch is omitted; a device
            // asin(1) one ulp above pi/2 would otherwise flip it.
            const double s0 = sqrt((mpos - mu0 * mu0) / (mpos - mneg));
            const double ak = sqrt(1.0 - mm1);
            const double q0 = (1.0 - s0 * ak) * (1.0 + s0 * ak);
            const double rf0 = rf_d(1.0 - s0 * s0, q0, 1.0);
            const double rfc = rf_d(0.0, mm1, 1.0);
            double imum = 2.0 * rfc, imu0 = s0 * rf0;
            i1mu = imu0 / fac;
            i3mu = imum / fac / 2.0;
        }
        const double a1 = sm;
        if (sm == 1.0) {
            tpm = (int)((iu - i1mu) / i3mu) +
                  (sym ? (int)((1 + sgn1((iu - i1mu) / i3mu)) / 2.0)
                       : (int)fabs((1 + sgn1((iu - i1mu) / i3mu)) / 2.0));
        } else {
            tpm = (int)((iu + i1mu) / i3mu);
        }
        const double a2 = sm * neg1pow(tpm);
        const double a3 = two * (int)((two * (double)tpm + 3.0 - sm) / 4.0) - one;
        const double iarg = fabs(a) / a2 * (iu - a1 * i1mu - a3 * i3mu);
        double sn, cn, dn;
        if (sym) {
            uarg = sqrt(mpos - mneg) * iarg;
            m1 = -mneg / (mpos - mneg);
            sncndn(uarg, m1, sn, cn, dn);
            muf = muminus * cn;
        } else {
            uarg = fabs(muplus) * iarg;
            m1 = mneg / mpos;
            sncndn(uarg, m1, sn, cn, dn);
            muf = muminus / dn;
        }
    }
    GeomuResult res;
    res.muf = muf;
    res.iu = iu;
    res.tpm = tpm;
    res.ncase = ncase;
    return res;
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

__global__ void __launch_bounds__(256)
geomu_camera_kernel(const double* __restrict__ alpha, const double* __restrict__ beta,
                    const double* __restrict__ spin, const double* __restrict__ mu0,
                    const double* __restrict__ uf, const double u0, const int64_t n,
                    double* __restrict__ muf, int32_t* __restrict__ tpm, int32_t* __restrict__ ncase,
                    double* __restrict__ iu) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double L, Q2, a, su, sm;
        camera_constants(alpha[i], beta[i], mu0[i], spin[i], L, Q2, a, su, sm);
        const GeomuResult r = geomu(u0, uf[i], mu0[i], a, L, Q2, 0, su, sm);
        muf[i] = r.muf;
        if (tpm) tpm[i] = r.tpm;
        if (ncase) ncase[i] = r.ncase;
        if (iu) iu[i] = r.iu;
    }
}

__global__ void __launch_bounds__(256)
geomu_kernel(const double* __restrict__ u0, const double* __restrict__ uf,
             const double* __restrict__ mu0, const double* __restrict__ spin,
             const double* __restrict__ L, const double* __restrict__ Q2,
             const double* __restrict__ su, const double* __restrict__ sm,
             const int32_t* __restrict__ tpr, const int64_t n, double* __restrict__ muf,
             int32_t* __restrict__ tpm, int32_t* __restrict__ ncase, double* __restrict__ iu) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const GeomuResult r = geomu(u0[i], uf[i], mu0[i], spin[i], L[i], Q2[i], tpr ? tpr[i] : 0,
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
            case NP_CARLSON_RF: v = rf_d(x[i], y[i], z[i]); break;
            case NP_CARLSON_RC: v = rc_d(x[i], y[i]); break;
            case NP_CARLSON_RD: v = rd_d(x[i], y[i], z[i]); break;
            default: v = rj_d(x[i], y[i], z[i], p[i]); break;
        }
        out[i] = v;
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
    hipLaunchKernelGGL(np::geomu_camera_kernel, dim3(np::grid_for(n)), dim3(256), 0,
                       (hipStream_t)stream, d_alpha, d_beta, d_spin, d_mu0, d_uf, u0, n, d_muf,
                       d_tpm, d_ncase, d_iu);
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
