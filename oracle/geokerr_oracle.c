/*
 * geokerr_oracle.c -- FP64 restatement of GEOKERR's time/azimuth/affine
 * integrals, its sampling driver and the geokerr_ref_api entry points:
 *   GEOPHITIME                    geokerr_original.f:1648-2360
 *   calcphitmuasym / calcphitmusym 2361-2434 / 2435-2507
 *   ellphitmu                     2508-2592
 *   phifnkerr / tfnkerr           2593-2628 / 2629-2671
 *   GEOKERR driver                357-604 (USEGEOR = .false. branch as
 *                                 geokerr_eval_lambda runs it; the whole
 *                                 driver with FINDMROOTS 607-642 and
 *                                 INDEP_MUF 645-706 in orc_gk_geokerr_full)
 *   geokerr_eval_lambda, geokerr_estimate_periods, geokerr_phi_omega
 *                                 geokerr_eval_lambda.f90:62-306
 *
 * The restatement follows the Fortran formula by formula.  The
 * [-1,-1,-1,-4] and [-1,-1,-1,-1,-4] Carlson reductions it takes the first
 * term of Eq. (47) and the affine parameter (Eq. 49) from are exact for
 * U(u), which has no linear term (pinned against quadrature on U's own
 * roots in the tests).  Where the Fortran reads a variable it never set,
 * the choice made here is stated at the spot.
 *
 * TEST INFRASTRUCTURE ONLY (see np_oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "np_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

static double sgn1(double x) { return copysign(1.0, x); }
static double neg1pow(unsigned n) { return (n & 1u) ? -1.0 : 1.0; }
/* Fortran INT() as the reference's x86-64 build executes it: INT_MIN for
 * NaN and out-of-range values. */
static int f_int(double x) { return (x > -2147483649.0 && x < 2147483648.0) ? (int)x : (int)0x80000000u; }

/* Eq. (47): t_u from the four u integrals (geokerr_original.f:1796-1798). */
static double eq47(double a, double l, double uplus, double umi, double t1, double t2, double t3,
                   double t4) {
    return (umi - 1.0 / uplus) * t1 + (umi * umi - 1.0 / (uplus * uplus)) * t2 -
           (2.0 * a * (a - l) + a * a / uplus + 1.0 / (uplus * uplus * uplus)) * t3 +
           (2.0 * a * (a - l) + a * a * umi + umi * umi * umi) * t4;
}

/* Eq. (48) for one of the u0 / uf pieces (geokerr_original.f:1806-1808). */
static double eq48(double a, double l, double uplus, double umi, double p1, double p2) {
    return -(l / uplus + 2.0 * (a - l)) * p1 + (l * umi + 2.0 * (a - l)) * p2;
}

/* ---------------- equal-root cubic closed forms ---------------- */

/* phifnkerr (geokerr_original.f:2593-2628). */
static double phifnkerr(double u, double u1, double u2, double l, double a) {
    double up = 1.0 / (1.0 + sqrt(1.0 - a * a));
    double umi = 1.0 - sqrt(1.0 - a * a);
    double s = sqrt(u - u1);
    double su21 = sqrt(u2 - u1);
    double sup = sqrt(1.0 - u1 / up);
    double sum = sqrt(1.0 - u1 * umi);
    return (l + 2.0 * a * u2 - 2.0 * l * u2) / (su21 * (u2 * umi - 1.0) * (u2 / up - 1.0)) *
               log((su21 + s) / (su21 - s)) +
           (2.0 * a + l * (-2.0 + umi)) * sqrt(umi) * log((sum + s * sqrt(umi)) / (sum - s * sqrt(umi))) /
               (sum * (u2 * umi - 1.0) * (umi - 1.0 / up)) +
           (2.0 * a + l * (-2.0 + 1.0 / up)) * sqrt(1.0 / up) *
               log((sup + sqrt(1.0 / up) * s) / (sup - sqrt(1.0 / up) * s)) / (1.0 / up - umi) * sup *
               (u2 / up - 1.0);
}

/* tfnkerr (geokerr_original.f:2629-2671). */
static double tfnkerr(double u0, double uf, double u1, double u2, double l, double a) {
    double up = 1.0 / (1.0 + sqrt(1.0 - a * a));
    double umi = 1.0 - sqrt(1.0 - a * a);
    double sf = sqrt(uf - u1), ss = sqrt(u0 - u1);
    double su21 = sqrt(u2 - u1), su1 = sqrt(-u1);
    double sup = sqrt(1.0 - u1 / up), sum = sqrt(1.0 - u1 * umi);
    double rup = sqrt(1.0 / up), rum = sqrt(umi);
    return sf / (uf * u1 * u2) - ss / (u0 * u1 * u2) -
           ((-u2 - 2.0 * u1 * (1.0 + u2 * umi + u2 / up)) *
            log((sf + su1) * (su1 - ss) / ((-sf + su1) * (ss + su1)))) /
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

/* ---------------- mu integrals of t and phi ---------------- */

/* ellphitmu (geokerr_original.f:2508-2592): Legendre F/E/Pi pieces from the
 * Carlson forms (Press et al. 6.11.20-21).  n arrives with the Numerical
 * Recipes sign convention (the callers pass -n). */
static void ellphitmu(double phi0, double phif, double m1, double n, int firstpt, double rf0,
                      double rff, double rfc, double* rdc, double* rjc, double* elle0, double* ellef,
                      double* ellec, double* ellpi0, double* ellpif, double* ellpic) {
    const double pi = acos(-1.0);
    double ak = sqrt(1.0 - m1);
    if (firstpt) {
        double s0 = sin(phi0);
        double q0 = (1.0 - s0 * ak) * (1.0 + s0 * ak);
        double rd0 = orc_rd(1.0 - s0 * s0, q0, 1.0);
        *rdc = orc_rd(0.0, m1, 1.0);
        *elle0 = s0 * (rf0 - ((s0 * ak) * (s0 * ak)) * rd0 / 3.0);
        *ellec = 2.0 * (rfc - ak * ak * (*rdc) / 3.0);
        if (n != 1.0) {
            double rj0 = orc_rj(1.0 - s0 * s0, q0, 1.0, 1.0 - n * s0 * s0);
            *rjc = orc_rj(0.0, m1, 1.0, 1.0 - n);
            *ellpi0 = s0 * (rf0 + n * s0 * s0 * rj0 / 3.0);
            *ellpic = 2.0 * (rfc + n * (*rjc) / 3.0);
        } else {
            /* n = 1 flags a pole-crossing orbit; the Fortran leaves RJC unset,
             * here it is zero (every later use is multiplied by L = 0). */
            *rjc = 0.0;
            *ellpi0 = 0.0;
            *ellpic = 0.0;
        }
        if (phi0 > pi / 2.0) {
            *ellpi0 = *ellpic - *ellpi0;
            *elle0 = *ellec - *elle0;
        }
    }
    double sf = sin(phif);
    double qf = (1.0 - sf * ak) * (1.0 + sf * ak);
    double rdf = orc_rd(1.0 - sf * sf, qf, 1.0);
    *ellef = sf * (rff - ((sf * ak) * (sf * ak)) * rdf / 3.0);
    if (n != 1.0) {
        double rjf = orc_rj(1.0 - sf * sf, qf, 1.0, 1.0 - n * sf * sf);
        *ellpif = sf * (rff + n * sf * sf * rjf / 3.0);
    } else {
        *ellpif = 0.0;
    }
    if (phif > pi / 2.0) {
        *ellpic = 2.0 * (rfc + n * (*rjc) / 3.0);
        *ellec = 2.0 * (rfc - ak * ak * (*rdc) / 3.0);
        *ellpif = *ellpic - *ellpif;
        *ellef = *ellec - *ellef;
    }
}

/* calcphitmuasym / calcphitmusym (geokerr_original.f:2361-2434 / 2435-2507).
 * The mu0 and turning-point pieces (phimu0, phimum, tmu0, tmum) are
 * written only when firstpt; the muf pieces on every call. */
static void calcphitmu(int sym, double a, double mneg, double mpos, double mu0, double muf,
                       double muplus, double* phimu0, double* phimuf, double* phimum, double* tmu0,
                       double* tmuf, double* tmum, double rf0, double rff, double rfc, double* rdc,
                       double* rjc, int firstpt) {
    double fac, m1, n, fac2, phi0, phif;
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
    if (sym) {
        phi0 = acos(mu0 / muplus);
        phif = acos(muf / muplus);
    } else {
        phi0 = asin(sqrt((mpos - mu0 * mu0) / (mpos - mneg)));
        phif = asin(sqrt((mpos - muf * muf) / (mpos - mneg)));
    }
    double elle0 = 0.0, ellef = 0.0, ellec = 0.0, ellpi0 = 0.0, ellpif = 0.0, ellpic = 0.0;
    ellphitmu(phi0, phif, m1, -n, firstpt, rf0, rff, rfc, rdc, rjc, &elle0, &ellef, &ellec, &ellpi0,
              &ellpif, &ellpic);
    /* The symmetric case integrates between -mu+ and mu+, the asymmetric
     * one between mu- and mu+: half the complete integral. */
    double half = sym ? 1.0 : 2.0;
    if (firstpt) {
        *tmu0 = fac * elle0;
        *tmum = fac * ellec / half;
        *phimu0 = ellpi0 / fac2 / fac;
        *phimum = ellpic / fac2 / fac / half;
    }
    *tmuf = fac * ellef;
    *phimuf = ellpif / fac2 / fac;
}

/* ---------------- GEOPHITIME ---------------- */

void orc_gk_geophitime(orc_gk_vars* v, double u0, double uf, double mu0, double muf, double a,
                       double l, double q2, int32_t tpm, int32_t tpr, double su, double sm,
                       double iu, int firstpt, double* phimu_o, double* tmu_o, double* phiu_o,
                       double* tu_o, double* lambda_o) {
    const double one = 1.0, two = 2.0;
    const double pi = acos(-one);
    int p[5] = {-1, -1, -1, 0, 0};
    double uplus = one / (one + sqrt(one - a * a));
    double umi = one - sqrt(one - a * a);
    double ur = -one / (two * sqrt(one - a * a));
    double qs = sgn1(q2);
    double dd = two * ((a - l) * (a - l) + q2);
    double ee = -a * a * q2;
    double l2 = l * l;
    double ql2 = q2 + l2;
    double a1 = sm;
    double a2 = sm * neg1pow(tpm);
    double a3 = two * f_int((two * (double)tpm + 3.0 - sm) / 4.0) - one;
    double tu = 0.0, phiu = 0.0, phimu = 0.0, tmu = 0.0, lambdau = 0.0;
    const int ncase = v->ncase;
    const double u1 = v->u1, u2 = v->u2, u3 = v->u3, u4 = v->u4;
    const double sgn_tpr = neg1pow(tpr);
    double tu1, tu2, tu3, tu4, tu11, tu12, tu13, tu14, phiu01, phiu02, phiu11, phiu12, phiu0, phiu1;

    if (ncase == 0) {
        /* all zero */
    } else if (ncase < 3) {
        p[4] = 0;
        if (fabs(u3 - u2) < 1e-12) {
            /* Equal roots: elementary forms.  The Fortran leaves lambdau unset
             * here; it is zero in this restatement. */
            tu = su * tfnkerr(u0, uf, u1, u2, l, a) / sqrt(dd);
            phiu = su * (phifnkerr(uf, u1, u2, l, a) - phifnkerr(u0, u1, u2, l, a)) / sqrt(dd);
            if (u0 >= u3) {
                phiu = -phiu;
                tu = -tu;
            }
        } else if (u0 <= u2) {
            /* Table 1 row 1 (geokerr_original.f:1745-1808) */
            if (firstpt && u0 != u2) {
                p[3] = -2;
                v->tu03 = orc_ellcubicreal(p, -u1, one, u2, -one, u3, -one, one, -one / uplus, &v->rffu0, u0, u2);
                v->tu04 = orc_ellcubicreal(p, -u1, one, u2, -one, u3, -one, one, -umi, &v->rffu0, u0, u2);
                v->tu02 = orc_ellcubicreal(p, -u1, one, u2, -one, u3, -one, 0.0, one, &v->rffu0, u0, u2);
                p[3] = -4;
                v->tu01 = orc_ellcubicreal(p, -u1, one, u2, -one, u3, -one, 0.0, one, &v->rffu0, u0, u2);
            } else if (u0 == u2) {
                v->tu01 = v->tu02 = v->tu03 = v->tu04 = 0.0;
            }
            phiu01 = v->tu03;
            phiu02 = v->tu04;
            if (fabs(uf - u2) > 1e-16) {
                p[3] = -2;
                phiu11 = orc_ellcubicreal(p, -u1, one, u2, -one, u3, -one, one, -one / uplus, &v->rffu1, uf, u2);
                phiu12 = orc_ellcubicreal(p, -u1, one, u2, -one, u3, -one, one, -umi, &v->rffu1, uf, u2);
                tu12 = orc_ellcubicreal(p, -u1, one, u2, -one, u3, -one, 0.0, one, &v->rffu1, uf, u2);
                p[3] = -4;
                tu11 = orc_ellcubicreal(p, -u1, one, u2, -one, u3, -one, 0.0, one, &v->rffu1, uf, u2);
            } else {
                tu11 = tu12 = phiu11 = phiu12 = 0.0;
            }
            tu13 = phiu11;
            tu14 = phiu12;
            double tu0 = eq47(a, l, uplus, umi, v->tu01, v->tu02, v->tu03, v->tu04);
            double tuf = eq47(a, l, uplus, umi, tu11, tu12, tu13, tu14);
            tu = su * (tu0 - sgn_tpr * tuf) / sqrt(dd) * ur;
            phiu0 = eq48(a, l, uplus, umi, phiu01, phiu02);
            phiu1 = eq48(a, l, uplus, umi, phiu11, phiu12);
            phiu = su * (phiu0 - sgn_tpr * phiu1) / sqrt(dd) * ur;
            lambdau = su * (v->tu01 - sgn_tpr * tu11) / sqrt(dd);
        } else if (u0 >= u3) {
            /* Table 1 row 2 (geokerr_original.f:1809-1870) */
            if (firstpt && u0 != u3) {
                p[3] = -2;
                v->tu03 = -orc_ellcubicreal(p, -u1, one, -u2, one, -u3, one, one, -one / uplus, &v->rffu0, u3, u0);
                v->tu04 = -orc_ellcubicreal(p, -u1, one, -u2, one, -u3, one, one, -umi, &v->rffu0, u3, u0);
                v->tu02 = -orc_ellcubicreal(p, -u1, one, -u2, one, -u3, one, 0.0, one, &v->rffu0, u3, u0);
                p[3] = -4;
                v->tu01 = -orc_ellcubicreal(p, -u1, one, -u2, one, -u3, one, 0.0, one, &v->rffu0, u3, u0);
            } else if (u0 == u3) {
                v->tu01 = v->tu02 = v->tu03 = v->tu04 = 0.0;
            }
            phiu01 = v->tu03;
            phiu02 = v->tu04;
            if (uf != u3) {
                p[3] = -2;
                phiu11 = -orc_ellcubicreal(p, -u1, one, -u2, one, -u3, one, one, -one / uplus, &v->rffu1, u3, uf);
                phiu12 = -orc_ellcubicreal(p, -u1, one, -u2, one, -u3, one, one, -umi, &v->rffu1, u3, uf);
                tu12 = -orc_ellcubicreal(p, -u1, one, -u2, one, -u3, one, 0.0, one, &v->rffu1, u3, uf);
                p[3] = -4;
                tu11 = -orc_ellcubicreal(p, -u1, one, -u2, one, -u3, one, 0.0, one, &v->rffu1, u3, uf);
            } else {
                tu11 = tu12 = phiu11 = phiu12 = 0.0;
            }
            tu13 = phiu11;
            tu14 = phiu12;
            double tu0 = eq47(a, l, uplus, umi, v->tu01, v->tu02, v->tu03, v->tu04);
            double tuf = eq47(a, l, uplus, umi, tu11, tu12, tu13, tu14);
            tu = su * (tu0 - sgn_tpr * tuf) / sqrt(dd) * ur;
            phiu0 = eq48(a, l, uplus, umi, phiu01, phiu02);
            phiu1 = eq48(a, l, uplus, umi, phiu11, phiu12);
            phiu = su * (phiu0 - sgn_tpr * phiu1) / sqrt(dd) * ur;
            lambdau = su * (v->tu01 - sgn_tpr * tu11) / sqrt(dd);
        }
    } else if (ncase == 3) {
        /* Table 1 row 3 (geokerr_original.f:1872-1924) */
        double f = -one / dd / u1, g = f / u1, h = one;
        if (u0 != uf) {
            double y = (u0 < uf) ? u0 : uf, x = (u0 < uf) ? uf : u0, sg = (u0 < uf) ? one : -one;
            p[3] = -4;
            tu1 = sg * orc_ellcubiccomplex(p, -u1, one, 0.0, one, f, g, h, &v->rffu0, y, x);
            p[3] = -2;
            tu2 = sg * orc_ellcubiccomplex(p, -u1, one, 0.0, one, f, g, h, &v->rffu0, y, x);
            tu3 = sg * orc_ellcubiccomplex(p, -u1, one, one, -one / uplus, f, g, h, &v->rffu0, y, x);
            tu4 = sg * orc_ellcubiccomplex(p, -u1, one, one, -umi, f, g, h, &v->rffu0, y, x);
            tu = su * ur / sqrt(dd) * eq47(a, l, uplus, umi, tu1, tu2, tu3, tu4);
            phiu0 = -tu3;
            phiu1 = -tu4;
            phiu = su * ur / sqrt(dd) * ((l / uplus + 2.0 * (a - l)) * phiu0 - (l * umi + 2.0 * (a - l)) * phiu1);
            lambdau = su * tu1 / sqrt(dd);
        }
    }

    if (q2 == 0.0) {
        /* mu part for q2 = 0 (geokerr_original.f:1925-1941) */
        double s1 = sgn1(mu0);
        double b1 = s1 * sm;
        double b2 = s1 * sm * neg1pow((unsigned)tpm + 1u);
        if (fabs(l) < fabs(a)) {
            double muplus = s1 * sqrt(one - l2 / a / a);
            double sq = sqrt(one - muplus * muplus);
            double t0h = tan(asin(mu0 / muplus) / two), tfh = tan(asin(muf / muplus) / two);
            double pm1 = sgn1(muplus * l / a) * (atan((muplus - t0h) / sq) + atan((muplus + t0h) / sq));
            double pm2 = sgn1(muplus * l / a) * (atan((muplus - tfh) / sq) + atan((muplus + tfh) / sq));
            v->phimu1 = pm1;
            phimu = b1 * pm1 + b2 * pm2;
            tmu = fabs(muplus * a) *
                  (b2 * sqrt(1.0 - (muf * muf) / (muplus * muplus)) + b1 * sqrt(1.0 - (mu0 * mu0) / (muplus * muplus)));
        }
    } else if (a == 0.0) {
        /* mu part for a = 0 (geokerr_original.f:1942-1969) */
        double muplus = sqrt(q2 / ql2);
        double mp2 = muplus * muplus;
        double vfm = (muf - mp2) / (one - muf) / muplus;
        double vfp = -(muf + mp2) / (one + muf) / muplus;
        if (fabs(vfm) > one) vfm = sgn1(vfm) * one;
        if (fabs(vfp) > one) vfp = sgn1(vfp) * one;
        double vsm = ((mu0 - mp2) == 0.0) ? -one : (mu0 - mp2) / (one - mu0) / muplus;
        double vsp = ((mu0 + mp2) == 0.0) ? -one : -(mu0 + mp2) / (one + mu0) / muplus;
        if (fabs(vsp) > one) vsp = sgn1(vsp) * one;
        if (fabs(vsm) > one) vsm = sgn1(vsm) * one;
        double pm1 = pi - asin(vsm) + asin(vsp);
        double pm2 = asin(vfm) - asin(vfp) + pi;
        double pm3 = two * pi;
        v->phimu1 = pm1;
        v->phimu3 = pm3;
        tmu = 0.0;
        phimu = -l * iu + sgn1(l) * 0.5 * (a1 * pm1 + a2 * pm2 + a3 * pm3);
    }

    if (ncase == 4) {
        /* q2 = 0, l = a: U(u) = 1 (geokerr_original.f:1970-1982) */
        phiu = su * ur * (l * log((uf / uplus - one) / (u0 / uplus - one)) -
                          l * umi * umi * log((uf * umi - one) / (u0 * umi - one)));
        tu = su * ur *
             ((umi - one / uplus) * (one / u0 - one / uf) + (umi * umi - one / (uplus * uplus)) * log(uf / u0) +
              (a * a / uplus + one / (uplus * uplus * uplus)) * uplus * log((uf / uplus - one) / (u0 / uplus - one)) -
              (a * a * umi + umi * umi * umi) * umi * log((uf * umi - one) / (u0 * umi - one)));
        lambdau = su * (one / u0 - one / uf);
    } else if (ncase == 5 || ncase == 6) {
        /* Table 1 rows 5 and 6 (geokerr_original.f:1983-2088).  Row 6 leaves
         * lambdau unset when u0 = uf; it is zero here. */
        p[3] = -1;
        double f = 0, g = 0, h = one, f1 = 0, g1 = 0, h1 = v->h1, f2 = 0, g2 = 0, h2 = 0;
        if (ncase == 5) {
            f = -qs * one / fabs(ee) / u1 / u4;
            g = (u4 + u1) / u1 / u4 * f;
        } else {
            h2 = one / h1;
            g1 = dd / ee / (h2 - h1);
            g2 = -g1;
            f1 = one / sqrt(ee);
            f2 = f1;
        }
        if (u0 != uf) {
            double y = (u0 < uf) ? u0 : uf, x = (u0 < uf) ? uf : u0, sg = (u0 < uf) ? one : -one;
            double b5[4] = {one, one, -one / uplus, -umi}, a5[4] = {0.0, 0.0, one, one};
            double tt[4];
            for (int k = 0; k < 4; ++k) {
                p[4] = (k == 0) ? -4 : -2;
                double a5k = (k == 0) ? 0.0 : a5[k], b5k = (k == 0) ? one : b5[k];
                if (ncase == 5)
                    tt[k] = sg * orc_ellquarticcomplex(p, -u1, one, u4 * qs, -one * qs, a5k, b5k, f, g, h,
                                                       &v->rffu0, y, x);
                else
                    tt[k] = sg * orc_elldoublecomplex(p, f1, g1, h1, f2, g2, h2, a5k, b5k, &v->rffu0, y, x);
            }
            tu1 = tt[0]; tu2 = tt[1]; tu3 = tt[2]; tu4 = tt[3];
            tu = su * ur / sqrt(-qs * ee) * eq47(a, l, uplus, umi, tu1, tu2, tu3, tu4);
            phiu0 = -tu3;
            phiu1 = -tu4;
            phiu = su * ur / sqrt(-qs * ee) * ((l / uplus + 2.0 * (a - l)) * phiu0 - (l * umi + 2.0 * (a - l)) * phiu1);
            lambdau = (ncase == 5) ? su * tu1 / sqrt(fabs(ee)) : su * tu1 / sqrt(-qs * ee);
        }
    } else if (ncase > 6) {
        p[3] = -1;
        if (fabs(u3 - u2) < 1e-12) {
            /* Equal-root quartic (geokerr_original.f:2093-2172).  The Fortran
             * combines an unset phiu0 with the wrong pieces and, in row 8
             * ascending, calls with p(5) unset; this restatement uses the
             * rows 5/6 pattern (phiu0 = -tu3, phiu1 = -tu4) and p(5) = -2. */
            int row7 = (u0 < u2), row8 = (u0 > u3);
            if (u0 != uf && (row7 || row8)) {
                double y = (u0 < uf) ? u0 : uf, x = (u0 < uf) ? uf : u0, sg = (u0 < uf) ? one : -one;
                double s2 = row7 ? one : -one;
                p[4] = -2;
                tu3 = sg * orc_ellquarticreal(p, -u1, one, s2 * u2, -s2, s2 * u3, -s2, u4, -one, one, -one / uplus, &v->rffu0, y, x);
                tu4 = sg * orc_ellquarticreal(p, -u1, one, s2 * u2, -s2, s2 * u3, -s2, u4, -one, one, -umi, &v->rffu0, y, x);
                tu2 = sg * orc_ellquarticreal(p, -u1, one, s2 * u2, -s2, s2 * u3, -s2, u4, -one, 0.0, one, &v->rffu0, y, x);
                p[4] = -4;
                tu1 = sg * orc_ellquarticreal(p, -u1, one, s2 * u2, -s2, s2 * u3, -s2, u4, -one, 0.0, one, &v->rffu0, y, x);
                tu = su * ur / sqrt(-qs * ee) * eq47(a, l, uplus, umi, tu1, tu2, tu3, tu4);
                phiu0 = -tu3;
                phiu1 = -tu4;
                phiu = su * ur / sqrt(-qs * ee) * ((l / uplus + 2.0 * (a - l)) * phiu0 - (l * umi + 2.0 * (a - l)) * phiu1);
                lambdau = tu1 / sqrt(-qs * ee);
            }
        } else if (u0 <= u2) {
            /* Table 1 row 7 (geokerr_original.f:2173-2234) */
            if (firstpt && u0 != u2) {
                p[4] = -2;
                v->tu03 = orc_ellquarticreal(p, -u1, one, u2, -one, u3, -one, u4, -one, one, -one / uplus, &v->rffu0, u0, u2);
                v->tu04 = orc_ellquarticreal(p, -u1, one, u2, -one, u3, -one, u4, -one, one, -umi, &v->rffu0, u0, u2);
                v->tu02 = orc_ellquarticreal(p, -u1, one, u2, -one, u3, -one, u4, -one, 0.0, one, &v->rffu0, u0, u2);
                p[4] = -4;
                v->tu01 = orc_ellquarticreal(p, -u1, one, u2, -one, u3, -one, u4, -one, 0.0, one, &v->rffu0, u0, u2);
            } else if (u0 == u2) {
                v->tu01 = v->tu02 = v->tu03 = v->tu04 = 0.0;
            }
            phiu01 = v->tu03;
            phiu02 = v->tu04;
            if (fabs(uf - u2) > 1e-12) {
                p[4] = -2;
                phiu11 = orc_ellquarticreal(p, -u1, one, u2, -one, u3, -one, u4, -one, one, -one / uplus, &v->rffu1, uf, u2);
                phiu12 = orc_ellquarticreal(p, -u1, one, u2, -one, u3, -one, u4, -one, one, -umi, &v->rffu1, uf, u2);
                tu12 = orc_ellquarticreal(p, -u1, one, u2, -one, u3, -one, u4, -one, 0.0, one, &v->rffu1, uf, u2);
                p[4] = -4;
                tu11 = orc_ellquarticreal(p, -u1, one, u2, -one, u3, -one, u4, -one, 0.0, one, &v->rffu1, uf, u2);
            } else {
                tu11 = tu12 = phiu11 = phiu12 = 0.0;
            }
            tu13 = phiu11;
            tu14 = phiu12;
            double tu0 = eq47(a, l, uplus, umi, v->tu01, v->tu02, v->tu03, v->tu04);
            double tuf = eq47(a, l, uplus, umi, tu11, tu12, tu13, tu14);
            tu = su * (tu0 - sgn_tpr * tuf) / sqrt(fabs(ee)) * ur;
            phiu0 = eq48(a, l, uplus, umi, phiu01, phiu02);
            phiu1 = eq48(a, l, uplus, umi, phiu11, phiu12);
            phiu = su * (phiu0 - sgn_tpr * phiu1) / sqrt(fabs(ee)) * ur;
            lambdau = su * (v->tu01 - sgn_tpr * tu11) / sqrt(fabs(ee));
        } else if (u0 >= u3) {
            /* Table 1 row 8 (geokerr_original.f:2235-2296) */
            if (firstpt && u0 != u3) {
                p[4] = -2;
                v->tu03 = -orc_ellquarticreal(p, -u1, one, -u2, one, -u3, one, u4, -one, one, -one / uplus, &v->rffu0, u3, u0);
                v->tu04 = -orc_ellquarticreal(p, -u1, one, -u2, one, -u3, one, u4, -one, one, -umi, &v->rffu0, u3, u0);
                v->tu02 = -orc_ellquarticreal(p, -u1, one, -u2, one, -u3, one, u4, -one, 0.0, one, &v->rffu0, u3, u0);
                p[4] = -4;
                v->tu01 = -orc_ellquarticreal(p, -u1, one, -u2, one, -u3, one, u4, -one, 0.0, one, &v->rffu0, u3, u0);
            } else if (u0 == u3) {
                v->tu01 = v->tu02 = v->tu03 = v->tu04 = 0.0;
            }
            phiu01 = v->tu03;
            phiu02 = v->tu04;
            if (uf != u3) {
                p[4] = -2;
                phiu11 = -orc_ellquarticreal(p, -u1, one, -u2, one, -u3, one, u4, -one, one, -one / uplus, &v->rffu1, u3, uf);
                phiu12 = -orc_ellquarticreal(p, -u1, one, -u2, one, -u3, one, u4, -one, one, -umi, &v->rffu1, u3, uf);
                tu12 = -orc_ellquarticreal(p, -u1, one, -u2, one, -u3, one, u4, -one, 0.0, one, &v->rffu1, u3, uf);
                p[4] = -4;
                tu11 = -orc_ellquarticreal(p, -u1, one, -u2, one, -u3, one, u4, -one, 0.0, one, &v->rffu1, u3, uf);
            } else {
                tu11 = tu12 = phiu11 = phiu12 = 0.0;
            }
            tu13 = phiu11;
            tu14 = phiu12;
            double tu0 = eq47(a, l, uplus, umi, v->tu01, v->tu02, v->tu03, v->tu04);
            double tuf = eq47(a, l, uplus, umi, tu11, tu12, tu13, tu14);
            tu = su * (tu0 - sgn_tpr * tuf) / sqrt(fabs(ee)) * ur;
            phiu0 = eq48(a, l, uplus, umi, phiu01, phiu02);
            phiu1 = eq48(a, l, uplus, umi, phiu11, phiu12);
            phiu = su * (phiu0 - sgn_tpr * phiu1) / sqrt(fabs(ee)) * ur;
            lambdau = su * (v->tu01 - sgn_tpr * tu11) / sqrt(fabs(ee));
        }
    }

    if (ncase > 4) {
        /* mu part from the roots of M(mu) (geokerr_original.f:2297-2349) */
        double amq = a * a - ql2;
        double yy = -0.5 * (amq + sgn1(amq) * sqrt(amq * amq + 4.0 * q2 * a * a));
        double mneg, mpos;
        if (amq < 0.0) {
            mneg = -yy / a / a;
            mpos = q2 / yy;
        } else {
            mneg = q2 / yy;
            mpos = -yy / a / a;
        }
        if (mpos > 1.0) mpos = 1.0;
        double muplus = sqrt(mpos);
        double c1 = sm;
        double c2 = sm * neg1pow((unsigned)tpm + 1u);
        double c3 = 2.0 * f_int((2.0 * (double)tpm - sm + 1.0) / 4.0);
        double phimu2 = 0.0, tmu2 = 0.0;
        if (mneg < 0.0) {
            if (muplus < mu0) {
                muplus = mu0;
                mpos = muplus * muplus;
            }
            calcphitmu(1, a, mneg, mpos, mu0, muf, muplus, &v->phimu1, &phimu2, &v->phimu3, &v->tmu1, &tmu2,
                       &v->tmu3, v->rffmu1, v->rffmu2, v->rffmu3, &v->rdc, &v->rjc, firstpt);
            tmu = a * a * mneg * iu + (c1 * v->tmu1 + c2 * tmu2 + c3 * v->tmu3);
            phimu = -l * iu + l * (c1 * v->phimu1 + c2 * phimu2 + c3 * v->phimu3);
        } else {
            if (sgn1(mu0) == -1.0) muplus = -muplus;
            if (fabs(muplus) < fabs(mu0)) {
                muplus = mu0;
                mpos = muplus * muplus;
            }
            if (fabs(mneg) > mu0 * mu0) mneg = mu0 * mu0;
            calcphitmu(0, a, mneg, mpos, mu0, muf, muplus, &v->phimu1, &phimu2, &v->phimu3, &v->tmu1, &tmu2,
                       &v->tmu3, v->rffmu1, v->rffmu2, v->rffmu3, &v->rdc, &v->rjc, firstpt);
            tmu = c1 * v->tmu1 + c2 * tmu2 + c3 * v->tmu3;
            phimu = -l * iu + l * (c1 * v->phimu1 + c2 * phimu2 + c3 * v->phimu3);
        }
    }
    *lambda_o = lambdau + tmu;
    if (l == 0.0) phiu = phiu - sgn1(a) * pi * tpm;
    *phimu_o = phimu;
    *tmu_o = tmu;
    *phiu_o = phiu;
    *tu_o = tu;
}

/* ---------------- GEOKERR driver ---------------- */

int orc_gk_geokerr(double* u0p, double* ufp, double uout, double mu0, double a, double l, double q2,
                   int32_t tpm, int32_t tpr, double su, double sm, int nup, double offset,
                   int uf_is_u0, double* ufi, double* mufi, double* dti, double* dphi,
                   int32_t* tpmi, int32_t* tpri, double* lami, int* nup_out) {
    orc_gk_vars v;
    memset(&v, 0, sizeof v);
    double u0_store = *u0p, uf_store = *ufp;
    double* U0 = &u0_store;
    double* UF = uf_is_u0 ? &u0_store : &uf_store;
    const double one = 1.0, two = 2.0;
    double uplus = one / (one + sqrt(one - a * a));
    double muf = 0.0, iu = 0.0;
    int written = 0;

    /* First solve on (u0, uf): roots, case, cached u0 / mu0 pieces. */
    orc_gk_geomu(&v, U0, UF, mu0, a, l, q2, &tpm, &tpr, su, sm, 1, 1, &muf, &iu, NULL);
    int ncase = v.ncase;
    if ((ncase > 2 && ncase < 7) || (ncase == 1 && su < 0.0) || (ncase == 7 && su < 0.0) ||
        (ncase == 2 && su > 0.0) || (ncase == 8 && su > 0.0))
        tpr = 0;
    double ub;
    if (ncase > 2 && ncase < 7) {
        ub = (su > 0.0) ? uplus : 0.0;
    } else if (ncase == 1 || ncase == 7) {
        ub = v.u2;
        if (tpr == 1 && *UF == ub) *UF = uout;
    } else {
        ub = v.u3;
        if (tpr == 1 && *UF == ub) *UF = uout;
    }
    int32_t tpr1 = 0;
    double du = sgn1(ub - uout) * su * ((ub - uout) + (two * tpr - one) * (ub - *UF));
    du = du / (double)nup;
    int kmax = 0;
    if (du != 0.0) {
        /* MIN(INT((UB-UOUT)/DU+OFFSET), NUP), kept inside int range */
        double xk = (ub - uout) / du + offset;
        if (!(xk < (double)nup)) kmax = nup;
        else if (xk < -1.0) kmax = -1;
        else kmax = f_int(xk);
    }
    if (sgn1(uout - ub) != sgn1(uout - *U0) || uout == *U0) {
        int32_t tpm_k = tpm;
        double mun = 0.0, phimu, tmu, phiu, tu, lambda;
        /* Samples 1..KMAX from u0 toward the turning point or uf; the first
         * one fills the u0 pieces of GEOPHITIME (FIRSTPT). */
        for (int k = 1; k <= kmax; ++k) {
            double un = uout + ((double)k - offset) * du;
            orc_gk_geomu(&v, U0, &un, mu0, a, l, q2, &tpm_k, &tpr1, su, sm, 1, 0, &mun, &iu, NULL);
            orc_gk_geophitime(&v, *U0, un, mu0, mun, a, l, q2, tpm_k, tpr1, su, sm, iu, k == 1, &phimu, &tmu,
                              &phiu, &tu, &lambda);
            ufi[k - 1] = un;
            mufi[k - 1] = mun;
            dti[k - 1] = tmu + tu;
            dphi[k - 1] = phimu + phiu;
            tpmi[k - 1] = tpm_k;
            tpri[k - 1] = 0;
            lami[k - 1] = lambda;
            written = k;
        }
        if (tpr == 1) {
            /* MUFILL with NPTS = 0 makes one GEOMU call whose results nothing
             * reads, so it is not repeated here.  Samples KMAX+1..NUP come
             * back from the turning point: u = 2 ub - (uout + (k - offset) du),
             * with UF in GEOMU's u0 slot and GEOKERR's cached u0 pieces. */
            int firstpt2 = (nup == 1);
            for (int k = (kmax > 0 ? kmax : 0) + 1; k <= nup; ++k) {
                double un = two * ub - (uout + ((double)k - offset) * du);
                orc_gk_geomu(&v, UF, &un, mu0, a, l, q2, &tpm_k, &tpr, su, sm, 1, 0, &mun, &iu, NULL);
                orc_gk_geophitime(&v, *UF, un, mu0, mun, a, l, q2, tpm_k, tpr, su, sm, iu, firstpt2, &phimu,
                                  &tmu, &phiu, &tu, &lambda);
                ufi[k - 1] = un;
                mufi[k - 1] = mun;
                dti[k - 1] = tmu + tu;
                dphi[k - 1] = phimu + phiu;
                tpmi[k - 1] = tpm_k;
                tpri[k - 1] = 1;
                lami[k - 1] = lambda;
                written = k;
            }
        }
    } else {
        nup = 0;
    }
    if (nup_out) *nup_out = nup;
    *u0p = *U0;
    *ufp = *UF;
    return written;
}

/* ---------------- geokerr_ref_api ---------------- */

/* geokerr_eval_lambda (geokerr_eval_lambda.f90:62-198).  When the driver
 * writes fewer than NUP samples (the reference then reads unset memory),
 * the resampling here uses only the samples written and NaN for a grid
 * that has none. */
void orc_geokerr_eval_lambda(double u0, double mu0, double uf, double a, double lz, double q2,
                             double t0, double lam0, double dlam, int s, double* u, double* mu,
                             double* t, double* phi, double* affine, int* tpmi_o, int* tpri_o,
                             int* info) {
    *info = 0;
    if (s <= 0) {
        *info = 1;
        return;
    }
    int nup = s;
    double uout = u0;
    double su = ((uf - u0) >= 0.0) ? 1.0 : -1.0;
    double sm = (mu0 <= 0.0) ? 1.0 : -1.0;
    double* ufi = (double*)malloc(sizeof(double) * (size_t)nup * 5);
    int32_t* tpi = (int32_t*)malloc(sizeof(int32_t) * (size_t)nup * 2);
    double *mufi = ufi + nup, *dti = ufi + 2 * (size_t)nup, *dph = ufi + 3 * (size_t)nup,
           *lami = ufi + 4 * (size_t)nup;
    int32_t *tpmi = tpi, *tpri = tpi + nup;
    int nup_after = nup;
    int n = orc_gk_geokerr(&u0, &uf, uout, mu0, a, lz, q2, 0, 1, su, sm, nup, 0.5, 0, ufi, mufi, dti, dph,
                           tpmi, tpri, lami, &nup_after);
    if (nup_after < n) n = nup_after;
    double lam_base = (n > 0) ? lami[0] : 0.0;
    int j = 0; /* 0-based */
    for (int i = 0; i < s; ++i) {
        double lamk = lam0 + dlam * (double)i;
        affine[i] = lamk;
        if (n <= 0) {
            u[i] = mu[i] = t[i] = phi[i] = NAN;
        } else {
            while (j < n - 1 && (lami[j + 1] - lam_base) < lamk - 1.0e-30) ++j;
            if (lamk <= (lami[0] - lam_base)) {
                u[i] = ufi[0]; mu[i] = mufi[0]; t[i] = t0 + dti[0]; phi[i] = dph[0];
            } else if (lamk >= (lami[n - 1] - lam_base)) {
                u[i] = ufi[n - 1]; mu[i] = mufi[n - 1]; t[i] = t0 + dti[n - 1]; phi[i] = dph[n - 1];
            } else {
                int j1 = (j + 1 < n) ? j + 1 : n - 1;
                double x0 = lami[j] - lam_base, x1 = lami[j1] - lam_base;
                double w = (x1 <= x0 + 1.0e-30) ? 0.0 : (lamk - x0) / (x1 - x0);
                u[i] = (1.0 - w) * ufi[j] + w * ufi[j1];
                mu[i] = (1.0 - w) * mufi[j] + w * mufi[j1];
                t[i] = t0 + ((1.0 - w) * dti[j] + w * dti[j1]);
                phi[i] = (1.0 - w) * dph[j] + w * dph[j1];
            }
        }
        if (i == 0) {
            u[i] = u0;
            mu[i] = mu0;
        }
    }
    int tp_mu = -1, tp_r = -1;
    for (int jj = 0; jj < n; ++jj) {
        double lj = lami[jj] - lam_base;
        if (tp_mu < 0 && tpmi[jj] > 0) {
            for (int i = 0; i < s; ++i)
                if (affine[i] >= lj) { tp_mu = i; break; }
        }
        if (tp_r < 0 && tpri[jj] > 0) {
            for (int i = 0; i < s; ++i)
                if (affine[i] >= lj) { tp_r = i; break; }
        }
        if (tp_mu >= 0 && tp_r >= 0) break;
    }
    *tpmi_o = tp_mu;
    *tpri_o = tp_r;
    free(ufi);
    free(tpi);
}

/* geokerr_estimate_periods (geokerr_eval_lambda.f90:200-295): GEOKERR with
 * NUP = 4096 and one variable passed as both U0 and UF.  The driver only
 * ever writes TPRI = 0 or 1, so Tr keeps its -1 sentinel. */
void orc_geokerr_estimate_periods(double u0, double mu0, double a, double lz, double q2,
                                  double* tr, double* ttheta, int* info) {
    const int nup = 4096;
    double ufi[4096], mufi[4096], dti[4096], dph[4096], lami[4096];
    int32_t tpmi[4096], tpri[4096];
    double sm = (mu0 <= 0.0) ? 1.0 : -1.0;
    double u0v = u0, ufv = u0;
    int nup_after = nup;
    int n = orc_gk_geokerr(&u0v, &ufv, u0, mu0, a, lz, q2, 0, 1, 1.0, sm, nup, 0.5, 1, ufi, mufi, dti, dph,
                           tpmi, tpri, lami, &nup_after);
    if (nup_after < n) n = nup_after;
    double lam_base = (n > 0) ? lami[0] : 0.0;
    int im1 = -1, im2 = -1, ir1 = -1, ir2 = -1, prev_mu = 0, prev_r = 0;
    for (int j = 0; j < n; ++j) {
        if (im1 < 0 && prev_mu < 1 && tpmi[j] >= 1) im1 = j;
        if (im2 < 0 && prev_mu < 2 && tpmi[j] >= 2) im2 = j;
        if (ir1 < 0 && prev_r < 1 && tpri[j] >= 1) ir1 = j;
        if (ir2 < 0 && prev_r < 2 && tpri[j] >= 2) ir2 = j;
        if (tpmi[j] > prev_mu) prev_mu = tpmi[j];
        if (tpri[j] > prev_r) prev_r = tpri[j];
        if (im2 >= 0 && ir2 >= 0) break;
    }
    *tr = (ir1 >= 0 && ir2 >= 0) ? 2.0 * fmax(1.0e-18, (lami[ir2] - lam_base) - (lami[ir1] - lam_base)) : -1.0;
    *ttheta = (im1 >= 0 && im2 >= 0) ? 2.0 * fmax(1.0e-18, (lami[im2] - lam_base) - (lami[im1] - lam_base))
                                     : -1.0;
    *info = 0;
}

/* geokerr_phi_omega (geokerr_eval_lambda.f90:297-306): dphi/dlambda at
 * (1/u0, mu0) from dphi_dlam (:41-51). */
double orc_geokerr_phi_omega(double u0, double mu0, double a, double lz, double q2) {
    double r = fmax(1.0e-12, 1.0 / fmax(1.0e-12, u0));
    double delta = fmax(1.0e-30, r * r - 2.0 * r + a * a);
    double sin2 = fmax(1.0e-30, 1.0 - mu0 * mu0);
    return a * (((r * r + a * a) - a * lz) / delta) + lz / sin2 - a;
}

void orc_geokerr_eval_lambda_batch(const double* u0, const double* mu0, const double* uf,
                                   const double* a, const double* lz, const double* q2,
                                   const double* t0, double lam0, double dlam, int s, int n,
                                   double* u, double* mu, double* t, double* phi, double* affine,
                                   int* tpmi, int* tpri, int* info, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int i = 0; i < n; ++i) {
        size_t o = (size_t)i * (size_t)s;
        orc_geokerr_eval_lambda(u0[i], mu0[i], uf[i], a[i], lz[i], q2[i], t0[i], lam0, dlam, s, u + o, mu + o,
                                t + o, phi + o, affine + o, tpmi + i, tpri + i, info + i);
    }
}

/* ---------------- the whole GEOKERR driver ---------------- */

/* FINDMROOTS (geokerr_original.f:607-642): the mu turning points.  With
 * a = 0 the Fortran divides by zero here; the IEEE results are kept. */
void orc_findmroots(double q2, double l, double a, double mu0, double* muminus, double* muplus) {
    double ql2 = q2 + l * l;
    double amq = a * a - ql2;
    double yy = -0.5 * (amq + sgn1(amq) * sqrt(amq * amq + 4.0 * q2 * a * a));
    double mneg, mpos;
    if (amq < 0.0) { mneg = -yy / a / a; mpos = q2 / yy; }
    else { mneg = q2 / yy; mpos = -yy / a / a; }
    if (mpos > 1.0) mpos = 1.0;
    if (mneg > 0.0) {
        *muplus = sgn1(mu0) * sqrt(mpos);
        *muminus = sgn1(mu0) * sqrt(mneg);
        if (fabs(*muminus) > fabs(mu0)) *muminus = mu0;
        if (fabs(mu0) > fabs(*muplus)) *muplus = mu0;
    } else {
        *muplus = sqrt(mpos);
        if (*muplus < mu0) *muplus = mu0;
        *muminus = -*muplus;
    }
}

/* (+-1) ** n for the Fortran's integer and default-real powers of -1. */
static int m1pow(int n) { return (n & 1) ? -1 : 1; }

/* INDEP_MUF (geokerr_original.f:645-706): the k-th of npts mu values spaced
 * evenly in path length sum |dmu| from mu0 (after tpm0 turning points) to
 * muf (after tpmf), and the turning points before it.  The Fortran mixes
 * default-real (single precision) constants into integer expressions; for
 * the small integers involved they are exact, and the float casts below
 * keep the same operations. */
void orc_indep_muf(double muminus, double muplus, double mu0, double muf, int32_t tpm0, int32_t tpmf,
                   double sm, int32_t k, int32_t npts, double offset, double* mun, int32_t* tpmk_out) {
    int tpmk;
    if (muf <= muplus && muf >= muminus) {
        int dtpm = tpmf - tpm0;
        int a1 = f_int(sm * (double)m1pow(tpm0));
        int a2 = f_int(sm * (double)m1pow(tpmf));
        int a3 = 2 * (int)floorf((2.0f * (float)dtpm + 3.0f - (float)a1) / 4.0f) - 1;
        double dmu = a1 * (muplus - mu0) + a2 * (muf - muminus) + a3 * (muplus - muminus);
        double mu1 = (double)(((float)a1 + 1.0f) / 2.0f) * muplus + (double)((1.0f - (float)a1) / 2.0f) * muminus;
        double mu2 = (double)((1.0f - (float)a1) / 2.0f) * muplus + (double)(((float)a1 + 1.0f) / 2.0f) * muminus;
        double dmu1 = fabs(mu1 - mu0);
        double dmu2 = fabs(mu2 - mu1);
        double dmuk = ((double)k - offset) * dmu / (double)(float)npts;
        int dmukdmu1 = f_int(dmuk / dmu1);
        if (dmukdmu1 < 0) dmukdmu1 = 1;
        int tp2 = f_int((dmuk - dmu1) / dmu2);
        tpmk = (dmukdmu1 < 1 ? dmukdmu1 : 1) + (tp2 > 0 ? tp2 : 0);
        int a1k = f_int(sm * (double)m1pow(tpm0));
        int a2k = f_int(sm * (double)m1pow(tpmk + tpm0));
        int a3k = (int)(2.0f * (float)(int)((2.0f * (float)tpmk + 3.0f - (float)a1k) / 4.0f) - 1.0f);
        *mun = muminus + (double)(1.0f / (float)a2k) * (dmuk - a1k * (muplus - mu0) - a3k * (muplus - muminus));
    } else {
        *mun = 0.0;
        tpmk = 0;
    }
    *tpmk_out = tpmk + tpm0;
}

void orc_gk_geokerr_full(const orc_gk_opts* op, orc_gk_args* g, double* ufi, double* mufi, double* dti,
                         double* dphi, int32_t* tpmi, int32_t* tpri, double* lami, int32_t* ncase_out,
                         int32_t* nup_out, int32_t* npts_out) {
    const double one = 1.0, two = 2.0;
    const int total = op->nup + op->npts - 2 * op->kext;
    for (int i = 0; i < total; ++i) {
        ufi[i] = mufi[i] = dti[i] = dphi[i] = lami[i] = NAN;
        tpmi[i] = tpri[i] = -1;
    }
    orc_gk_vars v;
    memset(&v, 0, sizeof v);
    int nup = op->nup, npts = op->npts;
    const int kext = op->kext, phit = op->phit;
    const double offset = op->offset, uout = g->uout, mu0 = g->mu0, a = g->a, l = g->l, q2 = g->q2;
    const double su = g->su, sm = g->sm;
    double U0 = g->u0, UF = g->uf, MUF = g->muf;
    int32_t TPM = g->tpm, TPR = g->tpr;
    const double uplus = one / (one + sqrt(one - a * a));
    /* GEOKERR's locals IU, TMU, TU, PHIMU, PHIU, LAMBDA start unset in the
     * Fortran (read before any write when PHIT is false); 0 here. */
    double IU = 0.0, TMU = 0.0, TU = 0.0, PHIMU = 0.0, PHIU = 0.0, LAMBDA = 0.0;
    double MUN = 0.0, UN = 0.0;
#define GK_STORE(idx1, u_, mu_, tpm_, tpr_)                    \
    do {                                                        \
        const int64_t j_ = (int64_t)(idx1)-1;                   \
        if (j_ >= 0 && j_ < (int64_t)total) {                   \
            ufi[j_] = (u_);                                     \
            mufi[j_] = (mu_);                                   \
            dti[j_] = TMU + TU;                                 \
            dphi[j_] = PHIMU + PHIU;                            \
            tpmi[j_] = (tpm_);                                  \
            tpri[j_] = (tpr_);                                  \
            lami[j_] = LAMBDA;                                  \
        }                                                       \
    } while (0)

    if (op->usegeor) {
        /* samples evenly in mu between mu0 and MUF (:421-467) */
        for (int k = 1; k <= nup; ++k) {
            MUN = mu0 + ((double)k - offset) * (MUF - mu0) / (double)nup;
            if (k >= 2) TPM = f_int((sgn1(mu0) * sm + one) / two);
            if (mu0 != 0.0 || g->beta != 0.0) {
                orc_gk_geor(&v, &U0, &UF, mu0, MUN, a, l, q2, &IU, TPM, &TPR, su, sm, phit, k == 1, NULL);
                if (phit)
                    orc_gk_geophitime(&v, U0, UF, mu0, MUN, a, l, q2, TPM, TPR, su, sm, IU, 1, &PHIMU, &TMU, &PHIU,
                                      &TU, &LAMBDA);
            } else {
                UF = U0;
            }
            GK_STORE(k, UF, MUN, TPM, TPR);
        }
    } else {
        /* first solve on (U0, UF): case, roots, cached u0 / mu0 pieces */
        orc_gk_geomu(&v, &U0, &UF, mu0, a, l, q2, &TPM, &TPR, su, sm, phit, 1, &MUF, &IU, NULL);
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
        int32_t TPR1 = 0;
        double du = sgn1(ub - uout) * su * ((ub - uout) + (two * TPR - one) * (ub - UF));
        du = du / (double)nup;
        int kmax = f_int((ub - uout) / du + offset);
        if (kmax > nup) kmax = nup;
        if (du == 0.0) kmax = 0;
        if (sgn1(uout - ub) != sgn1(uout - U0) || uout == U0) {
            /* samples 1..KMAX from UOUT toward the turning point or UF; the
             * first fills GEOPHITIME's u0 pieces (:495-537) */
            if (kmax != 0) {
                UN = uout + (one - offset) * du;
                orc_gk_geomu(&v, &U0, &UN, mu0, a, l, q2, &TPM, &TPR1, su, sm, phit, 0, &MUN, &IU, NULL);
                if (phit)
                    orc_gk_geophitime(&v, U0, UN, mu0, MUN, a, l, q2, TPM, TPR1, su, sm, IU, 1, &PHIMU, &TMU, &PHIU,
                                      &TU, &LAMBDA);
                GK_STORE(1, UN, MUN, TPM, 0);
            }
            for (int k = 2; k <= kmax; ++k) {
                UN = uout + ((double)k - offset) * du;
                orc_gk_geomu(&v, &U0, &UN, mu0, a, l, q2, &TPM, &TPR1, su, sm, phit, 0, &MUN, &IU, NULL);
                if (phit)
                    orc_gk_geophitime(&v, U0, UN, mu0, MUN, a, l, q2, TPM, TPR1, su, sm, IU, 0, &PHIMU, &TMU, &PHIU,
                                      &TU, &LAMBDA);
                GK_STORE(k, UN, MUN, TPM, 0);
            }
            if (TPR == 1) {
                const int firstpt = (nup == 1);
                if (op->mufill) {
                    /* mu as the independent variable across the u turning
                     * point (:539-575): MUF and TPM1 on the far side, then
                     * npts GEOR solves at INDEP_MUF's mu values.  The
                     * Fortran reads MUFI(KMAX-KEXT) and TPMI(KMAX-KEXT); an
                     * index before the array (KMAX <= KEXT) is taken as the
                     * start of the geodesic (mu0, 0 turning points).  TPM1
                     * is unset in the Fortran when GEOMU leaves TPM alone
                     * (q2 = 0); it starts from TPM here. */
                    int32_t TPM1 = TPM;
                    UN = two * ub - (uout + ((double)((int64_t)kmax + kext + 1) - offset) * du);
                    orc_gk_geomu(&v, &UF, &UN, mu0, a, l, q2, &TPM1, &TPR, su, sm, phit, 0, &MUN, &IU, NULL);
                    MUF = MUN;
                    double muminus, muplus;
                    orc_findmroots(q2, l, a, mu0, &muminus, &muplus);
                    const int64_t jb = (int64_t)kmax - kext;   /* 1-based */
                    const double mustart = (jb >= 1 && jb <= total) ? mufi[jb - 1] : mu0;
                    const int32_t tpmstart = (jb >= 1 && jb <= total) ? tpmi[jb - 1] : 0;
                    for (int k = 1; k <= npts; ++k) {
                        orc_indep_muf(muminus, muplus, mustart, MUF, tpmstart, TPM1, sm, k, npts, offset, &MUN, &TPM);
                        orc_gk_geor(&v, &U0, &UN, mu0, MUN, a, l, q2, &IU, TPM, &TPR, su, sm, phit, 0, NULL);
                        if (phit)
                            orc_gk_geophitime(&v, U0, UN, mu0, MUN, a, l, q2, TPM, TPR, su, sm, IU, 0, &PHIMU, &TMU,
                                              &PHIU, &TU, &LAMBDA);
                        GK_STORE((int64_t)k + kmax - kext, UN, MUN, TPM, TPR);
                    }
                    npts = npts - 2 * kext;
                }
                /* back from the turning point to UF (:576-597); samples the
                 * Fortran would store before its arrays are skipped */
                int64_t kfirst = (int64_t)kmax + 1 + kext;
                if (kfirst < 1 - (int64_t)npts) kfirst = 1 - (int64_t)npts;
                for (int64_t k = kfirst; k <= nup; ++k) {
                    UN = two * ub - (uout + ((double)k - offset) * du);
                    orc_gk_geomu(&v, &UF, &UN, mu0, a, l, q2, &TPM, &TPR, su, sm, phit, 0, &MUN, &IU, NULL);
                    if (phit)
                        orc_gk_geophitime(&v, UF, UN, mu0, MUN, a, l, q2, TPM, TPR, su, sm, IU, firstpt, &PHIMU, &TMU,
                                          &PHIU, &TU, &LAMBDA);
                    GK_STORE(k + npts, UN, MUN, TPM, 1);
                }
                nup = nup + npts;
            }
        } else {
            nup = 0;
        }
    }
#undef GK_STORE
    g->u0 = U0;
    g->uf = UF;
    g->muf = MUF;
    g->tpm = TPM;
    g->tpr = TPR;
    *ncase_out = v.ncase;
    *nup_out = nup;
    *npts_out = npts;
}

void orc_gk_geokerr_full_batch(const orc_gk_opts* o, orc_gk_args* g, int n, double* ufi, double* mufi,
                               double* dti, double* dphi, int32_t* tpmi, int32_t* tpri, double* lami,
                               int32_t* ncase, int32_t* nup_out, int32_t* npts_out, int nthreads) {
    const size_t total = (size_t)(o->nup + o->npts - 2 * o->kext);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int i = 0; i < n; ++i) {
        const size_t s = (size_t)i * total;
        orc_gk_geokerr_full(o, g + i, ufi + s, mufi + s, dti + s, dphi + s, tpmi + s, tpri + s, lami + s, ncase + i,
                            nup_out + i, npts_out + i);
    }
    (void)nthreads;
}
