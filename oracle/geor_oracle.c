/*
 * geor_oracle.c -- FP64 restatement of GEOKERR's GEOR (geokerr_original.f:
 * 2672-3134): the final inverse radius uf reached when the polar coordinate
 * arrives at muf, after tpm mu turning points, from (u0, mu0) with
 * constants of motion (a, L, Q2) and initial directions su, sm (M = E = 1).
 * It computes the mu integral IMU first and inverts the u integral with the
 * Jacobi elliptic functions of Dexter & Agol (2009) Table 2.
 *
 * orc_gk_geor keeps the Fortran's argument contract: the variables shared
 * with GEOMU/GEOPHITIME along one geodesic (orc_gk_vars) are written where
 * FIRSTPT asks for them and read otherwise, PHT forms the RF pieces
 * GEOPHITIME takes (RFFU0 in rows 3/5/6, RFFU1 in rows 1/2/7/8, RFFMU2
 * always), U0, UF and TPR are by reference, and IMU is written into the
 * caller's IU.  orc_geor is the fresh call (FIRSTPT on, PHT off).  Quirks
 * kept as the Fortran has them:
 *   - the equal-root cubic branch compares against the UF the caller
 *     passed in (:2836, :2847);
 *   - TPR is written only by Table-2 rows 1, 2, 7, 8 and the cases that
 *     set it to 0; where the Fortran leaves it, it stays as passed;
 *   - row 2 with FIRSTPT off sets IU0 = 0 whenever u0 != u3 (:2795-2799:
 *     the ELSEIF tests u0 .ne. u3 where row 1 tests u0 .eq. u2);
 *   - the asymmetric-roots exit for |muf| < mu_- (:2924-2929) leaves IMU
 *     (the caller's IU) as it was;
 *   - an initial u inside a forbidden region moves to u3 (su = 1) or u2 and
 *     the case test runs again (:2807-2814, :2972-2980); *warned reports it.
 *     The Fortran retries without bound, which for Q2 < 0 (U(u) with a
 *     positive quartic coefficient) can cycle forever; here, as in the
 *     device code, the test runs at most 3 times and the case stays 0;
 *   - rows 3, 5, 6 return uf = -1, ncase = 0 when IMU exceeds the u
 *     integral to the horizon (su = 1) or to infinity (su = -1);
 *   - the sextic root loops for h1 and mn2 take the first real, nonzero
 *     root; the Fortran's loop has no bound, this one stops at 6 (NaN if
 *     none).
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
static double neg1pow(unsigned n) { return (n & 1u) ? -1.0 : 1.0; }
/* Fortran INT() as the reference's x86 build executes it (cvttsd2si):
 * truncation, INT_MIN for NaN and out-of-range values. */
static int f_int(double x) { return (x > -2147483649.0 && x < 2147483648.0) ? (int)x : (int)0x80000000u; }

/* calcimuasymf (geokerr_original.f:3175-3204): the mu integral from muf to
 * the turning point muplus in the asymmetric-roots case; imum is unused. */
double orc_calcimuasymf(double a, double mneg, double mpos, double muf, double muplus, double* rff) {
    double fac = fabs(a) * muplus;
    double m1 = mneg / mpos;
    double sf = sqrt((mpos - muf * muf) / (mpos - mneg));
    double ak = sqrt(1.0 - m1);
    double qf = (1.0 - sf * ak) * (1.0 + sf * ak);
    *rff = orc_rf(1.0 - sf * sf, qf, 1.0);
    return sf * (*rff) / fac;
}

/* calcimusymf (geokerr_original.f:3244-3274): symmetric roots; imum is the
 * half-period integral from calcimusym (already divided by fac). */
double orc_calcimusymf(double a, double mneg, double mpos, double muf, double muplus, double imum,
                       double* rff) {
    double fac = fabs(a) * sqrt(mpos - mneg);
    double m1 = -mneg / (mpos - mneg);
    double phif = acos(muf / muplus);
    double sf = sin(phif);
    double ak = sqrt(1.0 - m1);
    double qf = (1.0 - sf * ak) * (1.0 + sf * ak);
    *rff = orc_rf(1.0 - sf * sf, qf, 1.0);
    double imuf = sf * (*rff);
    if (tan(phif) < 0.0) imuf = imum * fac - imuf;
    return imuf / fac;
}

void orc_gk_geor(orc_gk_vars* v, double* u0p, double* ufp, double mu0, double muf, double a, double l,
                 double q2, double* imup, int32_t tpm, int32_t* tprp, double su, double sm, int pht,
                 int firstpt, int32_t* warned) {
    const double one = 1.0, two = 2.0, half = 0.5, third = 0.3333333333333333, three = 3.0;
    double pi = acos(-one);
    double uplus = one / (one + sqrt(one - a * a));
    double l2 = l * l;
    double ql2 = q2 + l2;
    double cc = a * a - q2 - l2;
    double dd = two * ((a - l) * (a - l) + q2);
    double ee = -a * a * q2;
    double u0 = *u0p, uf = *ufp, imu = *imup, dummy = 0.0, iu_local = 0.0, iu1 = 0.0;
    double i2mu = 0.0, a1, a2, a3, muplus;
    double sn, cn, dn, m1, jarg;
    int32_t tpr = *tprp;
    int p[5];
    unsigned utpm = (unsigned)tpm;
    (void)iu_local;
    (void)iu1;

#define GEOR_RETURN()        \
    do {                     \
        *u0p = u0;           \
        *ufp = uf;           \
        *imup = imu;         \
        *tprp = tpr;         \
        return;              \
    } while (0)

    /* the mu integral IMU in the special cases (:2749-2770) */
    if (q2 == 0.0) {
        if (l2 >= a * a || mu0 == 0.0) {
            imu = 0.0;
            uf = -1.0;
            v->ncase = 0;
            GEOR_RETURN();
        }
        double s1 = sgn1(mu0);
        a1 = s1 * sm;
        a2 = s1 * sm * neg1pow(utpm + 1u);
        muplus = s1 * sqrt(one - l2 / a / a);
        v->i1mu = one / fabs(a * muplus) * orc_asech(mu0 / muplus);
        i2mu = one / fabs(a * muplus) * orc_asech(muf / muplus);
        imu = a1 * v->i1mu + a2 * i2mu;
    } else if (a == 0.0) {
        a1 = sm;
        muplus = sqrt(q2 / ql2);
        if (mu0 > muplus) muplus = mu0;
        v->i1mu = (half * pi - asin(mu0 / muplus)) / sqrt(ql2);
        v->i3mu = pi / sqrt(ql2);
        a2 = sm * neg1pow(utpm);
        a3 = two * f_int((two * (double)tpm + 3.0 - sm) / 4.0) - one;
        i2mu = (half * pi + asin(muf / muplus)) / sqrt(ql2);
        imu = a1 * v->i1mu + a2 * i2mu + a3 * v->i3mu;
    }

    if (ee == 0.0 && dd != 0.0) {
        /* cubic U(u) (:2772-2895) */
        p[0] = -1; p[1] = -1; p[2] = -1; p[3] = 0; p[4] = 0;
        double qq = cc * cc / dd / dd / 9.0;
        double rr = (two * (cc * cc * cc) / (dd * dd * dd) + 27.0 / dd) / 54.0;
        double dis = rr * rr - qq * qq * qq;
        if (dis < -1e-16) {
            double theta = acos(rr / pow(qq, 1.5));
            double u1 = -two * sqrt(qq) * cos(theta / 3.0) - cc / dd / 3.0;
            double u2 = -two * sqrt(qq) * cos((theta - two * pi) / 3.0) - cc / dd / 3.0;
            double u3 = -two * sqrt(qq) * cos((theta + two * pi) / 3.0) - cc / dd / 3.0;
            v->u1 = u1; v->u2 = u2; v->u3 = u3;
            for (int guard = 0; guard < 3; ++guard) {
                if (u0 <= u2) {
                    v->ncase = 1;
                    if (firstpt && u0 != u2)
                        v->iu0 = orc_ellcubicreal(p, -u1, one, u2, -one, u3, -one, 0.0, 0.0, &v->rffu0, u0, u2);
                    else if (u0 == u2)
                        v->iu0 = 0.0;
                    /* Table 2 row 1 */
                    m1 = (u3 - u2) / (u3 - u1);
                    jarg = sqrt((u3 - u1) * dd) * half * (imu - su * v->iu0 / sqrt(dd));
                    orc_sncndn(jarg, m1, &sn, &cn, &dn);
                    double cd2 = cn * cn / (dn * dn);
                    uf = u1 + (u2 - u1) * cd2;
                    tpr = f_int((sgn1(su * (imu - su * v->iu0 / sqrt(dd))) + one) / two);
                    if (pht) iu1 = orc_ellcubicreal(p, -u1, one, u2, -one, u3, -one, 0.0, 0.0, &v->rffu1, uf, u2);
                    break;
                } else if (u0 >= u3) {
                    v->ncase = 2;
                    if (firstpt && u0 != u3)
                        v->iu0 = orc_ellcubicreal(p, -u1, one, -u2, one, -u3, one, 0.0, 0.0, &v->rffu0, u3, u0);
                    else if (u0 != u3)
                        v->iu0 = 0.0;
                    /* Table 2 row 2 */
                    m1 = (u3 - u2) / (u3 - u1);
                    jarg = sqrt((u3 - u1) * dd) * half * (imu + su * v->iu0 / sqrt(dd));
                    orc_sncndn(jarg, m1, &sn, &cn, &dn);
                    double dc2 = dn * dn / (cn * cn);
                    uf = u1 + (u3 - u1) * dc2;
                    tpr = f_int((-sgn1(su * (imu + v->iu0 / sqrt(dd))) + one) / two);
                    if (pht) iu1 = orc_ellcubicreal(p, -u1, one, -u2, one, -u3, one, 0.0, 0.0, &v->rffu1, u3, uf);
                    break;
                } else {
                    if (warned) *warned = 1;
                    u0 = (su == 1.0) ? u3 : u2;
                }
            }
        } else if (fabs(dis) < 1e-16) {
            /* equal roots u2 = u3 (:2832-2860) */
            tpr = 0;
            double u1 = -two * sqrt(qq) - cc / dd / 3.0;
            double u2 = -two * sqrt(qq) * cos((two * pi) / 3.0) - cc / dd / 3.0;
            v->u1 = u1; v->u2 = u2; v->u3 = u2;
            if (u0 <= u2) {
                v->ncase = 1;
                if (uf > u2) {
                    uf = u2;
                } else {
                    double sarg = sqrt((u0 - u1) / (u2 - u1));
                    jarg = sqrt((u2 - u1) * dd) * half * su * imu + half * log((one + sarg) / (one - sarg));
                    double th = tanh(jarg);
                    uf = u1 + (u2 - u1) * th * th;
                }
            } else if (u0 >= u2) {
                v->ncase = 2;
                if (uf < u2) {
                    uf = u2;
                } else {
                    double sarg = sqrt((u2 - u1) / (u0 - u1));
                    jarg = -sqrt((u2 - u1) * dd) * half * su * imu + half * log((one + sarg) / (one - sarg));
                    double th = tanh(jarg);
                    uf = u1 + (u2 - u1) / (th * th);
                }
            }
        } else {
            /* one real root: Table 2 row 3 (:2861-2895) */
            v->ncase = 3;
            tpr = 0;
            double aa = -sgn1(rr) * pow(fabs(rr) + sqrt(dis), third);
            double bb = (aa != 0.0) ? qq / aa : 0.0;
            double u1 = (aa + bb) - cc / dd / 3.0;
            v->u1 = u1;
            double f = -one / dd / u1;
            double g = f / u1;
            double iut;
            if (su > 0.0)
                iut = orc_ellcubiccomplex(p, -u1, one, 0.0, 0.0, f, g, one, &dummy, u0, uplus) / sqrt(dd);
            else
                iut = orc_ellcubiccomplex(p, -u1, one, 0.0, 0.0, f, g, one, &dummy, 0.0, u0) / sqrt(dd);
            if (imu > iut) {
                uf = -1.0;
                v->ncase = 0;
                tpr = 0;
                GEOR_RETURN();
            }
            if (firstpt) v->iu0 = su * orc_ellcubiccomplex(p, -u1, one, 0.0, 0.0, f, g, one, &dummy, u1, u0) / sqrt(dd);
            double c3 = -one;
            if (a != 0.0 || l != 0.0) c3 = (a + l) / (a - l);
            double c2 = sqrt(u1 * (three * u1 + c3));
            double c1 = sqrt(c2 * dd);
            m1 = half + (6.0 * u1 + c3) / (8.0 * c2);
            jarg = c1 * (imu + v->iu0);
            orc_sncndn(jarg, m1, &sn, &cn, &dn);
            uf = (c2 + u1 - (c2 - u1) * cn) / (one + cn);
            if (pht) iu_local = orc_ellcubiccomplex(p, -u1, one, 0.0, 0.0, f, g, one, &v->rffu0, u0, uf) / sqrt(dd);
        }
    } else if (ee == 0.0 && dd == 0.0) {
        /* q2 = 0 and l = a: mu stays 0, nothing to invert (:2896-2899) */
        v->ncase = 4;
        uf = -1.0;
        tpr = 0;
    } else {
        /* roots of M(mu) and the mu integral, biquadratic case (:2901-2945) */
        double amq = a * a - ql2;
        double yy = -0.5 * (amq + sgn1(amq) * sqrt(amq * amq + 4.0 * q2 * a * a));
        double mneg, mpos;
        if (amq < 0.0) { mneg = -yy / a / a; mpos = q2 / yy; }
        else { mneg = q2 / yy; mpos = -yy / a / a; }
        muplus = sqrt(mpos);
        a1 = sm;
        a2 = sm * neg1pow(utpm + 1u);
        a3 = two * f_int((2.0 * (double)tpm - sm + 1.0) / 4.0);
        if (mneg < 0.0) {
            if (mu0 > muplus) { muplus = mu0; mpos = muplus * muplus; }
            if (firstpt) orc_calcimusym(a, mneg, mpos, mu0, muplus, &v->i1mu, &v->i3mu, &v->rffmu1, &v->rffmu3);
            i2mu = orc_calcimusymf(a, mneg, mpos, muf, muplus, v->i3mu, &v->rffmu2);
        } else {
            if (fabs(muf) < sqrt(mneg)) {
                uf = -1.0;
                v->ncase = 0;
                tpr = 0;
                GEOR_RETURN();
            }
            if (sgn1(mu0) == -1.0) muplus = -muplus;
            if (fabs(muplus) < fabs(mu0)) { muplus = mu0; mpos = muplus * muplus; }
            mneg = fmin(mu0 * mu0, mneg);
            if (firstpt) orc_calcimuasym(a, mneg, mpos, mu0, muplus, &v->i1mu, &v->i3mu, &v->rffmu1, &v->rffmu3);
            i2mu = orc_calcimuasymf(a, mneg, mpos, muf, muplus, &v->rffmu2);
        }
        imu = a1 * v->i1mu + a2 * i2mu + a3 * v->i3mu;

        /* quartic U(u): roots and case (:2947-2993) */
        p[0] = -1; p[1] = -1; p[2] = -1; p[3] = -1; p[4] = 0;
        if (firstpt) {
            double complex c[5], root[4];
            c[0] = one; c[1] = 0.0; c[2] = cc; c[3] = dd; c[4] = ee;
            orc_zroots(c, 4, root, 1);
            int nreal = 0;
            for (int i = 0; i < 4; ++i)
                if (cimag(root[i]) == 0.0) nreal++;
            if (nreal == 2) {
                v->ncase = 5;
                tpr = 0;
                v->u1 = creal(root[0]);
                v->u4 = (cimag(root[1]) == 0.0) ? creal(root[1]) : creal(root[3]);
            } else if (nreal == 0) {
                v->ncase = 6;
                tpr = 0;
            } else {
                v->u1 = creal(root[0]); v->u2 = creal(root[1]); v->u3 = creal(root[2]); v->u4 = creal(root[3]);
                for (int guard = 0; guard < 3; ++guard) {
                    if (v->u2 > uplus && v->u3 > uplus) { v->ncase = 5; break; }
                    else if (u0 <= v->u2) { v->ncase = 7; break; }
                    else if (u0 >= v->u3) { v->ncase = 8; break; }
                    if (warned) *warned = 1;
                    u0 = (su == 1.0) ? v->u3 : v->u2;
                }
            }
        }
        const double u1 = v->u1, u2 = v->u2, u3 = v->u3, u4 = v->u4;
        double sqe = sqrt(fabs(ee));
        if (v->ncase == 5) {
            /* Table 2 row 5 (:2995-3027) */
            double qs = sgn1(q2);
            double f = -qs * one / fabs(ee) / u1 / u4;
            double g = (u4 + u1) / u1 / u4 * f;
            double iut;
            if (su > 0.0)
                iut = orc_ellquarticcomplex(p, -u1, one, qs * u4, -qs * one, 0.0, 0.0, f, g, one, &dummy, u0, uplus) / sqe;
            else
                iut = orc_ellquarticcomplex(p, -u1, one, qs * u4, -qs * one, 0.0, 0.0, f, g, one, &dummy, 0.0, u0) / sqe;
            if (imu > iut) {
                uf = -1.0;
                v->ncase = 0;
                tpr = 0;
                GEOR_RETURN();
            }
            double ub = (qs == one) ? u1 : u4;
            double m = -g * half;
            double n2 = f - g * g / 4.0;
            double c4 = sqrt((m - u4) * (m - u4) + n2);
            double c5 = sqrt((m - u1) * (m - u1) + n2);
            double c1 = sqrt(fabs(ee * c4 * c5));
            if (firstpt)
                v->iu0 = su * orc_ellquarticcomplex(p, -u1, one, qs * u4, -qs * one, 0.0, 0.0, f, g, one, &dummy, ub, u0) / sqe;
            m1 = qs * ((c4 + qs * c5) * (c4 + qs * c5) - (u4 - u1) * (u4 - u1)) / (4.0 * c4 * c5);
            jarg = c1 * (imu + v->iu0);
            orc_sncndn(jarg, m1, &sn, &cn, &dn);
            uf = (u4 * c5 + qs * u1 * c4 - (qs * u4 * c5 - u1 * c4) * cn) / ((c4 - qs * c5) * cn + (qs * c4 + c5));
            if (pht)
                iu_local = orc_ellquarticcomplex(p, -u1, one, qs * u4, -qs * one, 0.0, 0.0, f, g, one, &v->rffu0, u0, uf) / sqe;
        } else if (v->ncase == 6) {
            /* Table 2 row 6 (:3028-3095) */
            double complex coefs[7], hroots[6];
            double se = sqrt(ee);
            coefs[0] = one; coefs[1] = -cc / se; coefs[2] = -one;
            coefs[3] = se * (two * cc / ee - (dd / ee) * (dd / ee));
            coefs[4] = -one; coefs[5] = -cc / se; coefs[6] = one;
            orc_zroots(coefs, 6, hroots, 1);
            double h1 = 0.0;
            int found = 0;
            for (int i = 0; i < 6 && h1 == 0.0; ++i)
                if (cimag(hroots[i]) == 0.0) { h1 = creal(hroots[i]); found = (h1 != 0.0); }
            if (!found) h1 = NAN;
            v->h1 = h1;
            double h2 = one / h1;
            double g1 = dd / ee / (h2 - h1);
            double g2 = -g1;
            double f1 = one / sqrt(ee);
            double f2 = f1;
            double iut;
            if (su > 0.0)
                iut = orc_elldoublecomplex(p, f1, g1, h1, f2, g2, h2, 0.0, 0.0, &dummy, u0, uplus) / sqe;
            else
                iut = orc_elldoublecomplex(p, f1, g1, h1, f2, g2, h2, 0.0, 0.0, &dummy, 0.0, u0) / sqe;
            if (imu > iut) {
                uf = -1.0;
                v->ncase = 0;
                tpr = 0;
                GEOR_RETURN();
            }
            coefs[0] = pow(ee, -three);
            coefs[1] = -cc / pow(ee, three);
            coefs[2] = -pow(ee, -two);
            coefs[3] = -pow(ee, -two) * (pow(dd, two) / ee - two * cc);
            coefs[4] = -one / ee;
            coefs[5] = -cc / ee;
            coefs[6] = one;
            orc_zroots(coefs, 6, hroots, 1);
            double mn2 = 0.0;
            found = 0;
            for (int i = 0; i < 6 && mn2 == 0.0; ++i)
                if (cimag(hroots[i]) == 0.0) { mn2 = creal(hroots[i]); found = (mn2 != 0.0); }
            if (!found) mn2 = NAN;
            double pp = dd / (two * (ee * ee) * (mn2 * mn2 - one / ee));
            double m = -half * dd / ee - pp;
            if (m < pp) { double t = pp; pp = m; m = t; }
            double pr2 = one / (mn2 * ee);
            double n = sqrt(mn2 - m * m);
            double r = sqrt(pr2 - pp * pp);
            double c4 = sqrt((m - pp) * (m - pp) + (n + r) * (n + r));
            double c5 = sqrt((m - pp) * (m - pp) + (n - r) * (n - r));
            double c1 = (c4 + c5) * half * sqe;
            double c2 = sqrt((4.0 * n * n - (c4 - c5) * (c4 - c5)) / ((c4 + c5) * (c4 + c5) - 4.0 * n * n));
            double c3 = m + c2 * n;
            if (firstpt)
                v->iu0 = sgn1(u0 - c3) * orc_elldoublecomplex(p, f1, g1, h1, f2, g2, h2, 0.0, 0.0, &dummy, c3, u0) / sqe;
            m1 = ((c4 - c5) / (c4 + c5)) * ((c4 - c5) / (c4 + c5));
            jarg = c1 * (su * imu + v->iu0);
            orc_sncndn(jarg, m1, &sn, &cn, &dn);
            double sc = sn / cn;
            uf = c3 + (n * (one + c2 * c2) * sc) / (one - c2 * sc);
            if (pht)
                iu_local = orc_elldoublecomplex(p, f1, g1, h1, f2, g2, h2, 0.0, 0.0, &v->rffu0, u0, uf) / sqrt(-ee);
        } else if (v->ncase == 7) {
            /* Table 2 row 7 (:3098-3108) */
            if (firstpt && u0 != u2)
                v->iu0 = orc_ellquarticreal(p, -u1, one, u2, -one, u3, -one, u4, -one, 0.0, 0.0, &v->rffu0, u0, u2);
            jarg = sqrt(fabs(ee) * (u3 - u1) * (u4 - u2)) * half * (imu - su * v->iu0 / sqrt(-ee));
            m1 = (u4 - u1) * (u3 - u2) / ((u4 - u2) * (u3 - u1));
            orc_sncndn(jarg, m1, &sn, &cn, &dn);
            uf = ((u2 - u1) * u3 * sn * sn - u2 * (u3 - u1)) / ((u2 - u1) * sn * sn - (u3 - u1));
            tpr = f_int((sgn1(su * (imu - su * v->iu0 / sqe)) + one) / two);
            if (pht)
                iu1 = orc_ellquarticreal(p, -u1, one, u2, -one, u3, -one, u4, -one, 0.0, 0.0, &v->rffu1, uf, u2);
        } else if (v->ncase == 8) {
            /* Table 2 row 8 (:3109-3118) */
            if (firstpt && u0 != u3)
                v->iu0 = orc_ellquarticreal(p, -u1, one, -u2, one, -u3, one, u4, -one, 0.0, 0.0, &v->rffu0, u3, u0);
            jarg = sqrt(fabs(ee) * (u3 - u1) * (u4 - u2)) * half * (imu + su * v->iu0 / sqrt(-ee));
            m1 = (u4 - u1) * (u3 - u2) / ((u4 - u2) * (u3 - u1));
            orc_sncndn(jarg, m1, &sn, &cn, &dn);
            uf = ((u4 - u3) * u2 * sn * sn - u3 * (u4 - u2)) / ((u4 - u3) * sn * sn - (u4 - u2));
            tpr = f_int((-sgn1(su * (imu + su * v->iu0 / sqrt(-ee))) + one) / two);
            if (pht)
                iu1 = orc_ellquarticreal(p, -u1, one, -u2, one, -u3, one, u4, -one, 0.0, 0.0, &v->rffu1, u3, uf);
        }
    }
    GEOR_RETURN();
#undef GEOR_RETURN
}

void orc_geor(double u0, double uf_in, double mu0, double muf, double a, double l, double q2, int32_t tpm,
              int32_t tpr_in, double su, double sm, orc_geor_out* o) {
    orc_gk_vars v;
    memset(&v, 0, sizeof v);
    double U0 = u0, UF = uf_in, imu = 0.0;
    int32_t tpr = tpr_in, warned = 0;
    orc_gk_geor(&v, &U0, &UF, mu0, muf, a, l, q2, &imu, tpm, &tpr, su, sm, 0, 1, &warned);
    memset(o, 0, sizeof *o);
    o->uf = UF;
    o->imu = imu;
    o->iu0 = v.iu0;
    o->ncase = v.ncase;
    o->tpr = tpr;
    o->u1 = v.u1; o->u2 = v.u2; o->u3 = v.u3; o->u4 = v.u4;
    o->h1 = v.h1;
    o->u0_used = U0;
    o->warned = warned;
}

void orc_geor_batch(const double* u0, const double* uf_in, const double* mu0, const double* muf,
                    const double* a, const double* l, const double* q2, const int32_t* tpm,
                    const double* su, const double* sm, int n, double* uf, double* imu, int32_t* tpr,
                    int32_t* ncase, int nthreads) {
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
#endif
    for (int i = 0; i < n; ++i) {
        orc_geor_out o;
        orc_geor(u0[i], uf_in ? uf_in[i] : -1.0, mu0[i], muf[i], a[i], l[i], q2[i], tpm[i], 0, su[i], sm[i], &o);
        uf[i] = o.uf;
        imu[i] = o.imu;
        tpr[i] = o.tpr;
        ncase[i] = o.ncase;
    }
    (void)nthreads;
}
