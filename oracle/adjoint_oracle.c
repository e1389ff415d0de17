/*
 * adjoint_oracle.c -- TEST INFRASTRUCTURE ONLY.  FP64 forward-mode (dual
 * number) reference for the differentiable disk render of libnullpath
 * (np_adjoint_forward / np_adjoint_backward): the same model evaluated in
 * plain FP64 with exact first derivatives with respect to
 * (spin, theta_obs, emiss_q, gain) carried alongside every value.
 *
 * Model pieces and where they come from:
 *   camera directions   render.cu:143-159 (orc_camera_dir, sample 0)
 *   ZAMO initial data   bh_kerr.h:897-964 (tetrad of orc_init_zamo_d)
 *   flow                Hamilton's equations of bh_kerr.h:600-727 on H = 0,
 *                       written in Carter form, parameter s with dl = r ds
 *   crossing            render.cu:211-226 (linear interpolation in theta)
 *   redshift            bh_kerr.h:833-862 at theta = pi/2 (Keplerian disk)
 *   ISCO                Bardeen, Press & Teukolsky (1972)
 * The smooth edge window and r^-q emissivity are this library's choices
 * (the reference uses a hard edge and max(r,1)^-2).
 *
 * Pins (tests/test_oracle_adjoint.py): derivatives against central finite
 * differences of the oracle's own values; the Carter field against the
 * reference Hamiltonian field (orc_kerr_rhs_d) on shell; the redshift
 * against orc_doppler_d at crossing states; ISCO against its closed-form
 * values; initial data against orc_init_zamo_d.
 */
#include <math.h>
#include <string.h>

#include "np_oracle.h"

#define ND 4  /* d/d(spin), d/d(theta_obs), d/d(q), d/d(gain) */
#define BH_PI_D_ORC 3.141592653589793238462643383279502884

typedef struct {
    double v, d[ND];
} dual;

static dual dk(double v) {
    dual x;
    x.v = v;
    for (int i = 0; i < ND; ++i) x.d[i] = 0.0;
    return x;
}
static dual dvar(double v, int k) {
    dual x = dk(v);
    x.d[k] = 1.0;
    return x;
}
static dual dadd(dual x, dual y) {
    for (int i = 0; i < ND; ++i) x.d[i] += y.d[i];
    x.v += y.v;
    return x;
}
static dual dsub(dual x, dual y) {
    for (int i = 0; i < ND; ++i) x.d[i] -= y.d[i];
    x.v -= y.v;
    return x;
}
static dual dmul(dual x, dual y) {
    dual o;
    o.v = x.v * y.v;
    for (int i = 0; i < ND; ++i) o.d[i] = x.d[i] * y.v + x.v * y.d[i];
    return o;
}
static dual dscale(double s, dual x) {
    x.v *= s;
    for (int i = 0; i < ND; ++i) x.d[i] *= s;
    return x;
}
static dual ddiv(dual x, dual y) {
    dual o;
    o.v = x.v / y.v;
    for (int i = 0; i < ND; ++i) o.d[i] = (x.d[i] - o.v * y.d[i]) / y.v;
    return o;
}
static dual dfun(dual x, double f, double df) {
    dual o;
    o.v = f;
    for (int i = 0; i < ND; ++i) o.d[i] = df * x.d[i];
    return o;
}
static dual dsqrt_(dual x) { double s = sqrt(x.v); return dfun(x, s, 0.5 / s); }
static dual dcbrt_(dual x) { double s = cbrt(x.v); return dfun(x, s, s / (3.0 * x.v)); }
static dual dsin_(dual x) { return dfun(x, sin(x.v), cos(x.v)); }
static dual dcos_(dual x) { return dfun(x, cos(x.v), -sin(x.v)); }
static dual dlog_(dual x) { return dfun(x, log(x.v), 1.0 / x.v); }
static dual dexp_(dual x) { double e = exp(x.v); return dfun(x, e, e); }
static dual dfabs_(dual x) { return x.v < 0 ? dscale(-1.0, x) : x; }
static dual dsmooth(dual x) {
    if (x.v <= 0.0) return dk(0.0);
    if (x.v >= 1.0) return dk(1.0);
    return dfun(x, x.v * x.v * (3.0 - 2.0 * x.v), 6.0 * x.v * (1.0 - x.v));
}

typedef struct {
    dual r, th, pr, pth;
} dstate;

/* Carter-form field on H = 0 times r (dl = r ds). */
static dstate rhs(const dstate* y, dual E, dual L, dual a) {
    const dual one = dk(1.0), two = dk(2.0);
    dual s = dsin_(y->th), c = dcos_(y->th);
    dual r = y->r, a2 = dmul(a, a), r2 = dmul(r, r);
    dual s2 = dmul(s, s);
    if (s2.v < 1e-12) s2 = dk(1e-12);
    dual Sigma = dadd(r2, dmul(a2, dmul(c, c)));
    dual Delta = dadd(dsub(r2, dscale(2.0, r)), a2);
    dual P = dsub(dmul(dadd(r2, a2), E), dmul(a, L));
    dual w = dsub(ddiv(L, s2), dmul(a, E));
    dual Q = dadd(dmul(y->pth, y->pth), dmul(s2, dmul(w, w)));
    dual X = dsub(dmul(dscale(2.0, E), dsub(r2, a2)), dmul(dscale(2.0, dmul(a, L)), dsub(r, one)));
    dstate d;
    d.r = ddiv(dmul(Delta, y->pr), Sigma);
    d.th = ddiv(y->pth, Sigma);
    d.pr = ddiv(dsub(ddiv(dmul(dsub(r, one), Q), Delta), ddiv(dmul(P, X), dmul(Delta, Delta))), Sigma);
    d.pth = ddiv(dmul(dmul(dmul(c, s), w), dadd(ddiv(L, s2), dmul(a, E))), Sigma);
    (void)two;
    d.r = dmul(r, d.r);
    d.th = dmul(r, d.th);
    d.pr = dmul(r, d.pr);
    d.pth = dmul(r, d.pth);
    return d;
}

static dstate axpy(const dstate* y, double h, const dstate* k) {
    dstate o;
    o.r = dadd(y->r, dscale(h, k->r));
    o.th = dadd(y->th, dscale(h, k->th));
    o.pr = dadd(y->pr, dscale(h, k->pr));
    o.pth = dadd(y->pth, dscale(h, k->pth));
    return o;
}

static dual comb(dual k1, dual k2, dual k3, dual k4) {
    return dadd(dadd(k1, dscale(2.0, dadd(k2, k3))), k4);
}

static void rk4(dstate* y, double h, dual E, dual L, dual a) {
    dstate k1 = rhs(y, E, L, a);
    dstate y2 = axpy(y, 0.5 * h, &k1);
    dstate k2 = rhs(&y2, E, L, a);
    dstate y3 = axpy(y, 0.5 * h, &k2);
    dstate k3 = rhs(&y3, E, L, a);
    dstate y4 = axpy(y, h, &k3);
    dstate k4 = rhs(&y4, E, L, a);
    const double h6 = h / 6.0;
    y->r = dadd(y->r, dscale(h6, comb(k1.r, k2.r, k3.r, k4.r)));
    y->th = dadd(y->th, dscale(h6, comb(k1.th, k2.th, k3.th, k4.th)));
    y->pr = dadd(y->pr, dscale(h6, comb(k1.pr, k2.pr, k3.pr, k4.pr)));
    y->pth = dadd(y->pth, dscale(h6, comb(k1.pth, k2.pth, k3.pth, k4.pth)));
}

/* Prograde ISCO (Bardeen, Press & Teukolsky 1972). */
static dual isco(dual a) {
    const dual one = dk(1.0), three = dk(3.0);
    dual a2 = dmul(a, a);
    dual z1 = dadd(one, dmul(dcbrt_(dsub(one, a2)), dadd(dcbrt_(dadd(one, a)), dcbrt_(dsub(one, a)))));
    dual z2 = dsqrt_(dadd(dscale(3.0, a2), dmul(z1, z1)));
    return dsub(dadd(three, z2), dsqrt_(dmul(dsub(three, z1), dadd(dadd(three, z1), dscale(2.0, z2)))));
}

/* Crossing contribution gain W(r) r^-q g^3 at equatorial radius r. */
static dual disk_term(dual r, dual E, dual L, dual a, dual q, dual gain, double rout, double ew) {
    const dual one = dk(1.0);
    dual rin = isco(a);
    dual W = dmul(dsmooth(dscale(1.0 / ew, dsub(r, rin))), dsmooth(dscale(1.0 / ew, dsub(dk(rout), r))));
    if (W.v == 0.0) return dk(0.0);
    dual a2 = dmul(a, a);
    dual Om = ddiv(one, dadd(dmul(r, dsqrt_(r)), a));
    dual g_tt = dadd(dk(-1.0), ddiv(dk(2.0), r));
    dual g_tph = dscale(-2.0, ddiv(a, r));
    dual g_pp = dadd(dadd(dmul(r, r), a2), dscale(2.0, ddiv(a2, r)));
    dual Dn = dscale(-1.0, dadd(dadd(g_tt, dscale(2.0, dmul(g_tph, Om))), dmul(g_pp, dmul(Om, Om))));
    dual g = ddiv(dsqrt_(Dn), dsub(one, dmul(Om, ddiv(L, E))));
    dual rq = dexp_(dmul(dscale(-1.0, q), dlog_(r)));
    return dmul(gain, dmul(W, dmul(rq, dmul(g, dmul(g, g)))));
}

/* ZAMO tetrad initial covariant momenta (bh_kerr.h:897-964) in dual
 * arithmetic: E = -p_t, L = p_phi, p_r, p_theta. */
static void zamo_init(dual a, dual th, double r, const float n[3], dual* E, dual* L, dual* pr, dual* pth) {
    const dual one = dk(1.0);
    dual ct = dcos_(th), st = dsin_(th);
    dual rr = dk(r), r2 = dk(r * r), a2 = dmul(a, a);
    dual Sigma = dadd(r2, dmul(a2, dmul(ct, ct)));
    dual Delta = dadd(dsub(r2, dk(2.0 * r)), a2);
    dual st2 = dmul(st, st);
    dual A = dsub(dmul(dadd(r2, a2), dadd(r2, a2)), dmul(a2, dmul(Delta, st2)));
    dual alpha = dsqrt_(ddiv(dmul(Delta, Sigma), A));
    dual omega = ddiv(dscale(2.0, dmul(a, rr)), A);
    dual g_tt = dscale(-1.0, dsub(one, ddiv(dk(2.0 * r), Sigma)));
    dual g_tph = dscale(-2.0, ddiv(dmul(dmul(a, rr), st2), Sigma));
    dual g_pp = ddiv(dmul(A, st2), Sigma);
    dual g_rr = ddiv(Sigma, Delta);
    dual g_thth = Sigma;
    dual pt_c = ddiv(one, alpha);
    dual pph_c = dadd(ddiv(omega, alpha), dscale((double)n[2], ddiv(dsqrt_(ddiv(Sigma, A)), dfabs_(st))));
    dual pr_c = dscale((double)n[0], dsqrt_(ddiv(Delta, Sigma)));
    dual pth_c = dscale((double)n[1], ddiv(one, dsqrt_(Sigma)));
    dual pt = dadd(dmul(g_tt, pt_c), dmul(g_tph, pph_c));
    *pr = dmul(g_rr, pr_c);
    *pth = dmul(g_thth, pth_c);
    *L = dadd(dmul(g_tph, pt_c), dmul(g_pp, pph_c));
    *E = dscale(-1.0, pt);
}

/* One pixel: radiance sum I and dI/d(spin, theta_obs, q, gain). */
static void adj_pixel(const np_adjoint_params* p, int x, int y, double* I, double dI[ND], int* ncross_out) {
    np_render_params rp;
    np_render_params_default(&rp);
    rp.width = p->width;
    rp.height = p->height;
    rp.fov_y_deg = p->fov_y_deg;
    rp.seed = p->seed;
    float n[3], cam[2];
    orc_camera_dir(&rp, x, y, 0u, n, cam);
    dual a = dvar(p->spin, 0), th = dvar(p->theta_obs, 1), q = dvar(p->emiss_q, 2), gain = dvar(p->gain, 3);
    dual E, L, pr, pth;
    zamo_init(a, th, p->r_cam, n, &E, &L, &pr, &pth);
    const double disc = sqrt(1.0 - p->spin * p->spin);
    const double rplus = 1.0 + disc;
    const double r_far = p->r_cam + 20.0;
    dstate s;
    s.r = dk(p->r_cam);
    s.th = th;
    s.pr = pr;
    s.pth = pth;
    dual sum = dk(0.0);
    int ncross = 0;
    for (int k = 0; k < p->n_steps; ++k) {
        dstate t = s;
        rk4(&t, p->h, E, L, a);
        const double d0 = s.th.v - 0.5 * BH_PI_D_ORC, d1 = t.th.v - 0.5 * BH_PI_D_ORC;
        if (ncross < p->max_orders && ((d0 > 0 && d1 < 0) || (d0 < 0 && d1 > 0)) && t.r.v > rplus + 1e-3) {
            dual dd0 = dsub(s.th, dk(0.5 * BH_PI_D_ORC)), dd1 = dsub(t.th, dk(0.5 * BH_PI_D_ORC));
            dual al = ddiv(dd0, dsub(dd0, dd1));
            dual rx = dadd(s.r, dmul(al, dsub(t.r, s.r)));
            sum = dadd(sum, disk_term(rx, E, L, a, q, gain, p->disk_rout, p->edge_width));
            ++ncross;
        }
        s = t;
        const double r = t.r.v, tt = t.th.v;
        if (!(r > rplus + 1e-3) || !(tt > 1e-6) || !(tt < BH_PI_D_ORC - 1e-6) || (r > r_far && k > 10)) break;
    }
    *I = sum.v;
    for (int i = 0; i < ND; ++i) dI[i] = sum.d[i];
    if (ncross_out) *ncross_out = ncross;
}

void orc_adj_pixels(const np_adjoint_params* p, int npix, const int32_t* xs, const int32_t* ys,
                    double* I, double* dI, int32_t* ncross, int nthreads) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads > 0 ? nthreads : 1)
    for (int i = 0; i < npix; ++i) {
        int nc = 0;
        adj_pixel(p, xs[i], ys[i], &I[i], &dI[ND * i], &nc);
        if (ncross) ncross[i] = nc;
    }
}

/* Pieces exposed for their own pins. */
double orc_adj_isco(double a, double* d_da) {
    dual r = isco(dvar(a, 0));
    if (d_da) *d_da = r.d[0];
    return r.v;
}

double orc_adj_disk_term(double r, double E, double L, double a, double q, double gain, double rout,
                         double ew) {
    return disk_term(dk(r), dk(E), dk(L), dk(a), dk(q), dk(gain), rout, ew).v;
}

void orc_adj_zamo_init(double a, double th, double r_cam, const float n[3], double out[4]) {
    dual E, L, pr, pth;
    zamo_init(dk(a), dk(th), r_cam, n, &E, &L, &pr, &pth);
    out[0] = E.v;
    out[1] = L.v;
    out[2] = pr.v;
    out[3] = pth.v;
}

void orc_adj_rhs(const double y[4], double E, double L, double a, double out[4]) {
    dstate s;
    s.r = dk(y[0]);
    s.th = dk(y[1]);
    s.pr = dk(y[2]);
    s.pth = dk(y[3]);
    dstate d = rhs(&s, dk(E), dk(L), dk(a));
    out[0] = d.r.v;
    out[1] = d.th.v;
    out[2] = d.pr.v;
    out[3] = d.pth.v;
}
