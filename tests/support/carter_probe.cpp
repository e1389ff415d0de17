// Host build of the Carter-form right-hand side and its vector-Jacobian
// product (nullpath_amd/csrc/np_carter.hpp) for the CPU test that checks the
// product against finite differences.
#include <cmath>
#include "../../nullpath_amd/csrc/np_carter.hpp"

extern "C" {

void probe_rhs(const double* y, double E, double L, double a, double* out) {
    const double rp = 1.0 + std::sqrt(1.0 - a * a), rm = 1.0 - std::sqrt(1.0 - a * a);
    np::CState<double> s{y[0], y[1], y[2], y[3]};
    np::CState<double> d = np::carter_rhs(s, E, L, a, rp, rm);
    out[0] = d.r; out[1] = d.th; out[2] = d.pr; out[3] = d.pth;
}

void probe_vjp(const double* y, double E, double L, double a, const double* v, double* out) {
    const double rp = 1.0 + std::sqrt(1.0 - a * a), rm = 1.0 - std::sqrt(1.0 - a * a);
    np::CState<double> s{y[0], y[1], y[2], y[3]};
    np::CState<double> vv{v[0], v[1], v[2], v[3]};
    double gE = 0, gL = 0, ga = 0;
    np::CState<double> u = np::carter_vjp(s, E, L, a, rp, rm, vv, gE, gL, ga);
    out[0] = u.r; out[1] = u.th; out[2] = u.pr; out[3] = u.pth;
    out[4] = gE; out[5] = gL; out[6] = ga;
}

// r-scaled flow F = r f and its VJP (scaled = 1), for the same checks.
void probe_rhs_r(const double* y, double E, double L, double a, double* out) {
    const double rp = 1.0 + std::sqrt(1.0 - a * a), rm = 1.0 - std::sqrt(1.0 - a * a);
    np::CState<double> s{y[0], y[1], y[2], y[3]};
    np::CState<double> d = np::carter_rhs_r(s, E, L, a, rp, rm);
    out[0] = d.r; out[1] = d.th; out[2] = d.pr; out[3] = d.pth;
}

void probe_vjp_r(const double* y, double E, double L, double a, const double* v, double* out) {
    const double rp = 1.0 + std::sqrt(1.0 - a * a), rm = 1.0 - std::sqrt(1.0 - a * a);
    np::CState<double> s{y[0], y[1], y[2], y[3]};
    np::CState<double> vv{v[0], v[1], v[2], v[3]};
    double gE = 0, gL = 0, ga = 0;
    np::CState<double> u = np::carter_vjp_r(s, E, L, a, rp, rm, vv, gE, gL, ga);
    out[0] = u.r; out[1] = u.th; out[2] = u.pr; out[3] = u.pth;
    out[4] = gE; out[5] = gL; out[6] = ga;
}

}
