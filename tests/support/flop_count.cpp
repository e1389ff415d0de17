// Operation count of the Carter-form field and its vector-Jacobian product
// (nullpath_amd/csrc/np_carter.hpp) with a counting scalar type: add, sub,
// mul, div, reciprocal and sqrt count one flop each; sin and cos are counted apart;
// max/min, negation and comparisons are not flops.  The adjoint benchmark's
// per-step flop figures are built from these counts.
#include <cmath>
#include <cstdio>

namespace cnt {
struct Tally {
    long flops = 0, trans = 0;
};
inline Tally& tally() {
    static Tally t;
    return t;
}
struct Counted {
    double v;
    Counted() : v(0) {}
    Counted(double x) : v(x) {}
};
inline Counted op(double x) { ++tally().flops; return Counted(x); }
inline Counted operator+(Counted a, Counted b) { return op(a.v + b.v); }
inline Counted operator-(Counted a, Counted b) { return op(a.v - b.v); }
inline Counted operator*(Counted a, Counted b) { return op(a.v * b.v); }
inline Counted operator/(Counted a, Counted b) { return op(a.v / b.v); }
inline Counted operator-(Counted a) { return Counted(-a.v); }
inline Counted& operator+=(Counted& a, Counted b) { a = a + b; return a; }
inline void carter_sincos(Counted x, Counted& s, Counted& c) {
    tally().trans += 2;
    s = Counted(std::sin(x.v));
    c = Counted(std::cos(x.v));
}
inline Counted carter_fmax(Counted a, Counted b) { return Counted(a.v > b.v ? a.v : b.v); }
inline Counted carter_rcp(Counted x) { return op(1.0 / x.v); }
}  // namespace cnt

#include "../../nullpath_amd/csrc/np_carter.hpp"

using cnt::Counted;

int main() {
    np::CState<Counted> y{Counted(30.0), Counted(1.2), Counted(-0.7), Counted(3.0)};
    const Counted E(1.0), L(2.5), a(0.7), rp(1.714), rm(0.286);
    auto& t = cnt::tally();
    t = cnt::Tally();
    (void)np::carter_rhs_r(y, E, L, a, rp, rm);
    const long rhs = t.flops, rhs_trans = t.trans;
    t = cnt::Tally();
    Counted gE(0), gL(0), ga(0);
    (void)np::carter_vjp_r(y, E, L, a, rp, rm, np::CState<Counted>{Counted(1), Counted(2), Counted(3), Counted(4)}, gE,
                           gL, ga);
    std::printf("rhs_r %ld %ld\nvjp_r %ld %ld\n", rhs, rhs_trans, t.flops, t.trans);
    return 0;
}
