// Host build of the library's quartic root finders (np_geokerr.hpp): the
// FP64 ZROOTS and the mixed-precision form GEOMU's compared mode uses, for
// the CPU test that checks them against the oracle's ZROOTS.
#include <hip/hip_runtime.h>   // tests/support/hostshim: HIP qualifiers as plain C++


#include "../../nullpath_amd/csrc/np_geokerr.hpp"

// coefficients of U(u) = 1 + cc u^2 + dd u^3 + ee u^4 per case (3 doubles);
// out: 8 doubles (re, im of 4 roots) and a flag (1 = mixed handed back)
extern "C" void probe_roots(const double* coef, int n, int mixed, double* roots, int* fallback) {
    for (int i = 0; i < n; ++i) {
        np::gk::cplx c[5] = {np::gk::mk(1.0), np::gk::mk(0.0), np::gk::mk(coef[3 * i]), np::gk::mk(coef[3 * i + 1]),
                             np::gk::mk(coef[3 * i + 2])};
        np::gk::cplx r[4];
        int fb = 0;
        if (mixed) {
            if (!np::gk::zroots4_mixed(c, r)) {
                np::gk::zroots<4>(c, r);
                fb = 1;
            }
        } else {
            np::gk::zroots<4>(c, r);
        }
        for (int j = 0; j < 4; ++j) {
            roots[8 * i + 2 * j] = r[j].re;
            roots[8 * i + 2 * j + 1] = r[j].im;
        }
        fallback[i] = fb;
    }
}
