// Host build of the library's FP64 and float sincos
// (nullpath_amd/csrc/np_math.hpp) for the CPU test that checks them against
// the C library.
#include "../../nullpath_amd/csrc/np_math.hpp"

extern "C" void probe_sincos(const double* x, int n, double* s, double* c) {
    for (int i = 0; i < n; ++i) np::sincos_d(x[i], s[i], c[i]);
}

extern "C" void probe_sincos_f(const float* x, int n, float* s, float* c) {
    for (int i = 0; i < n; ++i) np::sincos_f(x[i], s[i], c[i]);
}
