// Host compilation of this repository's own device headers for CPU tests:
// the HIP function qualifiers become plain (inline) C++.
#pragma once
#include <cmath>
#include <cstdint>
#define __device__
#define __host__
#define __forceinline__ inline
using std::isfinite;
