// np_device.hpp -- device-side building blocks shared by the gfx950 kernels:
// compensated summation, wave64 work-queue fetch, counter reduction.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "np_math.hpp"

#define NP_PI_D 3.141592653589793238462643383279502884
#define NP_PI_F 3.14159265358979323846f

namespace np {

// Second-order compensated (Kahan-Babuska-Klein) running sum with the same
// operation sequence as BhKleinSum::add (bh_common.h:451-483); the
// magnitude test becomes two selects, no branch.
template <typename Real>
struct Klein {
    Real sum, cs, ccs;
    __device__ __forceinline__ Klein() : sum(Real(0)), cs(Real(0)), ccs(Real(0)) {}
    __device__ __forceinline__ void add(Real x) {
        Real t = sum + x;
        Real c = (fabs(sum) >= fabs(x)) ? ((sum - t) + x) : ((x - t) + sum);
        sum = t;
        Real t2 = cs + c;
        Real cc = (fabs(cs) >= fabs(c)) ? ((cs - t2) + c) : ((c - t2) + cs);
        cs = t2;
        ccs += cc;
    }
    __device__ __forceinline__ Real result() const { return sum + cs + ccs; }
};

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Wave-level refill from a global work counter: lanes with `need` set get
// consecutive indices from one atomicAdd issued by the lowest such lane
// (ballot + prefix popcount), the others get UINT64_MAX.
__device__ __forceinline__ uint64_t wave_fetch(bool need, unsigned long long* counter) {
    const uint64_t mask = __ballot(need);
    if (mask == 0) return ~0ull;
    const int leader = __ffsll((long long)mask) - 1;
    unsigned long long base = 0;
    if (lane_id() == leader) base = atomicAdd(counter, (unsigned long long)__popcll(mask));
    base = __shfl(base, leader);
    const uint64_t below = mask & ((lane_id() == 0) ? 0ull : (~0ull >> (64 - lane_id())));
    return need ? (uint64_t)base + (uint64_t)__popcll(below) : ~0ull;
}

// Sum a per-lane counter over the wave and add it once to *dst.
__device__ __forceinline__ void wave_add_u64(uint64_t* dst, uint64_t v) {
    unsigned long long w = v;
    for (int off = 32; off > 0; off >>= 1) w += __shfl_xor(w, off);
    if (lane_id() == 0 && w) atomicAdd(reinterpret_cast<unsigned long long*>(dst), w);
}

// Add a wave-uniform total to *dst once (lane 0).
__device__ __forceinline__ void wave_add_uniform_u64(uint64_t* dst, uint64_t v) {
    if (lane_id() == 0 && v) atomicAdd(reinterpret_cast<unsigned long long*>(dst), (unsigned long long)v);
}

// Atomic maximum of a non-negative float through its bit pattern (the
// ordering of non-negative IEEE floats equals that of their bits).
__device__ __forceinline__ void atomic_max_nonneg(float* addr, float v) {
    if (!(v >= 0.0f)) return;  // skips NaN
    atomicMax(reinterpret_cast<unsigned int*>(addr), __float_as_uint(v));
}

}  // namespace np
