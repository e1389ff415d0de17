// np_kernels.hpp -- host-visible launchers of the gfx950 kernels (internal;
// the public surface is the C ABI in include/nullpath.h).
#pragma once

#include <hip/hip_runtime.h>
#include "np_frame.hpp"
#include "nullpath.h"

namespace np {

int launch_kerr_fp64(const FrameConst& F, bool adaptive, int y0, int rows, int spp,
                     int sample_offset, float* d_out, int accumulate, float* d_inv,
                     np_trace_counters* d_counters, np_ray_record* d_records,
                     unsigned long long* d_work,
                     hipStream_t stream, int grid);
int occupancy_kerr_fp64(bool adaptive, int* blocks_per_cu);

}  // namespace np
