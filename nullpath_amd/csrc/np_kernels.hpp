// np_kernels.hpp -- host-visible launchers of the gfx950 kernels (internal;
// the public surface is the C ABI in include/nullpath.h).
#pragma once

#include <hip/hip_runtime.h>
#include "np_frame.hpp"
#include "nullpath.h"

namespace np {

// One frame of a launch in device memory: its constants, the window of
// pixels to trace (npix pixels from absolute pixel index pix0 = y0 * width),
// the offsets of the window's colours (floats) and ray records in the packed
// outputs, and the id of the first sample traced per pixel.
struct BatchFrame {
    FrameConst F;
    uint64_t npix;
    uint64_t pix0;
    uint64_t out_off;
    uint64_t rec_off;
    uint32_t sample0;
    uint32_t reserved;
};

// The Kerr renderer over nframes frame records in device memory, in one
// persistent launch; d_work holds one cleared counter per frame.
// precision is an np_precision value; `chunk` is the number of consecutive
// pixels a lane takes per refill.
// d_times (optional): the launch runs the kernel that also integrates t and
// lambda and writes one np_ray_times per sample (inv must be false).
// defer: the FP32C kernel that defers FP64 attempts in the frames whose
// FrameConst::defer64 is set (adaptive steps, inv and d_times unset).
int launch_kerr(const BatchFrame* d_frames, int nframes, int precision, bool adaptive, bool inv,
                int variant, int spp, float* d_out, int accumulate, float* d_inv,
                np_trace_counters* d_counters, np_ray_record* d_records,
                unsigned long long* d_work, hipStream_t stream, int grid, int block, int chunk,
                np_ray_times* d_times = nullptr, bool defer = false);
int occupancy_kerr(int precision, bool adaptive, bool inv, int variant, int block, int* blocks_per_cu,
                   bool time = false, bool defer = false);

// A cleared work-queue counter on the current device (cleared by a kernel
// on `stream`, so the call may be graph-captured).
int take_work_slot(hipStream_t stream, unsigned long long** slot);

// Writes one frame record and clears its work counter on `stream` (a
// kernel, so the call may be graph-captured).
int write_frame_record(const BatchFrame& rec, BatchFrame* d_dst, unsigned long long* d_counter,
                       hipStream_t stream);

// The same for a progressive pass whose first sample id is read from
// *d_next_sample, which then advances by spp (all on `stream`, kernels only).
int write_pass_record(const BatchFrame& rec, BatchFrame* d_dst, unsigned long long* d_counter,
                      uint32_t* d_next_sample, int spp, hipStream_t stream);
int set_device_u32(uint32_t* d_p, uint32_t v, hipStream_t stream);

}  // namespace np
