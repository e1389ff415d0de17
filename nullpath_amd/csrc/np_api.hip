// np_api.hip -- the C ABI of libnullpath.so (include/nullpath.h): argument
// checks, per-device work-queue slots, persistent-grid sizing, and the
// progressive host driver.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "np_frame.hpp"
#include "np_kernels.hpp"
#include "nullpath.h"

namespace np {

namespace {

constexpr int kMaxDevices = 16;
constexpr int kWorkSlots = 256;  // queue counters per device, used round-robin

struct DeviceState {
    std::once_flag once;
    int status = NP_OK;
    unsigned long long* slots = nullptr;  // kWorkSlots counters
    std::atomic<unsigned> next{0};
    int cus = 0;
};

DeviceState g_dev[kMaxDevices];

__global__ void zero_u64(unsigned long long* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

int current_device(int* dev) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return NP_ENODEV;
    if (hipGetDevice(dev) != hipSuccess) return NP_ENODEV;
    if (*dev < 0 || *dev >= kMaxDevices) return NP_ENODEV;
    return NP_OK;
}

DeviceState* device_state(int* err) {
    int dev = 0;
    *err = current_device(&dev);
    if (*err != NP_OK) return nullptr;
    DeviceState& d = g_dev[dev];
    std::call_once(d.once, [&d, dev]() {
        if (hipDeviceGetAttribute(&d.cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            d.cus <= 0) {
            d.status = NP_ENODEV;
            return;
        }
        if (hipMalloc(&d.slots, kWorkSlots * sizeof(unsigned long long)) != hipSuccess) {
            d.status = NP_EALLOC;
            d.slots = nullptr;
        }
    });
    *err = d.status;
    return d.status == NP_OK ? &d : nullptr;
}

// Take a queue counter and clear it on `stream` (a kernel, not a memset,
// so the call can be captured into a graph).
int take_slot(DeviceState* d, hipStream_t stream, unsigned long long** slot) {
    unsigned k = d->next.fetch_add(1u) % kWorkSlots;
    *slot = d->slots + k;
    hipLaunchKernelGGL(zero_u64, dim3(1), dim3(64), 0, stream, *slot, 1);
    return hipGetLastError() == hipSuccess ? NP_OK : NP_ELAUNCH;
}

int persistent_grid(DeviceState* d, int precision, bool adaptive, bool inv, int variant,
                    uint64_t items) {
    int per_cu = 0;
    if (occupancy_kerr(precision, adaptive, inv, variant, &per_cu) != NP_OK || per_cu <= 0) per_cu = 1;
    uint64_t grid = (uint64_t)per_cu * (uint64_t)d->cus;
    uint64_t need = (items + 255) / 256;
    if (need < grid) grid = need;
    if (grid < 1) grid = 1;
    return (int)grid;
}

int render_launch(const np_render_params* p, int y0, int rows, int spp, int sample_offset,
                  float* d_rgb, int accumulate, float* d_inv, np_trace_counters* d_counters,
                  np_ray_record* d_records, hipStream_t stream) {
    if (!p || !d_rgb || spp < 0 || rows < 0 || y0 < 0 || sample_offset < 0) return NP_EINVAL;
    if (p->precision != NP_PREC_FP64 && p->precision != NP_PREC_FP32C) return NP_EINVAL;
    FrameConst F;
    int err = make_frame(p, &F);
    if (err != NP_OK) return err;
    DeviceState* d = device_state(&err);
    if (!d) return err;
    const int y_end = std::min(y0 + rows, p->height);
    const uint64_t npix = y_end > y0 ? (uint64_t)(y_end - y0) * (uint64_t)p->width : 0ull;
    if (npix == 0 || spp == 0) return NP_OK;
    unsigned long long* slot = nullptr;
    err = take_slot(d, stream, &slot);
    if (err != NP_OK) return err;
    const bool adaptive = p->integrator == NP_INTEGRATOR_ADAPTIVE;
    const int grid = persistent_grid(d, p->precision, adaptive, p->check_inv != 0, F.variant, npix);
    return launch_kerr(F, p->precision, adaptive, y0, rows, spp, sample_offset, d_rgb, accumulate,
                       d_inv, d_counters, d_records, slot, stream, grid);
}

}  // namespace

int take_work_slot(hipStream_t stream, unsigned long long** slot) {
    int err = NP_OK;
    DeviceState* d = device_state(&err);
    if (!d) return err;
    return take_slot(d, stream, slot);
}

}  // namespace np

extern "C" {

const char* np_status_string(int code) {
    switch (code) {
        case NP_OK: return "ok";
        case NP_EINVAL: return "invalid argument";
        case NP_EALLOC: return "allocation failed";
        case NP_ELAUNCH: return "kernel launch or execution failed";
        case NP_ENODEV: return "no HIP device";
        default: return "unknown status";
    }
}

int np_render_tile(const np_render_params* p, int y0, int rows, int spp, int sample_offset,
                   float* d_rgb, float* d_inv_max3, np_trace_counters* d_counters,
                   np_ray_record* d_records, void* stream) {
    return np::render_launch(p, y0, rows, spp, sample_offset, d_rgb, 0, d_inv_max3, d_counters,
                             d_records, (hipStream_t)stream);
}

int np_render_accumulate(const np_render_params* p, int spp_done, int samples_per_pass, int passes,
                         int tile_h, float* h_accum, float* inv_max) {
    (void)tile_h;  // the persistent grid covers the whole frame in one launch
    if (!p || !h_accum || samples_per_pass <= 0 || passes < 0 || spp_done < 0) return NP_EINVAL;
    np::FrameConst F;
    int err = np::make_frame(p, &F);
    if (err != NP_OK) return err;
    const size_t n = (size_t)p->width * (size_t)p->height * 3;
    float* d_acc = nullptr;
    float* d_inv = nullptr;
    if (hipMalloc(&d_acc, n * sizeof(float)) != hipSuccess) return NP_EALLOC;
    if (hipMalloc(&d_inv, 3 * sizeof(float)) != hipSuccess) {
        (void)hipFree(d_acc);
        return NP_EALLOC;
    }
    err = NP_OK;
    if (hipMemcpy(d_acc, h_accum, n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(d_inv, 0, 3 * sizeof(float)) != hipSuccess)
        err = NP_ELAUNCH;
    for (int k = 0; k < passes && err == NP_OK; ++k) {
        err = np::render_launch(p, 0, p->height, samples_per_pass,
                                spp_done + k * samples_per_pass, d_acc, 1, d_inv, nullptr, nullptr,
                                nullptr);
    }
    if (err == NP_OK && hipDeviceSynchronize() != hipSuccess) err = NP_ELAUNCH;
    if (err == NP_OK && hipMemcpy(h_accum, d_acc, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        err = NP_ELAUNCH;
    if (err == NP_OK && inv_max) {
        float h_inv[3];
        if (hipMemcpy(h_inv, d_inv, sizeof(h_inv), hipMemcpyDeviceToHost) != hipSuccess)
            err = NP_ELAUNCH;
        else
            for (int i = 0; i < 3; ++i) inv_max[i] = std::fmax(inv_max[i], h_inv[i]);
    }
    (void)hipFree(d_acc);
    (void)hipFree(d_inv);
    return err;
}

}  // extern "C"
