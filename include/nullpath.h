/*
 * nullpath.h -- C ABI of libnullpath.so, the MI355X (gfx950) null-geodesic
 * library.  Plain pointers and sizes only; device pointers are HIP device
 * memory, `stream` is a hipStream_t passed as void* (NULL = default stream).
 *
 * Every entry point returns NP_OK (0) or a negative np_status code and
 * never aborts the process.
 */
#ifndef NULLPATH_H
#define NULLPATH_H

#include <stddef.h>
#include <stdint.h>
#include "np_render.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum np_status {
    NP_OK = 0,
    NP_EINVAL = -1,    /* bad argument (sizes, pointers, parameters) */
    NP_EALLOC = -2,    /* device or pinned-host allocation failed */
    NP_ELAUNCH = -3,   /* kernel launch or execution failed */
    NP_ENODEV = -4     /* no HIP device */
} np_status;

/* Human-readable text for a status code (static storage). */
const char* np_status_string(int code);

/* Counters a tracer launch adds to (device memory, caller-zeroed, may be
 * NULL).  Steps are accepted integration steps; attempts include rejected
 * adaptive attempts; rhs_evals counts Hamilton right-hand-side evaluations
 * actually performed. */
typedef struct np_trace_counters {
    uint64_t rays;
    uint64_t steps;
    uint64_t attempts;
    uint64_t rhs_evals;
    uint64_t captured;
    uint64_t escaped;
    uint64_t stepcap;
    uint64_t disk_hits;
} np_trace_counters;

/* Optional per-sample diagnostic record (extension; the reference's kernel
 * keeps these in registers only). */
typedef struct np_ray_record {
    int32_t cls;       /* np_ray_class */
    int32_t orders;    /* disk crossings accumulated */
    int32_t steps;     /* accepted integration steps */
    int32_t attempts;  /* integrator attempts (rejections included) */
    double  min_r;     /* closest approach, absolute units */
    double  r_end;     /* r when the ray stopped */
} np_ray_record;

/*
 * Trace `spp` samples (sample ids sample_offset .. sample_offset+spp-1) of
 * every pixel in rows [y0, y0+rows) of the frame described by *p, and write
 * each pixel's colour sum to d_rgb (device, rows*width*3 floats, row-major,
 * RGB interleaved: the float3 tile of the reference).  Rows past p->height
 * are left untouched.
 *
 * Replaces the renderKernel launch of render.cu:118-123 / 540-550 (same
 * per-pixel sample loop and colour rules, render.cu:124-363).  d_inv_max3
 * (optional, 3 floats) receives atomic maxima of |H|, dE/E, dL/L when
 * p->check_inv is set (render.cu:233-244, 327-345).  d_records (optional,
 * rows*width*spp entries, pixel-major then sample) receives one
 * np_ray_record per traced sample.
 */
int np_render_tile(const np_render_params* p, int y0, int rows, int spp, int sample_offset,
                   float* d_rgb, float* d_inv_max3, np_trace_counters* d_counters,
                   np_ray_record* d_records, void* stream);

/*
 * Progressive render of a whole frame into a host accumulation buffer
 * h_accum (width*height*3 floats, caller-zeroed or holding earlier passes),
 * `passes` passes of `samples_per_pass` samples starting at sample id
 * spp_done, accumulated in pass order as render.cu:523-580 does.
 * inv_max (optional, 3 floats) receives the invariant maxima.
 */
int np_render_accumulate(const np_render_params* p, int spp_done, int samples_per_pass,
                         int passes, int tile_h, float* h_accum, float* inv_max);

#ifdef __cplusplus
}
#endif

#endif /* NULLPATH_H */
