/*
 * realflare_amd: C ABI of the MI355X (gfx950) flare render path.
 *
 * Every function takes caller-owned DEVICE pointers, plain sizes and a
 * stream (a hipStream_t; NULL = the default stream), enqueues its kernels
 * and returns without synchronizing.  Nothing here allocates device memory:
 * functions that need scratch space take a workspace whose size the matching
 * *_workspace_size() query returns.  Every function returns RF_OK (0) or an
 * RF_ERR_* code; rf_last_error() then holds a message for the calling thread.
 *
 * Each entry point names the reference interface it replaces
 * (beatreichenbach/realflare v0.6.4, paths relative to that repository).
 */
#ifndef REALFLARE_AMD_H
#define REALFLARE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RF_ABI_VERSION 1

typedef void *rf_stream_t;

enum {
    RF_OK = 0,
    RF_ERR_INVALID = 1,   /* bad argument; realflare raises RealflareError */
    RF_ERR_HIP = 2,       /* HIP runtime / launch failure */
    RF_ERR_WORKSPACE = 3, /* workspace too small */
};

/* LensElement, 64 bytes (realflare/api/tasks/opencl.py:30-39). */
typedef struct rf_lens_element {
    float radius, ior, height, center;
    int32_t coating;
    int32_t pad_[3];
    float coefficients[8];
} rf_lens_element;

/* Scalar arguments of the `raytrace` kernel (raytracing.cl:218-232),
 * as RaytracingTask.raytrace sets them (raytracing.py:202-211). */
typedef struct rf_trace_params {
    int32_t lens_count;       /* elements incl. stop and sensor, <= 64 */
    int32_t path_count;       /* ghosts P */
    int32_t wavelength_count; /* W */
    int32_t aperture_index;
    float coating_min_ior;
    int32_t grid_count;       /* g; R = g*g rays per (ghost, wavelength) */
    float grid_length;
    float direction[4];       /* update_direction (raytracing.py:92-106) */
} rf_trace_params;

/* Message of the last failing call on this thread ("" if none). */
const char *rf_last_error(void);
int rf_abi_version(void);

/*
 * Ghost ray trace.  Replaces the `raytrace` kernel (raytracing.cl:218-359)
 * launched by RaytracingTask.raytrace (raytracing.py:140-221).
 *   lens        rf_lens_element[lens_count]
 *   paths       int32[path_count][2]   (bounce1, bounce2)
 *   wavelengths int32[wavelength_count] (nm)
 *   rays        Ray[P][W][R], 48-byte records (opencl.py:41-49)
 */
int rf_trace(const rf_lens_element *lens, const int32_t *paths, const int32_t *wavelengths,
             const rf_trace_params *prm, void *rays, rf_stream_t stream);

/*
 * Same trace, writing only what the rasterizer reads, as six float planes
 * planes[6][P*W*R]: pos.x, pos.y, pos_apt.x, pos_apt.y, rrel, reflectance.
 * Same kernel as rf_trace; the layout the render path uses.
 */
int rf_trace_planes(const rf_lens_element *lens, const int32_t *paths,
                    const int32_t *wavelengths, const rf_trace_params *prm, float *planes,
                    rf_stream_t stream);

/*
 * Trace that also records every surface hit.  Replaces the STORE_INTERSECTIONS
 * build of `raytrace` (raytracing.cl:229-232,268-271) used by IntersectionsTask
 * (raytracing.py:250-380).  intersections: Intersection[P][W][R][3L-1],
 * 48-byte records (opencl.py:20-27); the caller clears it first (the
 * reference copies zeros in, raytracing.py:115-123).
 */
int rf_trace_intersections(const rf_lens_element *lens, const int32_t *paths,
                           const int32_t *wavelengths, const rf_trace_params *prm, void *rays,
                           void *intersections, rf_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* REALFLARE_AMD_H */
