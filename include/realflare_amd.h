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

#define RF_ABI_VERSION 3  /* 2: rf_trace_fit takes huber_delta; 3: its lost-ray count */

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

/*
 * rf_trace_planes with the coating wavelength of every element taken from
 * coating_lambda[lens_count] (float, nm) instead of the element records, so
 * the coating is a continuous parameter.  Same kernel as rf_trace_planes; the
 * forward pass of the differentiable trace.  The coating value is the
 * element's `coating` field that raytracing.cl:310-318 reads.
 */
int rf_trace_planes_coated(const rf_lens_element *lens, const float *coating_lambda,
                           const int32_t *paths, const int32_t *wavelengths,
                           const rf_trace_params *prm, float *planes, rf_stream_t stream);

/*
 * Adjoint (reverse mode) of the ghost trace.  No counterpart in the
 * reference, which has no gradients: it differentiates the `raytrace` kernel
 * (raytracing.cl:218-359) as rf_trace_planes_coated evaluates it.
 *   grad_pos_x, grad_pos_y, grad_reflectance   float[P*W*R] upstream gradients
 *     of the sensor position and reflectance of every ray (ignored for rays
 *     that do not reach the sensor)
 *   grad_out   float[3L + W*L]: d/d radius[L], d/d centre[L] (element.center,
 *     raytracing.cl:104-168), d/d coating_lambda[L], then d/d ior[w][L], the
 *     refractive index each element has at wavelength w
 *   forward_out  optional float[3][P*W*R]: x, y, reflectance of the forward
 *     pass (NaN where the ray does not reach the sensor)
 *   checkpoint_every  K in {1, 2, 3, 4, 6, 8}: ray state is stored every K
 *     surface events and the events in between are recomputed; K does not
 *     change the result
 * The result is deterministic.  rf_trace_adjoint runs the forward pass and the
 * reverse pass in one launch; rf_trace_forward_checkpointed followed by
 * rf_trace_backward splits them (the forward pass leaves its checkpoints in
 * the workspace, which must be passed unchanged, with the same arguments, to
 * rf_trace_backward) and gives bit-identical results.
 */
size_t rf_trace_adjoint_workspace_size(const rf_trace_params *prm, int checkpoint_every);
int rf_trace_adjoint(const rf_lens_element *lens, const float *coating_lambda,
                     const int32_t *paths, const int32_t *wavelengths,
                     const rf_trace_params *prm, int checkpoint_every, const float *grad_pos_x,
                     const float *grad_pos_y, const float *grad_reflectance, void *workspace,
                     size_t workspace_bytes, float *grad_out, float *forward_out,
                     rf_stream_t stream);
/* Forward half: forward_out (required) as above, checkpoints into workspace. */
int rf_trace_forward_checkpointed(const rf_lens_element *lens, const float *coating_lambda,
                                  const int32_t *paths, const int32_t *wavelengths,
                                  const rf_trace_params *prm, int checkpoint_every,
                                  void *workspace, size_t workspace_bytes, float *forward_out,
                                  rf_stream_t stream);
/* Reverse half, from the checkpoints rf_trace_forward_checkpointed left. */
int rf_trace_backward(const rf_lens_element *lens, const float *coating_lambda,
                      const int32_t *paths, const int32_t *wavelengths,
                      const rf_trace_params *prm, int checkpoint_every, const float *grad_pos_x,
                      const float *grad_pos_y, const float *grad_reflectance, void *workspace,
                      size_t workspace_bytes, float *grad_out, rf_stream_t stream);

/*
 * Batched form.  No counterpart in the reference, which renders one light
 * position with one grid size for every ghost: a batch holds many frames
 * (light positions), each ghost with its own grid, traced in one persistent
 * launch and rasterized per group of frames.
 *
 * A segment is one (frame, ghost) pair.  Its rays are stored contiguously from
 * ray_offset, wavelength-major ([W][g*g], the layout of one ghost of
 * rf_trace_planes), its vertices likewise from ray_offset ([g*g][W], the layout
 * of one ghost of rf_vertex_shader), its quads from quad_offset ((g-1)^2 of
 * them).  Segments are ordered by frame, then ghost; offsets increase.
 */
typedef struct rf_batch_segment {
    int64_t ray_offset;
    int64_t quad_offset;
    int32_t frame;        /* output frame */
    int32_t path;         /* index into the paths array */
    int32_t lens;         /* index into the lens table */
    int32_t grid_count;   /* g of this ghost, >= 2 */
    float grid_length;
    float direction[3];   /* the frame's light direction (update_direction) */
} rf_batch_segment;       /* 48 bytes */

/* A run of consecutive segments (all of them, or the frames of one raster
 * pass).  segments is a DEVICE pointer to the run's first segment. */
typedef struct rf_batch {
    const rf_batch_segment *segments;
    int32_t segment_count;
    int32_t frame_base;   /* frame of the run's first segment */
    int32_t frame_count;  /* frames frame_base .. frame_base + frame_count - 1 */
    int32_t reserved_;
    int64_t ray_base;     /* ray_offset of segments[0] */
    int64_t ray_count;    /* rays of the run (W * g^2 summed) */
    int64_t quad_base;    /* quad_offset of segments[0] */
    int64_t quad_count;   /* quads of the run */
    int64_t ray_total;    /* rays of the whole trace output (the plane stride) */
} rf_batch;

/*
 * Trace every segment of `batch` in one persistent launch: `workgroups`
 * workgroups of 256 lanes take chunks of 256 * rays_per_lane rays of one
 * (segment, wavelength) from a queue until it is empty.
 *   lenses       rf_lens_element records of all lens instances, concatenated
 *   lens_table   int32[instances][3]: first element, lens_count, aperture_index
 *   chunk_offsets uint64[segment_count + 1] (device): exclusive prefix sum of
 *                W * ceil(g^2 / (256 * rays_per_lane)) over the segments
 *   workspace    >= rf_trace_batch_workspace_size() bytes (the queue head; it is
 *                reset by the call, on the stream)
 *   planes       float[6][ray_total], as rf_trace_planes per segment
 * Same per-ray arithmetic as rf_trace_planes.
 */
size_t rf_trace_batch_workspace_size(void);
int rf_trace_batch(const rf_lens_element *lenses, const int32_t *lens_table,
                   const int32_t *paths, const int32_t *wavelengths, int32_t wavelength_count,
                   float coating_min_ior, const rf_batch *batch, const uint64_t *chunk_offsets,
                   int32_t rays_per_lane, int32_t workgroups, void *workspace,
                   size_t workspace_bytes, float *planes, rf_stream_t stream);

/*
 * One gradient step of a batched lens fit.  No counterpart in the reference
 * (no gradients): every segment's rays are traced as rf_trace_planes_coated
 * traces them, scored against targets, and differentiated, all in one
 * persistent launch.  Segments may use different lens instances
 * (rf_batch_segment.lens indexes lens_table); they must be ordered by
 * instance, instance i owning segments lens_segment_offsets[i] ..
 * lens_segment_offsets[i+1]-1 (device uint64[lens_instances + 1]).
 *   lenses / coatings   records and coating wavelengths of all instances,
 *                       concatenated; lens_table int32[instances][3] = first
 *                       element, lens_count, aperture_index
 *   chunk_offsets       uint64[segment_count + 1] (device): exclusive prefix
 *                       sum of W * ceil(g^2 / (64 * rays_per_lane)); chunk_count
 *                       its last value
 *   target_*            float[ray_total] per ray (NaN: no target)
 *   loss per ray        loss_scale * (f(dx) + f(dy) + f(dr / reflectance_scale))
 *                       for rays that reach the sensor in trace and target, with
 *                       f(r) = r^2, or with huber_delta > 0 and |r| > huber_delta
 *                       f(r) = 2 huber_delta |r| - huber_delta^2
 *   grad_out            float[instances][(3 + W) * max_elements + 2]: d/d radius,
 *                       d/d centre, d/d coating (max_elements each), d/d ior
 *                       for each wavelength, the instance's loss, then the
 *                       number of its rays that have a target but that the
 *                       trace loses (stopped, or at a non-finite position;
 *                       counted exactly, stored as the nearest float)
 *   forward_out         optional float[3][ray_total]: x, y, reflectance (NaN
 *                       where the ray does not reach the sensor)
 * Deterministic: partials are per chunk and reduced in a fixed order.
 */
size_t rf_trace_fit_workspace_size(int32_t max_elements, int32_t checkpoint_every,
                                   int32_t workgroups, uint64_t chunk_count);
int rf_trace_fit(const rf_lens_element *lenses, const float *coatings, const int32_t *lens_table,
                 const int32_t *paths, const int32_t *wavelengths, int32_t wavelength_count,
                 float coating_min_ior, const rf_batch *batch, const uint64_t *chunk_offsets,
                 uint64_t chunk_count, const uint64_t *lens_segment_offsets,
                 int32_t lens_instances, int32_t max_elements, int32_t rays_per_lane,
                 int32_t checkpoint_every, int32_t workgroups, const float *target_x,
                 const float *target_y, const float *target_reflectance, float reflectance_scale,
                 float loss_scale, float huber_delta, void *workspace, size_t workspace_bytes,
                 float *grad_out, float *forward_out, rf_stream_t stream);

/* Settings of the rasterizer kernels (rasterizing.cl:233-663), as
 * RasterizingTask.rasterize derives them (rasterizing.py:296-421). */
typedef struct rf_raster_params {
    int32_t path_count, wavelength_count, grid_count;
    int32_t width, height;        /* render.resolution */
    int32_t bin_size;             /* render.bin_size: which quads a pixel visits */
    int32_t anti_aliasing;        /* 1, 2, 4 or 8 samples per pixel */
    int32_t wavelength_sub_count; /* 1 when wavelength_count == 1 (py:390-392) */
    float area_orig;              /* (grid_length / (grid_count - 1))^2 (py:152-156) */
    float min_area;               /* lens.min_area * area_orig (py:329) */
    float screen_transform;       /* width / half sensor diagonal (py:158-167) */
    float intensity;              /* light.intensity * 1e3 (py:415) */
    float ghost_scale;            /* 1 - lens.fstop / 32 (py:393) */
    int32_t ghost_width, ghost_height; /* at most 2^28 texels, each side below 2^24 - 4 */
    int32_t spectrum_count;       /* texels of the light spectrum, RGBA each */
    int32_t accumulate;           /* 0: write the image (the reference); 1: add into it */
    int32_t super_block;          /* 8x8-pixel tiles per side of the groups the rasterizer
                                     keeps on one XCD: 1, 2, 4 or 8; 0 = 4.  No effect on
                                     the result. */
    int32_t reserved_;
} rf_raster_params;

/*
 * Per-quad intensities and bounds.  Replaces `prim_shader` (rasterizing.cl:233-320,
 * host rasterizing.py:321-344).
 *   planes      float[6][P*W*R] from rf_trace_planes
 *   bounds      float4[P][Q]   (x0, y0, x1, y1) in sensor mm, x0 = NaN if culled
 *   intensities float[P][Q][W] (entries the reference leaves unwritten are 0)
 */
int rf_prim_shader(const float *planes, const rf_raster_params *prm, float *bounds,
                   float *intensities, rf_stream_t stream);

/*
 * Screen-space vertices.  Replaces `vertex_shader` (rasterizing.cl:322-386,
 * host rasterizing.py:348-362).  vertexes: Vertex[P][R][W], 32-byte records
 * (opencl.py:51-61); fields of rays that missed the sensor are 0 except
 * reflectance = NaN.
 */
int rf_vertex_shader(const float *planes, const float *intensities, const rf_raster_params *prm,
                     void *vertexes, rf_stream_t stream);

/*
 * Binning + rasterization.  Replaces `binner` and `rasterizer`
 * (rasterizing.cl:388-663, host rasterizing.py:366-418, the image clear at
 * py:263-266 included).  Every pixel sums the fragments of the quads the
 * reference's binner lists for its bin, in the reference's order.
 *   ghost           float[ghost_height][ghost_width] (GhostTask output)
 *   light_spectrum  float[spectrum_count][4] (CIE XYZ, rasterizing.py:169-179)
 *   image           float[height][width][4], fully written
 * pair_capacity bounds the (8x8 tile, quad step) pairs a frame may bin; the
 * workspace (rf_rasterize_workspace_size) grows with it.  If a frame needs
 * more, the image is cleared to 0 (left as it was when accumulating) and
 * rf_raster_status reports overflow and the count needed; call again with a
 * larger capacity.
 */
size_t rf_rasterize_workspace_size(const rf_raster_params *prm, uint64_t pair_capacity);
int rf_rasterize(const void *vertexes, const float *bounds, const float *ghost,
                 const float *light_spectrum, const rf_raster_params *prm,
                 uint64_t pair_capacity, void *workspace, size_t workspace_bytes, float *image,
                 rf_stream_t stream);

/*
 * rf_rasterize in its two halves, on the same workspace: rf_raster_bin is
 * the `binner` (rasterizing.cl:388-469, host py:366-385) and rf_raster_draw
 * the `rasterizer` (rasterizing.cl:524-663, host py:387-418).  rf_raster_draw
 * reads what the last rf_raster_bin on that workspace left there.
 */
int rf_raster_bin(const void *vertexes, const float *bounds, const float *light_spectrum,
                  const rf_raster_params *prm, uint64_t pair_capacity, void *workspace,
                  size_t workspace_bytes, rf_stream_t stream);
int rf_raster_draw(const void *vertexes, const float *bounds, const float *ghost,
                   const rf_raster_params *prm, uint64_t pair_capacity, void *workspace,
                   size_t workspace_bytes, float *image, rf_stream_t stream);

/*
 * The whole RasterizingTask.run (rasterizing.py:423-439): prim_shader,
 * vertex_shader, binner and rasterizer from trace planes to image, with all
 * intermediates in the workspace (rf_flare_workspace_size).
 */
size_t rf_flare_workspace_size(const rf_raster_params *prm, uint64_t pair_capacity);
int rf_flare(const float *planes, const float *ghost, const float *light_spectrum,
             const rf_raster_params *prm, uint64_t pair_capacity, void *workspace,
             size_t workspace_bytes, float *image, rf_stream_t stream);

/*
 * rf_flare with no workspace argument and no allocation: the workspace is a
 * static device array of RF_FLARE_STATIC_SCRATCH_BYTES (256 MiB) that the
 * library holds on each device it runs on until the process ends.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  The pair capacity is the largest the array holds; a
 * frame whose workspace exceeds the array even with no pairs fails with
 * RF_ERR_WORKSPACE before any launch.  status (device memory, may be NULL)
 * receives {pairs the frame needed, 1 if they overflowed the capacity else
 * 0}; on overflow the image is cleared, as with rf_flare.  Every call on one
 * device shares that device's array, which keeps the last call's contents:
 * calls that overlap on one device (different streams or threads) race, so
 * serialize them.
 */
#define RF_FLARE_STATIC_SCRATCH_BYTES ((size_t)256 << 20)
int rf_flare_static(const float *planes, const float *ghost, const float *light_spectrum,
                    const rf_raster_params *prm, float *image, uint64_t *status,
                    rf_stream_t stream);

/*
 * Reads back (synchronizing `stream`) the pair count the last rf_rasterize /
 * rf_flare on this workspace needed and whether it overflowed the capacity.
 * No reference counterpart: the reference sizes its bin queues on the host
 * (rasterizing.py:190-208).
 */
int rf_raster_status(const void *workspace, uint64_t *required_pairs, int *overflow,
                     rf_stream_t stream);

/*
 * Batched rasterization of a run of segments (see rf_batch).  prm supplies
 * the frame settings; its path_count, grid_count and area_orig are ignored,
 * and min_area holds lens.min_area unscaled: each segment's area_orig is
 * (grid_length / (g - 1))^2 (rasterizing.py:152-156) and its minimum area
 * min_area * area_orig (py:329).  Arrays are indexed as the segments say:
 *   planes    float[6][ray_total] (rf_trace_batch)
 *   bounds    float4[quads], intensities float[quads][W], by global quad index
 *   vertexes  Vertex[rays] by global ray index ([g*g][W] per segment)
 *   images    float4[frame_count][height][width], the run's frames in order
 * The prim / vertex passes cover the run's quads / rays; the binner and the
 * rasterizer bin (frame, tile) pairs of all the run's frames together, with
 * rf_raster_status reporting pairs and overflow as for one frame.
 */
int rf_prim_shader_batch(const float *planes, const rf_raster_params *prm, const rf_batch *batch,
                         float *bounds, float *intensities, rf_stream_t stream);
int rf_vertex_shader_batch(const float *planes, const float *intensities,
                           const rf_raster_params *prm, const rf_batch *batch, void *vertexes,
                           rf_stream_t stream);
size_t rf_raster_batch_workspace_size(const rf_raster_params *prm, const rf_batch *batch,
                                      uint64_t pair_capacity);
int rf_raster_bin_batch(const void *vertexes, const float *bounds, const float *light_spectrum,
                        const rf_raster_params *prm, const rf_batch *batch,
                        uint64_t pair_capacity, void *workspace, size_t workspace_bytes,
                        rf_stream_t stream);
int rf_raster_draw_batch(const void *vertexes, const float *bounds, const float *ghost,
                         const rf_raster_params *prm, const rf_batch *batch,
                         uint64_t pair_capacity, void *workspace, size_t workspace_bytes,
                         float *images, rf_stream_t stream);

/* Aperture settings as ApertureTask.aperture passes them to its kernels
 * (aperture.py:80-153): angles in radians, softness / 10, widths and the
 * dust radius * 0.1; a mask is applied when its strength is > 0. */
typedef struct rf_aperture_params {
    int32_t width, height;
    float size[2];
    int32_t blades;
    float rotation, roundness, softness;
    float grating_strength, grating_density, grating_length, grating_width, grating_softness;
    float scratches_strength, scratches_density, scratches_length, scratches_rotation,
        scratches_rotation_variation, scratches_width, scratches_parallax[2], scratches_softness;
    float dust_strength, dust_density, dust_radius, dust_parallax[2], dust_softness;
} rf_aperture_params;

/*
 * Aperture image float[height][width].  Replaces aperture_shape,
 * aperture_grating, aperture_scratches and aperture_dust
 * (aperture.cl:5-174) as ApertureTask.aperture runs them (aperture.py:63-186).
 */
int rf_aperture(const rf_aperture_params *prm, float *image, rf_stream_t stream);

/*
 * In-place 2-D FFT of interleaved complex64 data[height][width], unnormalized,
 * then multiplied by `scale`: numpy's fft2 is scale 1, ifft2 is inverse with
 * 1/(w h), norm='ortho' is 1/sqrt(w h).  Each side is a power of two up to
 * 4096 or any other length up to 2048 (Bluestein's chirp-z transform on a
 * power-of-two FFT of at least twice the length); larger sides are rejected
 * with RF_ERR_INVALID.  rf_power_spectrum and rf_ghost take the same sizes.
 * Replaces the numpy.fft calls of starburst.py:71 and frft.py:27-60.
 */
size_t rf_fft_workspace_size(int width, int height);
int rf_fft2(void *data, int width, int height, int inverse, float scale, void *workspace,
            size_t workspace_bytes, rf_stream_t stream);

/*
 * Starburst power spectrum |fftshift(fft2(A exp(i pi/(lambda d)(x^2+y^2))))|^2
 * as float[height][width].  Replaces StarburstTask.update_fourier_spectrum
 * (starburst.py:45-81); `distance` is Flare.Starburst.distance (m).
 */
size_t rf_power_spectrum_workspace_size(int width, int height);
int rf_power_spectrum(const float *aperture, int width, int height, float distance,
                      void *workspace, size_t workspace_bytes, float *spectrum,
                      rf_stream_t stream);

/*
 * Ghost texture |fftshift(frft(fftshift(A), alpha))| * sqrt(w h), alpha =
 * 0.15 (560/400) (fstop/18).  Replaces GhostTask.ghost (ghost.py:13-36) and
 * frft.frft (frft.py:43-62).
 */
size_t rf_ghost_workspace_size(int width, int height);
int rf_ghost(const float *aperture, int width, int height, float fstop, void *workspace,
             size_t workspace_bytes, float *ghost, rf_stream_t stream);

/* Starburst kernel settings as StarburstTask.starburst prepares them
 * (starburst.py:96-106, 131-141). */
typedef struct rf_starburst_params {
    int32_t width, height; /* render.resolution */
    int32_t samples;
    float blur;            /* Starburst.blur / 100 */
    float rotation;        /* radians */
    float rotation_weight;
    float vignetting[2];   /* NaN, NaN when vignetting is off */
    float intensity;       /* Starburst.intensity * 1e-6 */
    float offset[2];       /* (light.x, -light.y) */
    float size[2];         /* (scale.w, scale.h * width / height) */
} rf_starburst_params;

/*
 * Starburst image float[height][width][4].  Replaces the `starburst` kernel
 * (starburst.cl:1-87).  spectrum: rf_power_spectrum output, at most 2^28
 * values with sides below 2^24 (RF_ERR_INVALID otherwise);
 * light_spectrum: float[spectrum_count][4] CIE XYZ.
 */
int rf_starburst(const float *spectrum, int spectrum_width, int spectrum_height,
                 const float *light_spectrum, int spectrum_count,
                 const rf_starburst_params *prm, float *image, rf_stream_t stream);

/*
 * FLARE_STARBURST composition, out = flare + (flare + starburst) over RGBA
 * pixels, the reference's `array = flare.copy(); array += flare + starburst`
 * (engine.py:174-187), which counts the flare twice.
 */
int rf_compose_flare_starburst(const float *flare, const float *starburst, long pixel_count,
                               float *out, rf_stream_t stream);

/*
 * Image-based light composition.  Replaces the accumulation loop of
 * Engine.image_flare (engine.py:120-151): out (float4[height][width]) gets,
 * for each of `count` frames (float4[count][height][width], the flare of one
 * light sample) and each q in 0..3, weights[frame][q] (RGBA) times the frame
 * as rendered (q = 0), flipped in rows (1), in rows and columns (2), in
 * columns (3), in that order; accumulate = 1 adds onto out's previous
 * contents (to compose in several calls), and the result is divided by
 * `divisor` at the end of the call (the reference's final division by the
 * sample count w * h, given with the last call; 1 elsewhere).
 */
int rf_image_light_compose(const float *frames, int32_t count, int32_t width, int32_t height,
                           const float *weights, float divisor, int32_t accumulate, float *out,
                           rf_stream_t stream);

/*
 * Lens diagram, first pass.  Replaces the `lenses` kernel
 * (diagram.cl:21-68) as DiagramTask.lenses launches it (diagram.py:113-158):
 * writes every pixel of image (float4[height][width]) with the optical axis
 * in red, flat elements (radius 0) as a column (the stop and the last
 * element green, others blue) and spherical elements as a half circle in
 * blue; scale is pixels per lens unit ((width - 10) / total element distance,
 * diagram.py:44-56).
 */
int rf_diagram_lenses(const rf_lens_element *lens, int32_t lens_count, int32_t aperture_index,
                      float scale, int32_t width, int32_t height, float *image,
                      rf_stream_t stream);

/*
 * Lens diagram, second pass.  Replaces the `intersections` kernel
 * (diagram.cl:71-138) as DiagramTask.intersections launches it
 * (diagram.py:72-111): draws over image the surface normals (overwriting the
 * pixel with (0.25, 0, 0, 0)) and the ray segments (adding 0.25 to R, G, B)
 * of ray_count rays.  intersections is Intersection[ray_count][records_per_ray]
 * (48-byte records as rf_trace_intersections writes them; the reference
 * passes the middle grid column of one ghost, diagram.py:58-70); a ray's
 * records stop at the first record with hit == 0.
 */
int rf_diagram_intersections(const void *intersections, int32_t records_per_ray,
                             int32_t ray_count, float scale, int32_t width, int32_t height,
                             float *image, rf_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* REALFLARE_AMD_H */
