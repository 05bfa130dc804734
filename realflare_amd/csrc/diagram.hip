// Lens diagram on MI355X: the DIAGRAM render element, a 2-D side view of the
// lens elements with the traced rays of one ghost drawn over it.
//
// Reference: the `lenses` and `intersections` kernels
// (realflare/api/tasks/cl/diagram.cl:21-68, 71-138) launched by DiagramTask
// (realflare/api/tasks/diagram.py:86-170).  One thread per pixel, as there;
// the per-pixel arithmetic follows the source operation by operation with
// contraction off, so the image is reproducible bit for bit on the CPU.
//
// The intersections pass visits the segments of every ray in order (a normal
// line overwrites the pixel, a ray line adds to it, so the order matters).  A
// segment's endpoints are the same for every pixel, so they come through
// scalar loads, and a workgroup skips a segment whose bounding box, widened by
// a margin that covers the rounding of the line test, misses all its pixels:
// such a segment cannot change any of them.
#include <cmath>

#include "rf_common.h"

namespace rf {

static constexpr int DG_BX = 32, DG_BY = 8;  // 256 threads, 4 waves of 32x2 pixels

struct DgRecord {  // Intersection, 48 B (opencl.py:20-27 as the device lays it out)
    float4 pos;
    float4 normal;
    float incident;
    int32_t hit;
    int32_t pad_[2];
};
static_assert(sizeof(DgRecord) == 48, "Intersection record is 48 bytes");

// diagram.cl:1-19: the pixel lies strictly between the segment's end caps and
// within width/2 of its line.
__device__ __forceinline__ bool is_line(float px, float py, float x0, float y0, float x1, float y1,
                                        float width) {
#pragma clang fp contract(off)
    const float vx = x1 - x0, vy = y1 - y0;
    const float ax = px - x0, ay = py - y0;
    const float bx = px - x1, by = py - y1;
    const bool between = (vx * ax + vy * ay) > 0.0f && (vx * bx + vy * by) < 0.0f;
    const float d = fabsf(vx * ay - ax * vy) / sqrtf(vx * vx + vy * vy);
    return between && d <= width / 2.0f;
}

// True when no pixel of the block [bx0, bx1] x [by0, by1] can pass is_line
// for this segment (width 1).  A pixel that passes lies within 1/2 of the
// segment up to rounding of a few ulp of its coordinates; the margin covers
// both.  Any NaN or infinity makes is_line false everywhere, so whatever
// this returns for them is safe.
__device__ __forceinline__ bool misses_block(float x0, float y0, float x1, float y1, float bx0,
                                             float bx1, float by0, float by1) {
    const float lo_x = x0 < x1 ? x0 : x1, hi_x = x0 < x1 ? x1 : x0;
    const float lo_y = y0 < y1 ? y0 : y1, hi_y = y0 < y1 ? y1 : y0;
    const float mag = fmaxf(fmaxf(fabsf(lo_x), fabsf(hi_x)), fmaxf(fabsf(lo_y), fabsf(hi_y)));
    const float m = 1.0f + 1e-4f * fmaxf(mag, fmaxf(fmaxf(fabsf(bx0), fabsf(bx1)),
                                                   fmaxf(fabsf(by0), fabsf(by1))));
    return lo_x - m > bx1 || hi_x + m < bx0 || lo_y - m > by1 || hi_y + m < by0;
}

__global__ __launch_bounds__(DG_BX* DG_BY) void k_diagram_lenses(
    const rf_lens_element* __restrict__ lens, int lens_count, int aperture_index, float scale,
    int width, int height, float4* __restrict__ image) {
#pragma clang fp contract(off)
    const int px = blockIdx.x * DG_BX + threadIdx.x;
    const int py = blockIdx.y * DG_BY + threadIdx.y;
    if (px >= width || py >= height) return;
    const float x = (float)px;
    const float y = (float)(height / 2 - py);
    float4 rgba = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if ((int)y == 0) rgba.x = 1.0f;  // optical axis
    for (int i = 0; i < lens_count; ++i) {
        const float r = lens[i].radius * scale;
        const float h = lens[i].height * scale;
        const float c = lens[i].center * scale;
        if (r == 0.0f && fabsf(y) < h && (int)x == (int)c) {
            // flat element: the stop and the sensor green, others blue
            if (i == aperture_index || i == lens_count - 1)
                rgba.y = 1.0f;
            else
                rgba.z = 1.0f;
            break;
        } else {
            const float dx = x - c;
            const float r_circle = sqrtf(dx * dx + y * y);
            const bool draw_half = (r > 0.0f) ? (x < c) : (x > c);
            if (draw_half && (int)fabsf(r) == (int)r_circle && fabsf(y) < h) rgba.z = 1.0f;
        }
    }
    image[(long)py * width + px] = rgba;
}

__global__ __launch_bounds__(DG_BX* DG_BY) void k_diagram_intersections(
    const DgRecord* __restrict__ recs, int records_per_ray, int ray_count, float scale, int width,
    int height, float4* __restrict__ image) {
#pragma clang fp contract(off)
    const int px = blockIdx.x * DG_BX + threadIdx.x;
    const int py = blockIdx.y * DG_BY + threadIdx.y;
    const bool inside = px < width && py < height;
    const float pos_x = (float)px;
    const float pos_y = (float)(height / 2 - py);
    // pixel-space bounds of this workgroup (identical for all its waves)
    const float bx0 = (float)(blockIdx.x * DG_BX);
    const float bx1 = (float)(blockIdx.x * DG_BX + DG_BX - 1);
    const float by1 = (float)(height / 2 - (int)(blockIdx.y * DG_BY));
    const float by0 = (float)(height / 2 - (int)(blockIdx.y * DG_BY + DG_BY - 1));
    float4 rgba = inside ? image[(long)py * width + px] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const long total = (long)ray_count * records_per_ray;
    for (int ray = 0; ray < ray_count; ++ray) {
        const long base = (long)ray * records_per_ray;
        for (int k = 0; k < records_per_ray; ++k) {
            const long i1 = base + k;
            // the record after a ray's last one is the next ray's first, as in
            // the reference's contiguous slice; past the slice there is no hit
            if (i1 + 1 >= total) break;
            const DgRecord& r1 = recs[i1];
            const DgRecord& r2 = recs[i1 + 1];
            if (!r2.hit) break;
            const float p0x = (r1.pos.z * scale) * -1.0f;
            const float p0y = (r1.pos.y * scale) * -1.0f;
            const float nx = r1.normal.z * scale;
            const float ny = r1.normal.y * scale;
            const float pnx = p0x - nx * 1.0f;
            const float pny = p0y - ny * 1.0f;
            if (!misses_block(p0x, p0y, pnx, pny, bx0, bx1, by0, by1) &&
                is_line(pos_x, pos_y, p0x, p0y, pnx, pny, 1.0f))
                rgba = make_float4(0.25f, 0.0f, 0.0f, 0.0f);
            const float p1x = (r2.pos.z * scale) * -1.0f;
            const float p1y = (r2.pos.y * scale) * -1.0f;
            if (p1x == 0.0f) break;
            if (!misses_block(p0x, p0y, p1x, p1y, bx0, bx1, by0, by1) &&
                is_line(pos_x, pos_y, p0x, p0y, p1x, p1y, 1.0f)) {
                rgba.x += 0.25f;
                rgba.y += 0.25f;
                rgba.z += 0.25f;
                rgba.w += 0.0f;
            }
        }
    }
    if (inside) image[(long)py * width + px] = rgba;
}

static dim3 dg_grid(int width, int height) {
    return dim3((unsigned)((width + DG_BX - 1) / DG_BX), (unsigned)((height + DG_BY - 1) / DG_BY));
}

}  // namespace rf

extern "C" int rf_diagram_lenses(const rf_lens_element* lens, int32_t lens_count,
                                 int32_t aperture_index, float scale, int32_t width,
                                 int32_t height, float* image, rf_stream_t stream) {
    using namespace rf;
    if (!image || width < 1 || height < 1 || lens_count < 0 || (lens_count > 0 && !lens))
        return set_error(RF_ERR_INVALID, "rf_diagram_lenses: bad arguments");
    hipLaunchKernelGGL(k_diagram_lenses, dg_grid(width, height), dim3(DG_BX, DG_BY), 0,
                       (hipStream_t)stream, lens, lens_count, aperture_index, scale, width,
                       height, (float4*)image);
    RF_LAUNCH_CHECK("rf_diagram_lenses");
    return RF_OK;
}

extern "C" int rf_diagram_intersections(const void* intersections, int32_t records_per_ray,
                                        int32_t ray_count, float scale, int32_t width,
                                        int32_t height, float* image, rf_stream_t stream) {
    using namespace rf;
    if (!image || width < 1 || height < 1 || records_per_ray < 0 || ray_count < 0 ||
        ((long)records_per_ray * ray_count > 0 && !intersections))
        return set_error(RF_ERR_INVALID, "rf_diagram_intersections: bad arguments");
    hipLaunchKernelGGL(k_diagram_intersections, dg_grid(width, height), dim3(DG_BX, DG_BY), 0,
                       (hipStream_t)stream, (const DgRecord*)intersections, records_per_ray,
                       ray_count, scale, width, height, (float4*)image);
    RF_LAUNCH_CHECK("rf_diagram_intersections");
    return RF_OK;
}
