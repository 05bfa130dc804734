// Procedural aperture image on MI355X.
//
// Reference: aperture_shape / aperture_grating / aperture_scratches /
// aperture_dust (realflare/api/tasks/cl/aperture.cl:5-174) with geometry.cl
// and noise.cl, launched in sequence by ApertureTask.aperture
// (realflare/api/tasks/aperture.py:63-186).  Here one kernel evaluates the
// shape and every enabled mask per pixel, multiplying in the same order.
// The per-shape values that do not depend on the pixel (rotation angles,
// noise-placed centres) are evaluated once per workgroup into LDS with the
// same float expressions the source evaluates per pixel.
#include <cmath>
#include <cfloat>

#include "rf_common.h"

namespace rf {

static constexpr int AP_MAX_SHAPES = 1000;

__device__ __forceinline__ float wrap1(float v) { return fmodf(1.0f + fmodf(v, 1.0f), 1.0f); }

__device__ __forceinline__ float rectangle_sdf(float px, float py, float hx, float hy) {
#pragma clang fp contract(off)
    float ex = fabsf(px) - hx, ey = fabsf(py) - hy;
    float ox = fmaxf(ex, 0.0f), oy = fmaxf(ey, 0.0f);
    return sqrtf(ox * ox + oy * oy) + fminf(fmaxf(ex, ey), 0.0f);
}

__global__ __launch_bounds__(256) void k_aperture(rf_aperture_params p, float* __restrict__ out) {
#pragma clang fp contract(off)
    __shared__ float s_a[AP_MAX_SHAPES], s_b[AP_MAX_SHAPES], s_c[AP_MAX_SHAPES],
        s_d[AP_MAX_SHAPES];
    const float pi = 3.14159265358979323846f;
    const int w = p.width, h = p.height;
    const int px = blockIdx.x * 16 + (threadIdx.x & 15);
    const int py = blockIdx.y * 16 + (threadIdx.x >> 4);
    const bool inside = px < w && py < h;
    // ndc and scale (geometry.cl:1-25)
    const float nx = 2.0f * ((float)px / (float)(w - 1)) - 1.0f;
    const float ny = 2.0f * ((float)py / (float)(h - 1)) - 1.0f;
    const float sx = fmaxf(p.size[0], FLT_MIN), sy = fmaxf(p.size[1], FLT_MIN);
    const float qx = nx / sx, qy = ny / sy;

    // aperture_shape (aperture.cl:5-47)
    float intensity;
    {
        float s = sinf(p.rotation), c = cosf(p.rotation);
        float x = qx * c + qy * s, y = qy * c - qx * s;
        float sdf = 0.0f;
        for (int i = 0; i < p.blades; ++i) {
            float angle = ((float)i / (float)p.blades + 0.25f) * 2 * pi;
            sdf = fmaxf(sdf, cosf(angle) * x + sinf(angle) * y);
        }
        float circular = atan2f(-x, -y) / (2 * pi) + 0.5f;
        float blades_gradient = fmodf(circular * (float)p.blades + 0.5f, 1.0f);
        sdf += sinf(blades_gradient * pi) * p.roundness;
        intensity = 1 - smoothstepf(1 - p.softness, 1 + p.softness, sdf);
    }

    // aperture_grating (aperture.cl:50-86)
    if (p.grating_strength > 0) {
        const int count = (int)(fminf(p.grating_density, 1.0f) * 360);
        const float offset = 1.0f / (float)count * (2 * pi);
        if (threadIdx.x == 0) {
            float angle = 0.0f;
            for (int i = 0; i < count; i++) {
                angle += offset;
                s_a[i] = angle;
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < count; i += blockDim.x) {
            s_b[i] = sinf(s_a[i]);
            s_c[i] = cosf(s_a[i]);
        }
        __syncthreads();
        const float hx = p.grating_width / 2, hy = p.grating_length * 2 / 2;
        float sdf = 1.0f;
        for (int i = 0; i < count; i++) {
            float x = qx * s_c[i] + qy * s_b[i];
            float y = qy * s_c[i] - qx * s_b[i];
            y = y - 1.5f;
            sdf = fminf(sdf, rectangle_sdf(x, y, hx, hy));
        }
        intensity *= 1 - (p.grating_strength * smoothstepf(p.grating_softness, -p.grating_softness, sdf));
        __syncthreads();
    }

    // aperture_scratches (aperture.cl:88-136)
    if (p.scratches_strength > 0) {
        const int count = (int)(fminf(p.scratches_density, 1.0f) * 1000);
        const float rv = p.scratches_rotation_variation * pi;
        for (int i = threadIdx.x; i < count; i += blockDim.x) {
            float cx = noise3((float)i, (float)count, 0) + p.scratches_parallax[0];
            float cy = noise3((float)i, (float)count, 1) + p.scratches_parallax[1];
            s_a[i] = wrap1(cx) * 2 - 1;
            s_b[i] = wrap1(cy) * 2 - 1;
            float angle = (noise3((float)i, (float)count, 2) - 0.5f) * rv;
            s_c[i] = sinf(angle);
            s_d[i] = cosf(angle);
        }
        __syncthreads();
        float s = sinf(p.scratches_rotation), c = cosf(p.scratches_rotation);
        float x0 = qx * c + qy * s, y0 = qy * c - qx * s;
        const float hx = p.scratches_width / 2, hy = p.scratches_length / 2;
        float sdf = 1.0f;
        for (int i = 0; i < count; i++) {
            float x = x0 * s_d[i] + y0 * s_c[i];
            float y = y0 * s_d[i] - x0 * s_c[i];
            x = x - s_a[i];
            y = y - s_b[i];
            sdf = fminf(sdf, rectangle_sdf(x, y, hx, hy));
        }
        intensity *= 1 - (p.scratches_strength *
                          smoothstepf(p.scratches_softness, -p.scratches_softness, sdf));
        __syncthreads();
    }

    // aperture_dust (aperture.cl:139-174)
    if (p.dust_strength > 0) {
        const int count = (int)(fminf(p.dust_density, 1.0f) * 1000);
        for (int i = threadIdx.x; i < count; i += blockDim.x) {
            float cx = noise3((float)i, (float)count, 0) + p.dust_parallax[0];
            float cy = noise3((float)i, (float)count, 1) + p.dust_parallax[1];
            s_a[i] = wrap1(cx) * 2 - 1;
            s_b[i] = wrap1(cy) * 2 - 1;
        }
        __syncthreads();
        float sdf = 1.0f;
        for (int i = 0; i < count; i++) {
            float dx = qx - s_a[i], dy = qy - s_b[i];
            sdf = fminf(sdf, sqrtf(dx * dx + dy * dy) - p.dust_radius);
        }
        intensity *= 1.0f - (p.dust_strength * smoothstepf(p.dust_softness, -p.dust_softness, sdf));
    }
    if (inside) out[(long)py * w + px] = intensity;
}

}  // namespace rf

extern "C" int rf_aperture(const rf_aperture_params* prm, float* image, rf_stream_t stream) {
    using namespace rf;
    if (!prm || !image) return set_error(RF_ERR_INVALID, "rf_aperture: NULL");
    if (prm->width < 1 || prm->height < 1)
        return set_error(RF_ERR_INVALID, "rf_aperture: resolution %dx%d", prm->width, prm->height);
    if (prm->blades < 0) return set_error(RF_ERR_INVALID, "rf_aperture: blades < 0");
    dim3 grid((prm->width + 15) / 16, (prm->height + 15) / 16);
    hipLaunchKernelGGL(k_aperture, grid, dim3(256), 0, (hipStream_t)stream, *prm, image);
    RF_LAUNCH_CHECK("rf_aperture");
    return RF_OK;
}
