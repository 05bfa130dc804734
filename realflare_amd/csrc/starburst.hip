// Starburst (diffraction) image on MI355X.
//
// Reference: the `starburst` kernel (realflare/api/tasks/cl/starburst.cl:1-87)
// with geometry.cl, noise.cl and color.cl, launched by
// StarburstTask.starburst (realflare/api/tasks/starburst.py:83-153).  One
// thread per output pixel takes `samples` jittered, wavelength-scaled
// bilinear lookups of the power spectrum, weights them by the CIE XYZ table
// and converts to ACEScg (AP1).  The XYZ weight of every sample depends only
// on the sample index, so it is evaluated once per workgroup into LDS.
#include <cfloat>
#include <cmath>

#include "rf_common.h"

namespace rf {

static constexpr int SB_MAX_CACHED = 1024;

// x / b correctly rounded, from ib = 1 / b correctly rounded (Markstein's
// correction: q = x ib is within an ulp of x / b, the remainder x - q b is
// exact with a fused multiply-add, and q + (x - q b) ib rounded once is the
// correctly rounded quotient).  Valid while no intermediate is subnormal or
// overflows: here b lies in [0.69, 1.31] (the wavelength scale) and x is 0
// or at least 2^-100 in magnitude.
__device__ __forceinline__ float div_rn(float x, float b, float ib) {
    const float q = x * ib;
    const float r = __builtin_fmaf(-q, b, x);
    return __builtin_fmaf(r, ib, q);
}


__global__ __launch_bounds__(256) void k_starburst(const float* __restrict__ spectrum, int sw, int sh,
                                                   const float* __restrict__ light, int light_n,
                                                   rf_starburst_params p, float4* __restrict__ out) {
#pragma clang fp contract(off)
    // per sample: the X, Y, Z weights and the wavelength scale, and the
    // scale's reciprocal (20 KiB at the cap, so eight workgroups fit a CU;
    // the table's fourth component never reaches the output)
    __shared__ float4 s_xyzs[SB_MAX_CACHED];
    __shared__ float s_isc[SB_MAX_CACHED];
    const float pi = 3.14159265358979323846f;
    const int samples = p.samples;
    const bool cached = samples <= SB_MAX_CACHED;
    if (cached) {
        for (int t = threadIdx.x; t < samples; t += blockDim.x) {
            const float step = (float)t / (float)samples;
            const float wavelength = step * (float)(730 - 390) + 390.0f;
            const float sc = fmaxf(1 - (1 - (wavelength / 560.0f)) * 1.0f, FLT_MIN);
            float4 v;
            v.x = sample_bilinear(light + 0, light_n, 1, 4, step, 0.0f, false);
            v.y = sample_bilinear(light + 1, light_n, 1, 4, step, 0.0f, false);
            v.z = sample_bilinear(light + 2, light_n, 1, 4, step, 0.0f, false);
            v.w = sc;
            s_xyzs[t] = v;
            s_isc[t] = 1.0f / sc;
        }
        __syncthreads();
    }
    const int w = p.width, h = p.height;
    const int px = blockIdx.x * 16 + (threadIdx.x & 15);
    const int py = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (px >= w || py >= h) return;
    float nx = 2.0f * ((float)px / (float)(w - 1)) - 1.0f;
    float ny = 2.0f * ((float)py / (float)(h - 1)) - 1.0f;
    nx = nx - p.offset[0];
    ny = ny - p.offset[1];
    nx = nx / fmaxf(p.size[0], FLT_MIN);
    ny = ny / fmaxf(p.size[1], FLT_MIN);
    const int lambda_delta = 730 - 390;
    float r0 = 0.0f, r1 = 0.0f, r2 = 0.0f;
    const float fx = (float)px, fy = (float)py;
    // With rotation 0 (the default, and a weight >= 0, so that the power
    // below is finite) the rotation angle is exactly +-0: cos = 1 and
    // sin = +-0 leave the position unchanged up to the sign of a zero, which
    // the texture coordinate 0.5 * x + 0.5 does not see, so the two noise,
    // power and trigonometric evaluations of the rotation are skipped.  With
    // blur 0 (the default) the jitter radius is exactly 0 and the jitter's two
    // noise and trigonometric evaluations are skipped as well.  Both skips
    // are bit-identical to evaluating them.
    const bool jitter = p.blur != 0.0f;
    const bool rotate = !(p.rotation == 0.0f && p.rotation_weight >= 0.0f);
    for (int t = 0; t < samples; ++t) {
        const float step = (float)t / (float)samples;
        float4 xyzs;  // X, Y, Z weights and the wavelength scale
        float isc;
        if (cached) {
            xyzs = s_xyzs[t];
            isc = s_isc[t];
        } else {
            const float wavelength = step * (float)lambda_delta + 390.0f;
            xyzs.w = fmaxf(1 - (1 - (wavelength / 560.0f)) * 1.0f, FLT_MIN);
            isc = 1.0f / xyzs.w;
            xyzs.x = sample_bilinear(light + 0, light_n, 1, 4, step, 0.0f, false);
            xyzs.y = sample_bilinear(light + 1, light_n, 1, 4, step, 0.0f, false);
            xyzs.z = sample_bilinear(light + 2, light_n, 1, 4, step, 0.0f, false);
        }
        const float sc = xyzs.w;
        const float seed = (float)(t * 4);
        float x = nx, y = ny;
        if (jitter) {
            const float radius = p.blur * sqrtf(noise3(fx, fy, seed));
            const float alpha = noise3(fx, fy, seed + 1) * 2 * pi;
            float sa, ca;
            sincos_reduced(alpha, &sa, &ca);
            x += ca * radius;
            y += sa * radius;
        }
        x = div_rn(x, sc, isc);
        y = div_rn(y, sc, isc);
        float rx = x, ry = y;
        if (rotate) {
            const float sign = (noise3(fx, fy, seed + 2) > 0.5f) ? 1.0f : -1.0f;
            const float angle =
                sign * p.rotation * (1.0f - powf(noise3(fx, fy, seed + 3), p.rotation_weight));
            float sa, ca;
            sincos_reduced(angle, &sa, &ca);
            rx = x * ca + y * sa;
            ry = y * ca - x * sa;
        }
        const float u = rx * 0.5f + 0.5f, v = ry * 0.5f + 0.5f;
        const float fi = sample_bilinear(spectrum, sw, sh, 1, u, v, false);
        r0 += xyzs.x * fi;
        r1 += xyzs.y * fi;
        r2 += xyzs.z * fi;
    }
    r0 /= (float)samples;
    r1 /= (float)samples;
    r2 /= (float)samples;
    if (!isnan(p.vignetting[0])) {
        const float sdf = sqrtf(nx * nx + ny * ny) - 0.0f;
        const float f = smoothstepf(p.vignetting[1], p.vignetting[0], sdf);
        r0 *= f;
        r1 *= f;
        r2 *= f;
    }
    r0 *= p.intensity;
    r1 *= p.intensity;
    r2 *= p.intensity;
    float o[4];
    xyz_to_ap1(r0, r1, r2, o);
    out[(long)py * w + px] = make_float4(o[0], o[1], o[2], o[3]);
}

}  // namespace rf

extern "C" int rf_starburst(const float* spectrum, int spectrum_width, int spectrum_height,
                            const float* light_spectrum, int spectrum_count,
                            const rf_starburst_params* prm, float* image, rf_stream_t stream) {
    using namespace rf;
    if (!prm || !spectrum || !light_spectrum || !image)
        return set_error(RF_ERR_INVALID, "rf_starburst: NULL");
    if (prm->width < 1 || prm->height < 1 || prm->samples < 1 || spectrum_width < 1 ||
        spectrum_height < 1 || spectrum_count < 1)
        return set_error(RF_ERR_INVALID, "rf_starburst: empty size or samples < 1");
    dim3 grid((prm->width + 15) / 16, (prm->height + 15) / 16);
    hipLaunchKernelGGL(k_starburst, grid, dim3(256), 0, (hipStream_t)stream, spectrum,
                       spectrum_width, spectrum_height, light_spectrum, spectrum_count, *prm,
                       (float4*)image);
    RF_LAUNCH_CHECK("rf_starburst");
    return RF_OK;
}

namespace rf {
// Engine.flare_starburst (engine.py:174-187): array = flare.copy();
// array += flare + starburst, i.e. flare + (flare + starburst).
__global__ __launch_bounds__(256) void k_compose(const float4* __restrict__ flare,
                                                 const float4* __restrict__ starburst, long n,
                                                 float4* __restrict__ out) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 f = flare[i], s = starburst[i];
    out[i] = make_float4(f.x + (f.x + s.x), f.y + (f.y + s.y), f.z + (f.z + s.z), f.w + (f.w + s.w));
}
}  // namespace rf

extern "C" int rf_compose_flare_starburst(const float* flare, const float* starburst,
                                          long pixel_count, float* out, rf_stream_t stream) {
    using namespace rf;
    if (!flare || !starburst || !out || pixel_count < 0)
        return set_error(RF_ERR_INVALID, "rf_compose_flare_starburst: bad arguments");
    if (pixel_count == 0) return RF_OK;
    hipLaunchKernelGGL(k_compose, dim3((unsigned)((pixel_count + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, (const float4*)flare, (const float4*)starburst,
                       pixel_count, (float4*)out);
    RF_LAUNCH_CHECK("rf_compose_flare_starburst");
    return RF_OK;
}

namespace rf {
// Engine.image_flare (engine.py:98-155): every light sample's flare is added
// four times, as rendered and mirrored (rows; rows and columns; columns),
// weighted by the RGBA values of the four image samples it stands for, in
// sample order, into the running image, which the last call divides by the
// sample count (sample width * sample height).
__global__ __launch_bounds__(256) void k_image_light(const float4* __restrict__ frames, int count,
                                                     int width, int height,
                                                     const float4* __restrict__ weights,
                                                     float divisor, int accumulate,
                                                     float4* __restrict__ out) {
#pragma clang fp contract(off)
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)width * height) return;
    const int y = (int)(i / width), x = (int)(i % width);
    const int yf = height - 1 - y, xf = width - 1 - x;
    const long n = (long)width * height;
    const long src[4] = {(long)y * width + x, (long)yf * width + x, (long)yf * width + xf,
                         (long)y * width + xf};
    float4 acc = make_float4(0, 0, 0, 0);
    for (int f = 0; f < count; f++) {
        const float4* fr = frames + f * n;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float4 v = weights[4 * f + q], p = fr[src[q]];
            acc = make_float4(acc.x + v.x * p.x, acc.y + v.y * p.y, acc.z + v.z * p.z,
                              acc.w + v.w * p.w);
        }
    }
    if (accumulate) {
        const float4 o = out[i];
        acc = make_float4(o.x + acc.x, o.y + acc.y, o.z + acc.z, o.w + acc.w);
    }
    if (divisor != 1.0f)
        acc = make_float4(acc.x / divisor, acc.y / divisor, acc.z / divisor, acc.w / divisor);
    out[i] = acc;
}
}  // namespace rf

extern "C" int rf_image_light_compose(const float* frames, int32_t count, int32_t width,
                                      int32_t height, const float* weights, float divisor,
                                      int32_t accumulate, float* out, rf_stream_t stream) {
    using namespace rf;
    if (!out || width < 1 || height < 1 || count < 0 || (count > 0 && (!frames || !weights)))
        return set_error(RF_ERR_INVALID, "rf_image_light_compose: bad arguments");
    const long n = (long)width * height;
    hipLaunchKernelGGL(k_image_light, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, (const float4*)frames, count, width, height,
                       (const float4*)weights, divisor, accumulate, (float4*)out);
    RF_LAUNCH_CHECK("rf_image_light_compose");
    return RF_OK;
}
