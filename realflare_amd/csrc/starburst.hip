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

// Samples whose XYZ weight, wavelength scale and its reciprocal are held in
// LDS at a time (20 KiB, so eight workgroups fit a CU); longer sample loops
// refill the table chunk by chunk.
static constexpr int SB_CHUNK = 1024;
static constexpr int SB_BUFFER_WORD3 = 0x00020000;  // raw 32-bit buffer, gfx9 descriptor
static constexpr uint32_t SB_OUT_OF_RANGE = 0x80000000u;  // past every record: the load returns 0

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

// The OpenCL 1.2 bilinear lookup with border 0 (sample_bilinear, one channel)
// of the power spectrum, without branches: the texel coordinate is clamped to
// [-2, size] before the conversion, and a texel outside the texture is read at
// an offset past the buffer's size, which a raw buffer load returns as 0.  The
// weights are the same arithmetic in the same order as sample_bilinear, so the
// value is the same bits.  w, h < 2^24 (exact as floats) and w * h * 4 bytes
// fits the descriptor's 31-bit size (both checked by the host), so j * w + i
// of a texel inside does not wrap.
__device__ __forceinline__ float spectrum_sample(__amdgpu_buffer_rsrc_t tex, int w, int h, float s,
                                                 float t) {
#pragma clang fp contract(off)
    const float u = s * (float)w - 0.5f, v = t * (float)h - 0.5f;
    const float fu = floorf(u), fv = floorf(v);
    // rf_to_int(fu) and the clamped index agree on which texels lie inside:
    // at or past w (or at or below -2) both texels of the axis are outside
    // either way, and a NaN coordinate gives NaN weights either way
    const int i0 = (int)__builtin_amdgcn_fmed3f(fu, -2.0f, (float)w);
    const int j0 = (int)__builtin_amdgcn_fmed3f(fv, -2.0f, (float)h);
    const int i1 = i0 + 1, j1 = j0 + 1;
    const float a = u - fu, b = v - fv;
    const bool ci0 = (uint32_t)i0 < (uint32_t)w, ci1 = (uint32_t)i1 < (uint32_t)w;
    const bool cj0 = (uint32_t)j0 < (uint32_t)h, cj1 = (uint32_t)j1 < (uint32_t)h;
    const uint32_t r0 = (uint32_t)j0 * (uint32_t)w, r1 = (uint32_t)j1 * (uint32_t)w;
    const uint32_t o00 = ci0 && cj0 ? (r0 + (uint32_t)i0) * 4u : SB_OUT_OF_RANGE;
    const uint32_t o10 = ci1 && cj0 ? (r0 + (uint32_t)i1) * 4u : SB_OUT_OF_RANGE;
    const uint32_t o01 = ci0 && cj1 ? (r1 + (uint32_t)i0) * 4u : SB_OUT_OF_RANGE;
    const uint32_t o11 = ci1 && cj1 ? (r1 + (uint32_t)i1) * 4u : SB_OUT_OF_RANGE;
    const float t00 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(tex, o00, 0, 0));
    const float t10 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(tex, o10, 0, 0));
    const float t01 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(tex, o01, 0, 0));
    const float t11 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(tex, o11, 0, 0));
    return (1.0f - a) * (1.0f - b) * t00 + a * (1.0f - b) * t10 + (1.0f - a) * b * t01 +
           a * b * t11;
}

// kJitter: blur != 0 (the sample position is jittered by two noise values);
// kRotate: the rotation is not exactly the identity.  With rotation 0 (the
// default, and a weight >= 0, so that the power below is finite) the rotation
// angle is exactly +-0: cos = 1 and sin = +-0 leave the position unchanged up
// to the sign of a zero, which the texture coordinate 0.5 * x + 0.5 does not
// see, so the two noise, power and trigonometric evaluations of the rotation
// are skipped.  With blur 0 (the default) the jitter radius is exactly 0 and
// the jitter's two noise and trigonometric evaluations are skipped as well.
// Both skips are bit-identical to evaluating them.
template <bool kJitter, bool kRotate>
__global__ __launch_bounds__(256) void k_starburst(const float* __restrict__ spectrum_values, int sw, int sh,
                                                   const float* __restrict__ light, int light_n,
                                                   rf_starburst_params p, float4* __restrict__ out) {
#pragma clang fp contract(off)
    // per sample: the X, Y, Z weights and the wavelength scale, and the
    // scale's reciprocal (the table's fourth component never reaches the
    // output)
    __shared__ float4 s_xyzs[SB_CHUNK];
    __shared__ float s_isc[SB_CHUNK];
    const float pi = 3.14159265358979323846f;
    const __amdgpu_buffer_rsrc_t spectrum = __builtin_amdgcn_make_buffer_rsrc(
        (void*)spectrum_values, (short)0, sw * sh * (int)sizeof(float), SB_BUFFER_WORD3);
    const int samples = p.samples;
    const int w = p.width, h = p.height;
    const int px = blockIdx.x * 16 + (threadIdx.x & 15);
    const int py = blockIdx.y * 16 + (threadIdx.x >> 4);
    float nx = 2.0f * ((float)px / (float)(w - 1)) - 1.0f;
    float ny = 2.0f * ((float)py / (float)(h - 1)) - 1.0f;
    nx = nx - p.offset[0];
    ny = ny - p.offset[1];
    nx = nx / fmaxf(p.size[0], FLT_MIN);
    ny = ny / fmaxf(p.size[1], FLT_MIN);
    float r0 = 0.0f, r1 = 0.0f, r2 = 0.0f;
    const float fx = (float)px, fy = (float)py;
    for (int t0 = 0; t0 < samples; t0 += SB_CHUNK) {
        const int tn = min(SB_CHUNK, samples - t0);
        if (t0 > 0) __syncthreads();  // the previous chunk has been read
        for (int k = threadIdx.x; k < tn; k += blockDim.x) {
            const float step = (float)(t0 + k) / (float)samples;
            const float wavelength = step * (float)(730 - 390) + 390.0f;
            const float sc = fmaxf(1 - (1 - (wavelength / 560.0f)) * 1.0f, FLT_MIN);
            float4 v;
            v.x = sample_bilinear(light + 0, light_n, 1, 4, step, 0.0f, false);
            v.y = sample_bilinear(light + 1, light_n, 1, 4, step, 0.0f, false);
            v.z = sample_bilinear(light + 2, light_n, 1, 4, step, 0.0f, false);
            v.w = sc;
            s_xyzs[k] = v;
            s_isc[k] = 1.0f / sc;
        }
        __syncthreads();
        for (int k = 0; k < tn; ++k) {
            const float4 xyzs = s_xyzs[k];  // X, Y, Z weights and the wavelength scale
            const float isc = s_isc[k];
            const float sc = xyzs.w;
            const float seed = (float)((t0 + k) * 4);
            float x = nx, y = ny;
            if (kJitter) {
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
            if (kRotate) {
                const float sign = (noise3(fx, fy, seed + 2) > 0.5f) ? 1.0f : -1.0f;
                const float angle =
                    sign * p.rotation * (1.0f - powf(noise3(fx, fy, seed + 3), p.rotation_weight));
                float sa, ca;
                sincos_reduced(angle, &sa, &ca);
                rx = x * ca + y * sa;
                ry = y * ca - x * sa;
            }
            const float u = rx * 0.5f + 0.5f, v = ry * 0.5f + 0.5f;
            const float fi = spectrum_sample(spectrum, sw, sh, u, v);
            r0 += xyzs.x * fi;
            r1 += xyzs.y * fi;
            r2 += xyzs.z * fi;
        }
    }
    if (px >= w || py >= h) return;
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
    if ((long)spectrum_width * spectrum_height > (1L << 28) || spectrum_width >= (1 << 24) ||
        spectrum_height >= (1 << 24))
        return set_error(RF_ERR_INVALID,
                         "rf_starburst: power spectrum of %dx%d exceeds 2^28 values or a side of 2^24",
                         spectrum_width, spectrum_height);
    const bool jitter = prm->blur != 0.0f;
    const bool rotate = !(prm->rotation == 0.0f && prm->rotation_weight >= 0.0f);
    decltype(&k_starburst<false, false>) kern = k_starburst<false, false>;
    if (jitter) kern = rotate ? k_starburst<true, true> : k_starburst<true, false>;
    else if (rotate) kern = k_starburst<false, true>;
    dim3 grid((prm->width + 15) / 16, (prm->height + 15) / 16);
    hipLaunchKernelGGL(kern, grid, dim3(256), 0, (hipStream_t)stream, spectrum, spectrum_width,
                       spectrum_height, light_spectrum, spectrum_count, *prm, (float4*)image);
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
