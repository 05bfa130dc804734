// Shared definitions for the MI355X (gfx950) flare render kernels.
//
// Numerics follow the OpenCL kernels of realflare (realflare/api/tasks/cl),
// with the behaviour that source leaves undefined fixed the same way the
// oracle fixes it:
//   * float -> int truncates toward zero, saturates, NaN -> 0 (rf_to_int);
//   * integer edge functions wrap modulo 2^32;
//   * float max/min are IEEE maxNum/minNum (fmaxf/fminf: NaN loses);
//   * unwritten Vertex fields of rays that missed the sensor are 0;
//   * image sampling is the OpenCL 1.2 bilinear formula in FP32, border 0;
//   * unsuffixed double literals of the OpenCL source are applied in double.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/realflare_amd.h"

#define RF_LMAX 64  // max lens elements (incl. stop and sensor) per lens

namespace rf {

// Records the error text for rf_last_error() and returns RF_ERR_*.
int set_error(int code, const char* fmt, ...);
int check_hip(hipError_t e, const char* what);

#define RF_HIP(call)                                  \
    do {                                              \
        hipError_t e_ = (call);                       \
        if (e_ != hipSuccess) return rf::check_hip(e_, #call); \
    } while (0)

#define RF_LAUNCH_CHECK(name)                                  \
    do {                                                       \
        hipError_t e_ = hipGetLastError();                     \
        if (e_ != hipSuccess) return rf::check_hip(e_, name);  \
    } while (0)

struct f2 {
    float x, y;
};
struct f3 {
    float x, y, z;
};

// The helpers below are compiled without FP contraction, like the kernels
// that use them, so every kernel evaluates a shared expression identically.
__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 add3(f3 a, f3 b) {
#pragma clang fp contract(off)
    return mk3(a.x + b.x, a.y + b.y, a.z + b.z);
}
__device__ __forceinline__ f3 sub3(f3 a, f3 b) {
#pragma clang fp contract(off)
    return mk3(a.x - b.x, a.y - b.y, a.z - b.z);
}
__device__ __forceinline__ f3 scl3(f3 a, float s) {
#pragma clang fp contract(off)
    return mk3(a.x * s, a.y * s, a.z * s);
}
__device__ __forceinline__ float dot3(f3 a, f3 b) {
#pragma clang fp contract(off)
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ __forceinline__ float len3(f3 a) {
#pragma clang fp contract(off)
    return sqrtf(dot3(a, a));
}
__device__ __forceinline__ f3 norm3(f3 a) {
#pragma clang fp contract(off)
    float l = len3(a);
    return mk3(a.x / l, a.y / l, a.z / l);
}

__device__ __forceinline__ float clampn(float x, float a, float b) { return fminf(fmaxf(x, a), b); }

__device__ __forceinline__ float smoothstepf(float e0, float e1, float x) {
#pragma clang fp contract(off)
    float t = clampn((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

__device__ __forceinline__ int rf_to_int(float f) {
    if (!(f == f)) return 0;
    if (f >= 2147483647.0f) return 2147483647;
    if (f <= -2147483648.0f) return (int)(-2147483647 - 1);
    return (int)f;
}

__device__ __forceinline__ int wsub(int a, int b) { return (int)((uint32_t)a - (uint32_t)b); }
__device__ __forceinline__ int wmul(int a, int b) { return (int)((uint32_t)a * (uint32_t)b); }
__device__ __forceinline__ int wadd(int a, int b) { return (int)((uint32_t)a + (uint32_t)b); }

// OpenCL 1.2 (section 8.2) bilinear filter, normalized coordinates, one
// channel at `stride` floats per texel.  edge=false: CLK_ADDRESS_CLAMP with
// border 0; edge=true: CLK_ADDRESS_CLAMP_TO_EDGE.
__device__ __forceinline__ float sample_bilinear(const float* __restrict__ tex, int w, int h,
                                                 int stride, float s, float t, bool edge) {
#pragma clang fp contract(off)
    float u = s * (float)w - 0.5f, v = t * (float)h - 0.5f;
    float fu = floorf(u), fv = floorf(v);
    int i0 = rf_to_int(fu), j0 = rf_to_int(fv);
    int i1 = wadd(i0, 1), j1 = wadd(j0, 1);
    float a = u - fu, b = v - fv;
    float T[4];
    int is[2] = {i0, i1}, js[2] = {j0, j1};
#pragma unroll
    for (int jj = 0; jj < 2; jj++) {
#pragma unroll
        for (int ii = 0; ii < 2; ii++) {
            int i = is[ii], j = js[jj];
            float val;
            if (edge) {
                i = i < 0 ? 0 : (i > w - 1 ? w - 1 : i);
                j = j < 0 ? 0 : (j > h - 1 ? h - 1 : j);
                val = tex[((long)j * w + i) * stride];
            } else if (i < 0 || i >= w || j < 0 || j >= h) {
                val = 0.0f;
            } else {
                val = tex[((long)j * w + i) * stride];
            }
            T[jj * 2 + ii] = val;
        }
    }
    return (1.0f - a) * (1.0f - b) * T[0] + a * (1.0f - b) * T[1] + (1.0f - a) * b * T[2] +
           a * b * T[3];
}

// sin of a float argument, evaluated in double: k = the integer nearest
// x / pi, r = x - k pi with pi split into two doubles (|r| <= pi/2 up to the
// rounding of k, with an error far below a float ulp of the result for
// |x| < 2^31), sin x = (-1)^k sin r, and sin r from its Taylor series through
// the r^17 term (the truncation is below 5e-14 at |r| = pi/2), rounded once
// to float: the correctly rounded sine except within ~1e-13 of a rounding
// midpoint.  The noise's arguments reach 10^6, where a single-precision
// sine would take its long reduction path.
__device__ __forceinline__ float sin_reduced(float x) {
    const double kPiHi = 3.141592653589793116;    // double nearest pi
    const double kPiLo = 1.2246467991473532e-16;  // pi - kPiHi
    const double xd = (double)x;
    const double k = rint(xd * 0.31830988618379067154);  // 1/pi
    const double r = fma(-k, kPiLo, fma(-k, kPiHi, xd));
    const double s = r * r;
    // (-1)^n / (2n + 1)! for n = 8 down to 1
    double p = 1.0 / 355687428096000.0;
    p = fma(p, s, -1.0 / 1307674368000.0);
    p = fma(p, s, 1.0 / 6227020800.0);
    p = fma(p, s, -1.0 / 39916800.0);
    p = fma(p, s, 1.0 / 362880.0);
    p = fma(p, s, -1.0 / 5040.0);
    p = fma(p, s, 1.0 / 120.0);
    p = fma(p, s, -1.0 / 6.0);
    const float v = (float)fma(r * s, p, r);
    return ((int)k & 1) ? -v : v;
}

// sin and cos of a float argument the same way, with the reduction by pi/2
// (|r| <= pi/4; Taylor series through r^13 for the sine and r^14 for the
// cosine, truncations below 3e-14) and the quadrant k mod 4 choosing
// +-sin r or +-cos r for each.
__device__ __forceinline__ void sincos_reduced(float x, float* sn, float* cs) {
    const double kPio2Hi = 1.5707963267948966;     // double nearest pi/2
    const double kPio2Lo = 6.123233995736766e-17;  // pi/2 - kPio2Hi
    const double xd = (double)x;
    const double k = rint(xd * 0.63661977236758134);  // 2/pi
    const double r = fma(-k, kPio2Lo, fma(-k, kPio2Hi, xd));
    const double s = r * r;
    // (-1)^n / (2n + 1)! for n = 6 down to 1, (-1)^(n+1) / (2n)! for n = 7 down to 1
    double ps = 1.0 / 6227020800.0;
    ps = fma(ps, s, -1.0 / 39916800.0);
    ps = fma(ps, s, 1.0 / 362880.0);
    ps = fma(ps, s, -1.0 / 5040.0);
    ps = fma(ps, s, 1.0 / 120.0);
    ps = fma(ps, s, -1.0 / 6.0);
    double pc = 1.0 / 87178291200.0;
    pc = fma(pc, s, -1.0 / 479001600.0);
    pc = fma(pc, s, 1.0 / 3628800.0);
    pc = fma(pc, s, -1.0 / 40320.0);
    pc = fma(pc, s, 1.0 / 720.0);
    pc = fma(pc, s, -1.0 / 24.0);
    pc = fma(pc, s, 0.5);
    const float sr = (float)fma(r * s, ps, r);
    const float cr = (float)fma(-s, pc, 1.0);
    const int q = (int)k;
    const float a = (q & 1) ? cr : sr, b = (q & 1) ? sr : cr;
    *sn = (q & 2) ? -a : a;
    *cs = ((q + 1) & 2) ? -b : b;
}

// noise (noise.cl:1-5): fract(sin(x 112.9898 + y 179.233 + z 237.212) 43758.5453),
// with the sine rounded correctly (its error is multiplied by 43758)
__device__ __forceinline__ float noise3(float x, float y, float z) {
#pragma clang fp contract(off)
    float v = sin_reduced(x * 112.9898f + y * 179.233f + z * 237.212f) * 43758.5453f;
    return fminf(v - floorf(v), 0x1.fffffep-1f);
}

// color.cl:3-10, with the matrix applied in double as the source's double
// literals ask for.
__device__ __forceinline__ void xyz_to_ap1(float x, float y, float z, float* o) {
#pragma clang fp contract(off)
    double dx = x, dy = y, dz = z;
    o[0] = (float)(1.6410233797 * dx + -0.3248032942 * dy + -0.2364246952 * dz);
    o[1] = (float)(-0.6636628587 * dx + 1.6153315917 * dy + 0.0167563477 * dz);
    o[2] = (float)(0.0117218943 * dx + -0.0082844420 * dy + 0.9883948585 * dz);
    o[3] = 0.0f;
}

}  // namespace rf
