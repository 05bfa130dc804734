// Ghost ray tracing on MI355X.
//
// One lane per (ghost path, wavelength, grid ray), the NDRange of the
// reference `raytrace` kernel (realflare/api/tasks/cl/raytracing.cl:218-359).
// A workgroup holds 256 rays of one (path, wavelength), so the surface
// sequence is wave-uniform: the loop over lens elements runs on scalar
// registers, element records are scalar loads, and a wave leaves the loop
// as soon as a ballot shows no live lane (vignetted rays stop costing issue
// slots).  The per-wavelength refractive index of every element (Sellmeier,
// raytracing.cl:107-120) is evaluated once per workgroup into LDS.
#include <cmath>

#include "rf_common.h"

namespace rf {

enum { OUT_AOS = 0, OUT_PLANES = 1, OUT_INTERSECTIONS = 2 };

// raytracing.cl:1-16
__device__ float fresnel(float theta0, float n1, float n2) {
#pragma clang fp contract(off)
    float theta1 = asinf(sinf(theta0) * n1 / n2);
    float cosi = cosf(theta0);
    float cost = cosf(theta1);
    float rs = ((n1 * cosi) - (n2 * cost)) / ((n1 * cosi) + (n2 * cost));
    float rp = ((n1 * cost) - (n2 * cosi)) / ((n1 * cost) + (n2 * cosi));
    return (rs * rs + rp * rp) / 2;
}

// raytracing.cl:18-59 (single-layer anti-reflective coating, Ritschel 2009)
__device__ float fresnel_ar(float theta0, float lambda, float d1, float n0, float n1, float n2) {
#pragma clang fp contract(off)
    const float pi = 3.14159265358979323846f;
    float s0 = sinf(theta0);
    float theta1 = asinf(s0 * n0 / n1);
    float theta2 = asinf(s0 * n0 / n2);
    float sp01 = sinf(theta0 + theta1);
    float rs01 = -sinf(theta0 - theta1) / sp01;
    float rp01 = tanf(theta0 - theta1) / tanf(theta0 + theta1);
    float ts01 = 2 * sinf(theta1) * cosf(theta0) / sp01;
    float tp01 = ts01 * cosf(theta0 - theta1);
    float rs12 = -sinf(theta1 - theta2) / sinf(theta1 + theta2);
    float rp12 = tanf(theta1 - theta2) / tanf(theta1 + theta2);
    float ris = ts01 * ts01 * rs12;
    float rip = tp01 * tp01 * rp12;
    float dy = d1 * n1;
    float dx = tanf(theta1) * dy;
    float delay = sqrtf(dx * dx + dy * dy);
    float rel_phase = 4 * pi / lambda * (delay - dx * s0);
    float c = cosf(rel_phase);
    float out_s2 = rs01 * rs01 + ris * ris + 2 * rs01 * ris * c;
    float out_p2 = rp01 * rp01 + rip * rip + 2 * rp01 * rip * c;
    return (out_s2 + out_p2) / 2;
}

// raytracing.cl:107-120; `lambda * 1e-3` is a double product in the source
__device__ float dispersion(float lambda, const float* c) {
#pragma clang fp contract(off)
    float l = (float)((double)lambda * 1e-3);
    float l2 = l * l;
    float d0 = (c[0] * l2) / (l2 - c[1]);
    float d1 = (c[2] * l2) / (l2 - c[3]);
    float d2 = (c[4] * l2) / (l2 - c[5]);
    return sqrtf(1 + d0 + d1 + d2);
}

struct Hit {
    f3 pos, normal;
    float incident;
    bool hit;
};

// raytracing.cl:123-175.  `want_incident` drops the acos on refraction
// events, where the source computes it but never reads it.
template <bool want_incident>
__device__ __forceinline__ Hit intersect(f3 pos, f3 dir, float radius, float center) {
#pragma clang fp contract(off)
    Hit i;
    i.pos = mk3(0, 0, 0);
    i.normal = mk3(0, 0, 0);
    i.incident = 0.0f;
    if (dir.z == 0) {
        i.hit = false;
    } else if (radius == 0) {
        float dz = -center - pos.z;
        i.pos = add3(pos, scl3(dir, dz / dir.z));
        i.normal = (dir.z < 0) ? mk3(0, 0, 1) : mk3(0, 0, -1);
        i.hit = true;
    } else {
        float r = fabsf(radius);
        f3 C = mk3(0, 0, -center);
        f3 U = sub3(C, pos);
        f3 U1 = scl3(dir, dot3(U, dir));
        float d = len3(sub3(U, U1));
        if (d > r) {
            i.hit = false;
            return i;
        }
        float sgn = (radius * dir.z) > 0 ? -1.0f : 1.0f;
        float m = sqrtf(r * r - d * d);
        f3 p = sub3(add3(pos, U1), scl3(scl3(dir, m), sgn));
        f3 n = scl3(norm3(sub3(p, C)), sgn);
        if (want_incident) i.incident = acosf(clampn(dot3(scl3(dir, -1.0f), n), -1.0f, 1.0f));
        i.hit = true;
        i.pos = p;
        i.normal = n;
    }
    return i;
}

template <int OUT>
__global__ __launch_bounds__(256) void k_trace(const rf_lens_element* __restrict__ lens,
                                               const int* __restrict__ paths,
                                               const int* __restrict__ wavelengths,
                                               rf_trace_params prm, void* __restrict__ out,
                                               void* __restrict__ inters,
                                               const float* __restrict__ coating_lambda) {
#pragma clang fp contract(off)
    __shared__ float s_n[RF_LMAX];
    const int L = prm.lens_count;
    const int w = blockIdx.y, p = blockIdx.z;
    const int g = prm.grid_count, R = g * g, W = prm.wavelength_count;
    const float wavelength = (float)wavelengths[w];
    const bool disperse = !isnan(lens[0].coefficients[0]);
    if ((int)threadIdx.x < L) {
        const rf_lens_element& e = lens[threadIdx.x];
        float n = e.ior;
        if (disperse && n > 1) n = dispersion(wavelength, e.coefficients);
        s_n[threadIdx.x] = n;
    }
    __syncthreads();

    const int ray_id = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = ray_id < R;
    const long idx = ((long)p * W + w) * R + ray_id;

    // init_ray, raytracing.cl:178-215
    f3 pos, dir;
    {
        int y = (int)truncf((float)ray_id / (float)g);
        int x = ray_id - y * g;
        float px = (float)((double)prm.grid_length *
                           ((double)((float)x / (float)(g - 1)) - 0.5));
        float py = (float)((double)prm.grid_length *
                           (0.5 - (double)((float)y / (float)(g - 1))));
        Hit h = intersect<false>(mk3(px, py, 1.0f), mk3(0, 0, -1.0f), lens[0].radius,
                                 lens[0].center);
        f3 d = norm3(mk3(-prm.direction[0], -prm.direction[1], -prm.direction[2]));
        pos = sub3(h.pos, d);
        dir = d;
    }
    float apt_x = 0.0f, apt_y = 0.0f, rrel = 0.0f, refl = 1.0f;

    const int b1 = paths[2 * p], b2 = paths[2 * p + 1];
    const int aperture = prm.aperture_index;
    bool alive = active, died = false;
    int lens_id = 0, step = 0, delta = 1, inter_id = 0;
    const int inter_count = 3 * L - 1;

    while (lens_id >= 0 && lens_id < L) {
        const float radius = lens[lens_id].radius;
        const float center = lens[lens_id].center;
        const float height = lens[lens_id].height;
        const int coating = lens[lens_id].coating;
        const float coating_f = coating_lambda ? coating_lambda[lens_id] : (float)coating;
        const bool is_stop = lens_id == aperture;
        const bool is_sensor = lens_id == L - 1;
        const bool do_reflect =
            !is_stop && !is_sensor && ((step == 0 && lens_id == b1) || (step == 1 && lens_id == b2));
        if (alive) {
            // refraction events never read the incidence angle, except that
            // the intersection records store it
            Hit h = (do_reflect || OUT == OUT_INTERSECTIONS)
                        ? intersect<true>(pos, dir, radius, center)
                        : intersect<false>(pos, dir, radius, center);
            if (!h.hit) {
                alive = false;
                died = true;
            } else {
                if (OUT == OUT_INTERSECTIONS && inter_id < inter_count) {
                    float4* rec = (float4*)inters + 3 * (idx * inter_count + inter_id);
                    rec[0] = make_float4(h.pos.x, h.pos.y, h.pos.z, 0.0f);
                    rec[1] = make_float4(h.normal.x, h.normal.y, h.normal.z, 0.0f);
                    rec[2] = make_float4(h.incident, __int_as_float(1), 0.0f, 0.0f);
                }
                pos = h.pos;
                if (is_stop) {
                    apt_x = h.pos.x / height;
                    apt_y = h.pos.y / height;
                } else {
                    rrel = fmaxf(rrel, sqrtf(h.pos.x * h.pos.x + h.pos.y * h.pos.y) / height);
                    if (!is_sensor) {
                        int n_index = (dir.z < 0) ? lens_id - 1 : lens_id + 1;
                        float n1 = (0 <= n_index && n_index < L) ? s_n[n_index] : 1.0f;
                        float n2 = s_n[lens_id];
                        if (do_reflect) {
                            // reflect, raytracing.cl:96-102
                            dir = add3(dir, scl3(h.normal, 2.0f * dot3(h.normal, scl3(dir, -1.0f))));
                            float theta = (float)((double)h.incident + 1e-9);
                            float r;
                            if (coating_f > 0) {
                                float lambda_c = coating_f;
                                float nc = fmaxf(sqrtf(n1 * n2), prm.coating_min_ior);
                                float d1 = lambda_c / (4.0f * nc);
                                r = fresnel_ar(theta, wavelength, d1, n1, nc, n2);
                            } else {
                                r = fresnel(theta, n1, n2);
                            }
                            if (r > 0) refl *= r;
                        } else {
                            // refract, raytracing.cl:70-86
                            float o = n1 / n2;
                            float cost = dot3(scl3(dir, -1.0f), h.normal);
                            float sint2 = (o * o) * (1 - cost * cost);
                            f3 t = add3(scl3(dir, o),
                                        scl3(h.normal, o * cost - sqrtf(fabsf(1 - sint2))));
                            dir = scl3(t, (float)(sint2 < 1));
                            if (dir.z == 0) {
                                alive = false;
                                died = true;
                            }
                        }
                    }
                }
            }
        }
        if (do_reflect) {
            step++;
            delta = -delta;
        }
        inter_id++;
        lens_id += delta;
        if (!__any(alive)) break;
    }
    if (died) refl = NAN;
    if (!active) return;

    if (OUT == OUT_PLANES) {
        const long N = (long)prm.path_count * W * R;
        float* o = (float*)out;
        o[idx] = pos.x;
        o[N + idx] = pos.y;
        o[2 * N + idx] = apt_x;
        o[3 * N + idx] = apt_y;
        o[4 * N + idx] = rrel;
        o[5 * N + idx] = refl;
    } else {
        float4* o = (float4*)out + 3 * idx;
        o[0] = make_float4(pos.x, pos.y, pos.z, 0.0f);
        o[1] = make_float4(dir.x, dir.y, dir.z, 0.0f);
        o[2] = make_float4(apt_x, apt_y, rrel, refl);
    }
}

static int check_trace(const rf_lens_element* lens, const int32_t* paths,
                       const int32_t* wavelengths, const rf_trace_params* prm, const void* out) {
    if (!prm) return set_error(RF_ERR_INVALID, "rf_trace: params is NULL");
    if (prm->lens_count < 1 || prm->lens_count > RF_LMAX)
        return set_error(RF_ERR_INVALID, "rf_trace: lens_count %d outside [1, %d]",
                         prm->lens_count, RF_LMAX);
    if (prm->path_count < 0 || prm->path_count > 65535)
        return set_error(RF_ERR_INVALID, "rf_trace: path_count %d outside [0, 65535]",
                         prm->path_count);
    if (prm->wavelength_count < 0 || prm->wavelength_count > 65535)
        return set_error(RF_ERR_INVALID, "rf_trace: wavelength_count %d outside [0, 65535]",
                         prm->wavelength_count);
    if (prm->grid_count < 1 || prm->grid_count > 4096)
        return set_error(RF_ERR_INVALID, "rf_trace: grid_count %d outside [1, 4096]",
                         prm->grid_count);
    if (prm->path_count && prm->wavelength_count && (!lens || !paths || !wavelengths || !out))
        return set_error(RF_ERR_INVALID, "rf_trace: NULL buffer");
    return RF_OK;
}

template <int OUT>
static int launch_trace(const rf_lens_element* lens, const int32_t* paths,
                        const int32_t* wavelengths, const rf_trace_params* prm, void* out,
                        void* inters, rf_stream_t stream, const float* coating_lambda = nullptr) {
    int rc = check_trace(lens, paths, wavelengths, prm, out);
    if (rc) return rc;
    if (prm->path_count == 0 || prm->wavelength_count == 0) return RF_OK;
    const int R = prm->grid_count * prm->grid_count;
    dim3 grid((R + 255) / 256, prm->wavelength_count, prm->path_count);
    hipLaunchKernelGGL(k_trace<OUT>, grid, dim3(256), 0, (hipStream_t)stream, lens, paths,
                       wavelengths, *prm, out, inters, coating_lambda);
    RF_LAUNCH_CHECK("k_trace");
    return RF_OK;
}

}  // namespace rf

extern "C" int rf_trace(const rf_lens_element* lens, const int32_t* paths,
                        const int32_t* wavelengths, const rf_trace_params* prm, void* rays,
                        rf_stream_t stream) {
    return rf::launch_trace<rf::OUT_AOS>(lens, paths, wavelengths, prm, rays, nullptr, stream);
}

extern "C" int rf_trace_planes(const rf_lens_element* lens, const int32_t* paths,
                               const int32_t* wavelengths, const rf_trace_params* prm,
                               float* planes, rf_stream_t stream) {
    return rf::launch_trace<rf::OUT_PLANES>(lens, paths, wavelengths, prm, planes, nullptr,
                                            stream);
}

extern "C" int rf_trace_intersections(const rf_lens_element* lens, const int32_t* paths,
                                      const int32_t* wavelengths, const rf_trace_params* prm,
                                      void* rays, void* intersections, rf_stream_t stream) {
    if (!intersections) return rf::set_error(RF_ERR_INVALID, "rf_trace_intersections: NULL");
    return rf::launch_trace<rf::OUT_INTERSECTIONS>(lens, paths, wavelengths, prm, rays,
                                                   intersections, stream);
}

extern "C" int rf_trace_planes_coated(const rf_lens_element* lens, const float* coating_lambda,
                                      const int32_t* paths, const int32_t* wavelengths,
                                      const rf_trace_params* prm, float* planes,
                                      rf_stream_t stream) {
    if (!coating_lambda) return rf::set_error(RF_ERR_INVALID, "rf_trace_planes_coated: NULL");
    return rf::launch_trace<rf::OUT_PLANES>(lens, paths, wavelengths, prm, planes, nullptr, stream,
                                            coating_lambda);
}
