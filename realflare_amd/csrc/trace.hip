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

#include "fresnel.h"
#include "rf_common.h"

namespace rf {

enum { OUT_AOS = 0, OUT_PLANES = 1, OUT_INTERSECTIONS = 2 };

// raytracing.cl:1-16 and 18-59, in the angle-sum form of fresnel.h
__device__ float fresnel(float theta0, float n1, float n2) {
    return fresnel_terms(theta0, n1, n2).r;
}

__device__ float fresnel_ar(float theta0, float lambda, float d1, float n0, float n1, float n2) {
    return fresnel_ar_terms(theta0, lambda, d1, n0, n1, n2).r;
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

// Per workgroup (one wavelength): s_n[i] = refractive index of element i
// (raytracing.cl:301-306), s_o[2 i] and s_o[2 i + 1] = the refraction ratio
// n1 / n2 at element i for a ray travelling towards -z (n1 of element i - 1)
// and towards +z (n1 of element i + 1), 1 outside the lens (raytracing.cl:
// 295-298, 345).  The same divisions as a per-ray evaluation, so the same
// values.  Every thread of the workgroup must call it.
__device__ __forceinline__ void lens_indices(const rf_lens_element* __restrict__ lens, int L,
                                             float wavelength, float* s_n, float* s_o) {
#pragma clang fp contract(off)
    const bool disperse = !isnan(lens[0].coefficients[0]);
    const int t = threadIdx.x;
    if (t < L) {
        const rf_lens_element& e = lens[t];
        float n = e.ior;
        if (disperse && n > 1) n = dispersion(wavelength, e.coefficients);
        s_n[t] = n;
    }
    __syncthreads();
    if (t < L) {
        const float n2 = s_n[t];
        s_o[2 * t] = (t - 1 >= 0 ? s_n[t - 1] : 1.0f) / n2;
        s_o[2 * t + 1] = (t + 1 < L ? s_n[t + 1] : 1.0f) / n2;
    }
    __syncthreads();
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

struct RayOut {
    f3 pos, dir;
    float apt_x, apt_y, rrel, refl;
};

// The state of one ray of the `raytrace` kernel between two elements.
struct RayState {
    f3 pos, dir;
    float apt_x, apt_y, rrel, refl;
    bool alive, died;
};

// init_ray (raytracing.cl:178-215) of grid ray `ray_id` of a g x g grid.
__device__ __forceinline__ RayState init_state(const rf_lens_element* __restrict__ lens, int g,
                                               float grid_length, const float* direction,
                                               int ray_id, bool active) {
#pragma clang fp contract(off)
    RayState s;
    int y = (int)truncf((float)ray_id / (float)g);
    int x = ray_id - y * g;
    float px = (float)((double)grid_length * ((double)((float)x / (float)(g - 1)) - 0.5));
    float py = (float)((double)grid_length * (0.5 - (double)((float)y / (float)(g - 1))));
    Hit h = intersect<false>(mk3(px, py, 1.0f), mk3(0, 0, -1.0f), lens[0].radius, lens[0].center);
    f3 d = norm3(mk3(-direction[0], -direction[1], -direction[2]));
    s.pos = sub3(h.pos, d);
    s.dir = d;
    s.apt_x = 0.0f;
    s.apt_y = 0.0f;
    s.rrel = 0.0f;
    s.refl = 1.0f;
    s.alive = active;
    s.died = false;
    return s;
}

// Where the raytrace loop stands: the next element, the bounce count and the
// direction of travel through the element list, the next intersection record.
struct Cursor {
    int lens_id, step, delta, inter_id;
};

// The events of the raytrace loop (raytracing.cl:258-350) from the cursor on,
// while the element index stays in [0, limit) (limit = L: until the ray
// leaves the lens), for ghost path (b1, b2); refractive indices of the
// wavelength in s_n, refraction ratios in s_o.  A wave stops as soon as a
// ballot shows no live lane: the remaining events would change nothing.
// Callers keep the path, wavelength, lens and cursor wave-uniform.
template <int OUT>
__device__ __forceinline__ void run_events(RayState& s, Cursor& c, int limit,
                                           const rf_lens_element* __restrict__ lens, int L,
                                           int aperture, int b1, int b2, float wavelength,
                                           const float* s_n, const float* s_o,
                                           float coating_min_ior,
                                           const float* __restrict__ coating_lambda,
                                           void* __restrict__ inters, long idx) {
#pragma clang fp contract(off)
    const int inter_count = 3 * L - 1;
    while (c.lens_id >= 0 && c.lens_id < limit) {
        const int lens_id = c.lens_id;
        const float radius = lens[lens_id].radius;
        const float center = lens[lens_id].center;
        const float height = lens[lens_id].height;
        const int coating = lens[lens_id].coating;
        const float coating_f = coating_lambda ? coating_lambda[lens_id] : (float)coating;
        const bool is_stop = lens_id == aperture;
        const bool is_sensor = lens_id == L - 1;
        const bool do_reflect = !is_stop && !is_sensor &&
                                ((c.step == 0 && lens_id == b1) || (c.step == 1 && lens_id == b2));
        if (s.alive) {
            // refraction events never read the incidence angle, except that
            // the intersection records store it
            Hit h = (do_reflect || OUT == OUT_INTERSECTIONS)
                        ? intersect<true>(s.pos, s.dir, radius, center)
                        : intersect<false>(s.pos, s.dir, radius, center);
            if (!h.hit) {
                s.alive = false;
                s.died = true;
            } else {
                if (OUT == OUT_INTERSECTIONS && c.inter_id < inter_count) {
                    float4* rec = (float4*)inters + 3 * (idx * inter_count + c.inter_id);
                    rec[0] = make_float4(h.pos.x, h.pos.y, h.pos.z, 0.0f);
                    rec[1] = make_float4(h.normal.x, h.normal.y, h.normal.z, 0.0f);
                    rec[2] = make_float4(h.incident, __int_as_float(1), 0.0f, 0.0f);
                }
                s.pos = h.pos;
                if (is_stop) {
                    s.apt_x = h.pos.x / height;
                    s.apt_y = h.pos.y / height;
                } else {
                    // rrel = max(rrel, |xy| / height) (raytracing.cl:280).  The
                    // square root and the quotient are skipped where |xy|^2 is
                    // below (rrel height)^2 (1 - 2^-20): then the rounded
                    // quotient is below rrel, whatever the roundings, and the
                    // maximum is rrel.
                    const float s2 = h.pos.x * h.pos.x + h.pos.y * h.pos.y;
                    const double bound = (double)s.rrel * (double)height;
                    if (!(height > 0 && (double)s2 < bound * bound * (1.0 - 0x1p-20)))
                        s.rrel = fmaxf(s.rrel, sqrtf(s2) / height);
                    if (!is_sensor) {
                        if (do_reflect) {
                            int n_index = (s.dir.z < 0) ? lens_id - 1 : lens_id + 1;
                            float n1 = (0 <= n_index && n_index < L) ? s_n[n_index] : 1.0f;
                            float n2 = s_n[lens_id];
                            // reflect, raytracing.cl:96-102
                            s.dir = add3(s.dir,
                                         scl3(h.normal, 2.0f * dot3(h.normal, scl3(s.dir, -1.0f))));
                            float theta = (float)((double)h.incident + 1e-9);
                            float r;
                            if (coating_f > 0) {
                                float lambda_c = coating_f;
                                float nc = fmaxf(sqrtf(n1 * n2), coating_min_ior);
                                float d1 = lambda_c / (4.0f * nc);
                                r = fresnel_ar(theta, wavelength, d1, n1, nc, n2);
                            } else {
                                r = fresnel(theta, n1, n2);
                            }
                            if (r > 0) s.refl *= r;
                        } else {
                            // refract, raytracing.cl:70-86; o = n1 / n2
                            float o = s_o[2 * lens_id + (s.dir.z < 0 ? 0 : 1)];
                            float cost = dot3(scl3(s.dir, -1.0f), h.normal);
                            float sint2 = (o * o) * (1 - cost * cost);
                            f3 t = add3(scl3(s.dir, o),
                                        scl3(h.normal, o * cost - sqrtf(fabsf(1 - sint2))));
                            s.dir = scl3(t, (float)(sint2 < 1));
                            if (s.dir.z == 0) {
                                s.alive = false;
                                s.died = true;
                            }
                        }
                    }
                }
            }
        }
        if (do_reflect) {
            c.step++;
            c.delta = -c.delta;
        }
        c.inter_id++;
        c.lens_id += c.delta;
        if (!__any(s.alive)) break;
    }
}

__device__ __forceinline__ RayOut ray_out(const RayState& s) {
    RayOut o;
    o.pos = s.pos;
    o.dir = s.dir;
    o.apt_x = s.apt_x;
    o.apt_y = s.apt_y;
    o.rrel = s.rrel;
    o.refl = s.died ? NAN : s.refl;
    return o;
}

// One ray of the `raytrace` kernel (raytracing.cl:178-359): grid ray `ray_id`
// of a g x g grid through lens[0..L), ghost path (b1, b2), refractive indices
// of the ray's wavelength in s_n.  Returns the final state; reflectance is NaN
// for a ray that left the lens or totally reflected.  Callers keep the path,
// wavelength and lens wave-uniform.
template <int OUT>
__device__ __forceinline__ RayOut trace_ray(const rf_lens_element* __restrict__ lens, int L,
                                            int aperture, int b1, int b2, float wavelength,
                                            const float* s_n, const float* s_o, int g,
                                            float grid_length,
                                            const float* direction, float coating_min_ior,
                                            const float* __restrict__ coating_lambda, int ray_id,
                                            bool active, void* __restrict__ inters, long idx) {
    RayState s = init_state(lens, g, grid_length, direction, ray_id, active);
    Cursor c{0, 0, 1, 0};
    run_events<OUT>(s, c, L, lens, L, aperture, b1, b2, wavelength, s_n, s_o, coating_min_ior,
                    coating_lambda, inters, idx);
    return ray_out(s);
}

// Ghosts per workgroup of k_trace (ray records and planes).  Every ghost path
// (b1, b2) runs the same events as the straight path through elements
// 0 .. b1 - 1: up to its first reflection at b1 it refracts through the same
// elements as every other ghost of the same grid ray and wavelength.  A
// workgroup traces a group of consecutive ghosts: it advances one
// straight-through state from element to element and continues a copy of it
// from each ghost's b1 along that ghost's own path, so the shared prefix is
// traced once per group instead of once per ghost.  The events are the same
// operations on the same values, so the outputs are the same bits.  Paths in
// increasing b1 (all_pair_paths' order) share the most; a ghost whose b1 lies
// behind the straight state restarts it from init_ray.
static constexpr int TRACE_GHOST_GROUP = 4;

// At least six waves per SIMD for the ray records and planes (an upper bound
// of 80 VGPRs); the intersection records (one ghost per workgroup: they
// record every event) keep the default bound.
template <int OUT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OUT == OUT_INTERSECTIONS ? 1 : 6)))
void k_trace(const rf_lens_element* __restrict__ lens,
                                               const int* __restrict__ paths,
                                               const int* __restrict__ wavelengths,
                                               rf_trace_params prm, void* __restrict__ out,
                                               void* __restrict__ inters,
                                               const float* __restrict__ coating_lambda) {
#pragma clang fp contract(off)
    __shared__ float s_n[RF_LMAX];
    __shared__ float s_o[2 * RF_LMAX];
    const int L = prm.lens_count;
    const int w = blockIdx.y;
    const int g = prm.grid_count, R = g * g, W = prm.wavelength_count, P = prm.path_count;
    const float wavelength = (float)wavelengths[w];
    lens_indices(lens, L, wavelength, s_n, s_o);

    const int ray_id = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = ray_id < R;
    const int group = OUT == OUT_INTERSECTIONS ? 1 : TRACE_GHOST_GROUP;
    const int p0 = blockIdx.z * group, p1 = min(P, p0 + group);
    // the straight-through state: elements [0, bc.lens_id) traced
    RayState base = init_state(lens, g, prm.grid_length, prm.direction, ray_id, active);
    Cursor bc{0, 0, 1, 0};
    for (int p = p0; p < p1; p++) {
        const int b1 = paths[2 * p], b2 = paths[2 * p + 1];
        // the ghost's events equal the straight path's below this element
        const int start = (b1 >= 0 && b1 < L) ? b1 : L;
        if (start < bc.lens_id) {
            base = init_state(lens, g, prm.grid_length, prm.direction, ray_id, active);
            bc = Cursor{0, 0, 1, 0};
        }
        const long idx = ((long)p * W + w) * R + ray_id;
        // (intersection records: one ghost per workgroup, so the prefix's
        // records are this ghost's)
        run_events<OUT>(base, bc, start, lens, L, prm.aperture_index, -1, -1, wavelength, s_n, s_o,
                        prm.coating_min_ior, coating_lambda, inters, idx);
        RayState st = base;
        Cursor c = bc;
        run_events<OUT>(st, c, L, lens, L, prm.aperture_index, b1, b2, wavelength, s_n, s_o,
                        prm.coating_min_ior, coating_lambda, inters, idx);
        if (!active) continue;
        const RayOut r = ray_out(st);
        if (OUT == OUT_PLANES) {
            const long N = (long)P * W * R;
            float* o = (float*)out;
            o[idx] = r.pos.x;
            o[N + idx] = r.pos.y;
            o[2 * N + idx] = r.apt_x;
            o[3 * N + idx] = r.apt_y;
            o[4 * N + idx] = r.rrel;
            o[5 * N + idx] = r.refl;
        } else {
            float4* o = (float4*)out + 3 * idx;
            o[0] = make_float4(r.pos.x, r.pos.y, r.pos.z, 0.0f);
            o[1] = make_float4(r.dir.x, r.dir.y, r.dir.z, 0.0f);
            o[2] = make_float4(r.apt_x, r.apt_y, r.rrel, r.refl);
        }
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
    const int group = OUT == OUT_INTERSECTIONS ? 1 : TRACE_GHOST_GROUP;
    dim3 grid((R + 255) / 256, prm->wavelength_count, (prm->path_count + group - 1) / group);
    hipLaunchKernelGGL(k_trace<OUT>, grid, dim3(256), 0, (hipStream_t)stream, lens, paths,
                       wavelengths, *prm, out, inters, coating_lambda);
    RF_LAUNCH_CHECK("k_trace");
    return RF_OK;
}

// ---------------------------------------------------------------------------
// Batched trace: a persistent grid drains a queue of chunks.  A chunk is
// 256 * RPL rays of one (segment, wavelength); the chunks of segment s are
// chunk_offsets[s] .. chunk_offsets[s + 1] - 1, wavelength-major.  Lane t of a
// workgroup traces rays t, t + 256, ... of the chunk, so each pass over the
// chunk's rays is coalesced and every pass keeps the path, wavelength and lens
// wave-uniform.

static constexpr int BATCH_THREADS = 256;

__global__ void k_queue_reset(uint64_t* __restrict__ head) { *head = 0; }

__device__ __forceinline__ int find_segment(const uint64_t* __restrict__ offsets, int n,
                                            uint64_t c) {
    // largest s with offsets[s] <= c
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (offsets[mid] <= c) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <int RPL>
__global__ __launch_bounds__(BATCH_THREADS) void k_trace_batch(
    const rf_lens_element* __restrict__ lenses, const int* __restrict__ lens_table,
    const int* __restrict__ paths, const int* __restrict__ wavelengths, int W,
    float coating_min_ior, const rf_batch_segment* __restrict__ segs, int nseg,
    const uint64_t* __restrict__ chunk_offsets, long ray_total, uint64_t* __restrict__ head,
    float* __restrict__ planes) {
#pragma clang fp contract(off)
    __shared__ float s_n[RF_LMAX];
    __shared__ float s_o[2 * RF_LMAX];
    __shared__ uint64_t s_chunk;
    const uint64_t total = chunk_offsets[nseg];
    int cur_lens = -1, cur_w = -1;
    for (;;) {
        if (threadIdx.x == 0) s_chunk = atomicAdd((unsigned long long*)head, 1ull);
        __syncthreads();
        const uint64_t c = s_chunk;
        __syncthreads();
        if (c >= total) break;
        const int si = find_segment(chunk_offsets, nseg, c);
        const rf_batch_segment sg = segs[si];
        const int g = sg.grid_count, R = g * g;
        const int per_w = (R + BATCH_THREADS * RPL - 1) / (BATCH_THREADS * RPL);
        const int local = (int)(c - chunk_offsets[si]);
        const int w = local / per_w, blk = local - w * per_w;
        const int first = lens_table[3 * sg.lens], L = lens_table[3 * sg.lens + 1];
        const int aperture = lens_table[3 * sg.lens + 2];
        const rf_lens_element* lens = lenses + first;
        const float wavelength = (float)wavelengths[w];
        if (sg.lens != cur_lens || w != cur_w) {
            // the previous chunk's lanes are done with s_n, s_o (barrier above)
            lens_indices(lens, L, wavelength, s_n, s_o);
            cur_lens = sg.lens;
            cur_w = w;
        }
        const int b1 = paths[2 * sg.path], b2 = paths[2 * sg.path + 1];
#pragma unroll
        for (int k = 0; k < RPL; k++) {
            const int ray_id = (blk * RPL + k) * BATCH_THREADS + threadIdx.x;
            const bool active = ray_id < R;
            const RayOut r = trace_ray<OUT_PLANES>(lens, L, aperture, b1, b2, wavelength, s_n, s_o, g,
                                                   sg.grid_length, sg.direction, coating_min_ior,
                                                   nullptr, ray_id, active, nullptr, 0);
            if (active) {
                const long idx = sg.ray_offset + (long)w * R + ray_id;
                planes[idx] = r.pos.x;
                planes[ray_total + idx] = r.pos.y;
                planes[2 * ray_total + idx] = r.apt_x;
                planes[3 * ray_total + idx] = r.apt_y;
                planes[4 * ray_total + idx] = r.rrel;
                planes[5 * ray_total + idx] = r.refl;
            }
        }
    }
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

extern "C" size_t rf_trace_batch_workspace_size(void) { return 256; }

extern "C" int rf_trace_batch(const rf_lens_element* lenses, const int32_t* lens_table,
                              const int32_t* paths, const int32_t* wavelengths,
                              int32_t wavelength_count, float coating_min_ior,
                              const rf_batch* batch, const uint64_t* chunk_offsets,
                              int32_t rays_per_lane, int32_t workgroups, void* workspace,
                              size_t workspace_bytes, float* planes, rf_stream_t stream) {
    using namespace rf;
    if (!batch || !lenses || !lens_table || !paths || !wavelengths || !chunk_offsets || !planes)
        return set_error(RF_ERR_INVALID, "rf_trace_batch: NULL argument");
    if (batch->segment_count < 0 || (batch->segment_count > 0 && !batch->segments))
        return set_error(RF_ERR_INVALID, "rf_trace_batch: bad segment list");
    if (wavelength_count < 1 || wavelength_count > 65535)
        return set_error(RF_ERR_INVALID, "rf_trace_batch: wavelength_count %d", wavelength_count);
    if (workgroups < 1 || workgroups > (1 << 20))
        return set_error(RF_ERR_INVALID, "rf_trace_batch: workgroups %d", workgroups);
    if (!workspace || workspace_bytes < rf_trace_batch_workspace_size())
        return set_error(RF_ERR_WORKSPACE, "rf_trace_batch: workspace %zu < %zu", workspace_bytes,
                         rf_trace_batch_workspace_size());
    if (batch->segment_count == 0) return RF_OK;
    hipStream_t s = (hipStream_t)stream;
    uint64_t* head = (uint64_t*)workspace;
    hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(1), 0, s, head);
    const dim3 grid(workgroups), block(BATCH_THREADS);
    switch (rays_per_lane) {
        case 1:
            hipLaunchKernelGGL(k_trace_batch<1>, grid, block, 0, s, lenses, lens_table, paths,
                               wavelengths, wavelength_count, coating_min_ior, batch->segments,
                               batch->segment_count, chunk_offsets, (long)batch->ray_total, head,
                               planes);
            break;
        case 2:
            hipLaunchKernelGGL(k_trace_batch<2>, grid, block, 0, s, lenses, lens_table, paths,
                               wavelengths, wavelength_count, coating_min_ior, batch->segments,
                               batch->segment_count, chunk_offsets, (long)batch->ray_total, head,
                               planes);
            break;
        case 4:
            hipLaunchKernelGGL(k_trace_batch<4>, grid, block, 0, s, lenses, lens_table, paths,
                               wavelengths, wavelength_count, coating_min_ior, batch->segments,
                               batch->segment_count, chunk_offsets, (long)batch->ray_total, head,
                               planes);
            break;
        default:
            return set_error(RF_ERR_INVALID, "rf_trace_batch: rays_per_lane %d not in {1, 2, 4}",
                             rays_per_lane);
    }
    RF_LAUNCH_CHECK("k_trace_batch");
    return RF_OK;
}
