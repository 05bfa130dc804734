/*
 * CPU restatement of realflare's OpenCL render kernels (test infrastructure).
 *
 * This file is included twice by rf_oracle.c, once with REAL=float (suffix
 * _f32) and once with REAL=double (suffix _f64).  It is the checker for the
 * HIP kernels and the CPU baseline of the benchmark; nothing in the product
 * path links it.
 *
 * Each function cites the reference source it restates (paths relative to
 * the realflare repository).  Every function is pinned by a check that does
 * not restate it (tests/test_oracle.py, tests/test_oracle_pins.py): the
 * reference's own vectors, closed forms, invariants or brute force.
 * Behaviour the reference leaves undefined is fixed here, identically to the
 * HIP kernels:
 *   - float -> int conversion truncates toward zero, saturates, NaN -> 0;
 *   - integer edge functions wrap modulo 2^32;
 *   - max/min/clamp on floats are IEEE maxNum/minNum (NaN loses);
 *   - fields of a Vertex whose ray did not reach the sensor are 0, except
 *     reflectance which is NaN;
 *   - image sampling uses the OpenCL 1.2 bilinear formula (section 8.2) in
 *     the working precision, border colour 0;
 *   - unsuffixed double literals of the OpenCL source (0.5, 1e-9, 1e-3, the
 *     AP1 matrix) are evaluated in double, as on an fp64-capable device.
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SFX)

typedef struct { REAL x, y, z; } FN(v3_t);
typedef struct { REAL x, y; } FN(v2_t);
#define V3 FN(v3_t)
#define V2 FN(v2_t)

typedef struct {
    V3 pos, dir;
    V2 pos_apt;
    REAL rrel, reflectance;
} FN(ray_t);
#define RAY FN(ray_t)

typedef struct {
    V3 pos, normal;
    REAL incident;
    int hit;
} FN(inter_t);
#define INTER FN(inter_t)

typedef struct {
    V2 pos, uv;
    REAL rrel, reflectance, intensity;
} FN(vertex_t);
#define VERTEX FN(vertex_t)

/* A lens element in the working precision (the f64 build takes double
 * parameters so finite differences can resolve small steps). */
typedef struct {
    REAL radius, ior, height, center, coating;
    REAL coefficients[6];
} FN(elem_t);
#define ELEM FN(elem_t)

static inline V3 FN(v3)(REAL x, REAL y, REAL z) { V3 r = {x, y, z}; return r; }
static inline V3 FN(v3_add)(V3 a, V3 b) { return FN(v3)(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 FN(v3_sub)(V3 a, V3 b) { return FN(v3)(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 FN(v3_scale)(V3 a, REAL s) { return FN(v3)(a.x * s, a.y * s, a.z * s); }
static inline REAL FN(v3_dot)(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline REAL FN(v3_length)(V3 a) { return SQRT(FN(v3_dot)(a, a)); }
static inline V3 FN(v3_normalize)(V3 a) {
    REAL l = FN(v3_length)(a);
    return FN(v3)(a.x / l, a.y / l, a.z / l);
}
static inline REAL FN(fmaxn)(REAL a, REAL b) { return FMAX(a, b); }
static inline REAL FN(fminn)(REAL a, REAL b) { return FMIN(a, b); }
static inline REAL FN(clampn)(REAL x, REAL a, REAL b) { return FMIN(FMAX(x, a), b); }
static inline REAL FN(smoothstep)(REAL e0, REAL e1, REAL x) {
    REAL t = FN(clampn)((x - e0) / (e1 - e0), (REAL)0, (REAL)1);
    return t * t * ((REAL)3 - (REAL)2 * t);
}

/* raytracing.cl:1-16 */
static REAL FN(fresnel)(REAL theta0, REAL n1, REAL n2) {
    REAL theta1 = ASIN(SIN(theta0) * n1 / n2);
    REAL cosi = COS(theta0);
    REAL cost = COS(theta1);
    REAL rs = ((n1 * cosi) - (n2 * cost)) / ((n1 * cosi) + (n2 * cost));
    REAL rp = ((n1 * cost) - (n2 * cosi)) / ((n1 * cost) + (n2 * cosi));
    return (rs * rs + rp * rp) / 2;
}

/* raytracing.cl:18-59 (Ritschel et al. 2009, single-layer AR coating) */
static REAL FN(fresnel_ar)(REAL theta0, REAL lambda, REAL d1, REAL n0, REAL n1, REAL n2) {
    REAL theta1 = ASIN(SIN(theta0) * n0 / n1);
    REAL theta2 = ASIN(SIN(theta0) * n0 / n2);
    REAL rs01 = -SIN(theta0 - theta1) / SIN(theta0 + theta1);
    REAL rp01 = TAN(theta0 - theta1) / TAN(theta0 + theta1);
    REAL ts01 = 2 * SIN(theta1) * COS(theta0) / SIN(theta0 + theta1);
    REAL tp01 = ts01 * COS(theta0 - theta1);
    REAL rs12 = -SIN(theta1 - theta2) / SIN(theta1 + theta2);
    REAL rp12 = TAN(theta1 - theta2) / TAN(theta1 + theta2);
    REAL ris = ts01 * ts01 * rs12;
    REAL rip = tp01 * tp01 * rp12;
    REAL dy = d1 * n1;
    REAL dx = TAN(theta1) * dy;
    REAL delay = SQRT(dx * dx + dy * dy);
    REAL rel_phase = 4 * PI_R / lambda * (delay - dx * SIN(theta0));
    REAL out_s2 = rs01 * rs01 + ris * ris + 2 * rs01 * ris * COS(rel_phase);
    REAL out_p2 = rp01 * rp01 + rip * rip + 2 * rp01 * rip * COS(rel_phase);
    return (out_s2 + out_p2) / 2;
}

/* raytracing.cl:70-86 */
static V3 FN(refract)(V3 i, V3 n, REAL o) {
    REAL cost = FN(v3_dot)(FN(v3_scale)(i, -1), n);
    REAL sint2 = (o * o) * (1 - cost * cost);
    V3 t = FN(v3_add)(FN(v3_scale)(i, o), FN(v3_scale)(n, o * cost - SQRT(FABS(1 - sint2))));
    return FN(v3_scale)(t, (REAL)(sint2 < 1));
}

/* raytracing.cl:96-102 */
static V3 FN(reflect)(V3 i, V3 n) {
    return FN(v3_add)(i, FN(v3_scale)(n, (REAL)2.0 * FN(v3_dot)(n, FN(v3_scale)(i, -1))));
}

/* raytracing.cl:107-120 (lambda in nm; `lambda * 1e-3` is a double product) */
static REAL FN(dispersion)(REAL lambda, const REAL *c) {
    REAL l = (REAL)((double)lambda * 1e-3);
    REAL l2 = l * l;
    REAL d0 = (c[0] * l2) / (l2 - c[1]);
    REAL d1 = (c[2] * l2) / (l2 - c[3]);
    REAL d2 = (c[4] * l2) / (l2 - c[5]);
    return SQRT(1 + d0 + d1 + d2);
}

/* raytracing.cl:123-175 */
static INTER FN(intersect)(RAY ray, const ELEM *lens) {
    INTER i;
    memset(&i, 0, sizeof(i));
    REAL radius = lens->radius, center = lens->center;
    if (ray.dir.z == 0) {
        i.hit = 0;
    } else if (radius == 0) {
        REAL dz = -center - ray.pos.z;
        i.pos = FN(v3_add)(ray.pos, FN(v3_scale)(ray.dir, dz / ray.dir.z));
        i.normal = (ray.dir.z < 0) ? FN(v3)(0, 0, 1) : FN(v3)(0, 0, -1);
        i.hit = 1;
        i.incident = 0;
    } else {
        REAL r = FABS(radius);
        V3 C = FN(v3)(0, 0, -center);
        V3 U = FN(v3_sub)(C, ray.pos);
        V3 U1 = FN(v3_scale)(ray.dir, FN(v3_dot)(U, ray.dir));
        REAL d = FN(v3_length)(FN(v3_sub)(U, U1));
        if (d > r) {
            i.hit = 0;
            return i;
        }
        REAL sgn = (radius * ray.dir.z) > 0 ? -1 : 1;
        REAL m = SQRT(r * r - d * d);
        V3 pos = FN(v3_sub)(FN(v3_add)(ray.pos, U1), FN(v3_scale)(FN(v3_scale)(ray.dir, m), sgn));
        V3 normal = FN(v3_scale)(FN(v3_normalize)(FN(v3_sub)(pos, C)), sgn);
        REAL incident = ACOS(FN(clampn)(FN(v3_dot)(FN(v3_scale)(ray.dir, -1), normal), -1, 1));
        i.hit = 1;
        i.incident = incident;
        i.pos = pos;
        i.normal = normal;
    }
    return i;
}

/* raytracing.cl:178-215 */
static RAY FN(init_ray)(const ELEM *lens, int ray_id, int grid_count,
                        REAL grid_length, const REAL *initial_direction) {
    int y = (int)TRUNC((REAL)ray_id / (REAL)grid_count);
    int x = ray_id - (y * grid_count);
    REAL px = (REAL)((double)grid_length * ((double)((REAL)x / (REAL)(grid_count - 1)) - 0.5));
    REAL py = (REAL)((double)grid_length * (0.5 - (double)((REAL)y / (REAL)(grid_count - 1))));
    RAY ray;
    memset(&ray, 0, sizeof(ray));
    ray.pos = FN(v3)(px, py, 1);
    ray.dir = FN(v3)(0, 0, -1);
    INTER i = FN(intersect)(ray, lens);
    V3 direction = FN(v3_normalize)(FN(v3)(-initial_direction[0], -initial_direction[1],
                                            -initial_direction[2]));
    ray.pos = FN(v3_sub)(i.pos, direction);
    ray.pos_apt.x = 0;
    ray.pos_apt.y = 0;
    ray.dir = direction;
    ray.rrel = 0;
    ray.reflectance = 1;
    return ray;
}

/* One lane of raytracing.cl:218-359.  `inters` (may be NULL) receives one
 * record per visited surface, as the STORE_INTERSECTIONS build does. */
static RAY FN(trace_one)(const ELEM *lens, int L, int disperse, int px, int py,
                         REAL wavelength, int aperture_index, REAL coating_min_ior,
                         int grid_count, REAL grid_length, const REAL *direction, int ray_id,
                         INTER *inters, int inters_count) {
    RAY ray = FN(init_ray)(&lens[0], ray_id, grid_count, grid_length, direction);
    int step = 0, inter_id = 0, delta = 1;
    /* size_t in the source: a stop listed as a bounce walks the index below 0,
     * which wraps and ends the loop. */
    long lens_id;
    for (lens_id = 0; lens_id >= 0 && lens_id < L; inter_id++, lens_id += delta) {
        const ELEM *el = &lens[lens_id];
        INTER inter = FN(intersect)(ray, el);
        if (!inter.hit) break;
        if (inters && inter_id < inters_count) inters[inter_id] = inter;
        ray.pos = inter.pos;
        if (lens_id == aperture_index) {
            ray.pos_apt.x = inter.pos.x / el->height;
            ray.pos_apt.y = inter.pos.y / el->height;
            continue;
        } else {
            REAL h = SQRT(inter.pos.x * inter.pos.x + inter.pos.y * inter.pos.y) / el->height;
            ray.rrel = FN(fmaxn)(ray.rrel, h);
        }
        if (lens_id == L - 1) continue;
        int do_reflect = 0;
        if ((step == 0 && lens_id == px) || (step == 1 && lens_id == py)) {
            do_reflect = 1;
            step++;
            delta = -delta;
        }
        long n_index = (ray.dir.z < 0) ? lens_id - 1 : lens_id + 1;
        REAL n1 = (0 <= n_index && n_index < L) ? lens[n_index].ior : 1;
        REAL n2 = el->ior;
        if (disperse && n1 > 1) n1 = FN(dispersion)(wavelength, lens[n_index].coefficients);
        if (disperse && n2 > 1) n2 = FN(dispersion)(wavelength, el->coefficients);
        if (do_reflect) {
            ray.dir = FN(reflect)(ray.dir, inter.normal);
            REAL theta = (REAL)((double)inter.incident + 1e-9);
            REAL reflectance;
            if (el->coating > 0) {
                REAL lambda = el->coating;
                REAL nc = FN(fmaxn)(SQRT(n1 * n2), coating_min_ior);
                REAL d1 = lambda / ((REAL)4.0 * nc);
                reflectance = FN(fresnel_ar)(theta, wavelength, d1, n1, nc, n2);
            } else {
                reflectance = FN(fresnel)(theta, n1, n2);
            }
            if (reflectance > 0) ray.reflectance *= reflectance;
        } else {
            REAL ratio = n1 / n2;
            ray.dir = FN(refract)(ray.dir, inter.normal, ratio);
            if (ray.dir.z == 0) break;
        }
    }
    if (lens_id >= 0 && lens_id < L) ray.reflectance = NAN;
    return ray;
}

static int FN(trace_elems)(const ELEM *el, int L, int disperse, const int *paths, int P,
                           const int *wavelengths, int W, int aperture_index,
                           double coating_min_ior, int grid_count, double grid_length,
                           const double *direction, RAY *out, INTER *inters, int inters_count) {
    int R = grid_count * grid_count;
    REAL dir[4] = {(REAL)(float)direction[0], (REAL)(float)direction[1],
                   (REAL)(float)direction[2], (REAL)(float)direction[3]};
    REAL cmi = (REAL)(float)coating_min_ior;
    REAL gl = (REAL)(float)grid_length;
    long total = (long)P * W * R;
#pragma omp parallel for schedule(dynamic, 1024)
    for (long idx = 0; idx < total; idx++) {
        int ray_id = (int)(idx % R);
        int w = (int)((idx / R) % W);
        int p = (int)(idx / ((long)R * W));
        INTER *rec = inters ? inters + idx * inters_count : NULL;
        out[idx] = FN(trace_one)(el, L, disperse, paths[2 * p], paths[2 * p + 1],
                                 (REAL)wavelengths[w], aperture_index, cmi, grid_count, gl, dir,
                                 ray_id, rec, inters_count);
    }
    return 0;
}

/* Whole NDRange of `raytrace`: out[P][W][R] rays.  `coating` (may be NULL)
 * overrides the elements' integer coating wavelengths. */
int FN(rfo_trace)(const rf_lens_element_t *lens, int L, const int *paths, int P,
                  const int *wavelengths, int W, int aperture_index, double coating_min_ior,
                  int grid_count, double grid_length, const double *direction, RAY *out,
                  INTER *inters, int inters_count) {
    ELEM el[256];
    if (L > 256) return -1;
    for (int i = 0; i < L; i++) {
        el[i].radius = (REAL)lens[i].radius;
        el[i].ior = (REAL)lens[i].ior;
        el[i].height = (REAL)lens[i].height;
        el[i].center = (REAL)lens[i].center;
        el[i].coating = (REAL)lens[i].coating;
        for (int k = 0; k < 6; k++) el[i].coefficients[k] = (REAL)lens[i].coefficients[k];
    }
    int disperse = !isnan(lens[0].coefficients[0]);
    return FN(trace_elems)(el, L, disperse, paths, P, wavelengths, W, aperture_index,
                           coating_min_ior, grid_count, grid_length, direction, out, inters,
                           inters_count);
}

/* Same trace from element parameters given in double: radius, ior, height,
 * centre, coating wavelength (each [L]) and Sellmeier coefficients [L][6]. */
int FN(rfo_trace_params)(const double *radius, const double *ior, const double *height,
                         const double *center, const double *coating, const double *coefficients,
                         int L, int disperse, const int *paths, int P, const int *wavelengths,
                         int W, int aperture_index, double coating_min_ior, int grid_count,
                         double grid_length, const double *direction, RAY *out) {
    ELEM el[256];
    if (L > 256) return -1;
    for (int i = 0; i < L; i++) {
        el[i].radius = (REAL)radius[i];
        el[i].ior = (REAL)ior[i];
        el[i].height = (REAL)height[i];
        el[i].center = (REAL)center[i];
        el[i].coating = (REAL)coating[i];
        for (int k = 0; k < 6; k++) el[i].coefficients[k] = (REAL)coefficients[i * 6 + k];
    }
    return FN(trace_elems)(el, L, disperse, paths, P, wavelengths, W, aperture_index,
                           coating_min_ior, grid_count, grid_length, direction, out, NULL, 0);
}

/* ------------------------------------------------------------------------ */
/* rasterizing.cl                                                           */
/* ------------------------------------------------------------------------ */

static inline REAL FN(edge_f)(V2 a, V2 b, V2 c) {
    return (a.x - b.x) * (c.y - a.y) - (a.y - b.y) * (c.x - a.x);
}

/* rasterizing.cl:218-230 */
static inline void FN(quad_vertexes)(int g, int quad_id, int q[4]) {
    int row = quad_id / (g - 1), column = quad_id % (g - 1);
    q[0] = row * g + column;
    q[1] = q[0] + 1;
    q[2] = q[1] + g;
    q[3] = q[0] + g;
}

/* rasterizing.cl:233-320.  bounds[P][Q] (x0,y0,x1,y1), intensities[P][Q][W]
 * (entries the kernel does not write keep their input value). */
int FN(rfo_prim_shader)(const RAY *rays, int P, int W, int g, double area_orig, double min_area,
                        REAL *bounds, REAL *intensities) {
    int R = g * g, Q = (g - 1) * (g - 1);
    REAL ao = (REAL)(float)area_orig, ma = (REAL)(float)min_area;
    long total = (long)P * Q;
#pragma omp parallel for schedule(static)
    for (long idx = 0; idx < total; idx++) {
        int path = (int)(idx / Q), quad = (int)(idx % Q);
        REAL b[4] = {INFINITY, INFINITY, -INFINITY, -INFINITY};
        int invalid_rrel = 0;
        int qv[4];
        FN(quad_vertexes)(g, quad, qv);
        for (int w = 0; w < W; w++) {
            if (isnan(b[0])) continue;
            long ro = ((long)path * W + w) * R;
            long prim_index = ((long)path * Q + quad) * W + w;
            RAY r[4];
            for (int i = 0; i < 4; i++) r[i] = rays[ro + qv[i]];
            V2 pos[4];
            int valid = 0;
            for (int i = 0; i < 4; i++) {
                if (!isnan(r[i].reflectance)) {
                    pos[valid].x = r[i].pos.x;
                    pos[valid].y = r[i].pos.y;
                    valid++;
                }
            }
            REAL area = 0;
            if (valid == 4) {
                REAL a0 = FN(edge_f)(pos[0], pos[1], pos[2]);
                REAL a1 = FN(edge_f)(pos[0], pos[2], pos[3]);
                area = FABS(a0 + a1) / 2;
            } else if (valid == 3) {
                area = FABS(FN(edge_f)(pos[0], pos[1], pos[2]));
                pos[3] = pos[2];
                r[3] = r[2];
            } else {
                b[0] = NAN;
                continue;
            }
            if (r[0].rrel > 1 && r[1].rrel > 1 && r[2].rrel > 1 && r[3].rrel > 1) invalid_rrel++;
            REAL area_actual = FN(fmaxn)(area, ma);
            intensities[prim_index] = area_actual > 0 ? ao / area_actual : 0;
            REAL mnx = FN(fminn)(FN(fminn)(pos[0].x, pos[1].x), FN(fminn)(pos[2].x, pos[3].x));
            REAL mny = FN(fminn)(FN(fminn)(pos[0].y, pos[1].y), FN(fminn)(pos[2].y, pos[3].y));
            REAL mxx = FN(fmaxn)(FN(fmaxn)(pos[0].x, pos[1].x), FN(fmaxn)(pos[2].x, pos[3].x));
            REAL mxy = FN(fmaxn)(FN(fmaxn)(pos[0].y, pos[1].y), FN(fmaxn)(pos[2].y, pos[3].y));
            b[0] = FN(fminn)(b[0], mnx);
            b[1] = FN(fminn)(b[1], mny);
            b[2] = FN(fmaxn)(b[2], mxx);
            b[3] = FN(fmaxn)(b[3], mxy);
        }
        if (invalid_rrel == W) b[0] = NAN;
        for (int k = 0; k < 4; k++) bounds[idx * 4 + k] = b[k];
    }
    return 0;
}

/* rasterizing.cl:322-386.  vertexes[P][R][W]. */
int FN(rfo_vertex_shader)(const RAY *rays, const REAL *intensities, int P, int W, int g,
                          double screen_transform, int res_x, int res_y, VERTEX *vertexes) {
    int R = g * g, Q = (g - 1) * (g - 1);
    REAL st = (REAL)(float)screen_transform;
    long total = (long)P * R * W;
#pragma omp parallel for schedule(static)
    for (long idx = 0; idx < total; idx++) {
        int w = (int)(idx % W);
        int ray_id = (int)((idx / W) % R);
        int path = (int)(idx / ((long)W * R));
        RAY r = rays[((long)path * W + w) * R + ray_id];
        VERTEX v;
        memset(&v, 0, sizeof(v));
        v.reflectance = r.reflectance;
        if (!isnan(r.reflectance)) {
            int x = ray_id % g, y = ray_id / g;
            int nb[4];
            nb[0] = (y - 1) * (g - 1) + (x - 1);
            nb[1] = nb[0] + 1;
            nb[2] = y * (g - 1) + (x - 1);
            nb[3] = nb[2] + 1;
            REAL intensity = 0;
            int count = 0;
            for (int i = 0; i < 4; i++) {
                if (nb[i] < 0 || nb[i] >= Q) continue;
                REAL it = intensities[((long)path * Q + nb[i]) * W + w];
                if (it > 0) {
                    intensity += it;
                    ++count;
                }
            }
            intensity /= (REAL)(count > 1 ? count : 1);
            v.intensity = intensity;
            v.pos.x = r.pos.x * st + (REAL)res_x / 2;
            v.pos.y = r.pos.y * st + (REAL)res_y / 2;
            v.uv = r.pos_apt;
            v.rrel = r.rrel;
        }
        vertexes[idx] = v;
    }
    return 0;
}

/* Saturating truncation, NaN -> 0 (the float->int conversion fixed above). */
static inline int FN(to_int)(REAL f) {
    if (isnan(f)) return 0;
    if (f >= (REAL)2147483647.0) return 2147483647;
    if (f <= (REAL)-2147483648.0) return (int)(-2147483647 - 1);
    return (int)f;
}

/* rasterizing.cl:388-469, as a predicate: does prim `b` land in bin (bx,by)? */
static inline int FN(bin_hit)(const REAL *b, REAL st, int res_x, int res_y, int bin_size,
                              int bx, int by) {
    if (isnan(b[0])) return 0;
    int sx0 = FN(to_int)(b[0] * st) + res_x / 2;
    int sy0 = FN(to_int)(b[1] * st) + res_y / 2;
    int sx1 = FN(to_int)(b[2] * st) + res_x / 2;
    int sy1 = FN(to_int)(b[3] * st) + res_y / 2;
    int x0 = bx * bin_size, y0 = by * bin_size;
    int x1 = x0 + bin_size, y1 = y0 + bin_size;
    return sx0 <= x1 && sx1 >= x0 && sy0 <= y1 && sy1 >= y0;
}

static inline int FN(edge_i)(int ax, int ay, int bx, int by, int cx, int cy) {
    uint32_t t0 = ((uint32_t)ax - (uint32_t)bx) * ((uint32_t)cy - (uint32_t)ay);
    uint32_t t1 = ((uint32_t)ay - (uint32_t)by) * ((uint32_t)cx - (uint32_t)ax);
    return (int)(t0 - t1);
}

static inline int FN(is_top_left)(int ax, int ay, int bx, int by) {
    int ex = (int)((uint32_t)bx - (uint32_t)ax), ey = (int)((uint32_t)by - (uint32_t)ay);
    return (ey == 0 && ex > 0) || ey > 0;
}

/* rasterizing.cl:55-93 (Hormann-Tarini quad coverage, top-left rule) */
static int FN(intersect_quad)(int px, int py, const int *x, const int *y) {
    int b01 = FN(is_top_left)(x[0], y[0], x[1], y[1]) ? 0 : -1;
    int b12 = FN(is_top_left)(x[1], y[1], x[2], y[2]) ? 0 : -1;
    int b23 = FN(is_top_left)(x[2], y[2], x[3], y[3]) ? 0 : -1;
    int b30 = FN(is_top_left)(x[3], y[3], x[0], y[0]) ? 0 : -1;
    int b20 = FN(is_top_left)(x[2], y[2], x[0], y[0]) ? 0 : -1;
    int l01 = (int)((uint32_t)FN(edge_i)(x[0], y[0], x[1], y[1], px, py) + (uint32_t)b01);
    int l12 = (int)((uint32_t)FN(edge_i)(x[1], y[1], x[2], y[2], px, py) + (uint32_t)b12);
    int l23 = (int)((uint32_t)FN(edge_i)(x[2], y[2], x[3], y[3], px, py) + (uint32_t)b23);
    int l30 = (int)((uint32_t)FN(edge_i)(x[3], y[3], x[0], y[0], px, py) + (uint32_t)b30);
    int l20 = (int)((uint32_t)FN(edge_i)(x[2], y[2], x[0], y[0], px, py) + (uint32_t)b20);
    int w01 = l01 >= 0, w12 = l12 >= 0, w23 = l23 >= 0, w30 = l30 >= 0;
    if (FN(edge_i)(x[0], y[0], x[1], y[1], x[2], y[2]) == 0 &&
        FN(edge_i)(x[0], y[0], x[2], y[2], x[3], y[3]) == 0)
        return 0;
    int w20 = l20 >= 0;
    int front = (w20 && w01 && w12 && (w23 || w30)) || (!w20 && w23 && w30 && (w01 || w12));
    int back = (!w20 && !w01 && !w12 && (!w23 || !w30)) || (w20 && !w23 && !w30 && (!w01 || !w12));
    return front || back;
}

/* rasterizing.cl:95-148 (mean-value coordinates of a quad) */
static void FN(barycentric_quad)(int px, int py, const int *x, const int *y, REAL w[4]) {
    V2 s[4];
    for (int i = 0; i < 4; i++) {
        s[i].x = (REAL)(int)((uint32_t)x[i] - (uint32_t)px);
        s[i].y = (REAL)(int)((uint32_t)y[i] - (uint32_t)py);
    }
    REAL a[4], d[4], r[4];
    for (int i = 0; i < 4; i++) {
        int j = (i + 1) & 3;
        a[i] = s[i].x * s[j].y - s[j].x * s[i].y;
        d[i] = s[i].x * s[j].x + s[i].y * s[j].y;
        r[i] = SQRT(s[i].x * s[i].x + s[i].y * s[i].y);
    }
    for (int i = 0; i < 4; i++) {
        if (r[i] == 0) {
            for (int k = 0; k < 4; k++) w[k] = (k == i) ? 1 : 0;
            return;
        }
    }
    for (int i = 0; i < 4; i++)
        if (a[i] == 0) a[i] = (REAL)1e-3;
    REAL t[4];
    for (int i = 0; i < 4; i++) {
        int j = (i + 1) & 3;
        t[i] = (r[i] * r[j] - d[i]) / a[i];
    }
    REAL ww[4];
    for (int i = 0; i < 4; i++) ww[i] = (t[(i + 3) & 3] + t[i]) / r[i];
    REAL sum = ww[0] + ww[1] + ww[2] + ww[3];
    for (int i = 0; i < 4; i++) w[i] = ww[i] / sum;
}

/* OpenCL 1.2 section 8.2 bilinear filter, normalized coordinates, one channel.
 * clamp_to_edge=0: CLK_ADDRESS_CLAMP with border 0; 1: CLK_ADDRESS_CLAMP_TO_EDGE. */
static REAL FN(sample_bilinear)(const float *tex, int w, int h, int stride, REAL s, REAL t,
                                int clamp_to_edge) {
    REAL u = s * (REAL)w - (REAL)0.5, v = t * (REAL)h - (REAL)0.5;
    REAL fu = FLOOR(u), fv = FLOOR(v);
    int i0 = FN(to_int)(fu), j0 = FN(to_int)(fv);
    int i1 = (int)((uint32_t)i0 + 1u), j1 = (int)((uint32_t)j0 + 1u);
    REAL a = u - fu, b = v - fv;
    REAL T[4];
    int is[2] = {i0, i1}, js[2] = {j0, j1};
    for (int jj = 0; jj < 2; jj++) {
        for (int ii = 0; ii < 2; ii++) {
            int i = is[ii], j = js[jj];
            REAL val;
            if (clamp_to_edge) {
                i = i < 0 ? 0 : (i > w - 1 ? w - 1 : i);
                j = j < 0 ? 0 : (j > h - 1 ? h - 1 : j);
                val = (REAL)tex[(long)j * w * stride + (long)i * stride];
            } else if (i < 0 || i >= w || j < 0 || j >= h) {
                val = 0;
            } else {
                val = (REAL)tex[(long)j * w * stride + (long)i * stride];
            }
            T[jj * 2 + ii] = val;
        }
    }
    return (1 - a) * (1 - b) * T[0] + a * (1 - b) * T[1] + (1 - a) * b * T[2] + a * b * T[3];
}

/* color.cl:3-10 (the matrix literals are double) */
static void FN(xyz_to_ap1)(const REAL *in, REAL *out) {
    double x = in[0], y = in[1], z = in[2];
    out[0] = (REAL)(1.6410233797 * x + -0.3248032942 * y + -0.2364246952 * z);
    out[1] = (REAL)(-0.6636628587 * x + 1.6153315917 * y + 0.0167563477 * z);
    out[2] = (REAL)(0.0117218943 * x + -0.0082844420 * y + 0.9883948585 * z);
    out[3] = 0;
}

/* rasterizing.cl:471-506 */
static REAL FN(fragment_shader)(const REAL *w, const VERTEX *v, const float *ghost, int gw, int gh,
                                REAL scale) {
    REAL ux = w[0] * v[0].uv.x + w[1] * v[1].uv.x + w[2] * v[2].uv.x + w[3] * v[3].uv.x;
    REAL uy = w[0] * v[0].uv.y + w[1] * v[1].uv.y + w[2] * v[2].uv.y + w[3] * v[3].uv.y;
    ux = (ux / scale + (REAL)1.0) / (REAL)2.0;
    uy = (uy / scale + (REAL)1.0) / (REAL)2.0;
    REAL ghost_intensity = FN(sample_bilinear)(ghost, gw, gh, 1, ux, uy, 0);
    REAL rrel = w[0] * v[0].rrel + w[1] * v[1].rrel + w[2] * v[2].rrel + w[3] * v[3].rrel;
    REAL rrel_intensity = FN(smoothstep)((REAL)1.0, (REAL)0.95f, rrel);
    REAL area = w[0] * v[0].intensity + w[1] * v[1].intensity + w[2] * v[2].intensity +
                w[3] * v[3].intensity;
    REAL coating = w[0] * v[0].reflectance + w[1] * v[1].reflectance +
                   w[2] * v[2].reflectance + w[3] * v[3].reflectance;
    area = FN(fmaxn)(area, 0);
    coating = FN(fmaxn)(coating, 0);
    return ghost_intensity * rrel_intensity * area * coating;
}

static inline REAL FN(mix)(REAL a, REAL b, REAL t) { return a + (b - a) * t; }

static const unsigned char FN(sub_offsets)[15] = {0, 1, 0, 1, 2, 0, 3, 4, 1, 6, 2, 5, 0, 3, 7};

/* rasterizing.cl:524-663, one pixel.  The prims the pixel visits are those
 * the binner (rasterizing.cl:388-469) sets in its bin's queues, visited in
 * ascending prim order, which is the order of the batch/bit walk. */
static void FN(raster_pixel)(int x, int y, int width, int height, const VERTEX *vertexes,
                             long vertex_total, const REAL *bounds, int P, int W, int Wsub, int g,
                             int AA, REAL intensity, REAL ghost_scale, const float *ghost, int gw,
                             int gh, const float *spectrum, int spectrum_n, REAL st, int bin_size,
                             REAL *out_rgba, const int *prim_list, long prim_list_n) {
    int Q = (g - 1) * (g - 1), R = g * g;
    int bx = x / bin_size, by = y / bin_size;
    int px = x * AA, py = y * AA;
    int pcx = px + AA / 2, pcy = py + AA / 2;
    int max_w = W - 1 > 1 ? W - 1 : 1;
    int total_samples = W * AA * Wsub;
    REAL sub_step = (REAL)1.0f / (REAL)Wsub;
    REAL w_step = sub_step / (REAL)W;
    REAL rgba[4] = {0, 0, 0, 0};
    long nprims = prim_list ? prim_list_n : (long)P * Q;
    for (long k = 0; k < nprims; k++) {
        long prim = prim_list ? prim_list[k] : k;
        if (!FN(bin_hit)(&bounds[prim * 4], st, width, height, bin_size, bx, by)) continue;
        int path = (int)(prim / Q), quad = (int)(prim % Q);
        int qv[4];
        FN(quad_vertexes)(g, quad, qv);
        long vi[4];
        for (int j = 0; j < 4; j++) vi[j] = ((long)path * R + qv[j]) * W;
        VERTEX src[8];
        for (int j = 0; j < 4; j++) src[4 + j] = vertexes[vi[j]];
        for (int j = 0; j < 4; j++) vi[j]++;
        for (int wl = 0; wl < max_w; wl++) {
            for (int j = 0; j < 4; j++) {
                src[j] = src[4 + j];
                if (vi[j] < vertex_total) {
                    src[4 + j] = vertexes[vi[j]];
                } else {
                    memset(&src[4 + j], 0, sizeof(VERTEX));
                }
                vi[j]++;
            }
            REAL sub_pos = 0;
            REAL w_pos = ((REAL)wl + (REAL)0.5f) / (REAL)W;
            for (int i = 0; i < Wsub; i++) {
                VERTEX v[4];
                int vx[4], vy[4];
                for (int j = 0; j < 4; j++) {
                    v[j].pos.x = FN(mix)(src[j].pos.x, src[j + 4].pos.x, sub_pos);
                    v[j].pos.y = FN(mix)(src[j].pos.y, src[j + 4].pos.y, sub_pos);
                    v[j].uv.x = FN(mix)(src[j].uv.x, src[j + 4].uv.x, sub_pos);
                    v[j].uv.y = FN(mix)(src[j].uv.y, src[j + 4].uv.y, sub_pos);
                    v[j].intensity = FN(mix)(src[j].intensity, src[j + 4].intensity, sub_pos);
                    v[j].rrel = FN(mix)(src[j].rrel, src[j + 4].rrel, sub_pos);
                    v[j].reflectance = FN(mix)(src[j].reflectance, src[j + 4].reflectance, sub_pos);
                    vx[j] = FN(to_int)(v[j].pos.x * (REAL)AA);
                    vy[j] = FN(to_int)(v[j].pos.y * (REAL)AA);
                }
                int hits = 0;
                for (int s = 0; s < AA; s++) {
                    int sx = px + s, sy = py + FN(sub_offsets)[AA - 1 + s];
                    if (FN(intersect_quad)(sx, sy, vx, vy)) hits++;
                }
                REAL fragment = 0;
                if (hits > 0) {
                    REAL wts[4];
                    FN(barycentric_quad)(pcx, pcy, vx, vy, wts);
                    fragment += FN(fragment_shader)(wts, v, ghost, gw, gh, ghost_scale) * (REAL)hits;
                }
                if (W > 1) {
                    REAL xyz[3];
                    for (int c = 0; c < 3; c++)
                        xyz[c] = FN(sample_bilinear)(spectrum + c, spectrum_n, 1, 4, w_pos, 0, 1);
                    for (int c = 0; c < 3; c++) rgba[c] += xyz[c] * fragment;
                } else {
                    for (int c = 0; c < 3; c++) rgba[c] += fragment;
                }
                sub_pos += sub_step;
                w_pos += w_step;
            }
        }
    }
    if (rgba[0] > 0 || rgba[1] > 0 || rgba[2] > 0) {
        REAL f = intensity / (REAL)total_samples;
        for (int c = 0; c < 4; c++) rgba[c] *= f;
        int yo = height - (y + 1);
        FN(xyz_to_ap1)(rgba, &out_rgba[((long)yo * width + x) * 4]);
    }
}

/* Full `rasterizer` NDRange over an image the caller has cleared to 0.
 * y0/y1 restrict the rows (unflipped) so a caller can raster a band. */
int FN(rfo_rasterize)(const VERTEX *vertexes, const REAL *bounds, int P, int W, int Wsub, int g,
                      int AA, double intensity, double ghost_scale, const float *ghost, int gw,
                      int gh, const float *spectrum, int spectrum_n, double screen_transform,
                      int bin_size, int width, int height, int y0, int y1, REAL *out_rgba) {
    int R = g * g, Q = (g - 1) * (g - 1);
    long vertex_total = (long)P * R * W;
    REAL st = (REAL)(float)screen_transform;
    REAL it = (REAL)(float)intensity, gs = (REAL)(float)ghost_scale;
    int bdx = (width + bin_size - 1) / bin_size, bdy = (height + bin_size - 1) / bin_size;
    long nb = (long)bdx * bdy;
    /* per-bin prim lists in ascending order: the binner's bitmasks as lists */
    int **lists = (int **)calloc((size_t)nb, sizeof(int *));
    long *counts = (long *)calloc((size_t)nb, sizeof(long));
    if (!lists || !counts) return -1;
    int ret = 0;
#pragma omp parallel for schedule(dynamic, 1)
    for (long b = 0; b < nb; b++) {
        int bx = (int)(b % bdx), by = (int)(b / bdx);
        if ((by + 1) * bin_size <= y0 || by * bin_size >= y1) continue;
        long cap = 1024, n = 0;
        int *l = (int *)malloc(sizeof(int) * cap);
        for (long pr = 0; pr < (long)P * Q; pr++) {
            if (FN(bin_hit)(&bounds[pr * 4], st, width, height, bin_size, bx, by)) {
                if (n == cap) {
                    cap *= 2;
                    l = (int *)realloc(l, sizeof(int) * cap);
                }
                l[n++] = (int)pr;
            }
        }
        lists[b] = l;
        counts[b] = n;
    }
    /* (row, pixel) pairs are shared among the threads, so a band of a few
     * rows uses every thread too */
#pragma omp parallel for collapse(2) schedule(dynamic, 64)
    for (int y = y0; y < y1; y++) {
        for (int x = 0; x < width; x++) {
            long b = (long)(y / bin_size) * bdx + x / bin_size;
            FN(raster_pixel)(x, y, width, height, vertexes, vertex_total, bounds, P, W, Wsub, g, AA,
                             it, gs, ghost, gw, gh, spectrum, spectrum_n, st, bin_size, out_rgba,
                             lists[b], counts[b]);
        }
    }
    for (long b = 0; b < nb; b++) free(lists[b]);
    free(lists);
    free(counts);
    return ret;
}

/* Unbinned brute force: every pixel visits every prim, filtered by the same
 * bin predicate.  Used to show the binning changes nothing. */
int FN(rfo_rasterize_unbinned)(const VERTEX *vertexes, const REAL *bounds, int P, int W, int Wsub,
                               int g, int AA, double intensity, double ghost_scale,
                               const float *ghost, int gw, int gh, const float *spectrum,
                               int spectrum_n, double screen_transform, int bin_size, int width,
                               int height, REAL *out_rgba) {
    int R = g * g;
    long vertex_total = (long)P * R * W;
    REAL st = (REAL)(float)screen_transform;
    REAL it = (REAL)(float)intensity, gs = (REAL)(float)ghost_scale;
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++)
            FN(raster_pixel)(x, y, width, height, vertexes, vertex_total, bounds, P, W, Wsub, g, AA,
                             it, gs, ghost, gw, gh, spectrum, spectrum_n, st, bin_size, out_rgba,
                             NULL, 0);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* geometry.cl, noise.cl, aperture.cl, starburst.cl                         */
/* ------------------------------------------------------------------------ */

static inline V2 FN(ndc)(int px, int py, int w, int h) {
    V2 r;
    r.x = (REAL)2.0f * ((REAL)px / (REAL)(w - 1)) - (REAL)1.0f;
    r.y = (REAL)2.0f * ((REAL)py / (REAL)(h - 1)) - (REAL)1.0f;
    return r;
}
static inline V2 FN(rot)(V2 p, REAL angle) {
    REAL s = SIN(angle), c = COS(angle);
    V2 r = {p.x * c + p.y * s, p.y * c - p.x * s};
    return r;
}
static inline V2 FN(scale2)(V2 p, REAL sx, REAL sy) {
    V2 r = {p.x / FMAX(sx, (REAL)FLT_MIN), p.y / FMAX(sy, (REAL)FLT_MIN)};
    return r;
}
static inline REAL FN(rectangle)(V2 p, REAL hx, REAL hy) {
    REAL ex = FABS(p.x) - hx, ey = FABS(p.y) - hy;
    REAL ox = FMAX(ex, 0), oy = FMAX(ey, 0);
    REAL outside = SQRT(ox * ox + oy * oy);
    REAL inside = FMIN(FMAX(ex, ey), 0);
    return outside + inside;
}
/* noise.cl:1-5 */
static inline REAL FN(noise)(REAL x, REAL y, REAL z) {
    REAL v = SIN(x * (REAL)112.9898f + y * (REAL)179.233f + z * (REAL)237.212f) * (REAL)43758.5453f;
    REAL f = v - FLOOR(v);
    REAL lim = sizeof(REAL) == 4 ? (REAL)0x1.fffffep-1f : (REAL)0x1.fffffffffffffp-1;
    return FMIN(f, lim);
}
static inline REAL FN(wrap1)(REAL v, REAL m) { return FMOD(m + FMOD(v, m), m); }

/* aperture.cl:5-47 */
int FN(rfo_aperture_shape)(int w, int h, double size_x, double size_y, int blades,
                           double rotation, double roundness, double softness, REAL *out) {
    REAL sx = (REAL)(float)size_x, sy = (REAL)(float)size_y, rot = (REAL)(float)rotation;
    REAL rnd = (REAL)(float)roundness, soft = (REAL)(float)softness;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            V2 pos = FN(rot)(FN(scale2)(FN(ndc)(x, y, w, h), sx, sy), rot);
            REAL sdf = 0;
            for (int i = 0; i < blades; i++) {
                REAL angle = ((REAL)i / (REAL)blades + (REAL)0.25f) * 2 * PI_R;
                sdf = FMAX(sdf, COS(angle) * pos.x + SIN(angle) * pos.y);
            }
            REAL circular = ATAN2(-pos.x, -pos.y) / (2 * PI_R) + (REAL)0.5f;
            REAL blades_gradient = FMOD(circular * (REAL)blades + (REAL)0.5f, 1);
            REAL roundness_gradient = SIN(blades_gradient * PI_R);
            sdf += roundness_gradient * rnd;
            out[(long)y * w + x] = 1 - FN(smoothstep)(1 - soft, 1 + soft, sdf);
        }
    }
    return 0;
}

/* aperture.cl:50-86 */
int FN(rfo_aperture_grating)(int w, int h, double size_x, double size_y, double strength,
                             double density, double length, double width, double softness,
                             REAL *img) {
    REAL sx = (REAL)(float)size_x, sy = (REAL)(float)size_y, st = (REAL)(float)strength;
    REAL hx = (REAL)(float)width / 2, hy = (REAL)(float)length * 2 / 2, soft = (REAL)(float)softness;
    int count = (int)(FMIN((REAL)(float)density, 1) * 360);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            V2 pos = FN(scale2)(FN(ndc)(x, y, w, h), sx, sy);
            REAL sdf = 1, angle = 0;
            REAL offset = (REAL)1.0f / (REAL)count * (2 * PI_R);
            for (int i = 0; i < count; i++) {
                angle += offset;
                V2 rp = FN(rot)(pos, angle);
                rp.y -= (REAL)1.5f;
                sdf = FMIN(sdf, FN(rectangle)(rp, hx, hy));
            }
            img[(long)y * w + x] *= 1 - (st * FN(smoothstep)(soft, -soft, sdf));
        }
    }
    return 0;
}

/* aperture.cl:88-136 */
int FN(rfo_aperture_scratches)(int w, int h, double size_x, double size_y, double strength,
                               double density, double length, double rotation,
                               double rotation_variation, double width, double parallax_x,
                               double parallax_y, double softness, REAL *img) {
    REAL sx = (REAL)(float)size_x, sy = (REAL)(float)size_y, st = (REAL)(float)strength;
    REAL hx = (REAL)(float)width / 2, hy = (REAL)(float)length / 2, soft = (REAL)(float)softness;
    REAL rot = (REAL)(float)rotation, rv = (REAL)(float)rotation_variation * PI_R;
    REAL plx = (REAL)(float)parallax_x, ply = (REAL)(float)parallax_y;
    int count = (int)(FMIN((REAL)(float)density, 1) * 1000);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            V2 pos = FN(rot)(FN(scale2)(FN(ndc)(x, y, w, h), sx, sy), rot);
            REAL sdf = 1;
            for (int i = 0; i < count; i++) {
                REAL cx = FN(noise)((REAL)i, (REAL)count, 0) + plx;
                REAL cy = FN(noise)((REAL)i, (REAL)count, 1) + ply;
                cx = FN(wrap1)(cx, 1) * 2 - 1;
                cy = FN(wrap1)(cy, 1) * 2 - 1;
                REAL angle = (FN(noise)((REAL)i, (REAL)count, 2) - (REAL)0.5f) * rv;
                V2 rp = FN(rot)(pos, angle);
                rp.x -= cx;
                rp.y -= cy;
                sdf = FMIN(sdf, FN(rectangle)(rp, hx, hy));
            }
            img[(long)y * w + x] *= 1 - (st * FN(smoothstep)(soft, -soft, sdf));
        }
    }
    return 0;
}

/* aperture.cl:139-174 */
int FN(rfo_aperture_dust)(int w, int h, double size_x, double size_y, double strength,
                          double density, double radius, double parallax_x, double parallax_y,
                          double softness, REAL *img) {
    REAL sx = (REAL)(float)size_x, sy = (REAL)(float)size_y, st = (REAL)(float)strength;
    REAL rad = (REAL)(float)radius, soft = (REAL)(float)softness;
    REAL plx = (REAL)(float)parallax_x, ply = (REAL)(float)parallax_y;
    int count = (int)(FMIN((REAL)(float)density, 1) * 1000);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            V2 pos = FN(scale2)(FN(ndc)(x, y, w, h), sx, sy);
            REAL sdf = 1;
            for (int i = 0; i < count; i++) {
                REAL cx = FN(noise)((REAL)i, (REAL)count, 0) + plx;
                REAL cy = FN(noise)((REAL)i, (REAL)count, 1) + ply;
                cx = FN(wrap1)(cx, 1) * 2 - 1;
                cy = FN(wrap1)(cy, 1) * 2 - 1;
                REAL dx = pos.x - cx, dy = pos.y - cy;
                sdf = FMIN(sdf, SQRT(dx * dx + dy * dy) - rad);
            }
            img[(long)y * w + x] *= (REAL)1.0f - (st * FN(smoothstep)(soft, -soft, sdf));
        }
    }
    return 0;
}

/* starburst.cl:1-87.  `spectrum_img` is the power spectrum (h_s x w_s, one
 * channel), `light` the CIE table as RGBA rows (n entries). */
int FN(rfo_starburst)(int w, int h, const float *spectrum_img, int sw, int sh, const float *light,
                      int light_n, int samples, double blur, double rotation,
                      double rotation_weight, double vig_x, double vig_y, double intensity,
                      double offset_x, double offset_y, double size_x, double size_y,
                      REAL *out_rgba) {
    REAL bl = (REAL)(float)blur, rot = (REAL)(float)rotation, rw = (REAL)(float)rotation_weight;
    REAL vx = (REAL)(float)vig_x, vy = (REAL)(float)vig_y, it = (REAL)(float)intensity;
    REAL ox = (REAL)(float)offset_x, oy = (REAL)(float)offset_y;
    REAL sx = (REAL)(float)size_x, sy = (REAL)(float)size_y;
    const int lambda_delta = 730 - 390;
#pragma omp parallel for schedule(dynamic, 4)
    for (int py = 0; py < h; py++) {
        for (int px = 0; px < w; px++) {
            V2 ndc = FN(ndc)(px, py, w, h);
            ndc.x -= ox;
            ndc.y -= oy;
            ndc = FN(scale2)(ndc, sx, sy);
            REAL rgba[4] = {0, 0, 0, 0};
            for (int t = 0; t < samples; ++t) {
                REAL step = (REAL)t / (REAL)samples;
                REAL wavelength = step * (REAL)lambda_delta + (REAL)390;
                REAL seed = (REAL)(t * 4);
                V2 pos = ndc;
                REAL radius = bl * SQRT(FN(noise)((REAL)px, (REAL)py, seed));
                REAL alpha = FN(noise)((REAL)px, (REAL)py, seed + 1) * 2 * PI_R;
                pos.x += COS(alpha) * radius;
                pos.y += SIN(alpha) * radius;
                REAL s = 1 - (1 - (wavelength / (REAL)560)) * (REAL)1.0f;
                pos = FN(scale2)(pos, s, s);
                REAL sign = (FN(noise)((REAL)px, (REAL)py, seed + 2) > (REAL)0.5f) ? 1 : -1;
                REAL angle = sign * rot * (1 - POW(FN(noise)((REAL)px, (REAL)py, seed + 3), rw));
                pos = FN(rot)(pos, angle);
                REAL u = pos.x * (REAL)0.5f + (REAL)0.5f, v = pos.y * (REAL)0.5f + (REAL)0.5f;
                REAL fi = FN(sample_bilinear)(spectrum_img, sw, sh, 1, u, v, 0);
                for (int c = 0; c < 4; c++) {
                    REAL xyz = FN(sample_bilinear)(light + c, light_n, 1, 4, step, 0, 0);
                    rgba[c] += xyz * fi;
                }
            }
            for (int c = 0; c < 4; c++) rgba[c] /= (REAL)samples;
            if (!isnan(vx)) {
                REAL sdf = SQRT(ndc.x * ndc.x + ndc.y * ndc.y) - 0;
                REAL f = FN(smoothstep)(vy, vx, sdf);
                for (int c = 0; c < 4; c++) rgba[c] *= f;
            }
            for (int c = 0; c < 4; c++) rgba[c] *= it;
            FN(xyz_to_ap1)(rgba, &out_rgba[((long)py * w + px) * 4]);
        }
    }
    return 0;
}

/* Single-function entry points for the unit tests of the oracle itself. */
double FN(rfo_fresnel)(double theta0, double n1, double n2) {
    return (double)FN(fresnel)((REAL)theta0, (REAL)n1, (REAL)n2);
}
double FN(rfo_fresnel_ar)(double theta0, double lambda, double d1, double n0, double n1,
                          double n2) {
    return (double)FN(fresnel_ar)((REAL)theta0, (REAL)lambda, (REAL)d1, (REAL)n0, (REAL)n1,
                                  (REAL)n2);
}
double FN(rfo_dispersion)(double lambda, const float *coefficients) {
    REAL c[6];
    for (int k = 0; k < 6; k++) c[k] = (REAL)coefficients[k];
    return (double)FN(dispersion)((REAL)lambda, c);
}
int FN(rfo_intersect_quad)(int px, int py, const int *x, const int *y) {
    return FN(intersect_quad)(px, py, x, y);
}
void FN(rfo_barycentric_quad)(int px, int py, const int *x, const int *y, double *w) {
    REAL ww[4];
    FN(barycentric_quad)(px, py, x, y, ww);
    for (int i = 0; i < 4; i++) w[i] = (double)ww[i];
}
double FN(rfo_sample_bilinear)(const float *tex, int w, int h, double s, double t, int edge) {
    return (double)FN(sample_bilinear)(tex, w, h, 1, (REAL)s, (REAL)t, edge);
}

#undef ELEM
#undef V3
#undef V2
#undef RAY
#undef INTER
#undef VERTEX
#undef FN
