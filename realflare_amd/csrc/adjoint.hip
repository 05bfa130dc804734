// Hand-written adjoint (reverse mode) of the ghost ray trace on MI355X.
//
// Forward semantics are those of the `raytrace` kernel
// (realflare/api/tasks/cl/raytracing.cl:218-359, as in trace.hip): the same
// per-event operations in the same order, so a ray's trajectory here is
// bit-identical to rf_trace_planes_coated's.  Differentiated outputs per ray:
// sensor position (x, y) and reflectance; parameters: element radius,
// element centre (the host maps thicknesses onto centres), the coating
// wavelength of every element, and the effective refractive index of every
// element at every wavelength (the host chains Sellmeier B through it).
//
// One lane per ray, a one-wave workgroup per (ghost, wavelength, 64 rays), so
// the surface sequence is wave-uniform.  The forward pass stores the ray state
// every K events (checkpoints) in a per-ray scratch laid out [slot][field][ray]
// (coalesced); the reverse pass walks the segments backwards, recomputes
// each segment's per-event states from its checkpoint, then for every event
// recomputes the event's intermediates (intersection, normal, Fresnel terms)
// and applies their vector-Jacobian products.  Each lane accumulates its
// parameter gradients in its own LDS column (radius, centre, n_eff) and in two
// registers (the coatings of the ghost's two reflecting elements); the
// workgroup sums the columns
// with a fixed butterfly into partials laid out [parameter][workgroup], which a
// second kernel reduces per parameter in a fixed order: results are
// deterministic and independent of scheduling.
#include <cmath>

#include "fresnel.h"
#include "primitives.h"

namespace rf {

static constexpr int ADJ_THREADS = 64;  // one wave per workgroup
static constexpr int ADJ_FIELDS = 7;  // pos.xyz, dir.xyz, reflectance

struct St {
    f3 pos, dir;
    float refl;
};

// --- forward pieces (same arithmetic as trace.hip) ---------------------------

__device__ __noinline__ float fresnel_f(float theta0, float n1, float n2) {
    return fresnel_terms(theta0, n1, n2).r;
}

__device__ __noinline__ float fresnel_ar_f(float theta0, float lambda, float d1, float n0, float n1,
                                           float n2) {
    return fresnel_ar_terms(theta0, lambda, d1, n0, n1, n2).r;
}

__device__ float dispersion_f(float lambda, const float* c) {
#pragma clang fp contract(off)
    float l = (float)((double)lambda * 1e-3);
    float l2 = l * l;
    float d0 = (c[0] * l2) / (l2 - c[1]);
    float d1 = (c[2] * l2) / (l2 - c[3]);
    float d2 = (c[4] * l2) / (l2 - c[5]);
    return sqrtf(1 + d0 + d1 + d2);
}

struct Inter {
    f3 p, n;
    float incident;
    bool hit;
};

__device__ __forceinline__ Inter intersect_fwd(f3 pos, f3 dir, float radius, float center,
                                               bool want_incident) {
#pragma clang fp contract(off)
    Inter I;
    I.p = mk3(0, 0, 0);
    I.n = mk3(0, 0, 0);
    I.incident = 0.0f;
    if (dir.z == 0) {
        I.hit = false;
        return I;
    }
    if (radius == 0) {
        const float dz = -center - pos.z;
        I.p = add3(pos, scl3(dir, dz / dir.z));
        I.n = (dir.z < 0) ? mk3(0, 0, 1) : mk3(0, 0, -1);
        I.hit = true;
        return I;
    }
    const float r = fabsf(radius);
    const f3 C = mk3(0, 0, -center);
    const f3 U = sub3(C, pos);
    const f3 U1 = scl3(dir, dot3(U, dir));
    const float d = len3(sub3(U, U1));
    if (d > r) {
        I.hit = false;
        return I;
    }
    const float sgn = (radius * dir.z) > 0 ? -1.0f : 1.0f;
    const float m = sqrtf(r * r - d * d);
    I.p = sub3(add3(pos, U1), scl3(scl3(dir, m), sgn));
    const f3 w = sub3(I.p, C);
    const float l = len3(w);
    I.n = scl3(mk3(w.x / l, w.y / l, w.z / l), sgn);
    if (want_incident) I.incident = acosf(clampn(dot3(scl3(dir, -1.0f), I.n), -1.0f, 1.0f));
    I.hit = true;
    return I;
}

// --- reverse pieces -----------------------------------------------------------
//
// The reverse pass evaluates derivatives with the hardware reciprocal, square
// root and reciprocal square root (about 1 ulp): a gradient needs no correctly
// rounded intermediates, and no trajectory is ever advanced with them.

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }

// Surface geometry at the hit point p of a ray travelling along dir: the
// normal as intersect() orients it and, for spheres, w = p - centre, 1/|w|.
struct Hit {
    f3 n, nh, w;
    float il, sgn;
    bool sphere;
};

__device__ __forceinline__ Hit hit_at(f3 dir, f3 p, float radius, float center) {
    Hit h;
    h.sphere = radius != 0;
    if (!h.sphere) {
        h.n = (dir.z < 0) ? mk3(0, 0, 1) : mk3(0, 0, -1);
        return h;
    }
    h.w = sub3(p, mk3(0, 0, -center));
    h.il = frsq(dot3(h.w, h.w));
    h.nh = scl3(h.w, h.il);
    h.sgn = (radius * dir.z) > 0 ? -1.0f : 1.0f;
    h.n = scl3(h.nh, h.sgn);
    return h;
}

// Adjoint of the intersection (raytracing.cl:104-168): hit point p, normal n
// and, on spheres, the incident angle acos(-dir . n), given their adjoints pb,
// nb, ib.  The sphere hit is differentiated implicitly: p = pos + t dir with
// |p - C| = |radius| gives dt = (|radius| d|radius| - w . (dpos + t ddir - dC)) / (w . dir).
__device__ __forceinline__ void hit_bwd(const Hit& h, f3 pos, f3 dir, f3 p, float radius,
                                        float center, f3 pb, f3 nb, float ib, f3& posb, f3& dirb,
                                        float& Rb, float& Cb) {
#pragma clang fp contract(off)
    if (!h.sphere) {
        // p = pos + dir tau, tau = (-center - pos.z) / dir.z
        const float idz = frcp(dir.z);
        const float tau = (-center - pos.z) * idz;
        posb = add3(posb, pb);
        dirb = add3(dirb, scl3(pb, tau));
        const float dzb = dot3(pb, dir) * idz;
        dirb.z += -dzb * tau;
        Cb += -dzb;
        posb.z += -dzb;
        return;
    }
    if (ib != 0.0f) {
        const float c = -dot3(dir, h.n);
        const float q = 1.0f - c * c;
        if (c > -1.0f && c < 1.0f && q > 1e-12f) {
            const float cb = -ib * frsq(q);
            dirb = add3(dirb, scl3(h.n, -cb));
            nb = add3(nb, scl3(dir, -cb));
        }
    }
    // n = sgn w / |w|
    const f3 wb = scl3(sub3(nb, scl3(h.nh, dot3(nb, h.nh))), h.sgn * h.il);
    // w = p - C, C = (0, 0, -center)
    pb = add3(pb, wb);
    Cb += wb.z;
    // p = pos + t dir
    const float t = dot3(sub3(p, pos), dir);
    const float lam = dot3(pb, dir) * frcp(dot3(h.w, dir));
    const f3 q = sub3(pb, scl3(h.w, lam));
    posb = add3(posb, q);
    dirb = add3(dirb, scl3(q, t));
    Cb += -lam * h.w.z;
    Rb += lam * radius;
}

// refract (raytracing.cl:70-86) for a ray that did not totally reflect
__device__ __forceinline__ void refract_bwd(f3 i, f3 n, float o, f3 tb, f3& ib, f3& nb, float& ob) {
    const float cost = -dot3(i, n);
    const float sint2 = (o * o) * (1 - cost * cost);
    const float s = fsqrt(fabsf(1 - sint2));
    const float k = o * cost - s;
    ib = add3(ib, scl3(tb, o));
    nb = add3(nb, scl3(tb, k));
    const float kb = dot3(tb, n);
    ob += dot3(tb, i) + kb * cost;
    float costb = kb * o;
    const float sb = -kb;
    if (s > 0) {
        const float qb = sb * 0.5f * frcp(s);
        const float sint2b = (1 - sint2) >= 0 ? -qb : qb;
        ob += sint2b * 2 * o * (1 - cost * cost);
        costb += sint2b * o * o * (-2 * cost);
    }
    ib = add3(ib, scl3(n, -costb));
    nb = add3(nb, scl3(i, -costb));
}

// reflect (raytracing.cl:96-102): r = i - 2 (n.i) n
__device__ __forceinline__ void reflect_bwd(f3 i, f3 n, f3 rb, f3& ib, f3& nb) {
    const float nr = dot3(n, rb), ni = dot3(n, i);
    ib = add3(ib, sub3(rb, scl3(n, 2 * nr)));
    nb = add3(nb, add3(scl3(i, -2 * nr), scl3(rb, -2 * ni)));
}

// The Fresnel adjoints run on the two reflection events of a ghost only.  Each
// evaluates the event's forward terms once, with the forward's operations in
// the forward's order (contraction off), so the reflectance it returns is
// bit-identical to fresnel_f / fresnel_ar_f's; the gradient then reuses those
// terms.  Derivatives take hardware reciprocals and square roots (about
// 1 ulp), and cosines that the terms already fix come from identities
// (cos asin a = sqrt(1 - a^2), cos(x +- y) from the sines and cosines of x, y).
// Kept out of line, returning the reflectance and partials by value.
struct FresnelGrad {
    float r, t0, d1, n0, n1, n2;
};

// fresnel (raytracing.cl:1-16): R and Rb * dR/d(theta0, n1, n2), for R > 0
__device__ __noinline__ FresnelGrad fresnel_grad(float theta0, float n1, float n2, float Rb) {
    FresnelGrad g{0, 0, 0, 0, 0, 0};
    const FresnelTerms T = fresnel_terms(theta0, n1, n2);
    const float s0 = T.s0, c0 = T.c0, a1 = T.a1, c1 = T.c1, A = T.A, B = T.B, Cc = T.Cc, D = T.D;
    const float rs = T.rs, rp = T.rp;
    g.r = T.r;
    if (!(g.r > 0)) return g;
    const float rsb = Rb * rs, rpb = Rb * rp;
    // rs = (A - B)/(A + B): d/dA = 2B/(A+B)^2, d/dB = -2A/(A+B)^2
    const float is = frcp(A + B), ip = frcp(Cc + D);
    const float Ab = rsb * 2 * B * is * is, Bb = -rsb * 2 * A * is * is;
    const float Cb2 = rpb * 2 * D * ip * ip, Db = -rpb * 2 * Cc * ip * ip;
    g.n1 = Ab * c0 + Cb2 * c1;
    g.n2 = Bb * c1 + Db * c0;
    const float c0b = Ab * n1 + Db * n2;
    const float c1b = Bb * n2 + Cb2 * n1;
    const float th1b = -c1b * a1;  // d cos(theta1) / d theta1 = -sin theta1 = -a1
    float s0b = 0.0f;
    const float q = 1 - a1 * a1;
    if (q > 1e-12f) {
        const float a1b = th1b * frsq(q), in2 = frcp(n2);
        s0b += a1b * n1 * in2;
        g.n1 += a1b * s0 * in2;
        g.n2 += -a1b * a1 * in2;
    }
    g.t0 = s0b * c0 - c0b * s0;
    return g;
}

// fresnel_ar (raytracing.cl:18-59): R and Rb * dR/d(theta0, d1, n0, n1, n2), for R > 0
__device__ __noinline__ FresnelGrad fresnel_ar_grad(float theta0, float lambda, float d1, float n0,
                                                    float n1, float n2, float Rb) {
    FresnelGrad g{0, 0, 0, 0, 0, 0};
    const FresnelArTerms T = fresnel_ar_terms(theta0, lambda, d1, n0, n1, n2);
    const float s0 = T.s0, c0 = T.c0, a1 = T.a1, a2 = T.a2, c1 = T.c1, c2 = T.c2;
    const float sdm = T.sdm, sdp = T.sdp, cdm = T.cdm, cdp = T.cdp, tdm = T.tdm, tdp = T.tdp;
    const float sem = T.sem, sep = T.sep, cem = T.cem, cep = T.cep, tem = T.tem, tep = T.tep;
    const float rs01 = T.rs01, rp01 = T.rp01, ts01 = T.ts01, tp01 = T.tp01, rs12 = T.rs12,
                rp12 = T.rp12, ris = T.ris, rip = T.rip;
    const float dy = T.dy, t1 = T.t1, dx = T.dx, delay = T.delay, k4 = T.k4, ph = T.ph, cph = T.cph;
    const float s1 = a1;  // sin theta1
    g.r = T.r;
    if (!(g.r > 0)) return g;
    const float osb = Rb * 0.5f, opb = Rb * 0.5f;
    const float rs01b = osb * (2 * rs01 + 2 * ris * cph);
    const float risb = osb * (2 * ris + 2 * rs01 * cph);
    const float rp01b = opb * (2 * rp01 + 2 * rip * cph);
    const float ripb = opb * (2 * rip + 2 * rp01 * cph);
    const float cphb = osb * 2 * rs01 * ris + opb * 2 * rp01 * rip;
    const float phb = -cphb * sinf(ph);
    const float delayb = phb * k4;
    float dxb = -phb * k4 * s0;
    float s0b = -phb * k4 * dx;
    float dyb = 0.0f;
    if (delay > 0) {
        const float idl = frcp(delay);
        dxb += delayb * dx * idl;
        dyb += delayb * dy * idl;
    }
    const float t1b = dxb * dy;
    dyb += dxb * t1;
    g.d1 += dyb * n1;
    g.n1 += dyb * d1;
    float th1b = t1b * (1 + t1 * t1);
    const float tp01b = ripb * 2 * tp01 * rp12;
    const float rp12b = ripb * tp01 * tp01;
    float ts01b = risb * 2 * ts01 * rs12;
    const float rs12b = risb * ts01 * ts01;
    const float itep = frcp(tep), isep = frcp(sep);
    const float temb = rp12b * itep, tepb = -rp12b * tem * itep * itep;
    const float semb = -rs12b * isep, sepb = rs12b * sem * isep * isep;
    const float q1 = 1 - a1 * a1, q2 = 1 - a2 * a2;
    const float emb = temb * (1 + tem * tem) + semb * cem;
    const float epb = tepb * (1 + tep * tep) + sepb * cep;
    th1b += emb + epb;
    const float th2b = -emb + epb;
    ts01b += tp01b * cdm;
    const float cdmb = tp01b * ts01;
    float dmb = -cdmb * sdm;
    const float isdp = frcp(sdp), itdp = frcp(tdp);
    const float s1b = ts01b * 2 * c0 * isdp;
    const float c0b = ts01b * 2 * s1 * isdp;
    float sdpb = -ts01b * ts01 * isdp;
    th1b += s1b * c1;
    const float tdmb = rp01b * itdp, tdpb = -rp01b * tdm * itdp * itdp;
    const float sdmb = -rs01b * isdp;
    sdpb += rs01b * sdm * isdp * isdp;
    dmb += tdmb * (1 + tdm * tdm) + sdmb * cdm;
    const float dpb = tdpb * (1 + tdp * tdp) + sdpb * cdp;
    float t0 = dmb + dpb;
    th1b += -dmb + dpb;
    if (q2 > 1e-12f) {
        const float a2b = th2b * frsq(q2), in2 = frcp(n2);
        s0b += a2b * n0 * in2;
        g.n0 += a2b * s0 * in2;
        g.n2 += -a2b * a2 * in2;
    }
    if (q1 > 1e-12f) {
        const float a1b = th1b * frsq(q1), in1 = frcp(n1);
        s0b += a1b * n0 * in1;
        g.n0 += a1b * s0 * in1;
        g.n1 += -a1b * a1 * in1;
    }
    t0 += s0b * c0 - c0b * s0;
    g.t0 = t0;
    return g;
}

// --- event bookkeeping ------------------------------------------------------------

struct Path {
    int b1, b2, E, L, stop;
};

__device__ __forceinline__ int event_element(const Path& P, int k) {
    if (P.b1 < 0) return k;
    if (k <= P.b1) return k;
    if (k <= 2 * P.b1 - P.b2) return 2 * P.b1 - k;
    return k - 2 * (P.b1 - P.b2);
}
__device__ __forceinline__ bool event_reflects(const Path& P, int k) {
    return P.b1 >= 0 && (k == P.b1 || k == 2 * P.b1 - P.b2);
}

struct LensF {
    const rf_lens_element* lens;
    const float* coating;  // per element, or NULL
    float cmin;
    float wavelength;
};

// s_n holds a workgroup's refractive indices in [0, L) and, from RF_LMAX on,
// the refraction ratio n1 / n2 of element i for a ray travelling towards -z
// (s_n[RF_LMAX + 2 i]) and towards +z (s_n[RF_LMAX + 2 i + 1]), as the
// tracer's per-workgroup table (the same divisions, so the same values).
// Fills the ratios; the indices must be written and visible to every thread.
__device__ __forceinline__ void refraction_ratios(float* s_n, int L, int t) {
#pragma clang fp contract(off)
    if (t < L) {
        const float n2 = s_n[t];
        s_n[RF_LMAX + 2 * t] = (t - 1 >= 0 ? s_n[t - 1] : 1.0f) / n2;
        s_n[RF_LMAX + 2 * t + 1] = (t + 1 < L ? s_n[t + 1] : 1.0f) / n2;
    }
}

// One event forward (trace.hip semantics).  Returns false when the ray dies.
__device__ __forceinline__ bool event_fwd(const LensF& Lf, const float* s_n, int L, int stop, int el,
                                          bool reflect, St& s) {
#pragma clang fp contract(off)
    const rf_lens_element& E = Lf.lens[el];
    Inter I = intersect_fwd(s.pos, s.dir, E.radius, E.center, reflect);
    if (!I.hit) return false;
    const f3 dir_in = s.dir;
    s.pos = I.p;
    if (el == stop || el == L - 1) return true;
    const int n_index = (dir_in.z < 0) ? el - 1 : el + 1;
    const float n1 = (0 <= n_index && n_index < L) ? s_n[n_index] : 1.0f;
    const float n2 = s_n[el];
    if (reflect) {
        s.dir = add3(dir_in, scl3(I.n, 2.0f * dot3(I.n, scl3(dir_in, -1.0f))));
        const float theta = (float)((double)I.incident + 1e-9);
        const float cf = Lf.coating ? Lf.coating[el] : (float)E.coating;
        float r;
        if (cf > 0) {
            const float nc = fmaxf(sqrtf(n1 * n2), Lf.cmin);
            const float d1 = cf / (4.0f * nc);
            r = fresnel_ar_f(theta, Lf.wavelength, d1, n1, nc, n2);
        } else {
            r = fresnel_f(theta, n1, n2);
        }
        if (r > 0) s.refl *= r;
        return true;
    }
    const float o = s_n[RF_LMAX + 2 * el + (dir_in.z < 0 ? 0 : 1)];  // n1 / n2
    const float cost = dot3(scl3(dir_in, -1.0f), I.n);
    const float sint2 = (o * o) * (1 - cost * cost);
    f3 t = add3(scl3(dir_in, o), scl3(I.n, o * cost - sqrtf(fabsf(1 - sint2))));
    s.dir = scl3(t, (float)(sint2 < 1));
    return s.dir.z != 0;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += (uint32_t)__shfl_xor((int)v, m, 64);
    return v;
}

// Per-lane parameter adjoints, summed over the wave once at the end.  Radius,
// centre and n_eff of every element: column `lane` of the workgroup's LDS
// array [3 L][lane] (rows radius 0..L-1, centre L..2L-1, n_eff 2L..3L-1).  A
// ghost reflects at exactly two elements (b1, then b2), so its coating
// adjoints are two registers.
struct Grad {
    float* lacc;
    float coat_b1, coat_b2;
};
static constexpr int ADJ_LDS_ROWS = 3;  // LDS rows per element
__device__ __forceinline__ void acc(Grad& G, int idx, float v) {
    G.lacc[idx * ADJ_THREADS + threadIdx.x] += v;
}
__device__ __forceinline__ void grad_clear(Grad& G, int L) {
    for (int i = 0; i < ADJ_LDS_ROWS * L; i++) G.lacc[i * ADJ_THREADS + threadIdx.x] = 0.0f;
    G.coat_b1 = 0.0f;
    G.coat_b2 = 0.0f;
}
// Wave sum of parameter i in the order [radius L][centre L][coating L][n_eff L].
__device__ __forceinline__ float grad_sum(const Grad& G, const Path& P, int i) {
    const int L = P.L;
    if (i < 2 * L) return wave_sum(G.lacc[i * ADJ_THREADS + threadIdx.x]);
    if (i >= 3 * L) return wave_sum(G.lacc[(i - L) * ADJ_THREADS + threadIdx.x]);
    const int e = i - 2 * L;
    float v = 0.0f;
    if (e == P.b1) v += wave_sum(G.coat_b1);
    if (e == P.b2) v += wave_sum(G.coat_b2);
    return v;
}

__device__ __forceinline__ void st_store(float* scratch, long N, long ray, int slot, const St& s) {
    float* b = scratch + (long)slot * ADJ_FIELDS * N;
    b[0 * N + ray] = s.pos.x;
    b[1 * N + ray] = s.pos.y;
    b[2 * N + ray] = s.pos.z;
    b[3 * N + ray] = s.dir.x;
    b[4 * N + ray] = s.dir.y;
    b[5 * N + ray] = s.dir.z;
    b[6 * N + ray] = s.refl;
}
__device__ __forceinline__ St st_load(const float* scratch, long N, long ray, int slot) {
    const float* b = scratch + (long)slot * ADJ_FIELDS * N;
    St s;
    s.pos = mk3(b[0 * N + ray], b[1 * N + ray], b[2 * N + ray]);
    s.dir = mk3(b[3 * N + ray], b[4 * N + ray], b[5 * N + ray]);
    s.refl = b[6 * N + ray];
    return s;
}

// Reverse of event k of a ray that reaches the sensor: `in` is the ray state
// before the event and p its hit point (the position after it); (posb, dirb,
// reflb) the adjoint of the state after the event on entry and before it on
// return.  Parameter adjoints go to the lane's LDS column.
__device__ __forceinline__ void event_rev(const LensF& Lf, const float* s_n, Grad& G,
                                          const Path& P, int k, const St& in, f3 p, f3& posb,
                                          f3& dirb, float& reflb) {
#pragma clang fp contract(off)
    const int L = P.L;
    const int el = event_element(P, k);
    const bool reflect = event_reflects(P, k);
    const rf_lens_element& Ei = Lf.lens[el];
    const Hit h = hit_at(in.dir, p, Ei.radius, Ei.center);
    float Rb = 0.0f, Cb = 0.0f;
    f3 nb = mk3(0, 0, 0), dirb_in = mk3(0, 0, 0);
    float ib = 0.0f;
    if (el == P.stop || el == L - 1) {
        dirb_in = dirb;
    } else {
        const int n_index = (in.dir.z < 0) ? el - 1 : el + 1;
        const bool inrange = 0 <= n_index && n_index < L;
        const float n1 = inrange ? s_n[n_index] : 1.0f;
        const float n2 = s_n[el];
        float n1b = 0.0f, n2b = 0.0f;
        if (reflect) {
            reflect_bwd(in.dir, h.n, dirb, dirb_in, nb);
            const float incident =
                h.sphere ? acosf(clampn(-dot3(in.dir, h.n), -1.0f, 1.0f)) : 0.0f;
            const float theta = (float)((double)incident + 1e-9);
            const float cf = Lf.coating ? Lf.coating[el] : (float)Ei.coating;
            float r, t0b = 0.0f;
            if (cf > 0) {
                const float q = sqrtf(n1 * n2);
                const float nc = fmaxf(q, Lf.cmin);
                const float d1 = cf / (4.0f * nc);
                const FresnelGrad fg =
                    fresnel_ar_grad(theta, Lf.wavelength, d1, n1, nc, n2, reflb * in.refl);
                r = fg.r;
                if (r > 0) {
                    t0b = fg.t0;
                    n1b = fg.n0;
                    n2b = fg.n2;
                    const float d1b = fg.d1;
                    float ncb = fg.n1;
                    if (k == P.b1) G.coat_b1 += d1b / (4.0f * nc);
                    else G.coat_b2 += d1b / (4.0f * nc);
                    ncb += -d1b * d1 / nc;
                    if (q > Lf.cmin && q > 0) {
                        n1b += ncb * n2 / (2 * q);
                        n2b += ncb * n1 / (2 * q);
                    }
                }
            } else {
                const FresnelGrad fg = fresnel_grad(theta, n1, n2, reflb * in.refl);
                r = fg.r;
                if (r > 0) {
                    t0b = fg.t0;
                    n1b = fg.n1;
                    n2b = fg.n2;
                }
            }
            if (r > 0) reflb = reflb * r;
            ib = t0b;
        } else {
            float ob = 0.0f;
            const float in2 = frcp(n2);
            refract_bwd(in.dir, h.n, n1 * in2, dirb, dirb_in, nb, ob);
            n1b += ob * in2;
            n2b += -ob * n1 * in2 * in2;
        }
        acc(G, 2 * L + el, n2b);
        if (inrange) acc(G, 2 * L + n_index, n1b);
    }
    f3 posb_in = mk3(0, 0, 0);
    hit_bwd(h, in.pos, in.dir, p, Ei.radius, Ei.center, posb, nb, ib, posb_in, dirb_in, Rb, Cb);
    posb = posb_in;
    dirb = dirb_in;
    acc(G, el, Rb);
    acc(G, L + el, Cb);
}

// init_ray (raytracing.cl:178-215) of grid ray `ray_id`: the vertical probe
// from the grid point onto surface 0, then one unit back along the light.
struct RayInit {
    f3 gpt, down;
    Inter I0;
    St s;
};

__device__ __forceinline__ RayInit init_ray_state(const rf_lens_element* __restrict__ lens, int g,
                                                  float grid_length, const float* direction,
                                                  int ray_id) {
#pragma clang fp contract(off)
    RayInit r;
    const int ray_y = (int)truncf((float)ray_id / (float)g);
    const int ray_x = ray_id - ray_y * g;
    r.gpt = mk3((float)((double)grid_length * ((double)((float)ray_x / (float)(g - 1)) - 0.5)),
                (float)((double)grid_length * (0.5 - (double)((float)ray_y / (float)(g - 1)))),
                1.0f);
    r.down = mk3(0, 0, -1.0f);
    r.I0 = intersect_fwd(r.gpt, r.down, lens[0].radius, lens[0].center, false);
    const f3 d0 = norm3(mk3(-direction[0], -direction[1], -direction[2]));
    r.s.pos = sub3(r.I0.p, d0);
    r.s.dir = d0;
    r.s.refl = 1.0f;
    return r;
}

// Forward pass of events [k0, P.E) of one ray, starting from s (the state
// before event k0; `alive` whether the ray still travels there).  Segments of
// K events start at k0: the current segment's input states are kept in st[]
// (on return, the last segment's); segment start states go to checkpoint slot
// = segment of the ray's record (sidx in a [slot][field][N] array): every
// segment's when store_all, else all but the last.  Returns whether the ray
// reaches the sensor; s is the final state.
template <int K>
__device__ __forceinline__ bool forward_span(const LensF& Lf, const float* s_n, const Path& P,
                                             int k0, St& s, bool alive, St (&st)[K],
                                             float* scratch, long N, long sidx, bool active,
                                             bool store_all) {
    const int nseg = (P.E - k0 + K - 1) / K;
    for (int seg = 0; seg < nseg; seg++) {
        const int kb = k0 + seg * K;
        if (active && (store_all || seg + 1 < nseg)) st_store(scratch, N, sidx, seg, s);
#pragma unroll
        for (int j = 0; j < K; j++) {
            if (kb + j < P.E) {
                st[j] = s;
                if (alive)
                    alive = event_fwd(Lf, s_n, P.L, P.stop, event_element(P, kb + j),
                                      event_reflects(P, kb + j), s);
            }
        }
    }
    return alive;
}

template <int K>
__device__ __forceinline__ bool forward_ray(const LensF& Lf, const float* s_n, const Path& P, St& s,
                                            St (&st)[K], float* scratch, long N, long sidx,
                                            bool active, bool store_all) {
    return forward_span<K>(Lf, s_n, P, 0, s, active, st, scratch, N, sidx, active, store_all);
}

// Reverse pass of events [k0, P.E) of a ray that reaches the sensor with final
// state `fin`: (posb, dirb, reflb) is the adjoint of the final state on entry,
// (posb, dirb) that of the state before event k0 on return.  Segments are
// walked backwards and rebuilt from their checkpoints (the last one too unless
// its states are still in st[]).
template <int K>
__device__ __forceinline__ void reverse_span(const LensF& Lf, const float* s_n, Grad& G,
                                             const Path& P, int k0, const St& fin, St (&st)[K],
                                             const float* scratch, long N, long sidx,
                                             bool last_in_registers, f3& posb, f3& dirb,
                                             float reflb) {
#pragma clang fp contract(off)
    const int nseg = (P.E - k0 + K - 1) / K;
    f3 hit = fin.pos;  // hit point of the event being reversed
    for (int seg = nseg - 1; seg >= 0; seg--) {
        const int kb = k0 + seg * K;
        if (!last_in_registers || seg + 1 < nseg) {
            st[0] = st_load(scratch, N, sidx, seg);
#pragma unroll
            for (int j = 1; j < K; j++) {
                if (kb + j < P.E) {
                    st[j] = st[j - 1];
                    event_fwd(Lf, s_n, P.L, P.stop, event_element(P, kb + j - 1),
                              event_reflects(P, kb + j - 1), st[j]);
                }
            }
        }
#pragma unroll
        for (int j = K - 1; j >= 0; j--) {
            if (kb + j < P.E) {
                event_rev(Lf, s_n, G, P, kb + j, st[j], hit, posb, dirb, reflb);
                hit = st[j].pos;
            }
        }
    }
}

// init_ray (pos0 = I0.p - d0 with the vertical probe ray; d0 is constant) in
// reverse, given the adjoint posb of the first state.
__device__ __forceinline__ void init_rev(const LensF& Lf, Grad& G, int L, const RayInit& ri,
                                         f3 posb) {
    if (ri.I0.hit) {
        float Rb = 0.0f, Cb = 0.0f;
        f3 gpb = mk3(0, 0, 0), ddb = mk3(0, 0, 0);
        const rf_lens_element& e0 = Lf.lens[0];
        const Hit h0 = hit_at(ri.down, ri.I0.p, e0.radius, e0.center);
        hit_bwd(h0, ri.gpt, ri.down, ri.I0.p, e0.radius, e0.center, posb, mk3(0, 0, 0), 0.0f, gpb,
                ddb, Rb, Cb);
        acc(G, 0, Rb);
        acc(G, L, Cb);
    }
}

// Reverse pass of a whole ray that reaches the sensor (events [0, P.E), then
// init_ray).
template <int K>
__device__ __forceinline__ void reverse_ray(const LensF& Lf, const float* s_n, Grad& G,
                                            const Path& P, const RayInit& ri, const St& fin,
                                            St (&st)[K], const float* scratch, long N, long sidx,
                                            bool last_in_registers, f3 posb, f3 dirb, float reflb) {
    reverse_span<K>(Lf, s_n, G, P, 0, fin, st, scratch, N, sidx, last_in_registers, posb, dirb,
                    reflb);
    init_rev(Lf, G, P.L, ri, posb);
}

// What one launch of k_trace_adjoint does.
enum AdjMode {
    ADJ_FUSED = 0,    // forward, then reverse (the last segment's states stay in registers)
    ADJ_FORWARD = 1,  // forward only: outputs, every segment's checkpoint and the final state
    ADJ_REVERSE = 2,  // reverse only, from what an ADJ_FORWARD launch left in the scratch
};

// Ghosts per workgroup of k_trace_adjoint.  A ghost path (b1, b2) runs the
// straight path's events through elements 0 .. b1 - 1 (the same refractions
// for every ghost of one grid ray and wavelength) before its first
// reflection.  The workgroup traces that straight chain once for its group of
// ghosts, keeping every state S_j (before element j) in the straight tape,
// and each ghost from S_b1 on.  In reverse, the loss is a sum over ghosts, so
// the straight chain's adjoint at S_j is the sum of the adjoints that the
// ghosts branching off at or after j hand down; the chain is reversed once,
// from its far end to S_0, taking each ghost's adjoint where the ghost branches
// off (ghosts in decreasing b1).  The straight events are the same operations
// on the same values as a per-ghost trace, so the outputs are the same bits.
static constexpr int ADJ_GHOST_GROUP = 4;

__device__ __forceinline__ Path ghost_path(const int* __restrict__ paths, int p, int L, int stop) {
    Path P;
    P.b1 = paths[2 * p];
    P.b2 = paths[2 * p + 1];
    P.L = L;
    P.stop = stop;
    P.E = P.b1 >= 0 ? L + 2 * (P.b1 - P.b2) : L;
    return P;
}

// The first event of ghost path P that differs from the straight path's
// (its first reflection); a path without reflection follows the straight
// path throughout; irregular paths share nothing.
__device__ __forceinline__ int branch_event(const Path& P) {
    if (P.b1 < 0) return P.E;
    if (P.b1 < P.L && P.b2 >= 0 && P.b2 < P.b1) return P.b1;
    return 0;
}

// Partials (floats) are laid out [parameter][workgroup], parameters ordered
// [radius L][centre L][coating L][n_eff L] (n_eff for the workgroup's wavelength).
//
// K events form a segment.  The forward pass keeps the current segment's input
// states in registers (st[], indexed by unrolled loops) and stores segment
// start states in the scratch (per ghost ray, slot = segment counted from the
// ghost's branch event; the final state, reflectance NaN for a ray that does
// not reach the sensor, in the slot after the last segment); the reverse pass
// rebuilds each segment from its checkpoint, walking segments backwards.
template <int K, int MODE>
__global__ __launch_bounds__(ADJ_THREADS) void k_trace_adjoint(
    const rf_lens_element* __restrict__ lens, const float* __restrict__ coating,
    const int* __restrict__ paths, const int* __restrict__ wavelengths, rf_trace_params prm,
    const float* __restrict__ gx, const float* __restrict__ gy, const float* __restrict__ gr,
    float* __restrict__ scratch, float* __restrict__ straight, float* __restrict__ partial,
    float* __restrict__ fwd_out) {
#pragma clang fp contract(off)
    __shared__ float s_n[3 * RF_LMAX];  // indices, then refraction ratios
    __shared__ float s_coat[RF_LMAX];   // the group's coating adjoints (wave sums)
    extern __shared__ float s_lacc[];   // [ADJ_LDS_ROWS L][ADJ_THREADS]
    const int L = prm.lens_count;
    const int w = blockIdx.y, grp = blockIdx.z;
    const int g = prm.grid_count, R = g * g, W = prm.wavelength_count, NP = prm.path_count;
    const float wavelength = (float)wavelengths[w];
    const bool disperse = !isnan(lens[0].coefficients[0]);
    if ((int)threadIdx.x < L) {
        const rf_lens_element& e = lens[threadIdx.x];
        float n = e.ior;
        if (disperse && n > 1) n = dispersion_f(wavelength, e.coefficients);
        s_n[threadIdx.x] = n;
        s_coat[threadIdx.x] = 0.0f;
    }
    Grad G{s_lacc, 0.0f, 0.0f};
    if (MODE != ADJ_FORWARD) grad_clear(G, L);
    __syncthreads();
    refraction_ratios(s_n, L, threadIdx.x);
    __syncthreads();

    const long N = (long)NP * W * R;  // ghost-ray records
    const long NS = (long)gridDim.z * W * R;  // straight-chain records
    const int ray_id = blockIdx.x * ADJ_THREADS + threadIdx.x;
    const bool active = ray_id < R;
    const int rid = active ? ray_id : 0;
    const long sidx = ((long)grp * W + w) * R + rid;
    LensF Lf{lens, coating, prm.coating_min_ior, wavelength};
    // the group's ghosts, in decreasing branch event (then increasing index):
    // an insertion sort of (branch event, index) keys over constant indices,
    // the order packed 4 bits a ghost
    const int p0 = grp * ADJ_GHOST_GROUP;
    const int np = min(ADJ_GHOST_GROUP, NP - p0);
    int key[ADJ_GHOST_GROUP];
    int jmax = 0;
#pragma unroll
    for (int i = 0; i < ADJ_GHOST_GROUP; i++) {
        const int p = p0 + min(i, np - 1);
        const Path P = ghost_path(paths, p, L, prm.aperture_index);
        const int k0 = i < np ? branch_event(P) : -1;
        jmax = max(jmax, k0);
        key[i] = (k0 + 1) * 16 + (ADJ_GHOST_GROUP - 1 - i);  // larger first
    }
#pragma unroll
    for (int i = 1; i < ADJ_GHOST_GROUP; i++)
#pragma unroll
        for (int j = i; j > 0; j--)
            if (key[j] > key[j - 1]) {
                const int t = key[j];
                key[j] = key[j - 1];
                key[j - 1] = t;
            }
    uint32_t order = 0;
#pragma unroll
    for (int i = 0; i < ADJ_GHOST_GROUP; i++)
        order |= (uint32_t)(ADJ_GHOST_GROUP - 1 - (key[i] & 15)) << (4 * i);
    const Path Ps{-1, -1, L, L, prm.aperture_index};  // the straight path
    St st[K];

    // straight chain S_0 .. S_jmax; died = the event that ends the ray (or jmax)
    int died = jmax;
    if (MODE != ADJ_REVERSE) {
        St s = init_ray_state(lens, g, prm.grid_length, prm.direction, rid).s;
        bool alive = active;
        for (int j = 0; j <= jmax; j++) {
            if (active) st_store(straight, NS, sidx, j, s);
            if (j == jmax) break;
            if (alive) {
                alive = event_fwd(Lf, s_n, L, prm.aperture_index, j, false, s);
                if (!alive) died = j;
            }
        }
    }
    // adjoint of the straight state at element jcur
    f3 Ap = mk3(0, 0, 0), Ad = mk3(0, 0, 0);
    int jcur = jmax;
    // straight events jcur - 1 .. j_to in reverse (lanes holding a non-zero
    // adjoint: a ghost branching off at or after jcur reached the sensor)
    auto straight_rev = [&](int j_to) {
        for (int j = jcur - 1; j >= j_to; j--) {
            const bool any = Ap.x != 0.0f || Ap.y != 0.0f || Ap.z != 0.0f || Ad.x != 0.0f ||
                             Ad.y != 0.0f || Ad.z != 0.0f;
            if (active && any) {
                const St in = st_load(straight, NS, sidx, j);
                const f3 hit = st_load(straight, NS, sidx, j + 1).pos;
                float rb = 0.0f;
                event_rev(Lf, s_n, G, Ps, j, in, hit, Ap, Ad, rb);
            }
        }
        jcur = min(jcur, j_to);
    };
#pragma unroll 1
    for (int oi = 0; oi < np; oi++) {
        const int i = (int)((order >> (4 * oi)) & 15u);
        const Path P = ghost_path(paths, p0 + i, L, prm.aperture_index);
        const int k0 = branch_event(P);
        const long idx = ((long)(p0 + i) * W + w) * R + rid;
        const int nseg = (P.E - k0 + K - 1) / K;
        if (MODE != ADJ_FORWARD) straight_rev(k0);
        St s;
        bool alive;
        if (MODE != ADJ_REVERSE) {
            s = St{mk3(0, 0, 0), mk3(0, 0, 0), 0.0f};  // lanes past the grid: unused
            if (active) s = st_load(straight, NS, sidx, k0);
            alive = active && died >= k0;
            alive = forward_span<K>(Lf, s_n, P, k0, s, alive, st, scratch, N, idx, active,
                                    MODE == ADJ_FORWARD);
            if (fwd_out && active) {
                fwd_out[idx] = alive ? s.pos.x : NAN;
                fwd_out[N + idx] = alive ? s.pos.y : NAN;
                fwd_out[2 * N + idx] = alive ? s.refl : NAN;
            }
            if (MODE == ADJ_FORWARD) {
                if (active) {
                    St fin = s;
                    if (!alive) fin.refl = NAN;
                    st_store(scratch, N, idx, nseg, fin);
                }
                continue;
            }
        } else {
            alive = false;
            if (active) {
                s = st_load(scratch, N, idx, nseg);
                alive = !isnan(s.refl);
            }
        }
        // the ghost's events in reverse (rays that do not reach the sensor
        // contribute nothing), then its adjoint joins the straight chain's
        G.coat_b1 = 0.0f;
        G.coat_b2 = 0.0f;
        if (alive) {
            f3 pb = mk3(gx[idx], gy[idx], 0.0f), db = mk3(0, 0, 0);
            reverse_span<K>(Lf, s_n, G, P, k0, s, st, scratch, N, idx, MODE == ADJ_FUSED, pb, db,
                            gr[idx]);
            Ap = add3(Ap, pb);
            Ad = add3(Ad, db);
        }
        // coating adjoints of the ghost's two reflecting elements
        const float c1 = wave_sum(G.coat_b1), c2 = wave_sum(G.coat_b2);
        if (threadIdx.x == 0) {
            if (P.b1 >= 0 && P.b1 < L) s_coat[P.b1] += c1;
            if (P.b2 >= 0 && P.b2 < L) s_coat[P.b2] += c2;
        }
    }
    if (MODE == ADJ_FORWARD) return;
    straight_rev(0);
    {
        const bool any = Ap.x != 0.0f || Ap.y != 0.0f || Ap.z != 0.0f;
        if (active && any)
            init_rev(Lf, G, L, init_ray_state(lens, g, prm.grid_length, prm.direction, rid), Ap);
    }
    __syncthreads();  // s_coat complete
    const long blk = ((long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    const long nblk = (long)gridDim.x * gridDim.y * gridDim.z;
    for (int i = 0; i < 4 * L; i++) {
        float v;
        if (i < 2 * L) v = wave_sum(G.lacc[i * ADJ_THREADS + threadIdx.x]);
        else if (i < 3 * L) v = s_coat[i - 2 * L];
        else v = wave_sum(G.lacc[(i - L) * ADJ_THREADS + threadIdx.x]);
        if (threadIdx.x == 0) partial[i * nblk + blk] = v;
    }
}

static constexpr int RED_THREADS = 256;

// out[0:L) radius, [L:2L) centre, [2L:3L) coating, [3L + w L + i) n_eff[w][i].
// One workgroup per output: lane j sums the output's workgroup partials
// j, j + 256, ... in order, then the lanes combine in a fixed tree.
__global__ __launch_bounds__(RED_THREADS) void k_adjoint_reduce(const float* __restrict__ partial,
                                                                 int L, int bx, int W, int P,  // P: ghost groups
                                                                 float* __restrict__ out) {
    __shared__ float s_w[RED_THREADS / 64];
    const int t = blockIdx.x;
    const long nblk = (long)bx * W * P;
    float v = 0.0f;
    if (t < 3 * L) {
        const float* src = partial + (long)t * nblk;
        for (long j = threadIdx.x; j < nblk; j += RED_THREADS) v += src[j];
    } else {
        const int w = (t - 3 * L) / L, i = (t - 3 * L) % L;
        const float* src = partial + (long)(3 * L + i) * nblk;
        const long n = (long)P * bx;  // the workgroups of wavelength w: bx per ghost group
        for (long j = threadIdx.x; j < n; j += RED_THREADS) {
            const long p = j / bx, x = j - p * bx;
            v += src[(p * W + w) * bx + x];
        }
    }
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = 0.0f;
        for (int k = 0; k < RED_THREADS / 64; k++) r += s_w[k];
        out[t] = r;
    }
}

typedef void (*AdjKernel)(const rf_lens_element*, const float*, const int*, const int*,
                          rf_trace_params, const float*, const float*, const float*, float*, float*,
                          float*, float*);

template <int MODE>
static AdjKernel adj_kernel_k(int K) {
    switch (K) {
        case 1: return k_trace_adjoint<1, MODE>;
        case 2: return k_trace_adjoint<2, MODE>;
        case 3: return k_trace_adjoint<3, MODE>;
        case 4: return k_trace_adjoint<4, MODE>;
        case 6: return k_trace_adjoint<6, MODE>;
        case 8: return k_trace_adjoint<8, MODE>;
        default: return nullptr;
    }
}

// The segment length is a compile-time constant so that a segment's states
// live in registers.
static AdjKernel adj_kernel(int K, int mode = ADJ_FUSED) {
    switch (mode) {
        case ADJ_FUSED: return adj_kernel_k<ADJ_FUSED>(K);
        case ADJ_FORWARD: return adj_kernel_k<ADJ_FORWARD>(K);
        case ADJ_REVERSE: return adj_kernel_k<ADJ_REVERSE>(K);
        default: return nullptr;
    }
}

// ---------------------------------------------------------------------------
// Batched fit step: lens instances x light positions x ghosts x wavelengths in
// one persistent launch, forward + loss + reverse per ray.
//
// A chunk is 64 * rpl rays of one (segment, wavelength), chunks wavelength-
// major within a segment (chunk_offsets as rf_trace_batch's for 64-lane
// chunks).  One wave per workgroup pops chunks; each lane runs its rays one
// after the other, keeping checkpoints in its own scratch column
// ([workgroup][slot][field][64]), so scratch does not grow with the batch.
// The loss of a ray that reaches the sensor in the trace and in the target is
// loss_scale * (f(dx) + f(dy) + f(dr / refl_scale)), f the square or its Huber
// form (fit_term); its gradient seeds the reverse pass in place.  Partials per chunk are laid out [chunk][4 LF + 1]:
// radius, centre, coating, n_eff of the chunk's wavelength (each LF wide),
// then the loss.

struct FitArgs {
    const rf_lens_element* lenses;
    const float* coatings;
    const int* lens_table;  // [instance][3]: first element, L, aperture index
    const int* paths;
    const int* wavelengths;
    const rf_batch_segment* segs;
    const uint64_t* chunk_offsets;
    const float *tx, *ty, *tr;
    uint64_t* head;
    float* scratch;
    float* partial;
    float* fwd_out;
    long ray_total;
    int nseg, W, LF, rpl, slots;
    float cmin, inv_refl_scale, loss_scale, huber;
};

// Per-component loss and its derivative: r^2 (derivative 2 r), or with a
// Huber threshold delta > 0 and |r| > delta, 2 delta |r| - delta^2 (derivative
// 2 delta sign r): large residuals (rays whose path changes between the trace
// and the target) then pull linearly instead of quadratically.
__device__ __forceinline__ float fit_term(float r, float delta, float* d) {
#pragma clang fp contract(off)
    const float ar = fabsf(r);
    if (delta > 0.0f && ar > delta) {
        *d = 2.0f * copysignf(delta, r);
        return 2.0f * delta * ar - delta * delta;
    }
    *d = 2.0f * r;
    return r * r;
}

template <int K>
__global__ __launch_bounds__(ADJ_THREADS) void k_trace_fit(FitArgs a) {
#pragma clang fp contract(off)
    __shared__ float s_n[3 * RF_LMAX];  // indices, then refraction ratios
    __shared__ uint64_t s_chunk;
    extern __shared__ float s_lacc[];  // [ADJ_LDS_ROWS L][ADJ_THREADS]
    const int lane = threadIdx.x;
    const uint64_t total = a.chunk_offsets[a.nseg];
    float* scratch = a.scratch + (long)blockIdx.x * a.slots * ADJ_FIELDS * ADJ_THREADS;
    const int pstride = 4 * a.LF + 2;
    for (;;) {
        if (lane == 0) s_chunk = atomicAdd((unsigned long long*)a.head, 1ull);
        __syncthreads();
        const uint64_t c = s_chunk;
        __syncthreads();
        if (c >= total) break;
        int lo = 0, hi = a.nseg - 1;  // segment of chunk c
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (a.chunk_offsets[mid] <= c) lo = mid;
            else hi = mid - 1;
        }
        const rf_batch_segment sg = a.segs[lo];
        const int g = sg.grid_count, R = g * g;
        const int per_w = (R + ADJ_THREADS * a.rpl - 1) / (ADJ_THREADS * a.rpl);
        const int local = (int)(c - a.chunk_offsets[lo]);
        const int w = local / per_w, blk = local - w * per_w;
        const int first = a.lens_table[3 * sg.lens], L = a.lens_table[3 * sg.lens + 1];
        const rf_lens_element* lens = a.lenses + first;
        const float wavelength = (float)a.wavelengths[w];
        const bool disperse = !isnan(lens[0].coefficients[0]);
        if (lane < L) {
            const rf_lens_element& e = lens[lane];
            float n = e.ior;
            if (disperse && n > 1) n = dispersion_f(wavelength, e.coefficients);
            s_n[lane] = n;
        }
        Grad G{s_lacc, 0.0f, 0.0f};
        grad_clear(G, L);
        __syncthreads();
        refraction_ratios(s_n, L, lane);
        __syncthreads();
        LensF Lf{lens, a.coatings + first, a.cmin, wavelength};
        Path P;
        P.b1 = a.paths[2 * sg.path];
        P.b2 = a.paths[2 * sg.path + 1];
        P.L = L;
        P.stop = a.lens_table[3 * sg.lens + 2];
        P.E = P.b1 >= 0 ? L + 2 * (P.b1 - P.b2) : L;
        float loss = 0.0f;
        uint32_t lost = 0;  // rays with a target that this trace loses
        for (int k = 0; k < a.rpl; k++) {
            const int ray_id = (blk * a.rpl + k) * ADJ_THREADS + lane;
            const bool active = ray_id < R;
            const long idx = sg.ray_offset + (long)w * R + (active ? ray_id : 0);
            const RayInit ri =
                init_ray_state(lens, g, sg.grid_length, sg.direction, active ? ray_id : 0);
            St s = ri.s;
            St st[K];
            const bool alive =
                forward_ray<K>(Lf, s_n, P, s, st, scratch, ADJ_THREADS, lane, active, false);
            if (a.fwd_out && active) {
                a.fwd_out[idx] = alive ? s.pos.x : NAN;
                a.fwd_out[a.ray_total + idx] = alive ? s.pos.y : NAN;
                a.fwd_out[2 * a.ray_total + idx] = alive ? s.refl : NAN;
            }
            const float rt = active ? a.tr[idx] : NAN;
            if (!isnan(rt) &&
                (!alive || isnan(s.refl) || !isfinite(s.pos.x) || !isfinite(s.pos.y)))
                lost++;
            if (!alive) continue;
            const float xt = a.tx[idx], yt = a.ty[idx];
            if (isnan(xt) || isnan(yt) || isnan(rt)) continue;
            const float dx = s.pos.x - xt, dy = s.pos.y - yt, dr = (s.refl - rt) * a.inv_refl_scale;
            // a ray reaching the sensor plane almost parallel to it lands at
            // an overflowing position: no finite residual, no loss term
            if (!isfinite(dx) || !isfinite(dy) || !isfinite(dr)) continue;
            float gx, gy, gr;
            const float lx = fit_term(dx, a.huber, &gx), ly = fit_term(dy, a.huber, &gy);
            const float lr = fit_term(dr, a.huber, &gr);
            loss += a.loss_scale * (lx + ly + lr);
            reverse_ray<K>(Lf, s_n, G, P, ri, s, st, scratch, ADJ_THREADS, lane, true,
                           mk3(a.loss_scale * gx, a.loss_scale * gy, 0.0f), mk3(0, 0, 0),
                           a.loss_scale * gr * a.inv_refl_scale);
        }
        float* out = a.partial + (long)c * pstride;
        for (int b = 0; b < 4; b++) {
            for (int e = 0; e < a.LF; e++) {
                const float v = e < L ? grad_sum(G, P, b * L + e) : 0.0f;
                if (lane == 0) out[b * a.LF + e] = v;
            }
        }
        const float lv = wave_sum(loss);
        const uint32_t lc = wave_sum_u32(lost);
        if (lane == 0) {
            out[4 * a.LF] = lv;
            out[4 * a.LF + 1] = __uint_as_float(lc);  // a count, stored as its bits
        }
        __syncthreads();  // s_n and s_lacc are rewritten for the next chunk
    }
}

// Per lens instance: out[i][b LF + e] for b < 3 (radius, centre, coating),
// out[i][(3 + w) LF + e] (n_eff at wavelength w), out[i][(3 + W) LF] (loss),
// out[i][(3 + W) LF + 1] (rays with a target that the trace loses: counted in
// 64 bits, written as the nearest float).
// Instance i owns segments seg_off[i] .. seg_off[i + 1] - 1.  Lane j sums a
// fixed subset of chunks (or of segments, for n_eff) in order; lanes combine
// in a fixed tree.
__global__ __launch_bounds__(RED_THREADS) void k_fit_reduce(const float* __restrict__ partial,
                                                            const uint64_t* __restrict__ chunk_offsets,
                                                            const uint64_t* __restrict__ seg_off,
                                                            int LF, int W,
                                                            float* __restrict__ out) {
    __shared__ float s_w[RED_THREADS / 64];
    const int t = blockIdx.x, inst = blockIdx.y;
    const int nout = (3 + W) * LF + 2;
    const int pstride = 4 * LF + 2;
    const uint64_t s0 = seg_off[inst], s1 = seg_off[inst + 1];
    float v = 0.0f;
    if (t == nout - 1) {
        // lost-ray counts: integers, summed exactly
        __shared__ unsigned long long s_n[RED_THREADS / 64];
        const uint64_t c0 = chunk_offsets[s0], c1 = chunk_offsets[s1];
        unsigned long long n = 0;
        for (uint64_t c = c0 + threadIdx.x; c < c1; c += RED_THREADS)
            n += __float_as_uint(partial[c * pstride + 4 * LF + 1]);
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) n += __shfl_xor(n, m, 64);
        if ((threadIdx.x & 63) == 0) s_n[threadIdx.x >> 6] = n;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long r = 0;
            for (int k = 0; k < RED_THREADS / 64; k++) r += s_n[k];
            out[(long)inst * nout + t] = (float)r;
        }
        return;
    }
    if (t < 3 * LF || t == nout - 2) {
        const int col = t == nout - 2 ? 4 * LF : t;
        const uint64_t c0 = chunk_offsets[s0], c1 = chunk_offsets[s1];
        for (uint64_t c = c0 + threadIdx.x; c < c1; c += RED_THREADS) v += partial[c * pstride + col];
    } else {
        const int w = (t - 3 * LF) / LF, e = (t - 3 * LF) % LF;
        for (uint64_t sg = s0 + threadIdx.x; sg < s1; sg += RED_THREADS) {
            const uint64_t c0 = chunk_offsets[sg], per_w = (chunk_offsets[sg + 1] - c0) / W;
            for (uint64_t c = c0 + w * per_w; c < c0 + (w + 1) * per_w; c++)
                v += partial[c * pstride + 3 * LF + e];
        }
    }
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = 0.0f;
        for (int k = 0; k < RED_THREADS / 64; k++) r += s_w[k];
        out[(long)inst * nout + t] = r;
    }
}

__global__ void k_queue_reset_fit(uint64_t* __restrict__ head) { *head = 0; }

typedef void (*FitKernel)(FitArgs);
static FitKernel fit_kernel(int K) {
    switch (K) {
        case 1: return k_trace_fit<1>;
        case 2: return k_trace_fit<2>;
        case 3: return k_trace_fit<3>;
        case 4: return k_trace_fit<4>;
        case 6: return k_trace_fit<6>;
        case 8: return k_trace_fit<8>;
        default: return nullptr;
    }
}

struct FitLayout {
    size_t head, scratch, partial, total;
    int slots;
};

static FitLayout fit_layout(int max_elements, int checkpoint_every, int workgroups,
                            uint64_t chunk_count) {
    FitLayout F;
    F.slots = (3 * max_elements + checkpoint_every - 1) / checkpoint_every;
    size_t o = 0;
    F.head = o;
    o += 256;
    F.scratch = o;
    o += align_up((size_t)workgroups * F.slots * ADJ_FIELDS * ADJ_THREADS * sizeof(float), 256);
    F.partial = o;
    o += align_up((size_t)chunk_count * (4 * max_elements + 2) * sizeof(float), 256);
    F.total = o;
    return F;
}

static int max_events(const int* host_paths, int P, int L) {
    (void)host_paths;
    (void)P;
    return 3 * L;
}

struct AdjLayout {
    size_t scratch, straight, partial, total;
    int slots, bx, groups;
};

static AdjLayout adj_layout(const rf_trace_params* prm, int K) {
    AdjLayout A;
    const int L = prm->lens_count;
    const long R = (long)prm->grid_count * prm->grid_count;
    const long N = (long)prm->path_count * prm->wavelength_count * R;
    const int E = max_events(nullptr, prm->path_count, L);
    A.slots = (E + K - 1) / K + 1;  // segment starts and the final state
    A.bx = (int)((R + ADJ_THREADS - 1) / ADJ_THREADS);
    A.groups = (prm->path_count + ADJ_GHOST_GROUP - 1) / ADJ_GHOST_GROUP;
    const long NS = (long)A.groups * prm->wavelength_count * R;
    size_t o = 0;
    A.scratch = o;
    o += align_up((size_t)A.slots * ADJ_FIELDS * N * sizeof(float), 256);
    A.straight = o;  // straight-chain states S_0 .. S_L per group ray
    o += align_up((size_t)(L + 1) * ADJ_FIELDS * NS * sizeof(float), 256);
    A.partial = o;
    o += align_up((size_t)A.bx * prm->wavelength_count * A.groups * 4 * L * sizeof(float), 256);
    A.total = o;
    return A;
}

}  // namespace rf

using namespace rf;

extern "C" size_t rf_trace_adjoint_workspace_size(const rf_trace_params* prm, int checkpoint_every) {
    if (!prm || !adj_kernel(checkpoint_every) || prm->lens_count < 1 || prm->lens_count > RF_LMAX) return 0;
    return adj_layout(prm, checkpoint_every).total;
}

// Argument checks shared by the three entry points; fills A.
static int adj_check(const char* fn, const rf_lens_element* lens, const int32_t* paths,
                     const int32_t* wavelengths, const rf_trace_params* prm, int K,
                     const void* workspace, size_t workspace_bytes, AdjLayout& A) {
    if (!prm || !lens || !paths || !wavelengths)
        return set_error(RF_ERR_INVALID, "%s: NULL argument", fn);
    if (prm->lens_count < 1 || prm->lens_count > RF_LMAX)
        return set_error(RF_ERR_INVALID, "%s: lens_count %d", fn, prm->lens_count);
    if (!adj_kernel(K))
        return set_error(RF_ERR_INVALID, "%s: checkpoint_every %d not in {1, 2, 3, 4, 6, 8}", fn, K);
    if (prm->path_count < 1 || prm->wavelength_count < 1 || prm->grid_count < 2)
        return set_error(RF_ERR_INVALID, "%s: empty trace", fn);
    A = adj_layout(prm, K);
    if (!workspace || workspace_bytes < A.total)
        return set_error(RF_ERR_WORKSPACE, "%s: workspace %zu < %zu", fn, workspace_bytes, A.total);
    return RF_OK;
}

static int adj_launch(int mode, const rf_lens_element* lens, const float* coating_lambda,
                      const int32_t* paths, const int32_t* wavelengths, const rf_trace_params* prm,
                      int K, const float* gx, const float* gy, const float* gr, const AdjLayout& A,
                      void* workspace, float* grad_out, float* forward_out, rf_stream_t stream) {
    hipStream_t s = (hipStream_t)stream;
    char* w = (char*)workspace;
    dim3 grid(A.bx, prm->wavelength_count, A.groups);
    const size_t lds =
        mode == ADJ_FORWARD ? 0 : (size_t)ADJ_LDS_ROWS * prm->lens_count * ADJ_THREADS * sizeof(float);
    hipLaunchKernelGGL(adj_kernel(K, mode), grid, dim3(ADJ_THREADS), lds, s, lens, coating_lambda,
                       paths, wavelengths, *prm, gx, gy, gr, (float*)(w + A.scratch),
                       (float*)(w + A.straight), (float*)(w + A.partial), forward_out);
    if (mode != ADJ_FORWARD) {
        const int L = prm->lens_count, nout = 3 * L + prm->wavelength_count * L;
        hipLaunchKernelGGL(k_adjoint_reduce, dim3(nout), dim3(RED_THREADS), 0, s,
                           (const float*)(w + A.partial), L, A.bx, prm->wavelength_count,
                           A.groups, grad_out);
    }
    RF_LAUNCH_CHECK("rf_trace_adjoint");
    return RF_OK;
}

extern "C" int rf_trace_adjoint(const rf_lens_element* lens, const float* coating_lambda,
                                const int32_t* paths, const int32_t* wavelengths,
                                const rf_trace_params* prm, int checkpoint_every,
                                const float* grad_pos_x, const float* grad_pos_y,
                                const float* grad_reflectance, void* workspace,
                                size_t workspace_bytes, float* grad_out, float* forward_out,
                                rf_stream_t stream) {
    AdjLayout A;
    int rc = adj_check("rf_trace_adjoint", lens, paths, wavelengths, prm, checkpoint_every,
                       workspace, workspace_bytes, A);
    if (rc) return rc;
    if (!grad_pos_x || !grad_pos_y || !grad_reflectance || !grad_out)
        return set_error(RF_ERR_INVALID, "rf_trace_adjoint: NULL gradient argument");
    return adj_launch(ADJ_FUSED, lens, coating_lambda, paths, wavelengths, prm, checkpoint_every,
                      grad_pos_x, grad_pos_y, grad_reflectance, A, workspace, grad_out, forward_out,
                      stream);
}

extern "C" int rf_trace_forward_checkpointed(const rf_lens_element* lens,
                                             const float* coating_lambda, const int32_t* paths,
                                             const int32_t* wavelengths,
                                             const rf_trace_params* prm, int checkpoint_every,
                                             void* workspace, size_t workspace_bytes,
                                             float* forward_out, rf_stream_t stream) {
    AdjLayout A;
    int rc = adj_check("rf_trace_forward_checkpointed", lens, paths, wavelengths, prm,
                       checkpoint_every, workspace, workspace_bytes, A);
    if (rc) return rc;
    if (!forward_out)
        return set_error(RF_ERR_INVALID, "rf_trace_forward_checkpointed: NULL forward_out");
    return adj_launch(ADJ_FORWARD, lens, coating_lambda, paths, wavelengths, prm, checkpoint_every,
                      nullptr, nullptr, nullptr, A, workspace, nullptr, forward_out, stream);
}

extern "C" int rf_trace_backward(const rf_lens_element* lens, const float* coating_lambda,
                                 const int32_t* paths, const int32_t* wavelengths,
                                 const rf_trace_params* prm, int checkpoint_every,
                                 const float* grad_pos_x, const float* grad_pos_y,
                                 const float* grad_reflectance, void* workspace,
                                 size_t workspace_bytes, float* grad_out, rf_stream_t stream) {
    AdjLayout A;
    int rc = adj_check("rf_trace_backward", lens, paths, wavelengths, prm, checkpoint_every,
                       workspace, workspace_bytes, A);
    if (rc) return rc;
    if (!grad_pos_x || !grad_pos_y || !grad_reflectance || !grad_out)
        return set_error(RF_ERR_INVALID, "rf_trace_backward: NULL gradient argument");
    return adj_launch(ADJ_REVERSE, lens, coating_lambda, paths, wavelengths, prm, checkpoint_every,
                      grad_pos_x, grad_pos_y, grad_reflectance, A, workspace, grad_out, nullptr,
                      stream);
}

extern "C" size_t rf_trace_fit_workspace_size(int32_t max_elements, int32_t checkpoint_every,
                                              int32_t workgroups, uint64_t chunk_count) {
    if (max_elements < 2 || max_elements > RF_LMAX || !fit_kernel(checkpoint_every) ||
        workgroups < 1)
        return 0;
    return fit_layout(max_elements, checkpoint_every, workgroups, chunk_count).total;
}

extern "C" int rf_trace_fit(const rf_lens_element* lenses, const float* coatings,
                            const int32_t* lens_table, const int32_t* paths,
                            const int32_t* wavelengths, int32_t wavelength_count,
                            float coating_min_ior, const rf_batch* batch,
                            const uint64_t* chunk_offsets, uint64_t chunk_count,
                            const uint64_t* lens_segment_offsets, int32_t lens_instances,
                            int32_t max_elements, int32_t rays_per_lane, int32_t checkpoint_every,
                            int32_t workgroups, const float* target_x, const float* target_y,
                            const float* target_reflectance, float reflectance_scale,
                            float loss_scale, float huber_delta, void* workspace,
                            size_t workspace_bytes, float* grad_out, float* forward_out,
                            rf_stream_t stream) {
    if (!lenses || !coatings || !lens_table || !paths || !wavelengths || !batch ||
        !chunk_offsets || !lens_segment_offsets || !target_x || !target_y ||
        !target_reflectance || !grad_out || !batch->segments)
        return set_error(RF_ERR_INVALID, "rf_trace_fit: NULL argument");
    if (batch->segment_count < 1 || lens_instances < 1 || wavelength_count < 1 ||
        rays_per_lane < 1 || workgroups < 1 || workgroups > (1 << 20))
        return set_error(RF_ERR_INVALID, "rf_trace_fit: empty or inconsistent batch");
    if (max_elements < 2 || max_elements > RF_LMAX)
        return set_error(RF_ERR_INVALID, "rf_trace_fit: max_elements %d", max_elements);
    const FitKernel kern = fit_kernel(checkpoint_every);
    if (!kern)
        return set_error(RF_ERR_INVALID, "rf_trace_fit: checkpoint_every %d not in {1, 2, 3, 4, 6, 8}",
                         checkpoint_every);
    if (!(reflectance_scale > 0))
        return set_error(RF_ERR_INVALID, "rf_trace_fit: reflectance_scale must be positive");
    if (!(huber_delta >= 0))
        return set_error(RF_ERR_INVALID, "rf_trace_fit: huber_delta must be >= 0");
    const FitLayout F = fit_layout(max_elements, checkpoint_every, workgroups, chunk_count);
    if (!workspace || workspace_bytes < F.total)
        return set_error(RF_ERR_WORKSPACE, "rf_trace_fit: workspace %zu < %zu", workspace_bytes,
                         F.total);
    hipStream_t s = (hipStream_t)stream;
    char* w = (char*)workspace;
    FitArgs a;
    a.lenses = lenses;
    a.coatings = coatings;
    a.lens_table = lens_table;
    a.paths = paths;
    a.wavelengths = wavelengths;
    a.segs = batch->segments;
    a.chunk_offsets = chunk_offsets;
    a.tx = target_x;
    a.ty = target_y;
    a.tr = target_reflectance;
    a.head = (uint64_t*)(w + F.head);
    a.scratch = (float*)(w + F.scratch);
    a.partial = (float*)(w + F.partial);
    a.fwd_out = forward_out;
    a.ray_total = batch->ray_total;
    a.nseg = batch->segment_count;
    a.W = wavelength_count;
    a.LF = max_elements;
    a.rpl = rays_per_lane;
    a.slots = F.slots;
    a.cmin = coating_min_ior;
    a.inv_refl_scale = 1.0f / reflectance_scale;
    a.loss_scale = loss_scale;
    a.huber = huber_delta;
    hipLaunchKernelGGL(k_queue_reset_fit, dim3(1), dim3(1), 0, s, a.head);
    const size_t lds = (size_t)ADJ_LDS_ROWS * max_elements * ADJ_THREADS * sizeof(float);
    hipLaunchKernelGGL(kern, dim3(workgroups), dim3(ADJ_THREADS), lds, s, a);
    const int nout = (3 + wavelength_count) * max_elements + 2;
    hipLaunchKernelGGL(k_fit_reduce, dim3(nout, lens_instances), dim3(RED_THREADS), 0, s, a.partial,
                       chunk_offsets, lens_segment_offsets, max_elements, wavelength_count,
                       grad_out);
    RF_LAUNCH_CHECK("rf_trace_fit");
    return RF_OK;
}
