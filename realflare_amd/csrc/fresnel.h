// Fresnel reflectance of the ghost tracer's reflection events, shared by the
// forward trace (trace.hip) and its adjoint (adjoint.hip), so both return the
// same bits for the same arguments.
//
// fresnel (raytracing.cl:1-16) and fresnel_ar (raytracing.cl:18-59) take
// the incidence angle theta0 and work with the refraction angles
// theta_k = asin(a_k), a_k = sin(theta0) n0 / n_k, through sines, cosines and
// tangents of their sums and differences.  Here those are written with the
// angle-sum identities over sin theta_k = a_k and cos theta_k =
// sqrt(1 - a_k^2) (theta_k lies in [-pi/2, pi/2], so its cosine is the
// non-negative root; |a_k| > 1, where asin is NaN, gives a NaN root): the
// functions are the same, with three transcendental calls (sin theta0,
// cos theta0, the cosine of the phase) instead of sixteen, and fewer
// roundings.  Operations run in the order written (no contraction).
#pragma once

#include "rf_common.h"

namespace rf {

struct FresnelTerms {
    float s0, c0, a1, c1, A, B, Cc, D, rs, rp, r;
};

// fresnel(theta0, n1, n2): rs, rp and r = (rs^2 + rp^2) / 2.
__device__ __forceinline__ FresnelTerms fresnel_terms(float theta0, float n1, float n2) {
#pragma clang fp contract(off)
    FresnelTerms t;
    t.s0 = sinf(theta0);
    t.c0 = cosf(theta0);
    t.a1 = t.s0 * n1 / n2;           // sin theta1
    t.c1 = sqrtf(1 - t.a1 * t.a1);   // cos theta1
    t.A = n1 * t.c0;
    t.B = n2 * t.c1;
    t.Cc = n1 * t.c1;
    t.D = n2 * t.c0;
    t.rs = (t.A - t.B) / (t.A + t.B);
    t.rp = (t.Cc - t.D) / (t.Cc + t.D);
    t.r = (t.rs * t.rs + t.rp * t.rp) / 2;
    return t;
}

struct FresnelArTerms {
    float s0, c0, a1, a2, c1, c2;
    float sdm, sdp, cdm, cdp, tdm, tdp;  // sin / cos / tan of theta0 -+ theta1
    float sem, sep, cem, cep, tem, tep;  // sin / cos / tan of theta1 -+ theta2
    float rs01, rp01, ts01, tp01, rs12, rp12, ris, rip;
    float dy, t1, dx, delay, k4, ph, cph, r;
};

// fresnel_ar(theta0, lambda, d1, n0, n1, n2): the single-layer coating of
// Ritschel et al. 2009 as the source evaluates it.
__device__ __forceinline__ FresnelArTerms fresnel_ar_terms(float theta0, float lambda, float d1,
                                                           float n0, float n1, float n2) {
#pragma clang fp contract(off)
    const float pi = 3.14159265358979323846f;
    FresnelArTerms t;
    t.s0 = sinf(theta0);
    t.c0 = cosf(theta0);
    t.a1 = t.s0 * n0 / n1;  // sin theta1
    t.a2 = t.s0 * n0 / n2;  // sin theta2
    t.c1 = sqrtf(1 - t.a1 * t.a1);
    t.c2 = sqrtf(1 - t.a2 * t.a2);
    t.sdm = t.s0 * t.c1 - t.c0 * t.a1;
    t.sdp = t.s0 * t.c1 + t.c0 * t.a1;
    t.cdm = t.c0 * t.c1 + t.s0 * t.a1;
    t.cdp = t.c0 * t.c1 - t.s0 * t.a1;
    t.tdm = t.sdm / t.cdm;
    t.tdp = t.sdp / t.cdp;
    t.sem = t.a1 * t.c2 - t.c1 * t.a2;
    t.sep = t.a1 * t.c2 + t.c1 * t.a2;
    t.cem = t.c1 * t.c2 + t.a1 * t.a2;
    t.cep = t.c1 * t.c2 - t.a1 * t.a2;
    t.tem = t.sem / t.cem;
    t.tep = t.sep / t.cep;
    // amplitudes of the outer reflection and transmission, the inner reflection
    t.rs01 = -t.sdm / t.sdp;
    t.rp01 = t.tdm / t.tdp;
    t.ts01 = 2 * t.a1 * t.c0 / t.sdp;
    t.tp01 = t.ts01 * t.cdm;
    t.rs12 = -t.sem / t.sep;
    t.rp12 = t.tem / t.tep;
    t.ris = t.ts01 * t.ts01 * t.rs12;
    t.rip = t.tp01 * t.tp01 * t.rp12;
    // phase difference of the outer and inner reflections
    t.dy = d1 * n1;
    t.t1 = t.a1 / t.c1;  // tan theta1
    t.dx = t.t1 * t.dy;
    t.delay = sqrtf(t.dx * t.dx + t.dy * t.dy);
    t.k4 = 4 * pi / lambda;
    t.ph = t.k4 * (t.delay - t.dx * t.s0);
    t.cph = cosf(t.ph);
    const float out_s2 = t.rs01 * t.rs01 + t.ris * t.ris + 2 * t.rs01 * t.ris * t.cph;
    const float out_p2 = t.rp01 * t.rp01 + t.rip * t.rip + 2 * t.rp01 * t.rip * t.cph;
    t.r = (out_s2 + out_p2) / 2;
    return t;
}

}  // namespace rf
