// Binned quad rasterization of the traced ghost grids on MI355X.
//
// Reference: realflare/api/tasks/cl/rasterizing.cl (prim_shader 233-320,
// vertex_shader 322-386, binner 388-469, rasterizer 471-663) driven by
// RasterizingTask.rasterize (realflare/api/tasks/rasterizing.py:296-421).
//
// Pipeline of one frame (all on one stream, no host synchronization):
//   k_prim     one thread per quad: energy-preserving intensity per
//              wavelength, bounds over wavelengths, culling (as prim_shader).
//   k_vertex   one thread per (ghost, ray, wavelength) vertex (vertex_shader).
//   k_steps    per (wavelength pair, sub-wavelength) step: blend factor and
//              CIE XYZ weight, accumulated in float exactly as the
//              rasterizer's loop counters are.
//   k_setup    one thread per (quad, step): the integer sample-space vertex
//              positions, top-left edge biases, and the rectangle of 8x8
//              pixel tiles the step can cover, cut to the reference's bins
//              (bin_size) that the binner would have listed the quad in.
//   scan+expand+stable radix sort: (tile, quad-step) pairs grouped by tile,
//              each tile's pairs in ascending (quad, step) order, which is the
//              order the reference's per-pixel loop visits them.
//   k_raster   one wave per 8x8 tile: every lane owns one pixel, walks the
//              tile's pairs, tests coverage (Hormann-Tarini, top-left rule,
//              n-rooks samples), shades covered pixels and accumulates in
//              registers; one coalesced write per pixel, y flipped.
// A pixel therefore sums exactly the fragments the reference sums, in the
// same order.
#include <mutex>

#include "primitives.h"

namespace rf {

static constexpr int TILE = 8;

struct Geom {
    int P, W, g, R, Q, S, Wsub, maxw, AA, width, height, bin, ntx, nty, nbx, nby;
    long N;       // P*W*R
    long NV;      // P*R*W vertices
    long NE;      // P*Q*S (quad, step) entries
    float st, area_orig, min_area, intensity, ghost_scale;
    int gw, gh, nspec;
    int bin_exact;
    int accumulate;  // add into the image instead of writing it
    int sb;          // super-block edge in tiles (XCD grouping of k_raster)
    // Batched runs (batch != 0): the run's segments.  Plain frames use the
    // uniform P x g layout with one frame.
    int batch, nseg, frame_base, frames;
    const rf_batch_segment* segs;
    long ray_base, quad_base;  // first ray / quad of the run
    long NQ;                   // quads of this pass (plain: P*Q)
    long NR;                   // vertices of this pass (plain: N)
    float min_area_factor;     // batch: lens.min_area (scaled per segment)
    // Ghost texture lookup, set by make_geom so the rasterizer reads them as
    // scalars: texel coordinate = uv * k + c, folding the (uv / ghost_scale
    // + 1) / 2 of rasterizing.cl and the texel-centre offset; gwf, ghf = the
    // texture size as floats (the upper clamp of the lookup).
    float gku, gcu, gkv, gcv, gwf, ghf;
    int aa_shift;   // log2(AA)
    int bin_shift;  // log2(bin) when bin is a power of two, else -1
};

// Where a ghost's data lives: first ray (planes [W][R], vertices [R][W]),
// first global quad, grid, and the frame relative to the run.
struct Loc {
    long ray0, quad0;
    int g, R, Q, frame;
    float area_orig, min_area;
};

__device__ __forceinline__ Loc loc_segment(const Geom& G, int s) {
    const rf_batch_segment sg = G.segs[s];
    Loc l;
    l.ray0 = sg.ray_offset;
    l.quad0 = sg.quad_offset;
    l.g = sg.grid_count;
    l.R = l.g * l.g;
    l.Q = (l.g - 1) * (l.g - 1);
    l.frame = sg.frame - G.frame_base;
    // rasterizing.py:152-156 and :329, per segment
    const double d = (double)sg.grid_length / (double)(l.g - 1);
    l.area_orig = (float)(d * d);
    l.min_area = (float)((double)G.min_area_factor * (double)l.area_orig);
    return l;
}

__device__ __forceinline__ Loc loc_ghost(const Geom& G, int path) {
    Loc l;
    l.ray0 = (long)path * G.W * G.R;
    l.quad0 = (long)path * G.Q;
    l.g = G.g;
    l.R = G.R;
    l.Q = G.Q;
    l.frame = 0;
    l.area_orig = G.area_orig;
    l.min_area = G.min_area;
    return l;
}

// segment holding global quad gq (largest s with quad_offset <= gq)
__device__ __forceinline__ int seg_by_quad(const Geom& G, long gq) {
    int lo = 0, hi = G.nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (G.segs[mid].quad_offset <= gq) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int seg_by_ray(const Geom& G, long gr) {
    int lo = 0, hi = G.nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (G.segs[mid].ray_offset <= gr) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ Loc loc_quad(const Geom& G, long gq) {
    return G.batch ? loc_segment(G, seg_by_quad(G, gq)) : loc_ghost(G, (int)(gq / G.Q));
}

struct Header {
    uint64_t required;  // (tile, quad-step) pairs the frame needs
    uint64_t n_sort;    // pairs binned (0 on overflow)
    uint32_t overflow;     // required > capacity: nothing was rastered
    uint32_t heavy;        // quad steps left to k_expand_heavy
    uint32_t setup_heavy;  // quad steps whose pairs k_count_heavy counts
    uint32_t pad_;
    uint64_t blocks;       // super-blocks of the run (frames x per frame), k_raster's order
};

// Setup.flags bit 5: every vertex coordinate lies within +-SMALL_COORD
// samples, so the coverage test may use 24-bit multiplies (intersect_quad_small).
static constexpr int FLAG_SMALL = 32;
static constexpr int SMALL_COORD = 1 << 21;

// Quad steps above HEAVY_ROWS tile rows or HEAVY_PAIRS pairs are counted and
// expanded by whole workgroups (k_count_heavy, k_expand_heavy): a few quad
// steps span hundreds of rows or thousands of tiles, and one lane walking
// them would hold its wave for as many iterations.
static constexpr uint32_t HEAVY_ROWS = 16;
static constexpr uint32_t HEAVY_PAIRS = 1024;
static constexpr int HEAVY_BLOCKS = 4096;  // persistent grids of the two kernels

struct alignas(16) Setup {
    int vx[4], vy[4];
    int flags;  // bits 0..4: top-left of edges 01,12,23,30,20; bit 5: FLAG_SMALL
    uint32_t r0, r1;  // tile rect (tx0 | ty0 << 16), (tx1 | ty1 << 16)
    uint32_t prim;
};

// Edge rows of one (quad, step) for the coverage test of rasterizing.cl:55-93:
// for the edges 01, 12, 23, 30, 20 (vertex i to j) the edge function at
// sample (px, py) is (x_i - x_j)(py - y_i) - (y_i - y_j)(px - x_i) =
// a py + nb px + (-a y_i - nb x_i), with a = x_i - x_j, nb = y_j - y_i, and the
// top-left bias (0, or -1 off a top-left edge) added into c.  All mod 2^32.
struct alignas(16) Edges {
    int a[5], nb[5], c[5];
    int flags;  // as Setup.flags
};
static_assert(sizeof(Edges) == 64, "one edge-row record per 64-byte line");

// Shading inputs of one (quad, step): the four corners' attributes blended
// between the step's two wavelengths (mix_vertex, rasterizing.cl:509-522)
// and the step's XYZ weight.  Evaluated once per quad step instead of once
// per (tile, quad step).
struct alignas(16) Attr {
    float u[4], v[4], rrel[4], refl[4], inten[4];
    float x, y, z, pad_;
};

// sub_offsets of rasterizing.cl (y offset of AA sample s at index AA - 1 + s:
// 0 | 1, 0 | 1, 2, 0, 3 | 4, 1, 6, 2, 5, 0, 3, 7), 3 bits per entry, kept in a
// register-resident immediate.
static constexpr uint64_t SUB_OFFSETS = 0x1d8ab18c2208ull;
__host__ __device__ __forceinline__ int sub_offset(int i) { return (int)((SUB_OFFSETS >> (3 * i)) & 7); }

__device__ __forceinline__ long floordiv(long a, long b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }
__device__ __forceinline__ long ceildiv(long a, long b) { return -floordiv(-a, b); }
// floor(a / 2^s) for either sign (an arithmetic shift)
__device__ __forceinline__ long floorshr(long a, int s) { return a >> s; }
static constexpr int TILE_SHIFT = 3;
static_assert(TILE == 1 << TILE_SHIFT, "tile edge is a power of two");

__device__ __forceinline__ void quad_vertexes(int g, int quad, int q[4]) {
    int row = quad / (g - 1), col = quad % (g - 1);
    q[0] = row * g + col;
    q[1] = q[0] + 1;
    q[2] = q[1] + g;
    q[3] = q[0] + g;
}

__device__ __forceinline__ float edge_f(float ax, float ay, float bx, float by, float cx, float cy) {
#pragma clang fp contract(off)
    return (ax - bx) * (cy - ay) - (ay - by) * (cx - ax);
}

__device__ __forceinline__ int edge_i(int ax, int ay, int bx, int by, int cx, int cy) {
    uint32_t t0 = ((uint32_t)ax - (uint32_t)bx) * ((uint32_t)cy - (uint32_t)ay);
    uint32_t t1 = ((uint32_t)ay - (uint32_t)by) * ((uint32_t)cx - (uint32_t)ax);
    return (int)(t0 - t1);
}

__device__ __forceinline__ bool is_top_left(int ax, int ay, int bx, int by) {
    int ex = wsub(bx, ax), ey = wsub(by, ay);
    return (ey == 0 && ex > 0) || ey > 0;
}

// ---------------------------------------------------------------------------
// prim_shader (rasterizing.cl:233-320).  Intensities the source leaves
// unwritten (wavelengths after the first culled one) are written as 0, the
// value a freshly cleared buffer holds.
__global__ __launch_bounds__(256) void k_prim(const float* __restrict__ planes, Geom G,
                                              float4* __restrict__ bounds,
                                              float* __restrict__ intens) {
#pragma clang fp contract(off)
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= G.NQ) return;
    const long idx = G.quad_base + t;  // global quad
    const Loc l = loc_quad(G, idx);
    const int quad = (int)(idx - l.quad0);
    int qv[4];
    quad_vertexes(l.g, quad, qv);
    const float* px_ = planes;
    const float* py_ = planes + G.N;
    const float* rr_ = planes + 4 * G.N;
    const float* rf_ = planes + 5 * G.N;
    float b0 = INFINITY, b1 = INFINITY, b2 = -INFINITY, b3 = -INFINITY;
    int invalid_rrel = 0;
    for (int w = 0; w < G.W; w++) {
        float* out_i = intens + idx * G.W + w;
        if (isnan(b0)) {
            *out_i = 0.0f;
            continue;
        }
        const long ro = l.ray0 + (long)w * l.R;
        float rx[4], ry[4], rrel[4], refl[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            rx[i] = px_[ro + qv[i]];
            ry[i] = py_[ro + qv[i]];
            rrel[i] = rr_[ro + qv[i]];
            refl[i] = rf_[ro + qv[i]];
        }
        float cx[4] = {0, 0, 0, 0}, cy[4] = {0, 0, 0, 0};
        int valid = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (!isnan(refl[i])) {
                // compaction into the first `valid` slots
                if (valid == 0) { cx[0] = rx[i]; cy[0] = ry[i]; }
                else if (valid == 1) { cx[1] = rx[i]; cy[1] = ry[i]; }
                else if (valid == 2) { cx[2] = rx[i]; cy[2] = ry[i]; }
                else { cx[3] = rx[i]; cy[3] = ry[i]; }
                valid++;
            }
        }
        float area;
        if (valid == 4) {
            float a0 = edge_f(cx[0], cy[0], cx[1], cy[1], cx[2], cy[2]);
            float a1 = edge_f(cx[0], cy[0], cx[2], cy[2], cx[3], cy[3]);
            area = fabsf(a0 + a1) / 2;
        } else if (valid == 3) {
            area = fabsf(edge_f(cx[0], cy[0], cx[1], cy[1], cx[2], cy[2]));
            cx[3] = cx[2];
            cy[3] = cy[2];
            rrel[3] = rrel[2];
        } else {
            b0 = NAN;
            *out_i = 0.0f;
            continue;
        }
        if (rrel[0] > 1 && rrel[1] > 1 && rrel[2] > 1 && rrel[3] > 1) invalid_rrel++;
        float area_actual = fmaxf(area, l.min_area);
        *out_i = area_actual > 0 ? l.area_orig / area_actual : 0.0f;
        b0 = fminf(b0, fminf(fminf(cx[0], cx[1]), fminf(cx[2], cx[3])));
        b1 = fminf(b1, fminf(fminf(cy[0], cy[1]), fminf(cy[2], cy[3])));
        b2 = fmaxf(b2, fmaxf(fmaxf(cx[0], cx[1]), fmaxf(cx[2], cx[3])));
        b3 = fmaxf(b3, fmaxf(fmaxf(cy[0], cy[1]), fmaxf(cy[2], cy[3])));
    }
    if (invalid_rrel == G.W) b0 = NAN;
    bounds[idx] = make_float4(b0, b1, b2, b3);
}

// vertex_shader (rasterizing.cl:322-386); Vertex = 32 bytes
// (pos.xy, uv.xy, rrel, reflectance, intensity, pad).
__global__ __launch_bounds__(256) void k_vertex(const float* __restrict__ planes,
                                                const float* __restrict__ intens, Geom G,
                                                float4* __restrict__ vertexes) {
#pragma clang fp contract(off)
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= G.NR) return;
    const long idx = G.ray_base + t;  // global vertex
    const Loc l = G.batch ? loc_segment(G, seg_by_ray(G, idx))
                          : loc_ghost(G, (int)(idx / ((long)G.W * G.R)));
    const long local = idx - l.ray0;
    const int w = (int)(local % G.W);
    const int ray = (int)(local / G.W);
    const long ri = l.ray0 + (long)w * l.R + ray;
    const float refl = planes[5 * G.N + ri];
    float4 a = make_float4(0, 0, 0, 0), b = make_float4(0, refl, 0, 0);
    if (!isnan(refl)) {
        const int x = ray % l.g, y = ray / l.g;
        int nb[4];
        nb[0] = (y - 1) * (l.g - 1) + (x - 1);
        nb[1] = nb[0] + 1;
        nb[2] = y * (l.g - 1) + (x - 1);
        nb[3] = nb[2] + 1;
        float it = 0.0f;
        int count = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (nb[i] < 0 || nb[i] >= l.Q) continue;
            float v = intens[(l.quad0 + nb[i]) * G.W + w];
            if (v > 0) {
                it += v;
                ++count;
            }
        }
        it /= (float)(count > 1 ? count : 1);
        a.x = planes[ri] * G.st + (float)G.width / 2;
        a.y = planes[G.N + ri] * G.st + (float)G.height / 2;
        a.z = planes[2 * G.N + ri];
        a.w = planes[3 * G.N + ri];
        b.x = planes[4 * G.N + ri];
        b.z = it;
    }
    vertexes[2 * idx] = a;
    vertexes[2 * idx + 1] = b;
}

// Step tables: blend factor and XYZ weight per (wavelength pair, sub step),
// accumulated in float like the rasterizer's loop variables
// (rasterizing.cl:556-557, 616-617, 644, 650-651).
__global__ void k_steps(const float* __restrict__ spectrum, Geom G, float4* __restrict__ table,
                        Header* __restrict__ h) {
#pragma clang fp contract(off)
    const int i = threadIdx.x + blockIdx.x * blockDim.x;
    if (i == 0) h->setup_heavy = 0;
    if (i >= G.maxw) return;
    const float sub_step = 1.0f / (float)G.Wsub;
    const float w_step = sub_step / (float)G.W;
    float sub_pos = 0.0f;
    float w_pos = ((float)i + 0.5f) / (float)G.W;
    for (int j = 0; j < G.Wsub; j++) {
        float x = 1.0f, y = 1.0f, z = 1.0f;
        if (G.W > 1) {
            x = sample_bilinear(spectrum + 0, G.nspec, 1, 4, w_pos, 0.0f, true);
            y = sample_bilinear(spectrum + 1, G.nspec, 1, 4, w_pos, 0.0f, true);
            z = sample_bilinear(spectrum + 2, G.nspec, 1, 4, w_pos, 0.0f, true);
        }
        table[i * G.Wsub + j] = make_float4(sub_pos, x, y, z);
        sub_pos += sub_step;
        w_pos += w_step;
    }
}

__device__ __forceinline__ void load_pos(const float4* __restrict__ vtx, long vi, long NV, float& x,
                                         float& y) {
    if (vi < NV) {
        float4 a = vtx[2 * vi];
        x = a.x;
        y = a.y;
    } else {
        x = 0.0f;
        y = 0.0f;
    }
}

// Convex hull of the four integer vertices (sample space), prepared for
// row queries.  Coverage of the Hormann-Tarini test lies inside triangles
// (0,1,2) and (2,3,0), hence inside this hull, so tiles it cannot reach need
// no binning.  A row's x-extent comes from the vertices inside the row's
// sample band and the crossings of all six vertex pairs with the band's two
// edges (double precision, widened by one sample on each side).
struct HullRows {
    double x[4], y[4];
    double slope[6];  // pairs (0,1) (0,2) (0,3) (1,2) (1,3) (2,3)
    bool valid[6];    // the pair's vertices differ in y
};

// vertex pairs, indexed by compile-time constants only so the hull stays in
// registers
__host__ __device__ constexpr int hull_pa(int s) { return s < 3 ? 0 : (s < 5 ? 1 : 2); }
__host__ __device__ constexpr int hull_pb(int s) { return s < 3 ? s + 1 : (s < 5 ? s - 1 : 3); }

// n / d to within an ulp or two: the hardware reciprocal, one Newton step and
// a residual correction (the spans are widened by a sample, so the quotient
// need not be the correctly rounded one; every kernel that counts or lists a
// quad step's tiles uses this same function, so they agree).
__device__ __forceinline__ double hull_quot(double n, double d) {
    double r = __builtin_amdgcn_rcp(d);
    r = fma(fma(-d, r, 1.0), r, r);
    const double q = n * r;
    return fma(fma(-d, q, n), r, q);
}

__device__ __forceinline__ void hull_prepare(const int* vx, const int* vy, HullRows* h) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
        h->x[i] = vx[i];
        h->y[i] = vy[i];
    }
#pragma unroll
    for (int s = 0; s < 6; s++) {
        const int i = hull_pa(s), j = hull_pb(s);
        h->valid[s] = vy[i] != vy[j];
        h->slope[s] = h->valid[s] ? hull_quot((double)vx[j] - vx[i], (double)vy[j] - vy[i]) : 0.0;
    }
}

// Tiles [*a, *b] of tile row `ty` the hull can reach (empty when *a > *b);
// aa_shift = log2(AA).
__device__ __forceinline__ void hull_row_span(const HullRows& h, int AA, int aa_shift, int ty, long* a,
                                              long* b) {
    const double Y0 = (double)ty * TILE * AA, Y1 = ((double)ty + 1) * TILE * AA - 1;
    double xmin = INFINITY, xmax = -INFINITY;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        if (h.y[i] >= Y0 && h.y[i] <= Y1) {
            xmin = fmin(xmin, h.x[i]);
            xmax = fmax(xmax, h.x[i]);
        }
    }
#pragma unroll
    for (int s = 0; s < 6; s++) {
        if (!h.valid[s]) continue;
        const double xa = h.x[hull_pa(s)], ya = h.y[hull_pa(s)], yb = h.y[hull_pb(s)];
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const double Y = k ? Y1 : Y0;
            if ((ya - Y) * (yb - Y) <= 0) {
                const double x = xa + (Y - ya) * h.slope[s];
                xmin = fmin(xmin, x);
                xmax = fmax(xmax, x);
            }
        }
    }
    if (!(xmin <= xmax)) {
        *a = 1;
        *b = 0;
        return;
    }
    const double lim = 4.0e18;
    const long sl = (long)fmax(fmin(floor(xmin) - 1, lim), -lim);
    const long sh = (long)fmax(fmin(ceil(xmax) + 1, lim), -lim);
    *a = floorshr(floorshr(sl, aa_shift), TILE_SHIFT);
    *b = floorshr(floorshr(sh, aa_shift), TILE_SHIFT);
}

// Quad steps whose every fragment is exactly 0 bin no pair; the image is
// unchanged, since adding an exact 0 leaves a pixel's sum as it was.  A
// fragment is ghost * smoothstep(rrel) * max(intensity, 0) * max(coating, 0)
// (rasterizing.cl:471-506) with each attribute a weighted sum of the four
// corners'.  Two cases are recognized:
//
// (1) A corner whose ray died (NaN reflectance; the reference leaves its
// position and attributes undefined, rasterizing.cl:347-351).  The weighted
// reflectance is then NaN at every pixel, whatever the weights, and its
// maximum with 0 is 0, so the fragment is 0 as long as the other factors are
// finite: corner attributes bounded (below) and finite weights.  The weights
// of barycentric_fast are finite unless their sum vanishes to the last bit at
// a sample, a rounding coincidence at which the fragment is NaN and poisons
// its pixel, here and in the reference alike, at positions that rounding
// (and, in the reference, the undefined corner) decides; such a pixel is not
// reproduced by any evaluation other than the reference's own.
//
// (2) The ghost texture lookup reads only its zero border (rays that pass the
// aperture stop outside the texture's extent: rasterizing.cl:484-485 samples
// the ghost with border colour 0).
//
// Why the fragments of (2) are exactly 0.  With one sample per pixel the shading
// point is the covered sample itself, so it lies in the quad (interior or
// boundary).  For a strictly convex quad the mean value coordinates of such a
// point are non-negative and sum to 1, so an interpolated texel coordinate
// lies between the corners' minimum and maximum; with integer vertices less
// than CULL_EXTENT samples apart the edge terms (cross and dot products of
// integer offsets) are exact in float32, and the rounding of the reciprocal
// square roots and reciprocals leaves a negative weight mass far below 1e-2
// (checked over random and sliver-shaped quads against a float32 emulation
// of barycentric_fast with one-ulp errors: at most 5e-4).  A margin of one
// texel plus the corners' spread therefore keeps every interpolated
// coordinate beyond the border: at or past size (texels size, size + 1 of
// the framed copy) or below -2 (texels -2, -1).  The lookup then blends four
// zero texels with finite weights, which is +0 when the other coordinate is
// finite; the remaining factors are finite (smoothstep of anything, maxima
// with 0 of sums of values bounded by 1e30 or NaN), so the fragment is +0.
static constexpr long CULL_EXTENT = 4096;
static constexpr float CULL_BOUND = 1.0e6f;  // texel coordinates taken as finite

__device__ __forceinline__ bool strictly_convex(const int* vx, const int* vy) {
    int sign = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int j = (i + 1) & 3, k = (i + 2) & 3;
        const long c = ((long)vx[j] - vx[i]) * ((long)vy[k] - vy[j]) -
                       ((long)vy[j] - vy[i]) * ((long)vx[k] - vx[j]);
        const int s = c > 0 ? 1 : (c < 0 ? -1 : 0);
        if (s == 0 || (sign != 0 && s != sign)) return false;
        sign = s;
    }
    return true;
}

// Every interpolation of the four texel coordinates t (with the margin above)
// lies past the texture's edge [0, size) into its zero border.
__device__ __forceinline__ bool past_border(const float* t, float size) {
    const float lo = fminf(fminf(t[0], t[1]), fminf(t[2], t[3]));
    const float hi = fmaxf(fmaxf(t[0], t[1]), fmaxf(t[2], t[3]));
    const float spread = hi - lo;
    return lo >= size + 1.0f + spread || hi <= -3.0f - spread;
}

__device__ __forceinline__ bool bounded(const float* t) {
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 4; j++) ok = ok && fabsf(t[j]) <= CULL_BOUND;  // false for NaN
    return ok;
}

// Per (quad, step): integer vertices, edge biases, tile rectangle, pair
// count, and the blended shading attributes.
__global__ __launch_bounds__(256) void k_setup(const float4* __restrict__ vtx,
                                               const float4* __restrict__ bounds,
                                               const float4* __restrict__ table, Geom G,
                                               Setup* __restrict__ setup, Attr* __restrict__ attr,
                                               Edges* __restrict__ edges,
                                               uint32_t* __restrict__ counts,
                                               uint32_t* __restrict__ heavy,
                                               Header* __restrict__ h) {
#pragma clang fp contract(off)
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= G.NE) return;
    const long prim = G.quad_base + e / G.S;  // global quad
    const int step = (int)(e % G.S);
    const int wl = step / G.Wsub;
    uint32_t count = 0;
    const float4 b = bounds[prim];
    Setup rec;
    rec.flags = 0;
    rec.r0 = rec.r1 = 0;
    rec.prim = (uint32_t)prim;
    for (int j = 0; j < 4; j++) rec.vx[j] = rec.vy[j] = 0;
    if (!isnan(b.x)) {
        const Loc l = loc_quad(G, prim);
        const int quad = (int)(prim - l.quad0);
        int qv[4];
        quad_vertexes(l.g, quad, qv);
        const float4 tab = table[step];
        const float blend = tab.x;
        int vx[4], vy[4];
        long vbase[4];
        float tu[4], tv[4];  // ghost texel coordinates of the corners
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const long base = l.ray0 + (long)qv[j] * G.W + wl;
            vbase[j] = base;
            const float4 a0 = vtx[2 * base];
            const float4 b0 = base + 1 < G.NV ? vtx[2 * (base + 1)] : make_float4(0, 0, 0, 0);
            const float mx = a0.x + (b0.x - a0.x) * blend;
            const float my = a0.y + (b0.y - a0.y) * blend;
            vx[j] = rf_to_int(mx * (float)G.AA);
            vy[j] = rf_to_int(my * (float)G.AA);
            rec.vx[j] = vx[j];
            rec.vy[j] = vy[j];
            // the corner's u, v exactly as the shading attributes blend them
            tu[j] = (a0.z + (b0.z - a0.z) * blend) * G.gku + G.gcu;
            tv[j] = (a0.w + (b0.w - a0.w) * blend) * G.gkv + G.gcv;
        }
        const bool collinear = edge_i(vx[0], vy[0], vx[1], vy[1], vx[2], vy[2]) == 0 &&
                               edge_i(vx[0], vy[0], vx[2], vy[2], vx[3], vy[3]) == 0;
        if (!collinear) {
            int flags = 0;
            flags |= is_top_left(vx[0], vy[0], vx[1], vy[1]) ? 1 : 0;
            flags |= is_top_left(vx[1], vy[1], vx[2], vy[2]) ? 2 : 0;
            flags |= is_top_left(vx[2], vy[2], vx[3], vy[3]) ? 4 : 0;
            flags |= is_top_left(vx[3], vy[3], vx[0], vy[0]) ? 8 : 0;
            flags |= is_top_left(vx[2], vy[2], vx[0], vy[0]) ? 16 : 0;
            bool small = true;
#pragma unroll
            for (int j = 0; j < 4; j++)
                small = small && abs(vx[j]) <= SMALL_COORD && abs(vy[j]) <= SMALL_COORD;
            flags |= small ? FLAG_SMALL : 0;
            rec.flags = flags;
            // covered samples lie inside the integer vertices' bounding box
            long sx0 = min(min(vx[0], vx[1]), min(vx[2], vx[3]));
            long sx1 = max(max(vx[0], vx[1]), max(vx[2], vx[3]));
            long sy0 = min(min(vy[0], vy[1]), min(vy[2], vy[3]));
            long sy1 = max(max(vy[0], vy[1]), max(vy[2], vy[3]));
            long x0 = floorshr(sx0, G.aa_shift), x1 = floorshr(sx1, G.aa_shift);
            long y0 = floorshr(sy0, G.aa_shift), y1 = floorshr(sy1, G.aa_shift);
            // every fragment is exactly 0: no pairs
            bool zero = false;
            if (bounded(tu) && bounded(tv)) {
                bool finite = true, dead = false;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const long base = vbase[j];
                    const float4 a1 = vtx[2 * base + 1];
                    const float4 b1 = base + 1 < G.NV ? vtx[2 * (base + 1) + 1] : make_float4(0, 0, 0, 0);
                    const float rr = a1.x + (b1.x - a1.x) * blend;
                    const float rf = a1.y + (b1.y - a1.y) * blend;
                    const float in = a1.z + (b1.z - a1.z) * blend;
                    // NaN passes (the maxima with 0 drop it), infinities do not
                    finite = finite && !(fabsf(rr) > 1e30f) && !(fabsf(rf) > 1e30f) && !(fabsf(in) > 1e30f);
                    dead = dead || isnan(rf);
                }
                zero = finite && (dead || (G.AA == 1 && sx1 - sx0 < CULL_EXTENT &&
                                           sy1 - sy0 < CULL_EXTENT &&
                                           (past_border(tu, G.gwf) || past_border(tv, G.ghf)) &&
                                           strictly_convex(vx, vy)));
            }
            if (zero) x1 = x0 - 1;  // an empty rectangle
            // bins the binner lists this quad in (rasterizing.cl:449-453)
            const int sbx0 = wadd(rf_to_int(b.x * G.st), G.width / 2);
            const int sby0 = wadd(rf_to_int(b.y * G.st), G.height / 2);
            const int sbx1 = wadd(rf_to_int(b.z * G.st), G.width / 2);
            const int sby1 = wadd(rf_to_int(b.w * G.st), G.height / 2);
            long bx0, bx1, by0, by1;
            if (G.bin_shift >= 0) {  // ceil(a / bin) = -floor(-a / bin)
                bx0 = -floorshr(-((long)sbx0 - G.bin), G.bin_shift);
                bx1 = floorshr(sbx1, G.bin_shift);
                by0 = -floorshr(-((long)sby0 - G.bin), G.bin_shift);
                by1 = floorshr(sby1, G.bin_shift);
            } else {
                bx0 = ceildiv((long)sbx0 - G.bin, G.bin);
                bx1 = floordiv(sbx1, G.bin);
                by0 = ceildiv((long)sby0 - G.bin, G.bin);
                by1 = floordiv(sby1, G.bin);
            }
            bx0 = max(bx0, 0L);
            by0 = max(by0, 0L);
            bx1 = min(bx1, (long)G.nbx - 1);
            by1 = min(by1, (long)G.nby - 1);
            x0 = max(x0, max(bx0 * G.bin, 0L));
            y0 = max(y0, max(by0 * G.bin, 0L));
            x1 = min(x1, min((bx1 + 1) * G.bin - 1, (long)G.width - 1));
            y1 = min(y1, min((by1 + 1) * G.bin - 1, (long)G.height - 1));
            if (x0 <= x1 && y0 <= y1) {
                const uint32_t tx0 = (uint32_t)(x0 / TILE), tx1 = (uint32_t)(x1 / TILE);
                const uint32_t ty0 = (uint32_t)(y0 / TILE), ty1 = (uint32_t)(y1 / TILE);
                rec.r0 = tx0 | (ty0 << 16);
                rec.r1 = tx1 | (ty1 << 16);
                const bool tall = ty1 - ty0 + 1 > HEAVY_ROWS;
                if (tall) {
                    heavy[atomicAdd(&h->setup_heavy, 1u)] = (uint32_t)e;  // counted by k_count_heavy
                } else {
                    HullRows hull;
                    hull_prepare(vx, vy, &hull);
                    for (uint32_t ty = ty0; ty <= ty1; ty++) {
                        long a, bb;
                        hull_row_span(hull, G.AA, G.aa_shift, (int)ty, &a, &bb);
                        a = max(a, (long)tx0);
                        bb = min(bb, (long)tx1);
                        if (a <= bb) count += (uint32_t)(bb - a + 1);
                    }
                }
                if (count || tall) {
                    // edge rows 01, 12, 23, 30, 20 for the rasterizer
                    Edges er;
#pragma unroll
                    for (int k = 0; k < 5; k++) {
                        const int i = k < 4 ? k : 2, j = k < 3 ? k + 1 : 0;
                        const uint32_t ea = (uint32_t)vx[i] - (uint32_t)vx[j];
                        const uint32_t enb = (uint32_t)vy[j] - (uint32_t)vy[i];
                        const uint32_t bias = ((flags >> k) & 1) ? 0u : 0xffffffffu;
                        er.a[k] = (int)ea;
                        er.nb[k] = (int)enb;
                        er.c[k] = (int)(bias - ea * (uint32_t)vy[i] - enb * (uint32_t)vx[i]);
                    }
                    er.flags = flags;
                    edges[e] = er;
                    // the records of quad steps that bin no pair are never read;
                    // the shading inputs are blended here, from vertices still
                    // in cache, so they are not held through the counting above
                    setup[e] = rec;
                    Attr at;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const long base = vbase[j];
                        const float4 a0 = vtx[2 * base], a1 = vtx[2 * base + 1];
                        float4 b0 = make_float4(0, 0, 0, 0), b1 = make_float4(0, 0, 0, 0);
                        if (base + 1 < G.NV) {
                            b0 = vtx[2 * (base + 1)];
                            b1 = vtx[2 * (base + 1) + 1];
                        }
                        at.u[j] = a0.z + (b0.z - a0.z) * blend;
                        at.v[j] = a0.w + (b0.w - a0.w) * blend;
                        at.rrel[j] = a1.x + (b1.x - a1.x) * blend;
                        at.refl[j] = a1.y + (b1.y - a1.y) * blend;
                        at.inten[j] = a1.z + (b1.z - a1.z) * blend;
                    }
                    at.x = tab.y;
                    at.y = tab.z;
                    at.z = tab.w;
                    at.pad_ = 0.0f;
                    attr[e] = at;
                }
            }
        }
    }
    counts[e] = count;
}

__global__ void k_finalize(Header* __restrict__ h, uint64_t capacity) {
    const uint64_t need = h->required;
    h->overflow = need > capacity ? 1u : 0u;
    h->n_sort = need > capacity ? 0 : need;
    h->heavy = 0;
}


// One thread per (quad, step): walks the tile rows of its rectangle with the
// hull span k_setup counted with.  A block's 256 entries own one contiguous
// run of the pair arrays (the offsets are an exclusive scan), so the block
// stages its pairs in LDS and stores the run coalesced; a run larger than the
// staging area is written directly.
static constexpr int EXPAND_STAGE = 4096;

__device__ __forceinline__ void expand_entry(const Setup& rec, const Geom& G, uint32_t e,
                                             uint64_t off, uint32_t* __restrict__ keys,
                                             uint32_t* __restrict__ vals) {
    const uint32_t tx0 = rec.r0 & 0xffff, ty0 = rec.r0 >> 16;
    const uint32_t tx1 = rec.r1 & 0xffff, ty1 = rec.r1 >> 16;
    HullRows hull;
    hull_prepare(rec.vx, rec.vy, &hull);
    const uint32_t frame_key =
        G.batch ? (uint32_t)loc_quad(G, rec.prim).frame * (uint32_t)(G.ntx * G.nty) : 0u;
    for (uint32_t ty = ty0; ty <= ty1; ty++) {
        long a, b;
        hull_row_span(hull, G.AA, G.aa_shift, (int)ty, &a, &b);
        a = max(a, (long)tx0);
        b = min(b, (long)tx1);
        for (long tx = a; tx <= b; tx++) {
            keys[off] = frame_key + ty * (uint32_t)G.ntx + (uint32_t)tx;
            vals[off] = e;
            off++;
        }
    }
}

__global__ __launch_bounds__(256) void k_expand(const Setup* __restrict__ setup,
                                                const uint32_t* __restrict__ counts,
                                                const uint64_t* __restrict__ offsets,
                                                Header* __restrict__ h, Geom G,
                                                uint32_t* __restrict__ keys,
                                                uint32_t* __restrict__ vals,
                                                uint32_t* __restrict__ heavy) {
    __shared__ uint32_t s_keys[EXPAND_STAGE], s_vals[EXPAND_STAGE];
    if (h->overflow) return;
    const long first = (long)blockIdx.x * 256;
    const long last = min(first + 256, G.NE) - 1;  // block's last entry
    const long e = first + threadIdx.x;
    const uint64_t begin = offsets[first];
    const uint64_t end = offsets[last] + counts[last];
    const bool staged = end - begin <= EXPAND_STAGE;  // block-uniform
    const uint32_t cnt = e <= last ? counts[e] : 0u;
    if (cnt) {
        const Setup rec = setup[e];
        if (cnt > HEAVY_PAIRS || (rec.r1 >> 16) - (rec.r0 >> 16) + 1 > HEAVY_ROWS) {
            // its run is written by k_expand_heavy, after the staged copy below
            heavy[atomicAdd(&h->heavy, 1u)] = (uint32_t)e;
        } else {
            uint32_t* kp = staged ? s_keys : keys;  // generic pointers: LDS or global
            uint32_t* vp = staged ? s_vals : vals;
            expand_entry(rec, G, (uint32_t)e, offsets[e] - (staged ? begin : 0), kp, vp);
        }
    }
    if (!staged) return;
    __syncthreads();
    const uint32_t n = (uint32_t)(end - begin);
    for (uint32_t i = threadIdx.x; i < n; i += 256) {
        keys[begin + i] = s_keys[i];
        vals[begin + i] = s_vals[i];
    }
}

// Both kernels below give each listed quad step to one wave (four per
// workgroup, independent of each other, so they synchronize at wave level
// only): 64 rows at a time, one row per lane.
static constexpr int HW_WAVES = 4;

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Tile-row span of row ty0 + c + lane of a quad step (an empty row or a row
// past the last has length 0) and the exclusive prefix of the lengths over
// the wave, in s_a / s_len / s_pre; returns the 64 rows' pair count.
__device__ __forceinline__ uint32_t heavy_rows(const Setup& rec, const Geom& G, uint32_t c, int lane,
                                               uint32_t* s_a, uint32_t* s_len, uint32_t* s_pre) {
    const uint32_t tx0 = rec.r0 & 0xffff, ty0 = rec.r0 >> 16;
    const uint32_t tx1 = rec.r1 & 0xffff, ty1 = rec.r1 >> 16;
    const uint32_t ty = ty0 + c + (uint32_t)lane;
    uint32_t len = 0, a32 = 0;
    if (ty <= ty1) {
        HullRows hull;
        hull_prepare(rec.vx, rec.vy, &hull);
        long a, b;
        hull_row_span(hull, G.AA, G.aa_shift, (int)ty, &a, &b);
        a = max(a, (long)tx0);
        b = min(b, (long)tx1);
        if (a <= b) {
            len = (uint32_t)(b - a + 1);
            a32 = (uint32_t)a;
        }
    }
    uint32_t inc = len;  // inclusive scan over the wave
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t v = (uint32_t)__shfl_up((int)inc, off, 64);
        if (lane >= off) inc += v;
    }
    const uint32_t total = (uint32_t)__shfl((int)inc, 63, 64);
    wave_sync();  // the previous chunk's tables have been read
    s_a[lane] = a32;
    s_len[lane] = len;
    s_pre[lane] = inc - len;
    wave_sync();
    return total;
}

// Pair counts of the quad steps k_setup listed (more than HEAVY_ROWS rows).
__global__ __launch_bounds__(256) void k_count_heavy(const Setup* __restrict__ setup,
                                                     const Header* __restrict__ h, Geom G,
                                                     const uint32_t* __restrict__ heavy,
                                                     uint32_t* __restrict__ counts) {
    __shared__ uint32_t s_a[HW_WAVES][64], s_len[HW_WAVES][64], s_pre[HW_WAVES][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t n = h->setup_heavy;
    for (uint32_t i = blockIdx.x * HW_WAVES + wave; i < n; i += gridDim.x * HW_WAVES) {
        const uint32_t e = heavy[i];
        const Setup rec = setup[e];
        const uint32_t rows = (rec.r1 >> 16) - (rec.r0 >> 16) + 1;
        uint32_t count = 0;
        for (uint32_t c = 0; c < rows; c += 64)
            count += heavy_rows(rec, G, c, lane, s_a[wave], s_len[wave], s_pre[wave]);
        if (lane == 0) counts[e] = count;
    }
}

// The quad steps k_expand deferred: per 64 rows, the rows' spans and prefix,
// then the lanes write the chunk's pairs, pair j finding its row by binary
// search over the prefix; a chunk of more than 64 * HEAVY_ROW_WALK pairs is
// written row by row instead (a search per 64 pairs costs more than a pass
// over the rows there: a quad step over the whole screen has 65536 pairs).
// Output positions are fixed by the scan, so the order the list was filled
// in does not matter.
static constexpr uint32_t HEAVY_ROW_WALK = 16;
__global__ __launch_bounds__(256) void k_expand_heavy(const Setup* __restrict__ setup,
                                                      const uint64_t* __restrict__ offsets,
                                                      const Header* __restrict__ h, Geom G,
                                                      const uint32_t* __restrict__ heavy,
                                                      uint32_t* __restrict__ keys,
                                                      uint32_t* __restrict__ vals) {
    __shared__ uint32_t s_a[HW_WAVES][64], s_len[HW_WAVES][64], s_pre[HW_WAVES][64];
    if (h->overflow) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t n = h->heavy;
    for (uint32_t i = blockIdx.x * HW_WAVES + wave; i < n; i += gridDim.x * HW_WAVES) {
        const uint32_t e = heavy[i];
        const Setup rec = setup[e];
        uint64_t off = offsets[e];
        const uint32_t ty0 = rec.r0 >> 16, rows = (rec.r1 >> 16) - ty0 + 1;
        const uint32_t frame_key =
            G.batch ? (uint32_t)loc_quad(G, rec.prim).frame * (uint32_t)(G.ntx * G.nty) : 0u;
        const uint32_t* pa = s_a[wave];
        const uint32_t* pl = s_len[wave];
        const uint32_t* pp = s_pre[wave];
        for (uint32_t c = 0; c < rows; c += 64) {
            const uint32_t total = heavy_rows(rec, G, c, lane, s_a[wave], s_len[wave], s_pre[wave]);
            if (total > 64 * HEAVY_ROW_WALK) {
                // many pairs per row: row by row, each row's run of tiles spread
                // over the lanes, no search
                const uint32_t nrows = min(64u, rows - c);
                for (uint32_t r = 0; r < nrows; r++) {
                    const uint32_t len = pl[r];
                    if (len == 0) continue;
                    const uint64_t o = off + pp[r];
                    const uint32_t key0 = frame_key + (ty0 + c + r) * (uint32_t)G.ntx + pa[r];
                    for (uint32_t j = (uint32_t)lane; j < len; j += 64) {
                        keys[o + j] = key0 + j;
                        vals[o + j] = e;
                    }
                }
                off += total;
                continue;
            }
            for (uint32_t j = (uint32_t)lane; j < total; j += 64) {
                int lo = 0, hi = 63;  // last row whose prefix is <= j
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (pp[mid] <= j) lo = mid;
                    else hi = mid - 1;
                }
                // rows of length 0 share their successor's prefix; step back to a real one
                while (lo > 0 && pl[lo] == 0) lo--;
                const uint32_t ty = ty0 + c + (uint32_t)lo;
                const uint32_t tx = pa[lo] + (j - pp[lo]);
                keys[off + j] = frame_key + ty * (uint32_t)G.ntx + tx;
                vals[off + j] = e;
            }
            off += total;
        }
    }
}

// Tile ranges of the sorted pairs: each thread looks at four consecutive keys
// (one 16-byte load; the key array is 16-byte aligned) and the key after them.
__global__ __launch_bounds__(256) void k_ranges(const uint32_t* __restrict__ keys,
                                                const Header* __restrict__ h,
                                                uint32_t* __restrict__ tstart,
                                                uint32_t* __restrict__ tend) {
    const uint64_t i0 = 4 * ((uint64_t)blockIdx.x * 256 + threadIdx.x);
    const uint64_t n = h->n_sort;
    if (i0 >= n) return;
    uint32_t k[5];
    if (i0 + 4 <= n) {
        const uint4 v = *reinterpret_cast<const uint4*>(keys + i0);
        k[0] = v.x;
        k[1] = v.y;
        k[2] = v.z;
        k[3] = v.w;
    } else {
        for (int j = 0; j < 4; j++) k[j] = i0 + j < n ? keys[i0 + j] : 0u;
    }
    k[4] = i0 + 4 < n ? keys[i0 + 4] : 0u;
    const uint32_t prev = i0 > 0 ? keys[i0 - 1] : 0u;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint64_t i = i0 + j;
        if (i >= n) break;
        if (i == 0 || (j == 0 ? prev : k[j - 1]) != k[j]) tstart[k[j]] = (uint32_t)i;
        if (i == n - 1 || k[j + 1] != k[j]) tend[k[j]] = (uint32_t)(i + 1);
    }
}

// rasterizing.cl:55-93 with the edge biases precomputed per quad step.
__device__ __forceinline__ bool intersect_quad(int px, int py, const int* x, const int* y, int flags) {
    const int b01 = (flags & 1) ? 0 : -1, b12 = (flags & 2) ? 0 : -1, b23 = (flags & 4) ? 0 : -1;
    const int b30 = (flags & 8) ? 0 : -1, b20 = (flags & 16) ? 0 : -1;
    const bool w01 = wadd(edge_i(x[0], y[0], x[1], y[1], px, py), b01) >= 0;
    const bool w12 = wadd(edge_i(x[1], y[1], x[2], y[2], px, py), b12) >= 0;
    const bool w23 = wadd(edge_i(x[2], y[2], x[3], y[3], px, py), b23) >= 0;
    const bool w30 = wadd(edge_i(x[3], y[3], x[0], y[0], px, py), b30) >= 0;
    const bool w20 = wadd(edge_i(x[2], y[2], x[0], y[0], px, py), b20) >= 0;
    const bool front = (w20 && w01 && w12 && (w23 || w30)) || (!w20 && w23 && w30 && (w01 || w12));
    const bool back = (!w20 && !w01 && !w12 && (!w23 || !w30)) || (w20 && !w23 && !w30 && (!w01 || !w12));
    return front || back;
}

// intersect_quad for a quad step with FLAG_SMALL, from its precomputed edge
// rows: edge k at sample (px, py) is a_k py + nb_k px + c_k (mod 2^32), the
// top-left bias folded into c_k.  The sample lies inside the image
// (coordinates below 2^22: width and height are at most 2^19 pixels, AA at
// most 8) and FLAG_SMALL bounds a_k and nb_k by 2^22, so both are signed
// 24-bit operands and a 24-bit multiply returns the low 32 bits of the exact
// product, the same wrapped value the 32-bit edge function computes.
__device__ __forceinline__ int edge_row(const Edges& er, int k, int px, int py) {
    return (int)((uint32_t)__mul24(er.a[k], py) + (uint32_t)__mul24(er.nb[k], px) + (uint32_t)er.c[k]);
}

// The front / back test of rasterizing.cl:55-93 over the signs w_k of the edges
// 01, 12, 23, 30, 20 (bit masks or bools): front = w20 w01 w12 (w23 | w30) |
// !w20 w23 w30 (w01 | w12), back = !w20 !w01 !w12 (!w23 | !w30) | w20 !w23
// !w30 (!w01 | !w12).  With A = w01 w12, O = w01 | w12, B = w23 w30,
// P = w23 | w30, front | back = w20 (A P | !A !P) | !w20 (B O | !B !O): an
// equality of A and P where w20 holds, of B and O where it does not.
template <typename T>
__device__ __forceinline__ T cover_terms(T w01, T w12, T w23, T w30, T w20) {
    const T a = w01 & w12, o = w01 | w12, b = w23 & w30, p = w23 | w30;
    return (w20 & ~(a ^ p)) | (~w20 & ~(b ^ o));
}

// cover_small for every lane of the wave at once, as a lane mask: one
// comparison per edge gives the edge's sign bits of all 64 samples, and the
// front / back tests combine the five masks with wave-uniform (scalar) bit
// operations.  Every lane of the wave must be active.
__device__ __forceinline__ uint64_t cover_small_mask(int px, int py, const Edges& er) {
    const uint64_t w01 = __ballot(edge_row(er, 0, px, py) >= 0);
    const uint64_t w12 = __ballot(edge_row(er, 1, px, py) >= 0);
    const uint64_t w23 = __ballot(edge_row(er, 2, px, py) >= 0);
    const uint64_t w30 = __ballot(edge_row(er, 3, px, py) >= 0);
    const uint64_t w20 = __ballot(edge_row(er, 4, px, py) >= 0);
    return cover_terms(w01, w12, w23, w30, w20);
}

__device__ __forceinline__ bool cover_small(int px, int py, const Edges& er) {
    const uint32_t w01 = edge_row(er, 0, px, py) >= 0, w12 = edge_row(er, 1, px, py) >= 0;
    const uint32_t w23 = edge_row(er, 2, px, py) >= 0, w30 = edge_row(er, 3, px, py) >= 0;
    const uint32_t w20 = edge_row(er, 4, px, py) >= 0;
    return cover_terms(w01, w12, w23, w30, w20) & 1u;
}

// Bit `lane` of a wave-uniform lane mask, for the calling lane: the mask is
// ANDed into the execution mask by one scalar instruction instead of a 64-bit
// shift and a test per lane.
__device__ __forceinline__ bool lane_bit(uint64_t m) { return __builtin_amdgcn_inverse_ballot_w64(m); }

// cover_small_mask at the kAA samples of every pixel: sample s of the pixel
// whose sample-grid origin is (px, py) lies at (px + s, py + sub_offset(kAA -
// 1 + s)) (rasterizing.cl:627-634).  Edge k there is the row at the origin
// plus the wave-uniform a_k oy + nb_k s, modulo 2^32 like the row itself, so
// the rows are evaluated once per pixel and a sample costs one add and one
// comparison per edge.  Returns the lanes with a covered sample and sets
// *hits to this lane's number of covered samples.  Every lane must be active.
template <int kAA>
__device__ __forceinline__ uint64_t cover_small_samples(int px, int py, const Edges& er, int* hits) {
    uint32_t base[5];
#pragma unroll
    for (int k = 0; k < 5; k++) base[k] = (uint32_t)edge_row(er, k, px, py);
    uint64_t any = 0;
    int n = 0;
#pragma unroll
    for (int s = 0; s < kAA; s++) {
        const uint32_t oy = (uint32_t)sub_offset(kAA - 1 + s);
        uint64_t w[5];
#pragma unroll
        for (int k = 0; k < 5; k++) {
            const uint32_t d = (uint32_t)er.a[k] * oy + (uint32_t)er.nb[k] * (uint32_t)s;
            w[k] = __ballot((int)(base[k] + d) >= 0);
        }
        const uint64_t ms = cover_terms(w[0], w[1], w[2], w[3], w[4]);
        any |= ms;
        n += lane_bit(ms) ? 1 : 0;
    }
    *hits = n;
    return any;
}

// rasterizing.cl:95-148 (mean value coordinates of a quad), shading precision:
// hardware reciprocal square root, reciprocal and fused multiply-adds (the
// weights only scale a fragment's colour; coverage is decided elsewhere in
// integers).  1 / r_i comes from one rsq and r_i = |s|^2 / r_i, so the four
// square roots and four divisions by r_i become four rsq.  A pixel on a
// vertex (r_i == 0, as the integer squared distance shows) takes that
// vertex's weight alone, as the source's early return does; it is handled
// off the common path so the common path has no branches.
__device__ __forceinline__ void barycentric_fast(int px, int py, const int* x, const int* y,
                                                 float* w) {
#pragma clang fp contract(fast)
    float sx[4], sy[4], q[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        sx[i] = (float)wsub(x[i], px);
        sy[i] = (float)wsub(y[i], py);
        q[i] = sx[i] * sx[i] + sy[i] * sy[i];
    }
    if (q[0] == 0.0f || q[1] == 0.0f || q[2] == 0.0f || q[3] == 0.0f) {
        const int z = q[0] == 0.0f ? 0 : (q[1] == 0.0f ? 1 : (q[2] == 0.0f ? 2 : 3));
#pragma unroll
        for (int k = 0; k < 4; k++) w[k] = (k == z) ? 1.0f : 0.0f;
        return;
    }
    float ir[4], r[4], t[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        ir[i] = __builtin_amdgcn_rsqf(q[i]);
        r[i] = q[i] * ir[i];
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int j = (i + 1) & 3;
        float a = sx[i] * sy[j] - sx[j] * sy[i];
        const float d = sx[i] * sx[j] + sy[i] * sy[j];
        // On the line through vertices i and j but outside their segment
        // (a == 0, d > 0) the term r_i r_j - d is exactly 0, as the source's
        // correctly rounded lengths of integer vectors give it wherever the
        // lengths are integers; the rsq-based lengths are not exact, and their
        // residual over the source's 1e-3 would be a spurious weight.
        const float num = (a == 0.0f && d > 0.0f) ? 0.0f : r[i] * r[j] - d;
        a = a == 0.0f ? 1e-3f : a;
        t[i] = num * __builtin_amdgcn_rcpf(a);
    }
    float ww[4];
#pragma unroll
    for (int i = 0; i < 4; i++) ww[i] = (t[(i + 3) & 3] + t[i]) * ir[i];
    const float inv = __builtin_amdgcn_rcpf(ww[0] + ww[1] + ww[2] + ww[3]);
#pragma unroll
    for (int i = 0; i < 4; i++) w[i] = ww[i] * inv;
}

// Bilinear lookup of the one-channel ghost texture with border 0 (the OpenCL
// 1.2 formula of sample_bilinear), for shading: no branches.  The lookup reads
// a copy of the texture framed by GHOST_PAD zero texels on every side (row
// pitch w + 2 GHOST_PAD texels, written by k_pad_ghost).  The texel
// coordinate is clamped to [-2, size] (one median-of-three) before the
// conversion: a clamped
// coordinate reads only border texels, as the unclamped one would (both
// texels of an axis lie outside the texture once the coordinate is at or past
// size, or at or below -2), every texel it reads lies in the frame, so no
// texel needs a range test, and NaN becomes a border texel (the weights are
// then NaN, as in the source).  The two texels of a row are one 8-byte load;
// the second row is the same offset plus the pitch, passed as the load's
// scalar offset.
static constexpr int RF_BUFFER_WORD3 = 0x00020000;  // raw 32-bit buffer, gfx9 descriptor
static constexpr int GHOST_PAD = 2;

typedef uint32_t rf_u32x2 __attribute__((ext_vector_type(2)));

// (u, v) in texel units: s * w - 0.5, t * h - 0.5; w, h = the texture size.
// Blended as two lerps along u and one along v (the source's four-product
// form, reassociated).  pitch = w + 2 GHOST_PAD; w + 2 GHOST_PAD and h + 2 GHOST_PAD are below 2^24,
// so the 24-bit multiply is exact.
__device__ __forceinline__ float ghost_sample(__amdgpu_buffer_rsrc_t tex, float w, float h, int pitch,
                                              float u, float v) {
#pragma clang fp contract(fast)
    const float fu = floorf(u), fv = floorf(v);
    const int i0 = (int)__builtin_amdgcn_fmed3f(fu, -2.0f, w);
    const int j0 = (int)__builtin_amdgcn_fmed3f(fv, -2.0f, h);
    const float a = u - fu, b = v - fv;
    const uint32_t off =
        ((uint32_t)__umul24((uint32_t)(j0 + GHOST_PAD), (uint32_t)pitch) + (uint32_t)(i0 + GHOST_PAD)) * 4u;
    const rf_u32x2 r0 = __builtin_amdgcn_raw_buffer_load_b64(tex, off, 0, 0);
    const rf_u32x2 r1 = __builtin_amdgcn_raw_buffer_load_b64(tex, off, pitch * 4, 0);
    const float t00 = __uint_as_float(r0.x), t10 = __uint_as_float(r0.y);
    const float t01 = __uint_as_float(r1.x), t11 = __uint_as_float(r1.y);
    const float top = t00 + a * (t10 - t00), bottom = t01 + a * (t11 - t01);
    return top + b * (bottom - top);
}

// The ghost texture copied into the centre of a frame of GHOST_PAD zero texels.
__global__ __launch_bounds__(256) void k_pad_ghost(const float* __restrict__ ghost, int w, int h,
                                                   float* __restrict__ padded) {
    const int pw = w + 2 * GHOST_PAD, ph = h + 2 * GHOST_PAD;
    const long n = (long)pw * ph;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const int x = (int)(i % pw) - GHOST_PAD, y = (int)(i / pw) - GHOST_PAD;
        padded[i] = (x >= 0 && x < w && y >= 0 && y < h) ? ghost[(long)y * w + x] : 0.0f;
    }
}

// Workgroups are dealt to the XCDs round-robin (workgroup i runs on XCD
// i % RF_XCDS) and each XCD has its own L2.  A quad's records are read by every
// tile it overlaps, so neighbouring tiles should share an L2: the screen is cut
// into SB x SB-tile super-blocks (SB = G.sb) and the SB * SB tiles of one
// super-block run back to back on one XCD.  Super-blocks are taken in the
// order `order` gives (the binner sorts them by pair count, largest first, so
// the longest tiles start early and the end of the launch is not left to a
// few of them): the r-th super-block goes to XCD r % RF_XCDS.
static constexpr int RF_XCDS = 8;

__host__ __device__ __forceinline__ long super_blocks(const Geom& G) {
    const int SB = G.sb;
    return (long)((G.ntx + SB - 1) / SB) * ((G.nty + SB - 1) / SB);
}

__host__ __device__ __forceinline__ long raster_grid(const Geom& G) {
    const long per_xcd = (super_blocks(G) * G.frames + RF_XCDS - 1) / RF_XCDS;
    return per_xcd * RF_XCDS * G.sb * G.sb;
}

__device__ __forceinline__ bool raster_tile(unsigned bid, const Geom& G,
                                            const uint32_t* __restrict__ order, int* frame,
                                            int* tx, int* ty) {
    const int SB = G.sb;
    const unsigned xcd = bid % RF_XCDS, j = bid / RF_XCDS;
    const unsigned rank = (j / (SB * SB)) * RF_XCDS + xcd, t = j % (SB * SB);
    const long nsb = super_blocks(G);
    if ((long)rank >= nsb * G.frames) return false;
    const unsigned v = order[rank];
    const int nsbx = (G.ntx + SB - 1) / SB;
    const unsigned sb = v % (unsigned)nsb;
    *frame = (int)(v / (unsigned)nsb);
    *tx = (int)(sb % nsbx) * SB + (int)(t % SB);
    *ty = (int)(sb / nsbx) * SB + (int)(t / SB);
    return *tx < G.ntx && *ty < G.nty;
}

// Cost of every super-block of the run (its tiles' pair counts, in units of 4
// and capped at 2^16 - 1) as a sort key that puts the largest first; value =
// frame * super_blocks + super-block.
__global__ __launch_bounds__(256) void k_block_cost(const uint32_t* __restrict__ tstart,
                                                    const uint32_t* __restrict__ tend, Geom G,
                                                    uint32_t* __restrict__ keys,
                                                    uint32_t* __restrict__ vals,
                                                    Header* __restrict__ h) {
    const long nsb = super_blocks(G);
    const long n = nsb * G.frames;
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i == 0) h->blocks = (uint64_t)n;
    if (i >= n) return;
    const int SB = G.sb, nsbx = (G.ntx + SB - 1) / SB;
    const long frame = i / nsb, sb = i % nsb;
    const int x0 = (int)(sb % nsbx) * SB, y0 = (int)(sb / nsbx) * SB;
    uint64_t total = 0;
    for (int ty = y0; ty < min(y0 + SB, G.nty); ty++)
        for (int tx = x0; tx < min(x0 + SB, G.ntx); tx++) {
            const long t = frame * G.ntx * G.nty + (long)ty * G.ntx + tx;
            total += tend[t] - tstart[t];
        }
    keys[i] = 0xffffu - (uint32_t)min(total >> 2, (uint64_t)0xffffu);  // 16-bit key: 2 passes
    vals[i] = (uint32_t)i;
}

// The super-block order in one workgroup, for runs of at most
// BLOCK_ORDER_MAX super-blocks: a counting sort on a logarithmic cost bucket
// (16 buckets per doubling of the pair count, the most expensive first).  The
// order only schedules the rasterizer's workgroups (every tile's pixels are
// computed alike in any order), so super-blocks of one bucket may come in any
// order.  Writes a permutation of the run's super-blocks to vals.
static constexpr int BLOCK_ORDER_MAX = 8192;
static constexpr int BLOCK_ORDER_THREADS = 1024;
static constexpr int BLOCK_BUCKETS = 256;

__global__ __launch_bounds__(BLOCK_ORDER_THREADS) void k_block_order(
    const uint32_t* __restrict__ tstart, const uint32_t* __restrict__ tend, Geom G,
    uint32_t* __restrict__ vals, Header* __restrict__ h) {
    __shared__ uint32_t s_bucket[BLOCK_ORDER_MAX];
    __shared__ uint32_t s_count[BLOCK_BUCKETS];
    const long nsb = super_blocks(G);
    const int n = (int)(nsb * G.frames);
    const int t = threadIdx.x;
    if (t == 0) h->blocks = (uint64_t)n;
    if (t < BLOCK_BUCKETS) s_count[t] = 0;
    __syncthreads();
    const int SB = G.sb, nsbx = (G.ntx + SB - 1) / SB;
    for (int i = t; i < n; i += BLOCK_ORDER_THREADS) {
        const long frame = i / nsb, sb = i % nsb;
        const int x0 = (int)(sb % nsbx) * SB, y0 = (int)(sb / nsbx) * SB;
        uint64_t total = 0;
        for (int ty = y0; ty < min(y0 + SB, G.nty); ty++)
            for (int tx = x0; tx < min(x0 + SB, G.ntx); tx++) {
                const long tl = frame * G.ntx * G.nty + (long)ty * G.ntx + tx;
                total += tend[tl] - tstart[tl];
            }
        const int level = min(BLOCK_BUCKETS - 1, (int)(16.0f * __log2f((float)total + 1.0f)));
        const uint32_t b = (uint32_t)(BLOCK_BUCKETS - 1 - level);
        s_bucket[i] = b;
        atomicAdd(&s_count[b], 1u);
    }
    __syncthreads();
    if (t == 0) {  // 256 buckets: an exclusive scan by one lane
        uint32_t run = 0;
        for (int b = 0; b < BLOCK_BUCKETS; b++) {
            const uint32_t c = s_count[b];
            s_count[b] = run;
            run += c;
        }
    }
    __syncthreads();
    for (int i = t; i < n; i += BLOCK_ORDER_THREADS) vals[atomicAdd(&s_count[s_bucket[i]], 1u)] = (uint32_t)i;
}

// tstart and tend cleared by one launch
__global__ __launch_bounds__(256) void k_clear_ranges(uint32_t* __restrict__ a, uint32_t* __restrict__ b,
                                                      long n) {
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        a[i] = 0u;
        b[i] = 0u;
    }
}

// fragment_shader (rasterizing.cl:471-506) for one covered pixel of one quad
// step, scaled by its covered AA samples; fused multiply-adds allowed.
__device__ __forceinline__ float shade_fragment(int pcx, int pcy, const Setup& rec, const Attr& at,
                                                int hits, __amdgpu_buffer_rsrc_t ghost,
                                                const Geom& G) {
#pragma clang fp contract(fast)
    float w[4];
    barycentric_fast(pcx, pcy, rec.vx, rec.vy, w);
    const float uu = w[0] * at.u[0] + w[1] * at.u[1] + w[2] * at.u[2] + w[3] * at.u[3];
    const float vv = w[0] * at.v[0] + w[1] * at.v[1] + w[2] * at.v[2] + w[3] * at.v[3];
    const float gi = ghost_sample(ghost, G.gwf, G.ghf, G.gw + 2 * GHOST_PAD, uu * G.gku + G.gcu,
                                  vv * G.gkv + G.gcv);
    const float rel = w[0] * at.rrel[0] + w[1] * at.rrel[1] + w[2] * at.rrel[2] + w[3] * at.rrel[3];
    // smoothstep(1, 0.95, rrel)
    constexpr float inv_edge = 1.0f / (0.95f - 1.0f);
    const float tt = clampn(rel * inv_edge - inv_edge, 0.0f, 1.0f);
    const float ri = tt * tt * (3.0f - 2.0f * tt);
    float ai = w[0] * at.inten[0] + w[1] * at.inten[1] + w[2] * at.inten[2] + w[3] * at.inten[3];
    float ci = w[0] * at.refl[0] + w[1] * at.refl[1] + w[2] * at.refl[2] + w[3] * at.refl[3];
    ai = fmaxf(ai, 0.0f);
    ci = fmaxf(ci, 0.0f);
    return gi * ri * ai * ci * (float)hits;
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
}

// Fragment queue capacity: two waves' worth, so a flush shades two passes and
// its fixed costs (barriers, the colour exchange) are paid half as often.
// 4.6 KB of LDS per tile keeps the occupancy the registers allow; more
// (256 entries, 9 KB) would lower it.
static constexpr uint32_t QCAP = 128;

// Pairs between a pair whose edge rows have arrived and the pair whose rows
// are then pulled into the scalar cache (k_raster's walk).
static constexpr uint32_t TOUCH = 2;

// rasterizer (rasterizing.cl:524-663): one wave per 8x8 tile, lane = pixel.
// The tile's (quad step) pairs are walked in binning order.  A pair's coverage
// is tested by every pixel lane; the covered pixels are appended to a QCAP-entry
// fragment queue in LDS (pixel lane, covered samples, pair slot).  When the
// queue is full, lane i shades fragments i, i + 64, ..., and the colours go back
// to their pixel lanes slot by slot, i.e. in pair order, so each pixel adds
// exactly the terms of the per-pixel loop in the same order.  Shading thus runs
// on full waves even when a quad covers few pixels of the tile.
// kMask: bins made of whole tiles (G.bin_exact, every bin size that is a
// multiple of 8), with kAA = G.AA samples per pixel known at compile time: its
// instantiations carry no bin re-test, and the coverage of a FLAG_SMALL quad
// step is computed as lane masks, one per sample.  kMask false is the generic
// form (any bin size, G.AA read at run time, kAA unused).
template <bool kMask, int kAA>
__global__ __launch_bounds__(64) void k_raster(const float4* __restrict__ bounds,
                                               const float* __restrict__ ghost_pad,
                                               const Setup* __restrict__ setup,
                                               const Edges* __restrict__ edges,
                                               const Attr* __restrict__ attr,
                                               const uint32_t* __restrict__ vals,
                                               const uint32_t* __restrict__ tstart,
                                               const uint32_t* __restrict__ tend,
                                               const uint32_t* __restrict__ order,
                                               const Header* __restrict__ h, Geom G,
                                               float4* __restrict__ image) {
#pragma clang fp contract(off)
    __shared__ uint32_t s_q[QCAP];  // pixel lane | hits << 6 | slot << 10
    __shared__ uint4 s_slot[QCAP];  // per queued pair: coverage mask lo, hi, first queue index, entry
    __shared__ float4 s_c[QCAP];    // shaded colour of each queue entry
    const int lane = threadIdx.x;
    if (h->overflow) {
        // the pairs were not binned: the image is cleared (an accumulated
        // image is left as it was), never left holding an earlier frame
        if (G.accumulate) return;
        const long px = (long)G.width * G.height * G.frames;
        for (long i = (long)blockIdx.x * 64 + lane; i < px; i += (long)gridDim.x * 64)
            image[i] = make_float4(0, 0, 0, 0);
        return;
    }
    int frame, tx, ty;
    if (!raster_tile(blockIdx.x, G, order, &frame, &tx, &ty)) return;
    const uint32_t tile = (uint32_t)frame * (uint32_t)(G.ntx * G.nty) + (uint32_t)(ty * G.ntx + tx);
    const int x = tx * TILE + (lane & (TILE - 1)), y = ty * TILE + (lane / TILE);
    const bool inside = x < G.width && y < G.height;
    const int AA = kMask ? kAA : G.AA;
    const int px = x * AA, py = y * AA;
    float r0 = 0.0f, r1 = 0.0f, r2 = 0.0f;
    // queue state (wave-uniform); slot s (the s-th queued pair) in s_slot[s]
    uint32_t qlen = 0, nslot = 0;
    const __amdgpu_buffer_rsrc_t ghost_rsrc = __builtin_amdgcn_make_buffer_rsrc(
        (void*)ghost_pad, (short)0,
        (G.gw + 2 * GHOST_PAD) * (G.gh + 2 * GHOST_PAD) * (int)sizeof(float), RF_BUFFER_WORD3);

    auto flush = [&]() {
        __syncthreads();  // queue entries written by other lanes are visible
        for (uint32_t q0 = 0; q0 < qlen; q0 += 64) {
            const uint32_t q = q0 + (uint32_t)lane;
            float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
            if (q < qlen) {
                const uint32_t ent = s_q[q];
                const int p = ent & 63, hits = (ent >> 6) & 15;
                const uint32_t e = s_slot[ent >> 10].w;
                const Setup rec = setup[e];
                const Attr at = attr[e];
                const int fx = tx * TILE + (p & (TILE - 1)), fy = ty * TILE + p / TILE;
                const float frag = shade_fragment(fx * AA + AA / 2, fy * AA + AA / 2, rec,
                                                  at, hits, ghost_rsrc, G);
                c0 = at.x * frag;
                c1 = at.y * frag;
                c2 = at.z * frag;
            }
            s_c[q] = make_float4(c0, c1, c2, 0.0f);
        }
        __syncthreads();  // colours visible to their pixel lanes
        for (uint32_t sl = 0; sl < nslot; sl++) {
            const uint4 si = s_slot[sl];
            const uint64_t m = ((uint64_t)si.y << 32) | si.x;
            if ((m >> lane) & 1) {
                const float4 v = s_c[si.z + lanes_below(m)];
                r0 += v.x;
                r1 += v.y;
                r2 += v.z;
            }
        }
        qlen = 0;
        nslot = 0;
        __syncthreads();  // every entry has been read before the queue is refilled
    };

    const uint32_t start = tstart[tile], end = tend[tile];
    const uint64_t inside_mask = __ballot(inside);
    // The pair ids are read 64 at a time by one vector load (lane l holds the
    // id of pair kb + l) and taken per pair with readlane, so no scalar load
    // of an id precedes the rows of a pair.  The scalar load of a pair's edge
    // rows is then the walk's only latency: once the rows of pair j have
    // arrived, one dword of the rows of pair j + TOUCH is loaded, which brings
    // that pair's 64-byte record into the scalar cache while pairs
    // j .. j + TOUCH - 1 are processed.  The touch sits in the branch on the
    // rows' flags, so it is issued after the rows' wait: scalar loads complete
    // out of order, so a wait covers every one in flight, and a touch issued
    // before that wait would be waited for with the rows.  The touched value
    // is consumed by an empty asm statement after the wait of a later pair, so
    // the load is kept and costs no wait.
    uint32_t touched = 0;
    for (uint32_t kb = start; kb < end; kb += 64) {
    const uint32_t ids = kb + (uint32_t)lane < end ? vals[kb + lane] : 0u;
    const uint32_t jn = (uint32_t)__builtin_amdgcn_readfirstlane((int)min(64u, end - kb));
    for (uint32_t j = 0; j < jn; j++) {
        const uint32_t e = (uint32_t)__builtin_amdgcn_readlane((int)ids, (int)j);
        const Edges er = edges[e];
        int hits = 0;
        uint64_t m;
        // bins made of whole tiles: the coverage of a FLAG_SMALL quad step is
        // computed as lane masks (cover_small_mask; with several samples per
        // pixel, one mask per sample)
        if (kMask && (er.flags & FLAG_SMALL)) {
            if (j + TOUCH < jn) {
                const uint32_t en = (uint32_t)__builtin_amdgcn_readlane((int)ids, (int)(j + TOUCH));
                asm volatile("" ::"s"(touched));
                touched = reinterpret_cast<const uint32_t*>(edges)[(size_t)en * (sizeof(Edges) / 4)];
            }
            if (kAA == 1) {
                m = cover_small_mask(px, py, er) & inside_mask;
                hits = 1;  // read by the covered lanes only
            } else {
                m = cover_small_samples<kAA>(px, py, er, &hits) & inside_mask;
            }
        } else {
            bool ok = inside;
            if (!kMask && ok && !G.bin_exact) {
                const float4 b = bounds[setup[e].prim];
                const int sbx0 = wadd(rf_to_int(b.x * G.st), G.width / 2);
                const int sby0 = wadd(rf_to_int(b.y * G.st), G.height / 2);
                const int sbx1 = wadd(rf_to_int(b.z * G.st), G.width / 2);
                const int sby1 = wadd(rf_to_int(b.w * G.st), G.height / 2);
                const int bx0 = x / G.bin * G.bin, by0 = y / G.bin * G.bin;
                ok = sbx0 <= bx0 + G.bin && sbx1 >= bx0 && sby0 <= by0 + G.bin && sby1 >= by0;
            }
            if (ok) {
                if (er.flags & FLAG_SMALL) {
                    if (AA == 1) {
                        hits = cover_small(px, py, er) ? 1 : 0;
                    } else {
                        for (int s = 0; s < AA; s++)
                            hits += cover_small(px + s, py + sub_offset(AA - 1 + s), er) ? 1 : 0;
                    }
                } else {
                    const Setup rec = setup[e];
                    for (int s = 0; s < AA; s++) {
                        const int sx = px + s, sy = py + sub_offset(AA - 1 + s);
                        if (intersect_quad(sx, sy, rec.vx, rec.vy, rec.flags)) hits++;
                    }
                }
            }
            m = __ballot(hits > 0);
        }
        if (m == 0) continue;
        const uint32_t n = (uint32_t)__popcll(m);
        if (qlen + n > QCAP) flush();
        if (lane_bit(m))
            s_q[qlen + lanes_below(m)] = (uint32_t)lane | (uint32_t)hits << 6 | nslot << 10;
        if (lane == 0) s_slot[nslot] = make_uint4((uint32_t)m, (uint32_t)(m >> 32), qlen, e);
        qlen += n;
        nslot++;
    }
    }
    if (qlen > 0) flush();
    if (!inside) return;
    float4 out = make_float4(0, 0, 0, 0);
    if (r0 > 0 || r1 > 0 || r2 > 0) {
        const float f = G.intensity / (float)(G.W * AA * G.Wsub);
        float o[4];
        xyz_to_ap1(r0 * f, r1 * f, r2 * f, o);
        out = make_float4(o[0], o[1], o[2], o[3]);
    }
    float4* dst = image + (long)frame * G.width * G.height + (long)(G.height - 1 - y) * G.width + x;
    if (G.accumulate) {
        const float4 prev = *dst;
        out = make_float4(prev.x + out.x, prev.y + out.y, prev.z + out.z, prev.w + out.w);
    }
    *dst = out;
}

// ---------------------------------------------------------------------------

static int make_geom(const rf_raster_params* p, Geom* G) {
    if (!p) return set_error(RF_ERR_INVALID, "raster: params is NULL");
    if (p->path_count < 0 || p->wavelength_count < 1 || p->grid_count < 2)
        return set_error(RF_ERR_INVALID, "raster: need wavelength_count >= 1 and grid_count >= 2");
    if (p->width < 1 || p->height < 1 || p->width > 65536 * TILE || p->height > 65536 * TILE)
        return set_error(RF_ERR_INVALID, "raster: resolution %dx%d out of range", p->width, p->height);
    if (!(p->anti_aliasing == 1 || p->anti_aliasing == 2 || p->anti_aliasing == 4 ||
          p->anti_aliasing == 8))
        return set_error(RF_ERR_INVALID, "raster: anti_aliasing must be 1, 2, 4 or 8 (got %d)",
                         p->anti_aliasing);
    if (p->wavelength_sub_count < 1)
        return set_error(RF_ERR_INVALID, "raster: wavelength_sub_count must be >= 1");
    if (p->bin_size < 1) return set_error(RF_ERR_INVALID, "raster: bin_size must be >= 1");
    if (p->ghost_width < 1 || p->ghost_height < 1)
        return set_error(RF_ERR_INVALID, "raster: ghost texture size must be positive");
    if ((long)p->ghost_width * p->ghost_height > (1L << 28))
        return set_error(RF_ERR_INVALID, "raster: ghost texture of %dx%d texels exceeds 2^28",
                         p->ghost_width, p->ghost_height);
    if (p->ghost_width > (1 << 24) - 2 * GHOST_PAD - 1 || p->ghost_height > (1 << 24) - 2 * GHOST_PAD - 1)
        return set_error(RF_ERR_INVALID, "raster: ghost texture side of %dx%d reaches 2^24",
                         p->ghost_width, p->ghost_height);
    G->P = p->path_count;
    G->W = p->wavelength_count;
    G->g = p->grid_count;
    G->R = G->g * G->g;
    G->Q = (G->g - 1) * (G->g - 1);
    G->Wsub = p->wavelength_sub_count;
    G->maxw = G->W - 1 > 1 ? G->W - 1 : 1;
    G->S = G->maxw * G->Wsub;
    G->AA = p->anti_aliasing;
    G->width = p->width;
    G->height = p->height;
    G->bin = p->bin_size;
    G->ntx = (G->width + TILE - 1) / TILE;
    G->nty = (G->height + TILE - 1) / TILE;
    G->nbx = (G->width + G->bin - 1) / G->bin;
    G->nby = (G->height + G->bin - 1) / G->bin;
    G->N = (long)G->P * G->W * G->R;
    G->NV = G->N;
    G->NE = (long)G->P * G->Q * G->S;
    G->st = p->screen_transform;
    G->area_orig = p->area_orig;
    G->min_area = p->min_area;
    G->intensity = p->intensity;
    G->ghost_scale = p->ghost_scale;
    G->gw = p->ghost_width;
    G->gh = p->ghost_height;
    G->nspec = p->spectrum_count;
    G->bin_exact = (G->bin % TILE) == 0;
    G->accumulate = p->accumulate != 0;
    G->sb = p->super_block == 0 ? 4 : p->super_block;
    if (!(G->sb == 1 || G->sb == 2 || G->sb == 4 || G->sb == 8))
        return set_error(RF_ERR_INVALID, "raster: super_block must be 0, 1, 2, 4 or 8 (got %d)",
                         p->super_block);
    G->batch = 0;
    G->nseg = 0;
    G->frame_base = 0;
    G->frames = 1;
    G->segs = nullptr;
    G->ray_base = 0;
    G->quad_base = 0;
    G->NQ = (long)G->P * G->Q;
    G->NR = G->N;
    G->min_area_factor = 0.0f;
    const float half_inv = 0.5f / G->ghost_scale;
    G->gku = half_inv * (float)G->gw;
    G->gcu = 0.5f * (float)G->gw - 0.5f;
    G->gkv = half_inv * (float)G->gh;
    G->gcv = 0.5f * (float)G->gh - 0.5f;
    G->gwf = (float)G->gw;
    G->ghf = (float)G->gh;
    G->aa_shift = 0;
    while ((1 << G->aa_shift) < G->AA) G->aa_shift++;
    G->bin_shift = -1;
    if ((G->bin & (G->bin - 1)) == 0) {
        G->bin_shift = 0;
        while ((1 << G->bin_shift) < G->bin) G->bin_shift++;
    }
    if (G->W > 1 && G->nspec < 1)
        return set_error(RF_ERR_INVALID, "raster: spectrum_count must be >= 1");
    if (G->NE > 0xffffffffL)
        return set_error(RF_ERR_INVALID, "raster: %ld quad steps exceed 2^32", G->NE);
    return RF_OK;
}

static inline unsigned blocks_for(long n, int threads) { return (unsigned)((n + threads - 1) / threads); }

struct Layout {
    size_t header, table, setup, attr, edges, heavy, order_keys, order_vals, counts, offsets,
        scan_ws, keys, vals, radix_ws, tstart, tend, ghost_pad, total;
};

static Layout layout(const Geom& G, uint64_t cap) {
    Layout L;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        size_t at = o;
        o += align_up(bytes, 256);
        return at;
    };
    L.header = take(sizeof(Header));
    L.table = take((size_t)G.S * sizeof(float4));
    L.setup = take((size_t)G.NE * sizeof(Setup));
    L.attr = take((size_t)G.NE * sizeof(Attr));
    L.edges = take((size_t)G.NE * sizeof(Edges));
    L.heavy = take((size_t)G.NE * sizeof(uint32_t));
    const uint64_t nblk = (uint64_t)(super_blocks(G) * G.frames);
    L.order_keys = take(nblk * sizeof(uint32_t));
    L.order_vals = take(nblk * sizeof(uint32_t));
    L.counts = take((size_t)G.NE * sizeof(uint32_t));
    L.offsets = take((size_t)G.NE * sizeof(uint64_t));
    L.scan_ws = take(scan_workspace_bytes((uint64_t)G.NE));
    L.keys = take(cap * sizeof(uint32_t));
    L.vals = take(cap * sizeof(uint32_t));
    L.radix_ws = take(radix_workspace_bytes(cap > nblk ? cap : nblk));
    const size_t ntiles = (size_t)G.ntx * G.nty * G.frames;
    L.tstart = take(ntiles * sizeof(uint32_t));
    L.tend = take(ntiles * sizeof(uint32_t));
    L.ghost_pad = take((size_t)(G.gw + 2 * GHOST_PAD) * (G.gh + 2 * GHOST_PAD) * sizeof(float));
    L.total = o;
    return L;
}

struct RasterPtrs {
    Header* h;
    float4* table;
    Setup* setup;
    Attr* attr;
    Edges* edges;
    uint32_t* heavy;
    uint32_t *order_keys, *order_vals;
    uint32_t *counts, *keys, *vals, *tstart, *tend;
    uint64_t* offsets;
    void *scan_ws, *radix_ws;
    float* ghost_pad;
};

static int raster_ptrs(const Geom& G, uint64_t cap, void* ws, size_t ws_bytes, RasterPtrs* r,
                       const char* who) {
    const Layout L = layout(G, cap);
    if (!ws || ws_bytes < L.total)
        return set_error(RF_ERR_WORKSPACE, "%s: workspace %zu bytes < %zu needed", who, ws_bytes,
                         L.total);
    char* w = (char*)ws;
    r->h = (Header*)(w + L.header);
    r->table = (float4*)(w + L.table);
    r->setup = (Setup*)(w + L.setup);
    r->attr = (Attr*)(w + L.attr);
    r->edges = (Edges*)(w + L.edges);
    r->heavy = (uint32_t*)(w + L.heavy);
    r->order_keys = (uint32_t*)(w + L.order_keys);
    r->order_vals = (uint32_t*)(w + L.order_vals);
    r->counts = (uint32_t*)(w + L.counts);
    r->offsets = (uint64_t*)(w + L.offsets);
    r->scan_ws = w + L.scan_ws;
    r->keys = (uint32_t*)(w + L.keys);
    r->vals = (uint32_t*)(w + L.vals);
    r->radix_ws = w + L.radix_ws;
    r->tstart = (uint32_t*)(w + L.tstart);
    r->tend = (uint32_t*)(w + L.tend);
    r->ghost_pad = (float*)(w + L.ghost_pad);
    return RF_OK;
}

// binner: step tables, quad-step setup, pair expansion, stable sort, tile ranges
static int bin_impl(const float4* vtx, const float4* bounds, const float* spectrum, const Geom& G,
                    uint64_t cap, const RasterPtrs& r, hipStream_t s) {
    const long ntiles = (long)G.ntx * G.nty * G.frames;
    hipLaunchKernelGGL(k_steps, dim3(blocks_for(G.maxw, 64)), dim3(64), 0, s, spectrum, G, r.table,
                       r.h);
    if (G.NE > 0) {
        hipLaunchKernelGGL(k_setup, dim3(blocks_for(G.NE, 256)), dim3(256), 0, s, vtx, bounds,
                           r.table, G, r.setup, r.attr, r.edges, r.counts, r.heavy, r.h);
        hipLaunchKernelGGL(k_count_heavy, dim3(HEAVY_BLOCKS), dim3(256), 0, s, r.setup, r.h, G,
                           r.heavy, r.counts);
    }
    int rc = scan_exclusive_u32(r.counts, r.offsets, nullptr, (uint64_t)G.NE, &r.h->required,
                                r.scan_ws, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1), 0, s, r.h, cap);
    if (G.NE > 0) {
        hipLaunchKernelGGL(k_expand, dim3(blocks_for(G.NE, 256)), dim3(256), 0, s, r.setup, r.counts,
                           r.offsets, r.h, G, r.keys, r.vals, r.heavy);
        hipLaunchKernelGGL(k_expand_heavy, dim3(HEAVY_BLOCKS), dim3(256), 0, s, r.setup, r.offsets,
                           r.h, G, r.heavy, r.keys, r.vals);
    }
    int bits = 1;
    while ((1L << bits) < ntiles) bits++;
    rc = radix_sort_pairs(r.keys, r.vals, &r.h->n_sort, cap, bits, r.radix_ws, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_clear_ranges, dim3(blocks_for(ntiles, 256)), dim3(256), 0, s, r.tstart,
                       r.tend, ntiles);
    if (cap > 0) {
        hipLaunchKernelGGL(k_ranges, dim3(blocks_for(((long)cap + 3) / 4, 256)), dim3(256), 0, s,
                           r.keys, r.h, r.tstart, r.tend);
    }
    // the rasterizer's super-block order: largest pair count first (stable)
    const long nblk = super_blocks(G) * G.frames;
    if (nblk <= BLOCK_ORDER_MAX) {
        hipLaunchKernelGGL(k_block_order, dim3(1), dim3(BLOCK_ORDER_THREADS), 0, s, r.tstart, r.tend,
                           G, r.order_vals, r.h);
    } else {
        hipLaunchKernelGGL(k_block_cost, dim3(blocks_for(nblk, 256)), dim3(256), 0, s, r.tstart,
                           r.tend, G, r.order_keys, r.order_vals, r.h);
        rc = radix_sort_pairs(r.order_keys, r.order_vals, &r.h->blocks, (uint64_t)nblk, 16,
                              r.radix_ws, s);
        if (rc) return rc;
    }
    RF_LAUNCH_CHECK("binner");
    return RF_OK;
}

static int draw_impl(const float4* vtx, const float4* bounds, const float* ghost, const Geom& G,
                     const RasterPtrs& r, float4* image, hipStream_t s) {
    const long ntiles = (long)G.ntx * G.nty;
    (void)vtx;
    (void)ntiles;
    const long texels = (long)(G.gw + 2 * GHOST_PAD) * (G.gh + 2 * GHOST_PAD);
    hipLaunchKernelGGL(k_pad_ghost, dim3((unsigned)min(blocks_for(texels, 256), 4096u)), dim3(256), 0,
                       s, ghost, G.gw, G.gh, r.ghost_pad);
    decltype(&k_raster<false, 0>) kern = k_raster<false, 0>;
    if (G.bin_exact) {
        switch (G.AA) {
            case 1: kern = k_raster<true, 1>; break;
            case 2: kern = k_raster<true, 2>; break;
            case 4: kern = k_raster<true, 4>; break;
            case 8: kern = k_raster<true, 8>; break;
            default: break;
        }
    }
    hipLaunchKernelGGL(kern, dim3((unsigned)raster_grid(G)), dim3(64), 0, s, bounds, r.ghost_pad,
                       r.setup, r.edges, r.attr, r.vals, r.tstart, r.tend, r.order_vals, r.h, G,
                       image);
    RF_LAUNCH_CHECK("rasterizer");
    return RF_OK;
}

static int raster_impl(const float4* vtx, const float4* bounds, const float* ghost,
                       const float* spectrum, const Geom& G, uint64_t cap, void* ws, size_t ws_bytes,
                       float4* image, hipStream_t s) {
    RasterPtrs r;
    int rc = raster_ptrs(G, cap, ws, ws_bytes, &r, "rf_rasterize");
    if (rc) return rc;
    rc = bin_impl(vtx, bounds, spectrum, G, cap, r, s);
    if (rc) return rc;
    return draw_impl(vtx, bounds, ghost, G, r, image, s);
}

struct FlareLayout {
    size_t bounds, intens, vertexes, raster, total;
};

static FlareLayout flare_layout(const Geom& G, uint64_t cap) {
    FlareLayout F;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        size_t at = o;
        o += align_up(bytes, 256);
        return at;
    };
    F.raster = take(layout(G, cap).total);  // first: its header sits at offset 0
    F.bounds = take((size_t)G.P * G.Q * sizeof(float4));
    F.intens = take((size_t)G.P * G.Q * G.W * sizeof(float));
    F.vertexes = take((size_t)G.NV * 2 * sizeof(float4));
    F.total = o;
    return F;
}

}  // namespace rf

using namespace rf;

extern "C" int rf_prim_shader(const float* planes, const rf_raster_params* prm, float* bounds,
                              float* intensities, rf_stream_t stream) {
    Geom G;
    int rc = make_geom(prm, &G);
    if (rc) return rc;
    if ((long)G.P * G.Q == 0) return RF_OK;
    if (!planes || !bounds || !intensities) return set_error(RF_ERR_INVALID, "rf_prim_shader: NULL");
    hipLaunchKernelGGL(k_prim, dim3(blocks_for((long)G.P * G.Q, 256)), dim3(256), 0,
                       (hipStream_t)stream, planes, G, (float4*)bounds, intensities);
    RF_LAUNCH_CHECK("rf_prim_shader");
    return RF_OK;
}

extern "C" int rf_vertex_shader(const float* planes, const float* intensities,
                                const rf_raster_params* prm, void* vertexes, rf_stream_t stream) {
    Geom G;
    int rc = make_geom(prm, &G);
    if (rc) return rc;
    if (G.NV == 0) return RF_OK;
    if (!planes || !intensities || !vertexes)
        return set_error(RF_ERR_INVALID, "rf_vertex_shader: NULL");
    hipLaunchKernelGGL(k_vertex, dim3(blocks_for(G.NV, 256)), dim3(256), 0, (hipStream_t)stream,
                       planes, intensities, G, (float4*)vertexes);
    RF_LAUNCH_CHECK("rf_vertex_shader");
    return RF_OK;
}

extern "C" size_t rf_rasterize_workspace_size(const rf_raster_params* prm, uint64_t pair_capacity) {
    Geom G;
    if (make_geom(prm, &G)) return 0;
    return layout(G, pair_capacity).total;
}

extern "C" int rf_rasterize(const void* vertexes, const float* bounds, const float* ghost,
                            const float* light_spectrum, const rf_raster_params* prm,
                            uint64_t pair_capacity, void* workspace, size_t workspace_bytes,
                            float* image, rf_stream_t stream) {
    Geom G;
    int rc = make_geom(prm, &G);
    if (rc) return rc;
    if (!image || !ghost || (G.W > 1 && !light_spectrum) || (G.NE > 0 && (!vertexes || !bounds)))
        return set_error(RF_ERR_INVALID, "rf_rasterize: NULL buffer");
    return raster_impl((const float4*)vertexes, (const float4*)bounds, ghost, light_spectrum, G,
                       pair_capacity, workspace, workspace_bytes, (float4*)image,
                       (hipStream_t)stream);
}

extern "C" int rf_raster_bin(const void* vertexes, const float* bounds,
                             const float* light_spectrum, const rf_raster_params* prm,
                             uint64_t pair_capacity, void* workspace, size_t workspace_bytes,
                             rf_stream_t stream) {
    Geom G;
    int rc = make_geom(prm, &G);
    if (rc) return rc;
    if ((G.W > 1 && !light_spectrum) || (G.NE > 0 && (!vertexes || !bounds)))
        return set_error(RF_ERR_INVALID, "rf_raster_bin: NULL buffer");
    RasterPtrs r;
    rc = raster_ptrs(G, pair_capacity, workspace, workspace_bytes, &r, "rf_raster_bin");
    if (rc) return rc;
    return bin_impl((const float4*)vertexes, (const float4*)bounds, light_spectrum, G,
                    pair_capacity, r, (hipStream_t)stream);
}

extern "C" int rf_raster_draw(const void* vertexes, const float* bounds, const float* ghost,
                              const rf_raster_params* prm, uint64_t pair_capacity, void* workspace,
                              size_t workspace_bytes, float* image, rf_stream_t stream) {
    Geom G;
    int rc = make_geom(prm, &G);
    if (rc) return rc;
    if (!image || !ghost || (G.NE > 0 && (!vertexes || !bounds)))
        return set_error(RF_ERR_INVALID, "rf_raster_draw: NULL buffer");
    RasterPtrs r;
    rc = raster_ptrs(G, pair_capacity, workspace, workspace_bytes, &r, "rf_raster_draw");
    if (rc) return rc;
    return draw_impl((const float4*)vertexes, (const float4*)bounds, ghost, G, r, (float4*)image,
                     (hipStream_t)stream);
}

extern "C" size_t rf_flare_workspace_size(const rf_raster_params* prm, uint64_t pair_capacity) {
    Geom G;
    if (make_geom(prm, &G)) return 0;
    return flare_layout(G, pair_capacity).total;
}

extern "C" int rf_flare(const float* planes, const float* ghost, const float* light_spectrum,
                        const rf_raster_params* prm, uint64_t pair_capacity, void* workspace,
                        size_t workspace_bytes, float* image, rf_stream_t stream) {
    Geom G;
    int rc = make_geom(prm, &G);
    if (rc) return rc;
    const FlareLayout F = flare_layout(G, pair_capacity);
    if (!workspace || workspace_bytes < F.total)
        return set_error(RF_ERR_WORKSPACE, "rf_flare: workspace %zu bytes < %zu needed",
                         workspace_bytes, F.total);
    if (!image || !ghost || (G.W > 1 && !light_spectrum) || (G.NV > 0 && !planes))
        return set_error(RF_ERR_INVALID, "rf_flare: NULL buffer");
    char* w = (char*)workspace;
    float4* bounds = (float4*)(w + F.bounds);
    float* intens = (float*)(w + F.intens);
    float4* vtx = (float4*)(w + F.vertexes);
    hipStream_t s = (hipStream_t)stream;
    if ((long)G.P * G.Q > 0) {
        hipLaunchKernelGGL(k_prim, dim3(blocks_for((long)G.P * G.Q, 256)), dim3(256), 0, s, planes, G,
                           bounds, intens);
        hipLaunchKernelGGL(k_vertex, dim3(blocks_for(G.NV, 256)), dim3(256), 0, s, planes, intens, G,
                           vtx);
    }
    return raster_impl(vtx, bounds, ghost, light_spectrum, G, pair_capacity, w + F.raster,
                       F.bounds - F.raster, (float4*)image, s);
}

// ---------------------------------------------------------------------------
// Batched runs

static int make_geom_batch(const rf_raster_params* p, const rf_batch* b, const char* who,
                           Geom* G) {
    if (!p || !b) return set_error(RF_ERR_INVALID, "%s: NULL params or batch", who);
    rf_raster_params q = *p;
    q.path_count = 0;  // the segments give ghosts and grids
    q.grid_count = 2;
    int rc = make_geom(&q, G);
    if (rc) return rc;
    if (b->segment_count < 1 || !b->segments)
        return set_error(RF_ERR_INVALID, "%s: empty segment list", who);
    if (b->frame_count < 1 || b->frame_count > 65535 ||
        (long)b->frame_count * G->ntx * G->nty > 0xffffffffL)
        return set_error(RF_ERR_INVALID, "%s: frame_count %d", who, b->frame_count);
    if (b->ray_base < 0 || b->ray_count < 0 || b->quad_base < 0 || b->quad_count < 0 ||
        b->ray_base + b->ray_count > b->ray_total)
        return set_error(RF_ERR_INVALID, "%s: inconsistent ray / quad ranges", who);
    G->batch = 1;
    G->segs = b->segments;
    G->nseg = b->segment_count;
    G->frame_base = b->frame_base;
    G->frames = b->frame_count;
    G->ray_base = b->ray_base;
    G->quad_base = b->quad_base;
    G->NQ = b->quad_count;
    G->NR = b->ray_count;
    G->N = b->ray_total;  // plane stride
    G->NV = b->ray_total;  // vertices of the whole trace
    G->NE = b->quad_count * G->S;
    G->min_area_factor = p->min_area;
    if (G->NE > 0xffffffffL)
        return set_error(RF_ERR_INVALID, "%s: %ld quad steps exceed 2^32", who, G->NE);
    return RF_OK;
}

extern "C" int rf_prim_shader_batch(const float* planes, const rf_raster_params* prm,
                                    const rf_batch* batch, float* bounds, float* intensities,
                                    rf_stream_t stream) {
    Geom G;
    int rc = make_geom_batch(prm, batch, "rf_prim_shader_batch", &G);
    if (rc) return rc;
    if (!planes || !bounds || !intensities)
        return set_error(RF_ERR_INVALID, "rf_prim_shader_batch: NULL");
    if (G.NQ == 0) return RF_OK;
    hipLaunchKernelGGL(k_prim, dim3(blocks_for(G.NQ, 256)), dim3(256), 0, (hipStream_t)stream,
                       planes, G, (float4*)bounds, intensities);
    RF_LAUNCH_CHECK("rf_prim_shader_batch");
    return RF_OK;
}

extern "C" int rf_vertex_shader_batch(const float* planes, const float* intensities,
                                      const rf_raster_params* prm, const rf_batch* batch,
                                      void* vertexes, rf_stream_t stream) {
    Geom G;
    int rc = make_geom_batch(prm, batch, "rf_vertex_shader_batch", &G);
    if (rc) return rc;
    if (!planes || !intensities || !vertexes)
        return set_error(RF_ERR_INVALID, "rf_vertex_shader_batch: NULL");
    if (G.NR == 0) return RF_OK;
    hipLaunchKernelGGL(k_vertex, dim3(blocks_for(G.NR, 256)), dim3(256), 0, (hipStream_t)stream,
                       planes, intensities, G, (float4*)vertexes);
    RF_LAUNCH_CHECK("rf_vertex_shader_batch");
    return RF_OK;
}

extern "C" size_t rf_raster_batch_workspace_size(const rf_raster_params* prm,
                                                 const rf_batch* batch, uint64_t pair_capacity) {
    Geom G;
    if (make_geom_batch(prm, batch, "rf_raster_batch_workspace_size", &G)) return 0;
    return layout(G, pair_capacity).total;
}

extern "C" int rf_raster_bin_batch(const void* vertexes, const float* bounds,
                                   const float* light_spectrum, const rf_raster_params* prm,
                                   const rf_batch* batch, uint64_t pair_capacity, void* workspace,
                                   size_t workspace_bytes, rf_stream_t stream) {
    Geom G;
    int rc = make_geom_batch(prm, batch, "rf_raster_bin_batch", &G);
    if (rc) return rc;
    if ((G.W > 1 && !light_spectrum) || (G.NE > 0 && (!vertexes || !bounds)))
        return set_error(RF_ERR_INVALID, "rf_raster_bin_batch: NULL buffer");
    RasterPtrs r;
    rc = raster_ptrs(G, pair_capacity, workspace, workspace_bytes, &r, "rf_raster_bin_batch");
    if (rc) return rc;
    return bin_impl((const float4*)vertexes, (const float4*)bounds, light_spectrum, G,
                    pair_capacity, r, (hipStream_t)stream);
}

extern "C" int rf_raster_draw_batch(const void* vertexes, const float* bounds, const float* ghost,
                                    const rf_raster_params* prm, const rf_batch* batch,
                                    uint64_t pair_capacity, void* workspace,
                                    size_t workspace_bytes, float* images, rf_stream_t stream) {
    Geom G;
    int rc = make_geom_batch(prm, batch, "rf_raster_draw_batch", &G);
    if (rc) return rc;
    if (!images || !ghost || (G.NE > 0 && (!vertexes || !bounds)))
        return set_error(RF_ERR_INVALID, "rf_raster_draw_batch: NULL buffer");
    RasterPtrs r;
    rc = raster_ptrs(G, pair_capacity, workspace, workspace_bytes, &r, "rf_raster_draw_batch");
    if (rc) return rc;
    return draw_impl((const float4*)vertexes, (const float4*)bounds, ghost, G, r, (float4*)images,
                     (hipStream_t)stream);
}

extern "C" int rf_raster_status(const void* workspace, uint64_t* required_pairs, int* overflow,
                                rf_stream_t stream) {
    if (!workspace || !required_pairs || !overflow)
        return set_error(RF_ERR_INVALID, "rf_raster_status: NULL");
    Header h;
    hipError_t e = hipMemcpyAsync(&h, workspace, sizeof(Header), hipMemcpyDeviceToHost,
                                  (hipStream_t)stream);
    if (e != hipSuccess) return check_hip(e, "rf_raster_status");
    e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) return check_hip(e, "rf_raster_status");
    *required_pairs = h.required;
    *overflow = (int)h.overflow;
    return RF_OK;
}

// ---------------------------------------------------------------------------
// rf_flare_static: rf_flare whose workspace is a file-scope device array, for
// callers that cannot allocate.  The host refers to the array only through
// the device address hipGetSymbolAddress returns, kept per device (each
// device holds its own copy of the array).

static __device__ __align__(256) unsigned char scratch[RF_FLARE_STATIC_SCRATCH_BYTES];

static constexpr int RF_STATIC_MAX_DEVICES = 64;
static std::mutex g_scratch_mutex;
static void* g_scratch_ptr[RF_STATIC_MAX_DEVICES];

// The current device's address of `scratch`, looked up on its first use on
// that device.
static int static_scratch(void** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return check_hip(e, "rf_flare_static: hipGetDevice");
    if (dev < 0 || dev >= RF_STATIC_MAX_DEVICES)
        return set_error(RF_ERR_INVALID, "rf_flare_static: device %d beyond the %d supported", dev,
                         RF_STATIC_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return check_hip(e, "rf_flare_static: hipGetSymbolAddress");
        if (!p) return set_error(RF_ERR_HIP, "rf_flare_static: no device address for the scratch");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return RF_OK;
}

__global__ void k_flare_status(const Header* __restrict__ h, uint64_t* __restrict__ status) {
    if (threadIdx.x == 0) {
        status[0] = h->required;
        status[1] = h->overflow;
    }
}

// Largest pair capacity whose rf_flare workspace fits `bytes` (-1 if none).
static long static_capacity(const Geom& G, size_t bytes) {
    if (flare_layout(G, 0).total > bytes) return -1;
    uint64_t lo = 0, hi = bytes / 8 + 1;  // a pair takes at least 8 bytes
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo + 1) / 2;
        if (flare_layout(G, mid).total <= bytes) lo = mid;
        else hi = mid - 1;
    }
    return (long)lo;
}

extern "C" int rf_flare_static(const float* planes, const float* ghost, const float* light_spectrum,
                               const rf_raster_params* prm, float* image, uint64_t* status,
                               rf_stream_t stream) {
    Geom G;
    int rc = make_geom(prm, &G);
    if (rc) return rc;
    const long cap = static_capacity(G, RF_FLARE_STATIC_SCRATCH_BYTES);
    if (cap < 0)
        return set_error(RF_ERR_WORKSPACE,
                         "rf_flare_static: the frame needs %zu bytes before any binning pair, "
                         "above the %zu-byte static scratch; use rf_flare with a workspace",
                         flare_layout(G, 0).total, (size_t)RF_FLARE_STATIC_SCRATCH_BYTES);
    if (!image || !ghost || (G.W > 1 && !light_spectrum) || (G.NV > 0 && !planes))
        return set_error(RF_ERR_INVALID, "rf_flare_static: NULL buffer");
    void* ws = nullptr;
    rc = static_scratch(&ws);
    if (rc) return rc;
    rc = rf_flare(planes, ghost, light_spectrum, prm, (uint64_t)cap, ws,
                  RF_FLARE_STATIC_SCRATCH_BYTES, image, stream);
    if (rc) return rc;
    if (status) {
        // the raster header sits at the start of the flare workspace
        hipLaunchKernelGGL(k_flare_status, dim3(1), dim3(64), 0, (hipStream_t)stream,
                           (const Header*)ws, status);
        RF_LAUNCH_CHECK("rf_flare_static");
    }
    return RF_OK;
}
