// 2-D FFT of any size, starburst power spectrum and ghost fractional Fourier
// transform.
//
// The reference runs these on the host in numpy complex128:
//   * StarburstTask.update_fourier_spectrum (realflare/api/tasks/starburst.py:45-81):
//     |fftshift(fft2(A * exp(i pi / (lambda d) (x^2 + y^2))))|^2;
//   * GhostTask.ghost (realflare/api/tasks/ghost.py:13-36) with
//     realflare/utils/frft.py:8-62: |fftshift(frft(fftshift(A), alpha))| * sqrt(w h).
// Here they are complex64 on the device.  The FFT is a radix-2
// decimation-in-time transform held entirely in LDS: a row pass (one
// workgroup per row) and a column pass (a workgroup per group of adjacent
// columns, so every global access reads contiguous runs of a row); a side
// that is not a power of two goes through Bluestein's chirp-z convolution on
// that transform, as numpy (pocketfft) transforms any length.  Twiddles
// and chirp phases are evaluated in double and rounded once to float.
#include <cmath>

#include "primitives.h"

namespace rf {

static constexpr int FFT_THREADS = 256;
static constexpr int FFT_LDS_BYTES = 65536;
static constexpr int FFT_MAX_N = 4096;

static int ilog2(int n) {
    int l = 0;
    while ((1 << l) < n) l++;
    return l;
}
static bool pow2(int n) { return n >= 1 && (n & (n - 1)) == 0; }

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
#pragma clang fp contract(off)
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// tw[k] = exp(-2 pi i k / n), k < n/2
__global__ void k_twiddles(float2* __restrict__ tw, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n / 2) return;
    double s, c;
    sincospi(-2.0 * (double)k / (double)n, &s, &c);
    tw[k] = make_float2((float)c, (float)s);
}

// In-LDS radix-2 transform of `count` sequences of length n = 2^logn stored
// contiguously at s[seq * n], already in bit-reversed order.  Butterfly b of
// a stage: sequence b >> (logn - 1), group and offset from the low bits
// (powers of two throughout, so shifts and masks).
__device__ __forceinline__ void lds_fft(float2* s, int n, int logn, int count,
                                        const float2* __restrict__ tw, bool inverse) {
#pragma clang fp contract(off)
    if (logn == 0) return;
    const int total = count << (logn - 1), hmask = (n >> 1) - 1;
    for (int lh = 0; lh < logn; lh++) {  // half = 2^lh, len = 2^(lh + 1)
        const int half = 1 << lh, tshift = logn - 1 - lh;  // tstride = n / len
        for (int b = threadIdx.x; b < total; b += blockDim.x) {
            const int seq = b >> (logn - 1), r = b & hmask;
            const int grp = r >> lh, k = r & (half - 1);
            const int i0 = (seq << logn) + (grp << (lh + 1)) + k, i1 = i0 + half;
            float2 w = tw[k << tshift];
            if (inverse) w.y = -w.y;
            const float2 a = s[i0];
            const float2 t = cmul(s[i1], w);
            s[i0] = make_float2(a.x + t.x, a.y + t.y);
            s[i1] = make_float2(a.x - t.x, a.y - t.y);
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(FFT_THREADS) void k_fft_rows(float2* __restrict__ data, int n, int logn,
                                                           const float2* __restrict__ tw,
                                                           int inverse, float scale) {
    extern __shared__ float2 s[];
    float2* row = data + (long)blockIdx.x * n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int r = logn ? (int)(__brev((unsigned)i) >> (32 - logn)) : 0;
        s[r] = row[i];
    }
    __syncthreads();
    lds_fft(s, n, logn, 1, tw, inverse != 0);
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        row[i] = make_float2(s[i].x * scale, s[i].y * scale);
}

// columns [c0, c0 + cols) of a (h x w) array; LDS layout s[col * h + row]
__global__ __launch_bounds__(FFT_THREADS) void k_fft_cols(float2* __restrict__ data, int w, int h,
                                                           int logh, int cols,
                                                           const float2* __restrict__ tw,
                                                           int inverse, float scale) {
    extern __shared__ float2 s[];
    const int c0 = blockIdx.x * cols;
    const int total = cols * h;
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        const int row = i / cols, c = i % cols;
        const int r = logh ? (int)(__brev((unsigned)row) >> (32 - logh)) : 0;
        s[c * h + r] = data[(long)row * w + c0 + c];
    }
    __syncthreads();
    lds_fft(s, h, logh, cols, tw, inverse != 0);
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        const int row = i / cols, c = i % cols;
        const float2 v = s[c * h + row];
        data[(long)row * w + c0 + c] = make_float2(v.x * scale, v.y * scale);
    }
}

// A length that is not a power of two is transformed by Bluestein's
// chirp-z identity, jk = (j^2 + k^2 - (k - j)^2) / 2:
//   X_k = a_k sum_j (x_j a_j) b_(k-j),  a_j = exp(-i pi j^2 / n), b = conj(a),
// the sum a circular convolution of length M = the power of two >= 2n - 1,
// done with the radix-2 transform above: FFT, multiply by the transform of b
// wrapped to length M (bhat), inverse FFT.  The inverse transform is the
// conjugate of the forward transform of the conjugate.
struct FftDim {
    int n, M, logM;
    bool blue;
    float2 *tw;         // exp(-2 pi i k / M), k < M / 2
    float2 *a, *bhat;   // Bluestein tables: a[n], bhat[M]
};

struct FftPlan {
    FftDim x, y;
    int cols;       // columns per workgroup of the column pass (over y.M rows)
    float2* tmp;    // Bluestein staging: max(h * x.M, y.M * w) complex
};

static int next_pow2(int n) {
    int m = 1;
    while (m < n) m <<= 1;
    return m;
}

static FftDim fft_dim(int n) {
    FftDim d{};
    d.n = n;
    d.blue = !pow2(n);
    d.M = d.blue ? next_pow2(2 * n - 1) : n;
    d.logM = ilog2(d.M);
    return d;
}

static size_t dim_table_bytes(const FftDim& d) {
    size_t b = align_up((size_t)(d.M / 2 + 1) * sizeof(float2), 256);
    if (d.blue) b += align_up((size_t)d.n * sizeof(float2), 256) +
                     align_up((size_t)d.M * sizeof(float2), 256);
    return b;
}

static size_t tmp_bytes(const FftDim& x, const FftDim& y) {
    size_t n = 0;
    if (x.blue) n = (size_t)y.n * x.M;
    if (y.blue && (size_t)y.M * x.n > n) n = (size_t)y.M * x.n;
    return align_up(n * sizeof(float2), 256);
}

size_t fft_workspace_bytes(int w, int h) {
    if (w < 1 || h < 1) return 0;
    const FftDim x = fft_dim(w), y = fft_dim(h);
    return dim_table_bytes(x) + dim_table_bytes(y) + tmp_bytes(x, y);
}

// a[j] = exp(-i pi j^2 / n) (j^2 reduced mod 2n, exactly, before the sine)
__global__ void k_blue_chirp(float2* __restrict__ a, int n) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const long r = ((long)j * j) % (2L * n);
    double sn, cs;
    sincospi(-(double)r / (double)n, &sn, &cs);
    a[j] = make_float2((float)cs, (float)sn);
}

// b = conj(a) wrapped to length M: b[m] = conj(a[m]), b[M - m] = conj(a[m]),
// zero between
__global__ void k_blue_kernel(const float2* __restrict__ a, int n, int M, float2* __restrict__ b) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    float2 v = make_float2(0.0f, 0.0f);
    if (m < n) v = make_float2(a[m].x, -a[m].y);
    else if (M - m < n) v = make_float2(a[M - m].x, -a[M - m].y);
    b[m] = v;
}

// Rows (axis x): tmp[y][m] = (conj?)(data[y][m]) a[m] for m < n, 0 up to M.
__global__ __launch_bounds__(256) void k_blue_pre_rows(const float2* __restrict__ data, int n,
                                                       int rows, int M, const float2* __restrict__ a,
                                                       int inverse, float2* __restrict__ tmp) {
#pragma clang fp contract(off)
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)rows * M) return;
    const int m = (int)(i % M);
    const long r = i / M;
    float2 v = make_float2(0.0f, 0.0f);
    if (m < n) {
        float2 x = data[r * n + m];
        if (inverse) x.y = -x.y;
        v = cmul(x, a[m]);
    }
    tmp[i] = v;
}

// Columns (axis y): tmp[m][c] = (conj?)(data[m][c]) a[m] for m < n, 0 up to M.
__global__ __launch_bounds__(256) void k_blue_pre_cols(const float2* __restrict__ data, int w, int n,
                                                       int M, const float2* __restrict__ a,
                                                       int inverse, float2* __restrict__ tmp) {
#pragma clang fp contract(off)
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)M * w) return;
    const int m = (int)(i / w);
    float2 v = make_float2(0.0f, 0.0f);
    if (m < n) {
        float2 x = data[i];
        if (inverse) x.y = -x.y;
        v = cmul(x, a[m]);
    }
    tmp[i] = v;
}

// tmp *= bhat along the transformed axis (rows: index i % M; columns: i / w)
__global__ __launch_bounds__(256) void k_blue_mul(float2* __restrict__ tmp, long total, int M, int w,
                                                  int along_rows, const float2* __restrict__ bhat) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= total) return;
    const int m = along_rows ? (int)(i % M) : (int)(i / w);
    tmp[i] = cmul(tmp[i], bhat[m]);
}

// data[y][k] = (conj?)(a[k] tmp[y][k]) * scale, k < n (tmp holds M times the
// circular convolution: the inverse transform is unnormalized)
__global__ __launch_bounds__(256) void k_blue_post_rows(const float2* __restrict__ tmp, int n, int rows,
                                                        int M, const float2* __restrict__ a,
                                                        int inverse, float scale,
                                                        float2* __restrict__ data) {
#pragma clang fp contract(off)
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)rows * n) return;
    const int k = (int)(i % n);
    const long r = i / n;
    float2 v = cmul(a[k], tmp[r * M + k]);
    if (inverse) v.y = -v.y;
    data[i] = make_float2(v.x * scale, v.y * scale);
}

__global__ __launch_bounds__(256) void k_blue_post_cols(const float2* __restrict__ tmp, int w, int n,
                                                        const float2* __restrict__ a, int inverse,
                                                        float scale, float2* __restrict__ data) {
#pragma clang fp contract(off)
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)n * w) return;
    const int k = (int)(i / w);
    float2 v = cmul(a[k], tmp[i]);
    if (inverse) v.y = -v.y;
    data[i] = make_float2(v.x * scale, v.y * scale);
}

static inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

static int cols_for(int M, int w) {
    int cols = FFT_LDS_BYTES / (int)(M * sizeof(float2));
    if (cols > 16) cols = 16;
    if (cols > w) cols = w;
    while (cols > 1 && w % cols) cols--;  // whole groups of columns
    return cols < 1 ? 1 : cols;
}

static void run_rows_pow2(const FftDim& d, float2* data, int rows, bool inverse, float scale,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_fft_rows, dim3(rows), dim3(FFT_THREADS), d.M * sizeof(float2), s, data, d.M,
                       d.logM, d.tw, (int)inverse, scale);
}

static void run_cols_pow2(const FftDim& d, float2* data, int w, bool inverse, float scale,
                          hipStream_t s) {
    const int cols = cols_for(d.M, w);
    hipLaunchKernelGGL(k_fft_cols, dim3(w / cols), dim3(FFT_THREADS), cols * d.M * sizeof(float2), s,
                       data, w, d.M, d.logM, cols, d.tw, (int)inverse, scale);
}

static int carve_dim(FftDim* d, char** w, hipStream_t s) {
    d->tw = (float2*)*w;
    *w += align_up((size_t)(d->M / 2 + 1) * sizeof(float2), 256);
    if (d->M >= 2)
        hipLaunchKernelGGL(k_twiddles, dim3((d->M / 2 + 255) / 256), dim3(256), 0, s, d->tw, d->M);
    d->a = d->bhat = nullptr;
    if (d->blue) {
        d->a = (float2*)*w;
        *w += align_up((size_t)d->n * sizeof(float2), 256);
        d->bhat = (float2*)*w;
        *w += align_up((size_t)d->M * sizeof(float2), 256);
        hipLaunchKernelGGL(k_blue_chirp, dim3(nblk(d->n)), dim3(256), 0, s, d->a, d->n);
        hipLaunchKernelGGL(k_blue_kernel, dim3(nblk(d->M)), dim3(256), 0, s, d->a, d->n, d->M,
                           d->bhat);
        run_rows_pow2(*d, d->bhat, 1, false, 1.0f, s);  // bhat = FFT_M(b)
    }
    return RF_OK;
}

static int fft_plan(int w, int h, void* ws, hipStream_t s, FftPlan* p) {
    if (w < 1 || h < 1)
        return set_error(RF_ERR_INVALID, "fft: empty size %dx%d", w, h);
    p->x = fft_dim(w);
    p->y = fft_dim(h);
    if (p->x.M > FFT_MAX_N || p->y.M > FFT_MAX_N)
        return set_error(RF_ERR_INVALID,
                         "fft: a side must be a power of two up to %d or any length up to %d "
                         "(got %dx%d)",
                         FFT_MAX_N, FFT_MAX_N / 2, w, h);
    char* c = (char*)ws;
    carve_dim(&p->x, &c, s);
    carve_dim(&p->y, &c, s);
    p->tmp = (float2*)c;
    p->cols = cols_for(p->y.M, w);
    RF_LAUNCH_CHECK("fft plan");
    return RF_OK;
}

// unnormalized transform times `scale` (numpy: forward scale 1, inverse 1/N)
static int fft_run(const FftPlan& p, float2* data, bool inverse, float scale, hipStream_t s) {
    const int w = p.x.n, h = p.y.n;
    if (!p.x.blue) {
        run_rows_pow2(p.x, data, h, inverse, scale, s);
    } else {
        const FftDim& d = p.x;
        const long total = (long)h * d.M;
        hipLaunchKernelGGL(k_blue_pre_rows, dim3(nblk(total)), dim3(256), 0, s, data, w, h, d.M, d.a,
                           (int)inverse, p.tmp);
        run_rows_pow2(d, p.tmp, h, false, 1.0f, s);
        hipLaunchKernelGGL(k_blue_mul, dim3(nblk(total)), dim3(256), 0, s, p.tmp, total, d.M, 0, 1,
                           d.bhat);
        run_rows_pow2(d, p.tmp, h, true, 1.0f / (float)d.M, s);
        hipLaunchKernelGGL(k_blue_post_rows, dim3(nblk((long)h * w)), dim3(256), 0, s, p.tmp, w, h,
                           d.M, d.a, (int)inverse, scale, data);
    }
    if (!p.y.blue) {
        run_cols_pow2(p.y, data, w, inverse, 1.0f, s);
    } else {
        const FftDim& d = p.y;
        const long total = (long)d.M * w;
        hipLaunchKernelGGL(k_blue_pre_cols, dim3(nblk(total)), dim3(256), 0, s, data, w, h, d.M, d.a,
                           (int)inverse, p.tmp);
        run_cols_pow2(d, p.tmp, w, false, 1.0f, s);
        hipLaunchKernelGGL(k_blue_mul, dim3(nblk(total)), dim3(256), 0, s, p.tmp, total, d.M, w, 0,
                           d.bhat);
        run_cols_pow2(d, p.tmp, w, true, 1.0f / (float)d.M, s);
        hipLaunchKernelGGL(k_blue_post_cols, dim3(nblk((long)h * w)), dim3(256), 0, s, p.tmp, w, h,
                           d.a, (int)inverse, 1.0f, data);
    }
    RF_LAUNCH_CHECK("fft");
    return RF_OK;
}

// ---------------------------------------------------------------------------
// starburst power spectrum

// c = A * exp(i pi / (lambda d) (x^2 + y^2)) on linspace(-1, 1) grids
__global__ __launch_bounds__(256) void k_starburst_field(const float* __restrict__ ap, int w, int h,
                                                         double k, float2* __restrict__ out) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)w * h) return;
    const int x = (int)(i % w), y = (int)(i / w);
    const double xs = w > 1 ? -1.0 + (2.0 / (w - 1)) * x : -1.0;
    const double ys = h > 1 ? -1.0 + (2.0 / (h - 1)) * y : -1.0;
    const double xv = (w > 1 && x == w - 1) ? 1.0 : xs;  // numpy linspace ends exactly at 1
    const double yv = (h > 1 && y == h - 1) ? 1.0 : ys;
    double sn, cs;
    sincos(k * (xv * xv + yv * yv), &sn, &cs);
    const double a = (double)ap[i];
    out[i] = make_float2((float)(a * cs), (float)(a * sn));
}

// out = |fftshift(F)|^2 (squares summed in double, then rounded to float);
// fftshift moves element (i - n // 2) mod n to i, for odd n as well
__global__ __launch_bounds__(256) void k_power_shift(const float2* __restrict__ F, int w, int h,
                                                     float* __restrict__ out) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)w * h) return;
    const int x = (int)(i % w), y = (int)(i / w);
    const int sx = (x + (w + 1) / 2) % w, sy = (y + (h + 1) / 2) % h;
    const float2 v = F[(long)sy * w + sx];
    out[i] = (float)((double)v.x * v.x + (double)v.y * v.y);
}

// ---------------------------------------------------------------------------
// ghost: fractional Fourier transform (frft.py)

// in: real image; out: complex fftshift(image)
__global__ __launch_bounds__(256) void k_shift_real(const float* __restrict__ in, int w, int h,
                                                    float2* __restrict__ out) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)w * h) return;
    const int x = (int)(i % w), y = (int)(i / w);
    const int sx = (x + (w + 1) / 2) % w, sy = (y + (h + 1) / 2) % h;
    out[i] = make_float2(in[(long)sy * w + sx], 0.0f);
}

__global__ __launch_bounds__(256) void k_flip(const float2* __restrict__ in, int w, int h,
                                              float2* __restrict__ out) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)w * h) return;
    const int x = (int)(i % w), y = (int)(i / w);
    out[i] = in[(long)(h - 1 - y) * w + (w - 1 - x)];
}

// chirp phase of frft.py:8-17 at (x, y): pi * (ky^2 / h + kx^2 / w) with k
// the element of fftshift(arange(-n // 2, n // 2)) (floor division), which is
// x for x < n // 2 and x - n from there (for odd n the grid runs from
// -(n + 1) / 2, so it is not the FFT-order frequency index)
__device__ __forceinline__ double chirp_phase(int x, int y, int w, int h) {
    const double kx = x < w / 2 ? x : x - w;
    const double ky = y < h / 2 ? y : y - h;
    return 3.14159265358979323846 * (ky * ky / h + kx * kx / w);
}

// f2 = chirp1 * x, chirp2 in c2 (both complex64)
__global__ __launch_bounds__(256) void k_frft_pre(const float2* __restrict__ x, int w, int h,
                                                  double c1, double c2, float2* __restrict__ f2,
                                                  float2* __restrict__ ch2) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)w * h) return;
    const double ph = chirp_phase((int)(i % w), (int)(i / w), w, h);
    double s1, k1, s2, k2;
    sincos(ph * c1, &s1, &k1);
    sincos(ph * c2, &s2, &k2);
    const double xr = x[i].x, xi = x[i].y;
    f2[i] = make_float2((float)(k1 * xr - s1 * xi), (float)(k1 * xi + s1 * xr));
    ch2[i] = make_float2((float)k2, (float)s2);
}

__global__ __launch_bounds__(256) void k_cmul_inplace(float2* __restrict__ a,
                                                      const float2* __restrict__ b, long n) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) a[i] = cmul(a[i], b[i]);
}

// out = |fftshift(scale * chirp1 * y)| * amp
__global__ __launch_bounds__(256) void k_frft_post(const float2* __restrict__ y, int w, int h,
                                                   double c1, double sr, double si, double amp,
                                                   float* __restrict__ out) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)w * h) return;
    const int x = (int)(i % w), yy = (int)(i / w);
    const int sx = (x + (w + 1) / 2) % w, sy = (yy + (h + 1) / 2) % h;
    const long j = (long)sy * w + sx;
    double s1, k1;
    sincos(chirp_phase(sx, sy, w, h) * c1, &s1, &k1);
    const double ar = sr * k1 - si * s1, ai = sr * s1 + si * k1;  // scale * chirp1
    const double vr = y[j].x, vi = y[j].y;
    const double re = ar * vr - ai * vi, im = ar * vi + ai * vr;
    out[i] = (float)sqrt(re * re + im * im) * (float)amp;
}

}  // namespace rf

using namespace rf;

extern "C" size_t rf_fft_workspace_size(int width, int height) {
    return fft_workspace_bytes(width, height);
}

extern "C" int rf_fft2(void* data, int width, int height, int inverse, float scale, void* workspace,
                       size_t workspace_bytes, rf_stream_t stream) {
    if (!data) return set_error(RF_ERR_INVALID, "rf_fft2: NULL data");
    if (!workspace || workspace_bytes < fft_workspace_bytes(width, height))
        return set_error(RF_ERR_WORKSPACE, "rf_fft2: workspace too small");
    FftPlan p;
    int rc = fft_plan(width, height, workspace, (hipStream_t)stream, &p);
    if (rc) return rc;
    return fft_run(p, (float2*)data, inverse != 0, scale, (hipStream_t)stream);
}

extern "C" size_t rf_power_spectrum_workspace_size(int width, int height) {
    return align_up((size_t)width * height * sizeof(float2), 256) + fft_workspace_bytes(width, height);
}

extern "C" int rf_power_spectrum(const float* aperture, int width, int height, float distance,
                                 void* workspace, size_t workspace_bytes, float* spectrum,
                                 rf_stream_t stream) {
    if (!aperture || !spectrum) return set_error(RF_ERR_INVALID, "rf_power_spectrum: NULL");
    if (!workspace || workspace_bytes < rf_power_spectrum_workspace_size(width, height))
        return set_error(RF_ERR_WORKSPACE, "rf_power_spectrum: workspace too small");
    hipStream_t s = (hipStream_t)stream;
    float2* field = (float2*)workspace;
    FftPlan p;
    int rc = fft_plan(width, height,
                      (char*)workspace + align_up((size_t)width * height * sizeof(float2), 256), s,
                      &p);
    if (rc) return rc;
    // starburst.py:60-63: lambda = 560 nm in mm, d = max(1e-9, distance * 1e3) mm
    const double wavelength = ((390.0 + 730.0) / 2.0) * 1e-6;
    double d = (double)distance * 1e3;
    if (d < 1e-9) d = 1e-9;
    const double k = 3.14159265358979323846 / (wavelength * d);
    const long n = (long)width * height;
    hipLaunchKernelGGL(k_starburst_field, dim3(nblk(n)), dim3(256), 0, s, aperture, width, height, k,
                       field);
    rc = fft_run(p, field, false, 1.0f, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_power_shift, dim3(nblk(n)), dim3(256), 0, s, field, width, height, spectrum);
    RF_LAUNCH_CHECK("rf_power_spectrum");
    return RF_OK;
}

extern "C" size_t rf_ghost_workspace_size(int width, int height) {
    return 3 * align_up((size_t)width * height * sizeof(float2), 256) +
           fft_workspace_bytes(width, height);
}

extern "C" int rf_ghost(const float* aperture, int width, int height, float fstop,
                        void* workspace, size_t workspace_bytes, float* ghost,
                        rf_stream_t stream) {
    if (!aperture || !ghost) return set_error(RF_ERR_INVALID, "rf_ghost: NULL");
    if (!workspace || workspace_bytes < rf_ghost_workspace_size(width, height))
        return set_error(RF_ERR_WORKSPACE, "rf_ghost: workspace too small");
    hipStream_t s = (hipStream_t)stream;
    const long n = (long)width * height;
    const size_t plane = align_up((size_t)n * sizeof(float2), 256);
    float2* a = (float2*)workspace;
    float2* b = (float2*)((char*)workspace + plane);
    float2* c = (float2*)((char*)workspace + 2 * plane);
    FftPlan p;
    int rc = fft_plan(width, height, (char*)workspace + 3 * plane, s, &p);
    if (rc) return rc;
    const float ortho = (float)(1.0 / sqrt((double)n));

    // ghost.py:24: alpha = 0.15 * (lambda / 400) * (fstop / 18), lambda = 560
    double alpha = 0.15 * (560.0 / 400.0) * ((double)fstop / 18.0);
    hipLaunchKernelGGL(k_shift_real, dim3(nblk(n)), dim3(256), 0, s, aperture, width, height, a);
    // frft.py:20-40: bring alpha into [0.5, 1.5)
    double rel = fmod(alpha, 4.0);
    if (rel < 0) rel += 4.0;
    float2* x = a;
    if (rel < 0.5) {
        alpha = rel + 1;
        rc = fft_run(p, a, true, ortho, s);
    } else if (rel < 1.5) {
        alpha = rel;
    } else if (rel < 2.5) {
        alpha = rel - 1;
        rc = fft_run(p, a, false, ortho, s);
    } else if (rel < 3.5) {
        alpha = rel - 2;
        hipLaunchKernelGGL(k_flip, dim3(nblk(n)), dim3(256), 0, s, a, width, height, b);
        x = b;
    } else {
        alpha = rel - 3;
        rc = fft_run(p, a, true, ortho, s);
    }
    if (rc) return rc;
    // frft.py:43-62
    const double phi = alpha * 3.14159265358979323846 / 2.0;
    const double cot = 1.0 / tan(phi);
    const double sq = sqrt(1.0 + cot * cot);
    // scale = sqrt(1 - i cot) / sqrt(N): principal square root
    const double mod = sqrt(sqrt(1.0 + cot * cot));
    const double arg = 0.5 * atan2(-cot, 1.0);
    const double sr = mod * cos(arg) / sqrt((double)n), si = mod * sin(arg) / sqrt((double)n);
    float2* f2 = (x == a) ? b : a;
    hipLaunchKernelGGL(k_frft_pre, dim3(nblk(n)), dim3(256), 0, s, x, width, height, cot - sq, sq, f2,
                       c);
    rc = fft_run(p, c, false, ortho, s);  // fft1 = fftn(chirp2, ortho)
    if (rc) return rc;
    rc = fft_run(p, f2, false, ortho, s);  // fft2 = fftn(chirp1 * x, ortho)
    if (rc) return rc;
    hipLaunchKernelGGL(k_cmul_inplace, dim3(nblk(n)), dim3(256), 0, s, f2, c, n);
    rc = fft_run(p, f2, true, ortho, s);  // ifftn(fft1 * fft2, ortho)
    if (rc) return rc;
    // ghost.py:27-29: fftshift, abs, float32, * sqrt(w h)
    hipLaunchKernelGGL(k_frft_post, dim3(nblk(n)), dim3(256), 0, s, f2, width, height, cot - sq, sr,
                       si, sqrt((double)width * height), ghost);
    RF_LAUNCH_CHECK("rf_ghost");
    return RF_OK;
}
