/*
 * The oracle: a CPU restatement of realflare's OpenCL render kernels
 * (realflare/api/tasks/cl/{raytracing,rasterizing,aperture,starburst,
 * geometry,noise,color}.cl) in float and double.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, the smoke check and the CPU-baseline
 * leg of bench.py load this library, and only as the checker or the timed CPU
 * baseline.  The render path (realflare_amd) never links or calls it.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* LensElement as packed by the host (realflare/api/tasks/opencl.py:18-61). */
typedef struct {
    float radius, ior, height, center;
    int coating;
    int pad_[3];
    float coefficients[8];
} rf_lens_element_t;

_Static_assert(sizeof(rf_lens_element_t) == 64, "LensElement is 64 bytes");

int rfo_threads(void) {
#ifdef _OPENMP
    extern int omp_get_max_threads(void);
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---- float ---- */
#define REAL float
#define SFX _f32
#define SQRT sqrtf
#define SIN sinf
#define COS cosf
#define TAN tanf
#define ASIN asinf
#define ACOS acosf
#define ATAN2 atan2f
#define FABS fabsf
#define FLOOR floorf
#define TRUNC truncf
#define FMOD fmodf
#define FMAX fmaxf
#define FMIN fminf
#define POW powf
#define PI_R ((float)3.14159265358979323846)
#include "rf_oracle_impl.h"
#undef REAL
#undef SFX
#undef SQRT
#undef SIN
#undef COS
#undef TAN
#undef ASIN
#undef ACOS
#undef ATAN2
#undef FABS
#undef FLOOR
#undef TRUNC
#undef FMOD
#undef FMAX
#undef FMIN
#undef POW
#undef PI_R

/* ---- double ---- */
#define REAL double
#define SFX _f64
#define SQRT sqrt
#define SIN sin
#define COS cos
#define TAN tan
#define ASIN asin
#define ACOS acos
#define ATAN2 atan2
#define FABS fabs
#define FLOOR floor
#define TRUNC trunc
#define FMOD fmod
#define FMAX fmax
#define FMIN fmin
#define POW pow
#define PI_R 3.14159265358979323846
#include "rf_oracle_impl.h"
