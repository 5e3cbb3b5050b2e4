/*
 * Oracle traversal entry points.  TEST INFRASTRUCTURE ONLY (see xo.h).
 * Build with: gcc -O2 -std=c11 -ffp-contract=off -fno-fast-math -fopenmp
 *
 * Parity status (tests/test_oracle_trace.py, tests/test_oracle_projection.py,
 * and tests/test_oracle_mutants.py, which plants one mistake at a time and
 * shows the pins below fail under it):
 *   ray() / xo_ray_dir                 pinned by an independent pinhole model
 *                                      (tests/helpers.py): chord field from an
 *                                      eye inside a uniform cube, centre pixel
 *                                      = forward, edge pixel at atan(1/2),
 *                                      lit pixels of one voxel = its forward
 *                                      projection (orientation, pixel corner)
 *   emission coefficient, voxel_ratio  pinned: pixel = c*e*(camera-space chord
 *                                      in [0, voxel_ratio])
 *   unorm8 store                       pinned: Vulkan rule (clamp, *255, round
 *                                      to nearest) on every pixel
 *   DDA (dda.comp:13-53)               pinned: empty/miss closed forms, uniform
 *                                      cube, FP64 brute-force voxel-box integral
 *                                      on 2^3..16^3 and ragged 3x7x5 grids
 *   svo-naive, svo-df, esvo, svo-rope  pinned: the same closed forms and
 *                                      brute-force integrals on the lossless
 *                                      tree, cross-traversal invariant
 *   FP64 twin                          pinned: agrees with FP32 to 2e-5 rel.
 *   word counters                      pinned: hand-derived known answers for
 *                                      all five traversals (a 16-voxel DDA
 *                                      column; a hand-built two-level tree)
 *   xo_iteration_cap                   pinned: never reached on well-formed input
 */
#include "xo.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define XO_STACK_CAP 64
#define XO_MAX_LEVELS 64

static inline uint32_t xo_f2u(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
static inline float xo_u2f(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

/* GLSL ivec3(vec3): truncation toward zero.  Out-of-range and NaN inputs are
 * undefined in GLSL; they saturate (NaN -> 0) as on AMD hardware. */
static inline int32_t xo_f2i(double x) {
    if (x != x)
        return 0;
    if (x >= 2147483647.0)
        return 2147483647;
    if (x <= -2147483648.0)
        return (-2147483647 - 1);
    return (int32_t)x;
}

/* FP32 instantiation */
#define REAL float
#define SFX f32
#define SQRT sqrtf
#define FLOOR floorf
#define LDEXP ldexpf
#define XO_SQRT3 sqrtf(3.0f)
#define XO_PRECISION 32
#include "xo_trace_body.h"
#undef REAL
#undef SFX
#undef SQRT
#undef FLOOR
#undef LDEXP
#undef XO_SQRT3
#undef XO_PRECISION

/* FP64 twin */
#define REAL double
#define SFX f64
#define SQRT sqrt
#define FLOOR floor
#define LDEXP ldexp
#define XO_SQRT3 sqrt(3.0)
#define XO_PRECISION 64
#include "xo_trace_body.h"
#undef REAL
#undef SFX
#undef SQRT
#undef FLOOR
#undef LDEXP
#undef XO_SQRT3
#undef XO_PRECISION

uint64_t xo_iteration_cap(int traversal, const xo_volume* vol, const uint32_t model_dim[3]) {
    if (traversal == XO_DDA) {
        uint64_t s = 0;
        for (int i = 0; i < 3; ++i) {
            uint64_t d = vol->dims[i];
            if (model_dim && model_dim[i] > d)
                d = model_dim[i];
            s += d;
        }
        return 2 * s + 16;
    }
    uint32_t depth = vol->max_depth > 24 ? 24 : vol->max_depth;
    uint64_t side = vol->side;
    if ((1ull << depth) > side)
        side = 1ull << depth;
    return 64 * side + (1ull << 18);
}

/* Quantisation of imageStore(rgba8): clamp to [0,1], scale by 255 and round
 * to nearest even; NaN stores 0. */
static inline uint8_t xo_unorm8(float v) {
    if (!(v > 0.0f))
        return 0;
    if (v >= 1.0f)
        return 255;
#if XO_MUTANT == 6
    return (uint8_t)(v * 255.0f);
#else
    return (uint8_t)rintf(v * 255.0f);
#endif
}
static inline uint8_t xo_unorm8d(double v) {
    if (!(v > 0.0))
        return 0;
    if (v >= 1.0)
        return 255;
#if XO_MUTANT == 6
    return (uint8_t)(v * 255.0);
#else
    return (uint8_t)rint(v * 255.0);
#endif
}

static int xo_check(int traversal, const xo_volume* vol) {
    if (!vol)
        return -1;
    if (traversal == XO_DDA)
        return vol->grid ? 0 : -2;
    if (traversal < 0 || traversal > XO_SVO_ROPE)
        return -3;
    return (vol->nodes && vol->node_count > 0) ? 0 : -2;
}

int xo_render(int traversal, const xo_volume* vol, const xo_camera* cam, const xo_uniforms* u, int precision,
              int threads, uint8_t* out_rgba, float* out_linear, uint64_t* out_words) {
    int rc = xo_check(traversal, vol);
    if (rc)
        return rc;
    if (precision != 32 && precision != 64)
        return -4;
    const uint32_t W = u->out_extent[0], H = u->out_extent[1];
    const uint64_t cap = xo_iteration_cap(traversal, vol, u->model_dim);
#ifdef _OPENMP
    if (threads > 0)
        omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    int64_t rows = (int64_t)H;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t y = 0; y < rows; ++y) {
        for (uint32_t x = 0; x < W; ++x) {
            uint64_t words = 0;
            uint64_t i = (uint64_t)y * W + x;
            float c[3];
            uint8_t q[3];
            if (precision == 32) {
                v3_f32 r = pixel_f32(traversal, vol, cam, u, x, (uint32_t)y, cap, &words);
                c[0] = r.x;
                c[1] = r.y;
                c[2] = r.z;
                q[0] = xo_unorm8(r.x);
                q[1] = xo_unorm8(r.y);
                q[2] = xo_unorm8(r.z);
            } else {
                v3_f64 r = pixel_f64(traversal, vol, cam, u, x, (uint32_t)y, cap, &words);
                c[0] = (float)r.x;
                c[1] = (float)r.y;
                c[2] = (float)r.z;
                q[0] = xo_unorm8d(r.x);
                q[1] = xo_unorm8d(r.y);
                q[2] = xo_unorm8d(r.z);
            }
            if (out_rgba) {
                out_rgba[i * 4 + 0] = q[0];
                out_rgba[i * 4 + 1] = q[1];
                out_rgba[i * 4 + 2] = q[2];
                out_rgba[i * 4 + 3] = 255;
            }
            if (out_linear) {
                out_linear[i * 3 + 0] = c[0];
                out_linear[i * 3 + 1] = c[1];
                out_linear[i * 3 + 2] = c[2];
            }
            if (out_words)
                out_words[i] = words;
        }
    }
    return 0;
}

int xo_trace_ray(int traversal, const xo_volume* vol, const double ro[3], const double rd[3], int precision,
                 double out_rgb[3], uint64_t* out_words) {
    int rc = xo_check(traversal, vol);
    if (rc)
        return rc;
    uint32_t md[3] = {(uint32_t)vol->dims[0], (uint32_t)vol->dims[1], (uint32_t)vol->dims[2]};
    uint64_t cap = xo_iteration_cap(traversal, vol, md);
    uint64_t words = 0;
    if (precision == 32) {
        v3_f32 r = trace_direct_f32(traversal, vol, mk_f32((float)ro[0], (float)ro[1], (float)ro[2]),
                                    mk_f32((float)rd[0], (float)rd[1], (float)rd[2]), cap, &words);
        out_rgb[0] = r.x;
        out_rgb[1] = r.y;
        out_rgb[2] = r.z;
    } else {
        v3_f64 r = trace_direct_f64(traversal, vol, mk_f64(ro[0], ro[1], ro[2]), mk_f64(rd[0], rd[1], rd[2]), cap,
                                    &words);
        out_rgb[0] = r.x;
        out_rgb[1] = r.y;
        out_rgb[2] = r.z;
    }
    if (out_words)
        *out_words = words;
    return 0;
}

void xo_ray_dir(const xo_camera* cam, const xo_uniforms* u, uint32_t px, uint32_t py, float out_rd[3]) {
    int32_t pxl = u->out_offset[0] + (int32_t)px;
    int32_t pyl = u->out_offset[1] + (int32_t)py;
    float uvx = (float)(pxl - u->disp_offset[0]) / (float)u->disp_extent[0];
    float uvy = (float)(pyl - u->disp_offset[1]) / (float)u->disp_extent[1];
    v3_f32 rd = ray_f32(cam, u, uvx, uvy);
    out_rd[0] = rd.x;
    out_rd[1] = rd.y;
    out_rd[2] = rd.z;
}

void xo_free(void* p) { free(p); }
