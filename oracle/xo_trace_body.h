/*
 * Oracle traversal body, instantiated twice by xo_trace.c:
 *   REAL=float  SFX=f32   -- restatement of the GLSL shaders in FP32
 *   REAL=double SFX=f64   -- FP64 twin of the same algorithms (adjudicator)
 *
 * TEST INFRASTRUCTURE ONLY (see xo.h).
 *
 * GLSL semantics followed (GLSL 4.50 spec definitions):
 *   min(x,y) = y < x ? y : x;   max(x,y) = x < y ? y : x
 *   dot(a,b) = a.x*b.x + a.y*b.y + a.z*b.z (left to right, no fusion)
 *   cross(a,b) = (a.y*b.z - b.y*a.z, a.z*b.x - b.z*a.x, a.x*b.y - b.x*a.y)
 *   normalize(v) = v / length(v), length(v) = sqrt(dot(v,v))
 *   mod(x,y) = x - y*floor(x/y);  sign(x) in {-1,0,1}
 *   unorm8 -> float: c / 255 (correctly rounded)
 * Expressions are evaluated left to right as written in the shader, one
 * rounding per operation.
 *
 * XO_MUTANT=k (0, the default, is the faithful oracle) plants one deliberate
 * mistake, so that the tests can show each pin of the oracle fails under the
 * mistake it is meant to catch (tests/test_oracle_mutants.py):
 *   1  right = cross(dir, up): the image mirrored left-right
 *   2  uv at the pixel centre (+0.5) instead of its corner
 *   3  aspect applied as uv.x *= W/H instead of uv.y *= H/W
 *   4  up = cross(dir, right): the image upside down
 *   5  emission coefficient with voxel_ratio instead of voxel_ratio^2
 *   6  the unorm8 store truncates instead of rounding to nearest
 *   7  ray() does not divide the direction by voxel_ratio
 *   8  the svo-df pop does not count its is_leaf_depth word
 *   9  esvo does not count the leaf colour word
 *   10 horizontal field of view 60 degrees instead of 2*atan(1/2)
 */
#ifndef XO_MUTANT
#define XO_MUTANT 0
#endif

#define XO_CAT2(a, b) a##_##b
#define XO_CAT(a, b) XO_CAT2(a, b)
#define V3 XO_CAT(v3, SFX)
#define F(name) XO_CAT(name, SFX)

typedef struct { REAL x, y, z; } V3;

static inline V3 F(mk)(REAL x, REAL y, REAL z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline V3 F(splat)(REAL s) { return F(mk)(s, s, s); }
static inline V3 F(add)(V3 a, V3 b) { return F(mk)(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 F(sub)(V3 a, V3 b) { return F(mk)(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 F(mul)(V3 a, V3 b) { return F(mk)(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline V3 F(divv)(V3 a, V3 b) { return F(mk)(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline V3 F(muls)(V3 a, REAL s) { return F(mk)(a.x * s, a.y * s, a.z * s); }
static inline V3 F(smul)(REAL s, V3 a) { return F(mk)(s * a.x, s * a.y, s * a.z); }
static inline V3 F(neg)(V3 a) { return F(mk)(-a.x, -a.y, -a.z); }
static inline REAL F(gmin)(REAL x, REAL y) { return y < x ? y : x; }
static inline REAL F(gmax)(REAL x, REAL y) { return x < y ? y : x; }
static inline V3 F(vmin)(V3 a, V3 b) { return F(mk)(F(gmin)(a.x, b.x), F(gmin)(a.y, b.y), F(gmin)(a.z, b.z)); }
static inline V3 F(vmax)(V3 a, V3 b) { return F(mk)(F(gmax)(a.x, b.x), F(gmax)(a.y, b.y), F(gmax)(a.z, b.z)); }
/* common.glsl:58-64 */
static inline REAL F(min_elem)(V3 v) { return F(gmin)(v.x, F(gmin)(v.y, v.z)); }
static inline REAL F(max_elem)(V3 v) { return F(gmax)(v.x, F(gmax)(v.y, v.z)); }
static inline REAL F(dot)(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 F(cross)(V3 a, V3 b) {
    return F(mk)(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
static inline V3 F(normalize)(V3 v) {
    REAL len = SQRT(F(dot)(v, v));
    return F(mk)(v.x / len, v.y / len, v.z / len);
}
static inline REAL F(rabs)(REAL x) { return x < 0 ? -x : x; }
static inline REAL F(rsign)(REAL x) { return x > 0 ? (REAL)1 : (x < 0 ? (REAL)-1 : (REAL)0); }
static inline REAL F(gmod)(REAL x, REAL y) { return x - y * FLOOR(x / y); }
static inline V3 F(vload)(const float* p) { return F(mk)((REAL)p[0], (REAL)p[1], (REAL)p[2]); }

/* unpackUnorm4x8(c).rgb (octree colours) and texelFetch().rgb of an RGBA8
 * UNORM texel: each channel c / 255. */
static inline V3 F(unorm_rgb)(uint32_t packed) {
    return F(mk)((REAL)(packed & 0xFFu) / (REAL)255,
                 (REAL)((packed >> 8) & 0xFFu) / (REAL)255,
                 (REAL)((packed >> 16) & 0xFFu) / (REAL)255);
}

/* common.glsl:40-43 adjust_ray: components with |rd| < 2^-23 become +2^-23. */
static inline V3 F(adjust_ray)(V3 rd) {
    const REAL eps = (REAL)(1.0 / 8388608.0);
    V3 r;
    r.x = F(rabs)(rd.x) < eps ? eps : rd.x;
    r.y = F(rabs)(rd.y) < eps ? eps : rd.y;
    r.z = F(rabs)(rd.z) < eps ? eps : rd.z;
    return r;
}

/* common.glsl:45-56 ray(uv).  uv is computed as in every main():
 * vec2(pixel - display_region.offset) / vec2(display_region.extent). */
static V3 F(ray)(const xo_camera* cam, const xo_uniforms* u, REAL uvx, REAL uvy) {
    uvx = uvx - (REAL)0.5;
    uvy = uvy - (REAL)0.5;
#if XO_MUTANT == 3
    uvx = uvx * ((REAL)u->disp_extent[0] / (REAL)u->disp_extent[1]);
#else
    uvy = uvy * ((REAL)u->disp_extent[1] / (REAL)u->disp_extent[0]);
#endif
#if XO_MUTANT == 10
    uvx = uvx * (REAL)1.1547005383792515; /* tan(30 deg) / (1/2) */
    uvy = uvy * (REAL)1.1547005383792515;
#endif
    V3 dir = F(vload)(cam->forward);
    V3 up = F(vload)(cam->up);
#if XO_MUTANT == 1
    V3 right = F(normalize)(F(cross)(dir, up));
#else
    V3 right = F(normalize)(F(cross)(up, dir));
#endif
#if XO_MUTANT == 4
    up = F(normalize)(F(cross)(dir, right));
#else
    up = F(normalize)(F(cross)(right, dir));
#endif
    V3 rd = F(normalize)(F(add)(F(add)(F(smul)(uvx, right), F(smul)(uvy, up)), dir));
#if XO_MUTANT == 7
    return F(adjust_ray)(F(normalize)(rd));
#else
    return F(adjust_ray)(F(normalize)(F(divv)(rd, F(vload)(u->voxel_ratio))));
#endif
}

/* common.glsl:68-72 voxel_emission_coeff */
static REAL F(emission_coeff)(const xo_uniforms* u, V3 rd) {
    V3 rd2 = F(mul)(rd, rd);
    V3 ratio = F(vload)(u->voxel_ratio);
#if XO_MUTANT == 5
    V3 dim2 = ratio;
#else
    V3 dim2 = F(mul)(ratio, ratio);
#endif
    return (REAL)u->emission_coeff * SQRT(F(dot)(rd2, dim2) / F(dot)(rd2, F(splat)(1)));
}

/* ------------------------------------------------------------------ */
/* DDA: resources/dda.comp:13-53                                       */
/* ------------------------------------------------------------------ */

static inline V3 F(texel)(const xo_volume* vol, int32_t px, int32_t py, int32_t pz) {
    /* texelFetch outside the image is undefined in Vulkan; with robust
     * image access it returns zero.  The restatement defines it as zero. */
    if (px < 0 || py < 0 || pz < 0 || (uint64_t)px >= vol->dims[0] || (uint64_t)py >= vol->dims[1] ||
        (uint64_t)pz >= vol->dims[2])
        return F(splat)(0);
    uint64_t i = ((uint64_t)pz * vol->dims[1] + (uint64_t)py) * vol->dims[0] + (uint64_t)px;
    const uint8_t* p = vol->grid + i * 4;
    return F(unorm_rgb)((uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16));
}

static V3 F(dda_trace)(const xo_volume* vol, const REAL model_dim[3], V3 ro, V3 rd, uint64_t cap, uint64_t* words) {
    V3 rrd = F(divv)(F(splat)(1), rd);
    V3 bias = F(mul)(rrd, ro);
    V3 box_min = F(neg)(bias);
    V3 box_max = F(sub)(F(mul)(F(mk)(model_dim[0], model_dim[1], model_dim[2]), rrd), bias);
    REAL t_min = F(max_elem)(F(vmin)(box_min, box_max));
    REAL t_max = F(min_elem)(F(vmax)(box_min, box_max));
    if (t_min > t_max)
        return F(splat)(0);
    t_min = F(gmax)(t_min, 0);
    ro = F(add)(ro, F(muls)(rd, t_min));
    int32_t pos[3] = {xo_f2i((double)ro.x), xo_f2i((double)ro.y), xo_f2i((double)ro.z)};
    V3 t_delta = F(mk)(F(rabs)(rrd.x), F(rabs)(rrd.y), F(rabs)(rrd.z));
    V3 sgn = F(mk)(F(rsign)(rd.x), F(rsign)(rd.y), F(rsign)(rd.z));
    int32_t step[3] = {(int32_t)sgn.x, (int32_t)sgn.y, (int32_t)sgn.z};
    V3 fl = F(mk)(FLOOR(ro.x), FLOOR(ro.y), FLOOR(ro.z));
    V3 side_dist = F(mul)(F(add)(F(mul)(sgn, F(add)(F(sub)(fl, ro), F(splat)((REAL)0.5))), F(splat)((REAL)0.5)), t_delta);
    REAL t = 0;
    V3 total = F(splat)(0);
    const REAL limit = t_max - t_min;
    uint64_t it = 0;
    while (t < limit) {
        if (++it > cap)
            break;
        int mx = side_dist.x <= F(gmin)(side_dist.y, side_dist.z);
        int my = side_dist.y <= F(gmin)(side_dist.z, side_dist.x);
        int mz = side_dist.z <= F(gmin)(side_dist.x, side_dist.y);
        REAL t0 = F(min_elem)(side_dist);
        V3 c = F(texel)(vol, pos[0], pos[1], pos[2]);
        ++*words;
        REAL dt = t0 - t;
        total = F(add)(total, F(muls)(c, dt));
        t = t0;
        side_dist = F(add)(side_dist, F(mk)(mx ? t_delta.x : 0, my ? t_delta.y : 0, mz ? t_delta.z : 0));
        pos[0] += mx ? step[0] : 0;
        pos[1] += my ? step[1] : 0;
        pos[2] += mz ? step[2] : 0;
    }
    return total;
}

/* ------------------------------------------------------------------ */
/* Octree access: resources/octree.glsl:6-17                           */
/* ------------------------------------------------------------------ */

#define XO_NODE_WORDS 10u
#define XO_LEAF_MASK 0x80000000u
#define XO_DEPTH_MASK 0x7FFFFFFFu

static inline uint32_t F(child_of)(const xo_volume* v, uint32_t n, uint32_t c, uint64_t* w) {
    ++*w;
    return v->nodes[(uint64_t)n * XO_NODE_WORDS + c];
}
static inline uint32_t F(color_of)(const xo_volume* v, uint32_t n, uint64_t* w) {
    ++*w;
    return v->nodes[(uint64_t)n * XO_NODE_WORDS + 8];
}
static inline uint32_t F(leafdepth_of)(const xo_volume* v, uint32_t n, uint64_t* w) {
    ++*w;
    return v->nodes[(uint64_t)n * XO_NODE_WORDS + 9];
}

/* svo_naive.comp:8-27 find() (svo_rope.comp:8-27 is the same function). */
static uint32_t F(svo_find)(const xo_volume* v, V3 pos, V3* base, REAL* side, uint64_t* w) {
    REAL extent = 1;
    uint32_t index = 0;
    V3 offset = F(splat)(0);
    for (uint32_t level = 0;; ++level) {
        if (F(leafdepth_of)(v, index, w) >= XO_LEAF_MASK || level > XO_MAX_LEVELS) {
            *base = offset;
            *side = extent;
            return index;
        }
        extent = extent * (REAL)0.5;
        int mx = pos.x >= offset.x + extent;
        int my = pos.y >= offset.y + extent;
        int mz = pos.z >= offset.z + extent;
        uint32_t child = (uint32_t)(mx * 4 + my * 2 + mz);
        offset = F(add)(offset, F(mul)(F(mk)((REAL)mx, (REAL)my, (REAL)mz), F(splat)(extent)));
        index = F(child_of)(v, index, child, w);
    }
}

/* svo_naive.comp:29-71 */
static V3 F(naive_trace)(const xo_volume* v, V3 ro, V3 rd, uint64_t cap, uint64_t* w) {
    const REAL min_step = (REAL)0.00001f;
    V3 rrd = F(divv)(F(splat)(1), rd);
    V3 bias = F(mul)(rrd, ro);
    V3 box_min = F(neg)(bias);
    V3 box_max = F(sub)(rrd, bias);
    REAL t_min = F(max_elem)(F(vmin)(box_min, box_max));
    REAL t_max = F(min_elem)(F(vmax)(box_min, box_max));
    if (t_min > t_max)
        return F(splat)(0);
    t_min = F(gmax)(t_min, 0);
    V3 total = F(splat)(0);
    REAL t = t_min + min_step;
    uint64_t it = 0;
    while (t < t_max) {
        if (++it > cap)
            break;
        V3 p = F(add)(F(smul)(t, rd), ro);
        V3 offset;
        REAL side;
        uint32_t node = F(svo_find)(v, p, &offset, &side, w);
        V3 node_min = F(sub)(F(mul)(offset, rrd), bias);
        V3 node_max = F(sub)(F(mul)(F(add)(offset, F(splat)(side)), rrd), bias);
        REAL u_min = F(max_elem)(F(vmin)(node_min, node_max));
        REAL u_max = F(min_elem)(F(vmax)(node_min, node_max));
        u_min = F(gmax)(u_min, 0);
        REAL step = F(gmax)(u_max - u_min, min_step);
        t = t + step;
        V3 color = F(unorm_rgb)(F(color_of)(v, node, w));
        total = F(add)(total, F(muls)(color, step));
    }
    return total;
}

/* svo_df.comp:6-66 */
static V3 F(df_trace)(const xo_volume* v, V3 ro, V3 rd, uint64_t cap, uint64_t* w) {
    V3 rrd = F(divv)(F(splat)(1), rd);
    V3 bias = F(mul)(rrd, ro);
    int sp = 0;
    uint32_t node_stack[XO_STACK_CAP];
    uint32_t child_index_stack[XO_STACK_CAP];
    uint32_t node = 0;
    uint32_t child_idx = 0;
    V3 pos = F(splat)(0);
    REAL side = (REAL)0.5;
    V3 total = F(splat)(0);
    uint64_t it = 0;
    for (;;) {
        if (++it > cap)
            break;
        uint32_t child = F(child_of)(v, node, child_idx, w);
        V3 box_min = F(sub)(F(mul)(pos, rrd), bias);
        V3 box_max = F(sub)(F(mul)(F(add)(pos, F(splat)(side)), rrd), bias);
        REAL t_min = F(max_elem)(F(vmin)(box_min, box_max));
        REAL t_max = F(min_elem)(F(vmax)(box_min, box_max));
        if (t_min < t_max && t_max > 0) {
            if (F(leafdepth_of)(v, child, w) >= XO_LEAF_MASK) {
                V3 color = F(unorm_rgb)(F(color_of)(v, child, w));
                total = F(add)(total, F(muls)(color, t_max - F(gmax)(t_min, 0)));
            } else if (child_idx == 7 || sp < XO_STACK_CAP) {
                if (child_idx != 7) {
                    node_stack[sp] = node;
                    child_index_stack[sp] = child_idx;
                    ++sp;
                }
                side = side * (REAL)0.5;
                node = child;
                child_idx = 0;
                continue;
            }
        }
        if (child_idx == 7) {
            --sp;
            if (sp < 0)
                break;
            node = node_stack[sp];
            child_idx = child_index_stack[sp];
#if XO_MUTANT == 8
            uint64_t uncounted = 0;
            side = LDEXP((REAL)1, -(int)(F(leafdepth_of)(v, node, &uncounted) & XO_DEPTH_MASK)) * (REAL)0.5;
#else
            side = LDEXP((REAL)1, -(int)(F(leafdepth_of)(v, node, w) & XO_DEPTH_MASK)) * (REAL)0.5;
#endif
        }
        REAL s2 = side * (REAL)2.0;
        pos = F(sub)(pos, F(mk)(F(gmod)(pos.x, s2), F(gmod)(pos.y, s2), F(gmod)(pos.z, s2)));
        ++child_idx;
        pos = F(add)(pos, F(mk)((child_idx & 4u) ? side : 0, (child_idx & 2u) ? side : 0, (child_idx & 1u) ? side : 0));
    }
    return total;
}

/* svo_rope.comp:29-48 */
static uint32_t F(find_relative)(const xo_volume* v, uint32_t parent, V3 offset, V3 pos, V3* base, REAL* side,
                                 uint64_t* w) {
    REAL extent = LDEXP((REAL)1, -(int)(F(leafdepth_of)(v, parent, w) & XO_DEPTH_MASK));
    offset = F(sub)(offset, F(mk)(F(gmod)(offset.x, extent), F(gmod)(offset.y, extent), F(gmod)(offset.z, extent)));
    for (uint32_t level = 0;; ++level) {
        if (F(leafdepth_of)(v, parent, w) >= XO_LEAF_MASK || level > XO_MAX_LEVELS) {
            *base = offset;
            *side = extent;
            return parent;
        }
        extent = extent * (REAL)0.5;
        int mx = pos.x >= offset.x + extent;
        int my = pos.y >= offset.y + extent;
        int mz = pos.z >= offset.z + extent;
        uint32_t child = (uint32_t)(mx * 4 + my * 2 + mz);
        offset = F(add)(offset, F(mul)(F(mk)((REAL)mx, (REAL)my, (REAL)mz), F(splat)(extent)));
        parent = F(child_of)(v, parent, child, w);
    }
}

/* svo_rope.comp:50-63 */
static inline uint32_t F(neighbor_index)(const uint32_t nb[3], V3 far, V3* mask) {
    if (far.x < F(gmin)(far.y, far.z)) {
        *mask = F(mk)(1, 0, 0);
        return nb[0];
    } else if (far.y < far.z) {
        *mask = F(mk)(0, 1, 0);
        return nb[1];
    }
    *mask = F(mk)(0, 0, 1);
    return nb[2];
}

/* svo_rope.comp:65-135 */
static V3 F(rope_trace)(const xo_volume* v, V3 ro, V3 rd, uint64_t cap, uint64_t* w) {
    V3 sgn = F(mk)(F(rsign)(rd.x), F(rsign)(rd.y), F(rsign)(rd.z));
    uint32_t nb[3] = {1u - (sgn.x > 0 ? 1u : 0u), 3u - (sgn.y > 0 ? 1u : 0u), 5u - (sgn.z > 0 ? 1u : 0u)};
    sgn = F(add)(sgn, F(splat)((REAL)0.1f));
    V3 rrd = F(divv)(F(splat)(1), rd);
    V3 bias = F(mul)(rrd, ro);
    V3 box_min = F(neg)(bias);
    V3 box_max = F(sub)(rrd, bias);
    REAL t_min = F(max_elem)(F(vmin)(box_min, box_max));
    REAL t_max = F(min_elem)(F(vmax)(box_min, box_max));
    if (t_min > t_max)
        return F(splat)(0);
    t_min = F(gmax)(t_min, 0);
    V3 pos = F(add)(ro, F(smul)(t_min, rd));
    V3 offset;
    REAL side;
    uint32_t node = F(svo_find)(v, pos, &offset, &side, w);
    V3 node_min = F(sub)(F(mul)(offset, rrd), bias);
    V3 node_max = F(sub)(F(mul)(F(add)(offset, F(splat)(side)), rrd), bias);
    V3 far = F(vmax)(node_min, node_max);
    REAL u_min = F(max_elem)(F(vmin)(node_min, node_max));
    REAL u_max = F(min_elem)(far);
    REAL step = u_max - F(gmax)(u_min, 0);
    V3 color = F(unorm_rgb)(F(color_of)(v, node, w));
    V3 total = F(muls)(color, step);
    V3 mask;
    uint32_t n = F(neighbor_index)(nb, far, &mask);
    node = F(child_of)(v, node, n, w);
    offset = F(add)(offset, F(muls)(F(mul)(mask, sgn), side));
    uint64_t it = 0;
    while (node != 0) {
        if (++it > cap)
            break;
        pos = F(add)(ro, F(smul)(u_max, rd));
        node = F(find_relative)(v, node, offset, pos, &offset, &side, w);
        node_min = F(sub)(F(mul)(offset, rrd), bias);
        node_max = F(sub)(F(mul)(F(add)(offset, F(splat)(side)), rrd), bias);
        far = F(vmax)(node_min, node_max);
        u_min = F(max_elem)(F(vmin)(node_min, node_max));
        u_max = F(min_elem)(far);
        step = u_max - F(gmax)(u_min, 0);
        color = F(unorm_rgb)(F(color_of)(v, node, w));
        total = F(add)(total, F(muls)(color, step));
        n = F(neighbor_index)(nb, far, &mask);
        node = F(child_of)(v, node, n, w);
        offset = F(add)(offset, F(muls)(F(mul)(mask, sgn), side));
    }
    return total;
}

/* ------------------------------------------------------------------ */
/* ESVO: resources/esvo.comp:10-136                                    */
/* ------------------------------------------------------------------ */

/* esvo.comp:10-21 */
static void F(aabb_intersect)(V3 bmin, V3 bmax, V3 ro, V3 rd, REAL* t0, REAL* t1) {
    V3 rrd = F(divv)(F(splat)(1), F(add)(rd, F(splat)((REAL)0.00000001f)));
    V3 tbot = F(mul)(F(sub)(bmin, ro), rrd);
    V3 ttop = F(mul)(F(sub)(bmax, ro), rrd);
    V3 tmin = F(vmin)(ttop, tbot);
    V3 tmax = F(vmax)(ttop, tbot);
    REAL a = F(gmax)(tmin.x, tmin.y), b = F(gmax)(tmin.x, tmin.z);
    *t0 = F(gmax)(a, b);
    a = F(gmin)(tmax.x, tmax.y);
    b = F(gmin)(tmax.x, tmax.z);
    *t1 = F(gmin)(a, b);
}

static inline uint32_t F(cxor)(uint32_t x, int ax, int ay, int az) {
    x ^= ax ? 4u : 0u;
    x ^= ay ? 2u : 0u;
    x ^= az ? 1u : 0u;
    return x;
}

#if XO_PRECISION == 32
/* esvo.comp:112-131 POP, on the IEEE-754 single-precision bit patterns. */
static int F(esvo_pop)(V3* pos, REAL* scale_exp2, int ax, int ay, int az, uint32_t* scale, uint32_t* sx,
                       uint32_t* sy, uint32_t* sz) {
    uint32_t bx = xo_f2u(pos->x), by = xo_f2u(pos->y), bz = xo_f2u(pos->z);
    uint32_t x0 = bx ^ xo_f2u(pos->x + *scale_exp2);
    uint32_t x1 = by ^ xo_f2u(pos->y + *scale_exp2);
    uint32_t x2 = bz ^ xo_f2u(pos->z + *scale_exp2);
    uint32_t dbits = (ax ? x0 : 0u) | (ay ? x1 : 0u) | (az ? x2 : 0u);
    uint32_t s = (xo_f2u((float)dbits) >> 23) - 127u;
    *scale = s;
    if (s >= 23u)
        return 1;
    *scale_exp2 = xo_u2f((s - 23u + 127u) << 23);
    bx >>= s;
    by >>= s;
    bz >>= s;
    pos->x = xo_u2f(bx << s);
    pos->y = xo_u2f(by << s);
    pos->z = xo_u2f(bz << s);
    *sx = bx;
    *sy = by;
    *sz = bz;
    return 0;
}
#else
/* FP64 twin of the POP: the same "highest differing bit" rule applied to the
 * positions as 52-bit fixed point (pos in [0.5, 2]), where bit 52 is 2^0 and
 * therefore float scale s corresponds to fixed-point bit s + 29. */
static int F(esvo_pop)(V3* pos, REAL* scale_exp2, int ax, int ay, int az, uint32_t* scale, uint32_t* sx,
                       uint32_t* sy, uint32_t* sz) {
    const double two52 = 4503599627370496.0;
    uint64_t px = (uint64_t)(pos->x * two52), py = (uint64_t)(pos->y * two52), pz = (uint64_t)(pos->z * two52);
    uint64_t qx = (uint64_t)((pos->x + *scale_exp2) * two52);
    uint64_t qy = (uint64_t)((pos->y + *scale_exp2) * two52);
    uint64_t qz = (uint64_t)((pos->z + *scale_exp2) * two52);
    uint64_t dbits = (ax ? px ^ qx : 0u) | (ay ? py ^ qy : 0u) | (az ? pz ^ qz : 0u);
    if (dbits == 0) {
        *scale = 23u;
        return 1;
    }
    int s = 63 - __builtin_clzll(dbits) - 29;
    if (s < 0)
        s = 0;
    *scale = (uint32_t)s;
    if (s >= 23)
        return 1;
    *scale_exp2 = ldexp(1.0, s - 23);
    int sh = s + 29;
    px >>= sh;
    py >>= sh;
    pz >>= sh;
    pos->x = (double)(px << sh) / two52;
    pos->y = (double)(py << sh) / two52;
    pos->z = (double)(pz << sh) / two52;
    *sx = (uint32_t)(px & 1u);
    *sy = (uint32_t)(py & 1u);
    *sz = (uint32_t)(pz & 1u);
    return 0;
}
#endif

static V3 F(esvo_trace)(const xo_volume* v, V3 ro, V3 rd, uint64_t cap, uint64_t* w) {
    const uint32_t cast_stack_depth = 23;
    uint32_t node_stack[24] = {0};
    REAL t_max_stack[24] = {0};
    V3 t_coeff = F(divv)(F(splat)(1), F(neg)(F(mk)(F(rabs)(rd.x), F(rabs)(rd.y), F(rabs)(rd.z))));
    V3 t_bias = F(mul)(t_coeff, ro);
    int gx = rd.x > 0, gy = rd.y > 0, gz = rd.z > 0;
    V3 alt = F(sub)(F(smul)((REAL)3.0, t_coeff), t_bias);
    t_bias = F(mk)(gx ? alt.x : t_bias.x, gy ? alt.y : t_bias.y, gz ? alt.z : t_bias.z);
    uint32_t octant_mask = F(cxor)(0, gx, gy, gz);
    REAL t_min = F(max_elem)(F(sub)(F(smul)((REAL)2.0, t_coeff), t_bias));
    REAL t_max = F(min_elem)(F(sub)(t_coeff, t_bias));
    REAL h = t_max;
    t_min = F(gmax)(t_min, 0);
    t_max = F(gmin)(t_max, XO_SQRT3);
    uint32_t parent = 0;
    uint32_t idx = 0;
    V3 pos = F(splat)(1);
    uint32_t scale = cast_stack_depth - 1;
    REAL scale_exp2 = (REAL)0.5;
    {
        V3 c = F(sub)(F(smul)((REAL)1.5, t_coeff), t_bias);
        int ax = c.x > t_min, ay = c.y > t_min, az = c.z > t_min;
        pos = F(mk)(ax ? (REAL)1.5 : pos.x, ay ? (REAL)1.5 : pos.y, az ? (REAL)1.5 : pos.z);
        idx = F(cxor)(idx, ax, ay, az);
    }
    V3 total = F(splat)(0);
    uint64_t it = 0;
    while (scale < cast_stack_depth) {
        if (++it > cap)
            break;
        V3 t_corner = F(sub)(F(mul)(pos, t_coeff), t_bias);
        REAL tc_max = F(min_elem)(t_corner);
        if (t_min <= t_max) {
            REAL tv_max = F(gmin)(t_max, tc_max);
            if (t_min <= tv_max) {
                uint32_t child = F(child_of)(v, parent, idx ^ octant_mask, w);
                if (F(leafdepth_of)(v, child, w) >= XO_LEAF_MASK) {
#if XO_MUTANT == 9
                    uint64_t uncounted = 0;
                    V3 color = F(unorm_rgb)(F(color_of)(v, child, &uncounted));
#else
                    V3 color = F(unorm_rgb)(F(color_of)(v, child, w));
#endif
                    total = F(add)(total, F(muls)(color, tv_max - t_min));
                } else {
                    /* PUSH (scale wraps below zero only past depth 22; the loop then ends) */
                    if (tc_max < h) {
                        node_stack[scale] = parent;
                        t_max_stack[scale] = t_max;
                    }
                    h = tc_max;
                    parent = child;
                    --scale;
                    scale_exp2 = scale_exp2 * (REAL)0.5;
                    V3 t_center = F(add)(F(smul)(scale_exp2, t_coeff), t_corner);
                    int ax = t_center.x > t_min, ay = t_center.y > t_min, az = t_center.z > t_min;
                    idx = F(cxor)(0, ax, ay, az);
                    pos = F(add)(pos, F(mk)(ax ? scale_exp2 : 0, ay ? scale_exp2 : 0, az ? scale_exp2 : 0));
                    t_max = tv_max;
                    continue;
                }
            }
        }
        /* ADVANCE */
        int ax = t_corner.x <= tc_max, ay = t_corner.y <= tc_max, az = t_corner.z <= tc_max;
        uint32_t step_mask = F(cxor)(0, ax, ay, az);
        pos = F(sub)(pos, F(mk)(ax ? scale_exp2 : 0, ay ? scale_exp2 : 0, az ? scale_exp2 : 0));
        t_min = tc_max;
        idx ^= step_mask;
        if ((idx & step_mask) != 0) {
            /* POP: highest differing bit of the position in the stepped axes */
            uint32_t sx, sy, sz;
            if (F(esvo_pop)(&pos, &scale_exp2, ax, ay, az, &scale, &sx, &sy, &sz))
                break; /* scale >= 23: the ray left the root cube */
            parent = node_stack[scale];
            t_max = t_max_stack[scale];
            idx = (sx & 1u) * 4u + (sy & 1u) * 2u + (sz & 1u);
            h = 0;
        }
    }
    return total;
}

/* ------------------------------------------------------------------ */
/* Per-pixel main() of every shader (dda.comp:55-73 etc.)              */
/* ------------------------------------------------------------------ */

static V3 F(pixel)(int traversal, const xo_volume* vol, const xo_camera* cam, const xo_uniforms* u, uint32_t ix,
                   uint32_t iy, uint64_t cap, uint64_t* words) {
    int32_t pxl = u->out_offset[0] + (int32_t)ix;
    int32_t pyl = u->out_offset[1] + (int32_t)iy;
#if XO_MUTANT == 2
    REAL uvx = ((REAL)(pxl - u->disp_offset[0]) + (REAL)0.5) / (REAL)u->disp_extent[0];
    REAL uvy = ((REAL)(pyl - u->disp_offset[1]) + (REAL)0.5) / (REAL)u->disp_extent[1];
#else
    REAL uvx = (REAL)(pxl - u->disp_offset[0]) / (REAL)u->disp_extent[0];
    REAL uvy = (REAL)(pyl - u->disp_offset[1]) / (REAL)u->disp_extent[1];
#endif
    V3 rd = F(ray)(cam, u, uvx, uvy);
    V3 tr = F(vload)(cam->translation);
    if (traversal == XO_DDA) {
        REAL dims[3] = {(REAL)vol->dims[0], (REAL)vol->dims[1], (REAL)vol->dims[2]};
        REAL side = F(max_elem)(F(mk)(dims[0], dims[1], dims[2]));
        V3 ro = F(muls)(tr, side);
        REAL ec = F(emission_coeff)(u, rd) / side;
        REAL md[3] = {(REAL)u->model_dim[0], (REAL)u->model_dim[1], (REAL)u->model_dim[2]};
        return F(muls)(F(dda_trace)(vol, md, ro, rd, cap, words), ec);
    }
    if (traversal == XO_ESVO) {
        V3 ro = F(add)(tr, F(splat)(1));
        REAL t0, t1;
        F(aabb_intersect)(F(splat)(1), F(splat)(2), ro, rd, &t0, &t1);
        ro = F(add)(ro, F(smul)(F(gmax)(t0, 0), rd));
        return F(muls)(F(esvo_trace)(vol, ro, rd, cap, words), F(emission_coeff)(u, rd));
    }
    V3 c;
    if (traversal == XO_SVO_NAIVE)
        c = F(naive_trace)(vol, tr, rd, cap, words);
    else if (traversal == XO_SVO_DF)
        c = F(df_trace)(vol, tr, rd, cap, words);
    else
        c = F(rope_trace)(vol, tr, rd, cap, words);
    return F(muls)(c, F(emission_coeff)(u, rd));
}

static V3 F(trace_direct)(int traversal, const xo_volume* vol, V3 ro, V3 rd, uint64_t cap, uint64_t* words) {
    switch (traversal) {
    case XO_DDA: {
        REAL md[3] = {(REAL)vol->dims[0], (REAL)vol->dims[1], (REAL)vol->dims[2]};
        return F(dda_trace)(vol, md, ro, rd, cap, words);
    }
    case XO_SVO_NAIVE:
        return F(naive_trace)(vol, ro, rd, cap, words);
    case XO_SVO_DF:
        return F(df_trace)(vol, ro, rd, cap, words);
    case XO_ESVO:
        return F(esvo_trace)(vol, ro, rd, cap, words);
    default:
        return F(rope_trace)(vol, ro, rd, cap, words);
    }
}

#undef V3
#undef F
