// Octree traversals in the reference shaders' own structure, one node word
// per read exactly where the GLSL reads it (resources/svo_naive.comp,
// svo_df.comp, svo_rope.comp, esvo.comp).  Templated on the node accessor:
//   RawNodes    -- the .svo layout (10 words per node)
//   TaggedNodes -- the derived device layout, read through its plain fields
// These run (a) with COUNT=true to count the reference algorithm's volume
// words, and (b) for trees too large for tagged child words.
#pragma once

#include "xd_internal.hpp"

namespace xd {

struct RawNodes {
    const uint32_t* w;
    uint32_t n;
    __device__ __forceinline__ uint32_t child(uint32_t n, uint32_t k) const { return w[(uint64_t)n * kNodeWords + k]; }
    __device__ __forceinline__ uint32_t rope(uint32_t n, uint32_t k) const { return w[(uint64_t)n * kNodeWords + k]; }
    __device__ __forceinline__ uint32_t color(uint32_t n) const { return w[(uint64_t)n * kNodeWords + 8]; }
    __device__ __forceinline__ uint32_t leafdepth(uint32_t n) const { return w[(uint64_t)n * kNodeWords + 9]; }
};

struct TaggedNodes {
    const uint32_t* child_w;  // tagged child words: 8 per node below leaf_base, then the leaves' ropes as planes
    const uint32_t* meta;     // node_count colours, then node_count is_leaf_depth words
    uint32_t n;
    uint32_t mask;  // index bits of a child word (the volume's index_mask)
    uint32_t leaf_base;
    __device__ __forceinline__ uint32_t child(uint32_t n, uint32_t k) const {
        return child_w[(uint64_t)n * 8 + k] & mask;
    }
    __device__ __forceinline__ uint32_t rope(uint32_t l, uint32_t k) const {
        const uint64_t i = l >= leaf_base ? (uint64_t)leaf_base * 8 + (uint64_t)k * (n - leaf_base) + (l - leaf_base)
                                          : (uint64_t)l * 8 + k;
        return child_w[i] & mask;
    }
    __device__ __forceinline__ uint32_t color(uint32_t i) const { return meta[i]; }
    __device__ __forceinline__ uint32_t leafdepth(uint32_t i) const { return meta[(uint64_t)n + i]; }
};

template <bool COUNT, class A>
__device__ __forceinline__ uint32_t rd_child(const A& a, uint32_t n, uint32_t k, uint64_t& w) {
    if (COUNT) ++w;
    return a.child(n, k);
}
// A leaf's rope k (children[k] of the leaf in the .svo words).
template <bool COUNT, class A>
__device__ __forceinline__ uint32_t rd_rope(const A& a, uint32_t n, uint32_t k, uint64_t& w) {
    if (COUNT) ++w;
    return a.rope(n, k);
}
template <bool COUNT, class A>
__device__ __forceinline__ uint32_t rd_color(const A& a, uint32_t n, uint64_t& w) {
    if (COUNT) ++w;
    return a.color(n);
}
template <bool COUNT, class A>
__device__ __forceinline__ uint32_t rd_leafdepth(const A& a, uint32_t n, uint64_t& w) {
    if (COUNT) ++w;
    return a.leafdepth(n);
}

// svo_naive.comp:8-27 / svo_rope.comp:8-27
template <bool COUNT, class A>
__device__ __forceinline__ uint32_t ref_find(const A& a, f3 pos, f3& base, float& side, uint64_t& w) {
    float extent = 1.0f;
    uint32_t index = 0;
    f3 offset = splat(0.0f);
    for (uint32_t level = 0;; ++level) {
        if (rd_leafdepth<COUNT>(a, index, w) >= kLeafMask || level > kMaxLevels) {
            base = offset;
            side = extent;
            return index;
        }
        extent = extent * 0.5f;
        bool mx = pos.x >= offset.x + extent;
        bool my = pos.y >= offset.y + extent;
        bool mz = pos.z >= offset.z + extent;
        uint32_t child = (mx ? 4u : 0u) + (my ? 2u : 0u) + (mz ? 1u : 0u);
        offset = add(offset, mul(mk(mx ? 1.0f : 0.0f, my ? 1.0f : 0.0f, mz ? 1.0f : 0.0f), splat(extent)));
        index = rd_child<COUNT>(a, index, child, w);
    }
}

// svo_naive.comp:29-71
template <bool COUNT, class A>
__device__ __forceinline__ f3 ref_naive(const A& a, uint64_t cap, f3 ro, f3 rd, uint64_t& w) {
    const float min_step = 0.00001f;
    f3 rrd = divv(splat(1.0f), rd);
    f3 bias = mul(rrd, ro);
    f3 box_min = neg(bias);
    f3 box_max = sub(rrd, bias);
    float t_min = max_elem(vmin(box_min, box_max));
    float t_max = min_elem(vmax(box_min, box_max));
    if (t_min > t_max) return splat(0.0f);
    t_min = gmax(t_min, 0.0f);
    f3 total = splat(0.0f);
    float t = t_min + min_step;
    uint64_t it = 0;
    while (t < t_max) {
        if (++it > cap) break;
        f3 pt = add(smul(t, rd), ro);
        f3 offset;
        float side;
        uint32_t node = ref_find<COUNT>(a, pt, offset, side, w);
        f3 node_min = sub(mul(offset, rrd), bias);
        f3 node_max = sub(mul(add(offset, splat(side)), rrd), bias);
        float u_min = max_elem(vmin(node_min, node_max));
        float u_max = min_elem(vmax(node_min, node_max));
        u_min = gmax(u_min, 0.0f);
        float step = gmax(u_max - u_min, min_step);
        t = t + step;
        f3 color = unorm_rgb(rd_color<COUNT>(a, node, w));
        total = add(total, muls(color, step));
    }
    return total;
}

// svo_df.comp:6-66; the stack lives in LDS, one column per lane
template <bool COUNT, class A>
__device__ __forceinline__ f3 ref_df(const A& a, uint64_t cap, uint32_t slots, f3 ro, f3 rd, uint32_t* stk_node,
                                     uint32_t* stk_child, uint64_t& w) {
    f3 rrd = divv(splat(1.0f), rd);
    f3 bias = mul(rrd, ro);
    const int cap_sp = (int)slots;
    int sp = 0;
    uint32_t node = 0, child_idx = 0;
    f3 pos = splat(0.0f);
    float side = 0.5f;
    f3 total = splat(0.0f);
    uint64_t it = 0;
    for (;;) {
        if (++it > cap) break;
        uint32_t child = rd_child<COUNT>(a, node, child_idx, w);
        f3 box_min = sub(mul(pos, rrd), bias);
        f3 box_max = sub(mul(add(pos, splat(side)), rrd), bias);
        float t_min = max_elem(vmin(box_min, box_max));
        float t_max = min_elem(vmax(box_min, box_max));
        if (t_min < t_max && t_max > 0.0f) {
            if (rd_leafdepth<COUNT>(a, child, w) >= kLeafMask) {
                f3 color = unorm_rgb(rd_color<COUNT>(a, child, w));
                total = add(total, muls(color, t_max - gmax(t_min, 0.0f)));
            } else if (child_idx == 7 || sp < cap_sp) {
                if (child_idx != 7) {
                    stk_node[sp * kWaveLanes] = node;
                    stk_child[sp * kWaveLanes] = child_idx;
                    ++sp;
                }
                side = side * 0.5f;
                node = child;
                child_idx = 0;
                continue;
            }
        }
        if (child_idx == 7) {
            --sp;
            if (sp < 0) break;
            node = min(stk_node[sp * kWaveLanes], a.n - 1u);
            child_idx = stk_child[sp * kWaveLanes] & 7u;
            side = __builtin_ldexpf(1.0f, -(int)(rd_leafdepth<COUNT>(a, node, w) & kDepthMask)) * 0.5f;
        }
        float s2 = side * 2.0f;
        pos = sub(pos, mk(gmod(pos.x, s2), gmod(pos.y, s2), gmod(pos.z, s2)));
        ++child_idx;
        pos = add(pos, mk((child_idx & 4u) ? side : 0.0f, (child_idx & 2u) ? side : 0.0f,
                          (child_idx & 1u) ? side : 0.0f));
    }
    return total;
}

// svo_rope.comp:29-48
template <bool COUNT, class A>
__device__ __forceinline__ uint32_t ref_find_relative(const A& a, uint32_t parent, f3 offset, f3 pos, f3& base,
                                                      float& side, uint64_t& w) {
    float extent = __builtin_ldexpf(1.0f, -(int)(rd_leafdepth<COUNT>(a, parent, w) & kDepthMask));
    offset = sub(offset, mk(gmod(offset.x, extent), gmod(offset.y, extent), gmod(offset.z, extent)));
    for (uint32_t level = 0;; ++level) {
        if (rd_leafdepth<COUNT>(a, parent, w) >= kLeafMask || level > kMaxLevels) {
            base = offset;
            side = extent;
            return parent;
        }
        extent = extent * 0.5f;
        bool mx = pos.x >= offset.x + extent;
        bool my = pos.y >= offset.y + extent;
        bool mz = pos.z >= offset.z + extent;
        uint32_t child = (mx ? 4u : 0u) + (my ? 2u : 0u) + (mz ? 1u : 0u);
        offset = add(offset, mul(mk(mx ? 1.0f : 0.0f, my ? 1.0f : 0.0f, mz ? 1.0f : 0.0f), splat(extent)));
        parent = rd_child<COUNT>(a, parent, child, w);
    }
}

// svo_rope.comp:50-63
__device__ __forceinline__ uint32_t neighbor_index(uint32_t nbx, uint32_t nby, uint32_t nbz, f3 far, f3& mask) {
    if (far.x < gmin(far.y, far.z)) {
        mask = mk(1.0f, 0.0f, 0.0f);
        return nbx;
    } else if (far.y < far.z) {
        mask = mk(0.0f, 1.0f, 0.0f);
        return nby;
    }
    mask = mk(0.0f, 0.0f, 1.0f);
    return nbz;
}

// svo_rope.comp:65-135
template <bool COUNT, class A>
__device__ __forceinline__ f3 ref_rope(const A& a, uint64_t cap, f3 ro, f3 rd, uint64_t& w) {
    f3 sgn = mk(fsign(rd.x), fsign(rd.y), fsign(rd.z));
    const uint32_t nbx = 1u - (sgn.x > 0.0f ? 1u : 0u);
    const uint32_t nby = 3u - (sgn.y > 0.0f ? 1u : 0u);
    const uint32_t nbz = 5u - (sgn.z > 0.0f ? 1u : 0u);
    sgn = add(sgn, splat(0.1f));
    f3 rrd = divv(splat(1.0f), rd);
    f3 bias = mul(rrd, ro);
    f3 box_min = neg(bias);
    f3 box_max = sub(rrd, bias);
    float t_min = max_elem(vmin(box_min, box_max));
    float t_max = min_elem(vmax(box_min, box_max));
    if (t_min > t_max) return splat(0.0f);
    t_min = gmax(t_min, 0.0f);
    f3 pos = add(ro, smul(t_min, rd));
    f3 offset;
    float side;
    uint32_t node = ref_find<COUNT>(a, pos, offset, side, w);
    f3 node_min = sub(mul(offset, rrd), bias);
    f3 node_max = sub(mul(add(offset, splat(side)), rrd), bias);
    f3 far = vmax(node_min, node_max);
    float u_min = max_elem(vmin(node_min, node_max));
    float u_max = min_elem(far);
    float step = u_max - gmax(u_min, 0.0f);
    f3 total = muls(unorm_rgb(rd_color<COUNT>(a, node, w)), step);
    f3 mask;
    uint32_t n = neighbor_index(nbx, nby, nbz, far, mask);
    node = rd_rope<COUNT>(a, node, n, w);
    offset = add(offset, muls(mul(mask, sgn), side));
    uint64_t it = 0;
    while (node != 0) {
        if (++it > cap) break;
        pos = add(ro, smul(u_max, rd));
        node = ref_find_relative<COUNT>(a, node, offset, pos, offset, side, w);
        node_min = sub(mul(offset, rrd), bias);
        node_max = sub(mul(add(offset, splat(side)), rrd), bias);
        far = vmax(node_min, node_max);
        u_min = max_elem(vmin(node_min, node_max));
        u_max = min_elem(far);
        step = u_max - gmax(u_min, 0.0f);
        total = add(total, muls(unorm_rgb(rd_color<COUNT>(a, node, w)), step));
        n = neighbor_index(nbx, nby, nbz, far, mask);
        node = rd_rope<COUNT>(a, node, n, w);
        offset = add(offset, muls(mul(mask, sgn), side));
    }
    return total;
}

// esvo.comp:10-21
__device__ __forceinline__ void aabb_intersect(f3 bmin, f3 bmax, f3 ro, f3 rd, float& t0, float& t1) {
    f3 rrd = divv(splat(1.0f), add(rd, splat(0.00000001f)));
    f3 tbot = mul(sub(bmin, ro), rrd);
    f3 ttop = mul(sub(bmax, ro), rrd);
    f3 tmin = vmin(ttop, tbot);
    f3 tmax = vmax(ttop, tbot);
    t0 = gmax(gmax(tmin.x, tmin.y), gmax(tmin.x, tmin.z));
    t1 = gmin(gmin(tmax.x, tmax.y), gmin(tmax.x, tmax.z));
}

// esvo.comp:23-28
__device__ __forceinline__ uint32_t cxor(uint32_t x, bool ax, bool ay, bool az) {
    return x ^ (ax ? 4u : 0u) ^ (ay ? 2u : 0u) ^ (az ? 1u : 0u);
}

// Ray setup of esvo.comp:37-61, shared by both ESVO variants.
struct EsvoRay {
    f3 t_coeff, t_bias;
    uint32_t octant_mask;
    float t_min, t_max, h;
    f3 pos;
    uint32_t idx;
};

__device__ __forceinline__ EsvoRay esvo_setup(f3 ro, f3 rd) {
    EsvoRay r;
    r.t_coeff = divv(splat(1.0f), neg(mk(fabs_(rd.x), fabs_(rd.y), fabs_(rd.z))));
    r.t_bias = mul(r.t_coeff, ro);
    bool gx = rd.x > 0.0f, gy = rd.y > 0.0f, gz = rd.z > 0.0f;
    f3 alt = sub(smul(3.0f, r.t_coeff), r.t_bias);
    r.t_bias = mk(gx ? alt.x : r.t_bias.x, gy ? alt.y : r.t_bias.y, gz ? alt.z : r.t_bias.z);
    r.octant_mask = cxor(0u, gx, gy, gz);
    r.t_min = max_elem(sub(smul(2.0f, r.t_coeff), r.t_bias));
    r.t_max = min_elem(sub(r.t_coeff, r.t_bias));
    r.h = r.t_max;
    r.t_min = gmax(r.t_min, 0.0f);
    r.t_max = gmin(r.t_max, __builtin_sqrtf(3.0f));
    r.pos = splat(1.0f);
    r.idx = 0;
    f3 c = sub(smul(1.5f, r.t_coeff), r.t_bias);
    bool ax = c.x > r.t_min, ay = c.y > r.t_min, az = c.z > r.t_min;
    r.pos = mk(ax ? 1.5f : r.pos.x, ay ? 1.5f : r.pos.y, az ? 1.5f : r.pos.z);
    r.idx = cxor(r.idx, ax, ay, az);
    return r;
}

// esvo.comp:112-131 POP.  Returns false when the ray left the root cube.
__device__ __forceinline__ bool esvo_pop(f3& pos, float& scale_exp2, uint32_t& scale, uint32_t& idx, bool ax, bool ay,
                                         bool az) {
    uint32_t bx = f2u(pos.x), by = f2u(pos.y), bz = f2u(pos.z);
    uint32_t x0 = bx ^ f2u(pos.x + scale_exp2);
    uint32_t x1 = by ^ f2u(pos.y + scale_exp2);
    uint32_t x2 = bz ^ f2u(pos.z + scale_exp2);
    uint32_t dbits = (ax ? x0 : 0u) | (ay ? x1 : 0u) | (az ? x2 : 0u);
    scale = (f2u((float)dbits) >> 23) - 127u;
    if (scale >= 23u) return false;
    scale_exp2 = u2f((scale - 23u + 127u) << 23);
    bx >>= scale;
    by >>= scale;
    bz >>= scale;
    pos = mk(u2f(bx << scale), u2f(by << scale), u2f(bz << scale));
    idx = (bx & 1u) * 4u + (by & 1u) * 2u + (bz & 1u);
    return true;
}

// esvo.comp:30-136.  Stack slot of scale s is 22 - s (depth of the pushed
// parent).
template <bool COUNT, class A>
__device__ __forceinline__ f3 ref_esvo(const A& a, uint64_t cap, uint32_t slots, f3 ro, f3 rd, uint32_t* stk_parent,
                                       float* stk_tmax, uint64_t& w) {
    EsvoRay r = esvo_setup(ro, rd);
    float t_min = r.t_min, t_max = r.t_max, h = r.h;
    f3 pos = r.pos;
    uint32_t idx = r.idx, parent = 0, scale = 22;
    float scale_exp2 = 0.5f;
    f3 total = splat(0.0f);
    uint64_t it = 0;
    while (scale < 23u) {
        if (++it > cap) break;
        f3 t_corner = sub(mul(pos, r.t_coeff), r.t_bias);
        float tc_max = min_elem(t_corner);
        if (t_min <= t_max) {
            float tv_max = gmin(t_max, tc_max);
            if (t_min <= tv_max) {
                uint32_t child = rd_child<COUNT>(a, parent, idx ^ r.octant_mask, w);
                if (rd_leafdepth<COUNT>(a, child, w) >= kLeafMask) {
                    f3 color = unorm_rgb(rd_color<COUNT>(a, child, w));
                    total = add(total, muls(color, tv_max - t_min));
                } else {
                    if (tc_max < h) {
                        uint32_t slot = 22u - scale;
                        if (slot < slots) {
                            stk_parent[slot * kWaveLanes] = parent;
                            stk_tmax[slot * kWaveLanes] = t_max;
                        }
                    }
                    h = tc_max;
                    parent = child;
                    --scale;
                    scale_exp2 = scale_exp2 * 0.5f;
                    f3 t_center = add(smul(scale_exp2, r.t_coeff), t_corner);
                    bool ax = t_center.x > t_min, ay = t_center.y > t_min, az = t_center.z > t_min;
                    idx = cxor(0u, ax, ay, az);
                    pos = add(pos, mk(ax ? scale_exp2 : 0.0f, ay ? scale_exp2 : 0.0f, az ? scale_exp2 : 0.0f));
                    t_max = tv_max;
                    continue;
                }
            }
        }
        bool ax = t_corner.x <= tc_max, ay = t_corner.y <= tc_max, az = t_corner.z <= tc_max;
        uint32_t step_mask = cxor(0u, ax, ay, az);
        pos = sub(pos, mk(ax ? scale_exp2 : 0.0f, ay ? scale_exp2 : 0.0f, az ? scale_exp2 : 0.0f));
        t_min = tc_max;
        idx ^= step_mask;
        if ((idx & step_mask) != 0) {
            if (!esvo_pop(pos, scale_exp2, scale, idx, ax, ay, az)) break;
            uint32_t slot = 22u - scale;
            parent = slot < slots ? min(stk_parent[slot * kWaveLanes], a.n - 1u) : 0u;
            t_max = slot < slots ? stk_tmax[slot * kWaveLanes] : 0.0f;
            h = 0.0f;
        }
    }
    return total;
}

}  // namespace xd
