// Traversal kernels for gfx950: one ray per lane, wave64 screen tiles.
//
// Each kernel restates one reference shader (resources/*.comp) operation by
// operation in FP32, so that frames are bit-identical to the CPU restatement
// of the same shader.  Shortcuts are taken only where they cannot change a
// result: a texel whose RGB is zero adds exactly +0 to the sum, so the DDA
// skips the fetch of voxels in bricks known to be empty (the side-distance
// additions, which decide the voxel sequence, still run for every step).
//
// Work decomposition: a 256-thread block is four wavefronts; each wavefront
// takes one 8x8 (or 16x4 / 32x2) pixel tile, so a wave's rays start together
// and walk neighbouring voxels/nodes (coherent gathers through L2).  The
// hardware deals blocks to the 8 XCDs round-robin; blocks are renumbered so
// each XCD renders one contiguous band of the image and its L2 keeps that
// band's part of the volume.
#include "xd_internal.hpp"

namespace xd {

template <int TW>
__device__ __forceinline__ void block_pixel(uint32_t W, uint32_t H, uint32_t& ix, uint32_t& iy) {
    constexpr uint32_t TH = 64 / TW;
    constexpr uint32_t BW = 2 * TW, BH = 2 * TH;
    const uint32_t nbx = (W + BW - 1) / BW;
    const uint32_t nblocks = nbx * ((H + BH - 1) / BH);
    const uint32_t b = blockIdx.x;
    const uint32_t full = nblocks & ~7u;
    const uint32_t lb = b < full ? (b & 7u) * (full >> 3) + (b >> 3) : b;
    const uint32_t bx = lb % nbx, by = lb / nbx;
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    ix = bx * BW + (w & 1u) * TW + (lane % TW);
    iy = by * BH + (w >> 1) * TH + (lane / TW);
}

__device__ __forceinline__ void add_words(unsigned long long* counter, uint64_t w) {
    // wave-wide sum, one atomic per wavefront
    for (int off = 32; off > 0; off >>= 1) w += __shfl_xor(w, off, 64);
    if ((threadIdx.x & 63u) == 0 && w) atomicAdd(counter, (unsigned long long)w);
}

// GLSL ivec3(vec3) truncation; out-of-range saturates and NaN gives 0.
__device__ __forceinline__ int32_t f2i_sat(float x) {
    if (!(x == x)) return 0;
    if (x >= 2147483647.0f) return 2147483647;
    if (x <= -2147483648.0f) return (-2147483647 - 1);
    return (int32_t)x;
}

// ---------------------------------------------------------------------------
// DDA: resources/dda.comp:13-53
// ---------------------------------------------------------------------------
template <bool COUNT>
__device__ __forceinline__ f3 dda_trace(const RenderParams& p, f3 ro, f3 rd, uint64_t& words) {
    f3 rrd = divv(splat(1.0f), rd);
    f3 bias = mul(rrd, ro);
    f3 box_min = neg(bias);
    f3 md = mk((float)p.view.model_dim[0], (float)p.view.model_dim[1], (float)p.view.model_dim[2]);
    f3 box_max = sub(mul(md, rrd), bias);
    float t_min = max_elem(vmin(box_min, box_max));
    float t_max = min_elem(vmax(box_min, box_max));
    if (t_min > t_max) return splat(0.0f);
    t_min = gmax(t_min, 0.0f);
    ro = add(ro, muls(rd, t_min));
    int32_t px = f2i_sat(ro.x), py = f2i_sat(ro.y), pz = f2i_sat(ro.z);
    f3 t_delta = mk(fabs_(rrd.x), fabs_(rrd.y), fabs_(rrd.z));
    f3 sgn = mk(fsign(rd.x), fsign(rd.y), fsign(rd.z));
    int32_t sx = (int32_t)sgn.x, sy = (int32_t)sgn.y, sz = (int32_t)sgn.z;
    f3 fl = mk(__builtin_floorf(ro.x), __builtin_floorf(ro.y), __builtin_floorf(ro.z));
    f3 sd = mul(add(mul(sgn, add(sub(fl, ro), splat(0.5f))), splat(0.5f)), t_delta);
    float t = 0.0f;
    f3 total = splat(0.0f);
    const float limit = t_max - t_min;
    const uint32_t X = p.dims[0], Y = p.dims[1], Z = p.dims[2];
    const uint32_t nbx = p.nb[0], nby = p.nb[1];
    uint64_t cur_b = ~0ull;
    bool cur_occ = false;
    uint64_t it = 0;
    while (t < limit) {
        if (++it > p.cap) break;
        bool mx = sd.x <= gmin(sd.y, sd.z);
        bool my = sd.y <= gmin(sd.z, sd.x);
        bool mz = sd.z <= gmin(sd.x, sd.y);
        float t0 = min_elem(sd);
        if (COUNT) {
            ++words;
        } else if ((uint32_t)px < X && (uint32_t)py < Y && (uint32_t)pz < Z) {
            uint64_t b = ((uint64_t)((uint32_t)pz >> 2) * nby + ((uint32_t)py >> 2)) * nbx + ((uint32_t)px >> 2);
            if (b != cur_b) {
                cur_b = b;
                cur_occ = p.occ[b] != 0;
            }
            if (cur_occ) {
                uint32_t tex = p.bricks[b * 64 + (px & 3) + ((py & 3) << 2) + ((pz & 3) << 4)];
                if (tex & 0xFFFFFFu) {
                    float dt = t0 - t;
                    total = add(total, muls(unorm_rgb(tex), dt));
                }
            }
        }
        t = t0;
        sd = add(sd, mk(mx ? t_delta.x : 0.0f, my ? t_delta.y : 0.0f, mz ? t_delta.z : 0.0f));
        px += mx ? sx : 0;
        py += my ? sy : 0;
        pz += mz ? sz : 0;
    }
    return total;
}

template <int TW, bool COUNT>
__global__ __launch_bounds__(256) void k_dda(RenderParams p) {
    uint32_t ix, iy;
    block_pixel<TW>(p.view.out_ext[0], p.view.out_ext[1], ix, iy);
    uint64_t words = 0;
    if (ix < p.view.out_ext[0] && iy < p.view.out_ext[1]) {
        f3 rd = ray_dir(p.cam, p.view, ix, iy);
        float side = max_elem(mk((float)p.dims[0], (float)p.dims[1], (float)p.dims[2]));
        f3 ro = muls(p.cam.translation, side);
        float ec = emission_coeff(p.view, rd) / side;
        f3 total = dda_trace<COUNT>(p, ro, rd, words);
        if (!COUNT) p.out[(uint64_t)iy * p.out_pitch + ix] = pack_rgba(muls(total, ec));
    }
    if (COUNT) add_words(p.words, words);
}

// ---------------------------------------------------------------------------
// Octree helpers: resources/octree.glsl:6-17
// ---------------------------------------------------------------------------
template <bool COUNT>
__device__ __forceinline__ uint32_t word(const uint32_t* nodes, uint32_t n, uint32_t k, uint64_t& w) {
    if (COUNT) ++w;
    return nodes[(uint64_t)n * kNodeWords + k];
}

// svo_naive.comp:8-27 / svo_rope.comp:8-27
template <bool COUNT>
__device__ __forceinline__ uint32_t svo_find(const uint32_t* nodes, f3 pos, f3& base, float& side, uint64_t& w) {
    float extent = 1.0f;
    uint32_t index = 0;
    f3 offset = splat(0.0f);
    for (uint32_t level = 0;; ++level) {
        if (word<COUNT>(nodes, index, 9, w) >= kLeafMask || level > kMaxLevels) {
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
        index = word<COUNT>(nodes, index, child, w);
    }
}

// svo_naive.comp:29-71
template <bool COUNT>
__device__ __forceinline__ f3 naive_trace(const RenderParams& p, f3 ro, f3 rd, uint64_t& w) {
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
        if (++it > p.cap) break;
        f3 pt = add(smul(t, rd), ro);
        f3 offset;
        float side;
        uint32_t node = svo_find<COUNT>(p.nodes, pt, offset, side, w);
        f3 node_min = sub(mul(offset, rrd), bias);
        f3 node_max = sub(mul(add(offset, splat(side)), rrd), bias);
        float u_min = max_elem(vmin(node_min, node_max));
        float u_max = min_elem(vmax(node_min, node_max));
        u_min = gmax(u_min, 0.0f);
        float step = gmax(u_max - u_min, min_step);
        t = t + step;
        f3 color = unorm_rgb(word<COUNT>(p.nodes, node, 8, w));
        total = add(total, muls(color, step));
    }
    return total;
}

// svo_df.comp:6-66; the stack lives in LDS, one column per lane
template <bool COUNT>
__device__ __forceinline__ f3 df_trace(const RenderParams& p, f3 ro, f3 rd, uint32_t* stk_node,
                                       uint32_t* stk_child, uint64_t& w) {
    const uint32_t* nodes = p.nodes;
    f3 rrd = divv(splat(1.0f), rd);
    f3 bias = mul(rrd, ro);
    const int cap_sp = (int)p.stack_slots;
    int sp = 0;
    uint32_t node = 0, child_idx = 0;
    f3 pos = splat(0.0f);
    float side = 0.5f;
    f3 total = splat(0.0f);
    uint64_t it = 0;
    for (;;) {
        if (++it > p.cap) break;
        uint32_t child = word<COUNT>(nodes, node, child_idx, w);
        f3 box_min = sub(mul(pos, rrd), bias);
        f3 box_max = sub(mul(add(pos, splat(side)), rrd), bias);
        float t_min = max_elem(vmin(box_min, box_max));
        float t_max = min_elem(vmax(box_min, box_max));
        if (t_min < t_max && t_max > 0.0f) {
            if (word<COUNT>(nodes, child, 9, w) >= kLeafMask) {
                f3 color = unorm_rgb(word<COUNT>(nodes, child, 8, w));
                total = add(total, muls(color, t_max - gmax(t_min, 0.0f)));
            } else if (child_idx == 7 || sp < cap_sp) {
                if (child_idx != 7) {
                    stk_node[sp * kBlockThreads] = node;
                    stk_child[sp * kBlockThreads] = child_idx;
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
            node = stk_node[sp * kBlockThreads];
            child_idx = stk_child[sp * kBlockThreads];
            side = __builtin_ldexpf(1.0f, -(int)(word<COUNT>(nodes, node, 9, w) & kDepthMask)) * 0.5f;
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
template <bool COUNT>
__device__ __forceinline__ uint32_t find_relative(const uint32_t* nodes, uint32_t parent, f3 offset, f3 pos,
                                                  f3& base, float& side, uint64_t& w) {
    float extent = __builtin_ldexpf(1.0f, -(int)(word<COUNT>(nodes, parent, 9, w) & kDepthMask));
    offset = sub(offset, mk(gmod(offset.x, extent), gmod(offset.y, extent), gmod(offset.z, extent)));
    for (uint32_t level = 0;; ++level) {
        if (word<COUNT>(nodes, parent, 9, w) >= kLeafMask || level > kMaxLevels) {
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
        parent = word<COUNT>(nodes, parent, child, w);
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
template <bool COUNT>
__device__ __forceinline__ f3 rope_trace(const RenderParams& p, f3 ro, f3 rd, uint64_t& w) {
    const uint32_t* nodes = p.nodes;
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
    uint32_t node = svo_find<COUNT>(nodes, pos, offset, side, w);
    f3 node_min = sub(mul(offset, rrd), bias);
    f3 node_max = sub(mul(add(offset, splat(side)), rrd), bias);
    f3 far = vmax(node_min, node_max);
    float u_min = max_elem(vmin(node_min, node_max));
    float u_max = min_elem(far);
    float step = u_max - gmax(u_min, 0.0f);
    f3 total = muls(unorm_rgb(word<COUNT>(nodes, node, 8, w)), step);
    f3 mask;
    uint32_t n = neighbor_index(nbx, nby, nbz, far, mask);
    node = word<COUNT>(nodes, node, n, w);
    offset = add(offset, muls(mul(mask, sgn), side));
    uint64_t it = 0;
    while (node != 0) {
        if (++it > p.cap) break;
        pos = add(ro, smul(u_max, rd));
        node = find_relative<COUNT>(nodes, node, offset, pos, offset, side, w);
        node_min = sub(mul(offset, rrd), bias);
        node_max = sub(mul(add(offset, splat(side)), rrd), bias);
        far = vmax(node_min, node_max);
        u_min = max_elem(vmin(node_min, node_max));
        u_max = min_elem(far);
        step = u_max - gmax(u_min, 0.0f);
        total = add(total, muls(unorm_rgb(word<COUNT>(nodes, node, 8, w)), step));
        n = neighbor_index(nbx, nby, nbz, far, mask);
        node = word<COUNT>(nodes, node, n, w);
        offset = add(offset, muls(mul(mask, sgn), side));
    }
    return total;
}

// ---------------------------------------------------------------------------
// ESVO: resources/esvo.comp:10-136
// ---------------------------------------------------------------------------
__device__ __forceinline__ void aabb_intersect(f3 bmin, f3 bmax, f3 ro, f3 rd, float& t0, float& t1) {
    f3 rrd = divv(splat(1.0f), add(rd, splat(0.00000001f)));
    f3 tbot = mul(sub(bmin, ro), rrd);
    f3 ttop = mul(sub(bmax, ro), rrd);
    f3 tmin = vmin(ttop, tbot);
    f3 tmax = vmax(ttop, tbot);
    t0 = gmax(gmax(tmin.x, tmin.y), gmax(tmin.x, tmin.z));
    t1 = gmin(gmin(tmax.x, tmax.y), gmin(tmax.x, tmax.z));
}

__device__ __forceinline__ uint32_t cxor(uint32_t x, bool ax, bool ay, bool az) {
    return x ^ (ax ? 4u : 0u) ^ (ay ? 2u : 0u) ^ (az ? 1u : 0u);
}

// Stack slot of scale s is 22 - s (depth of the pushed parent).
template <bool COUNT>
__device__ __forceinline__ f3 esvo_trace(const RenderParams& p, f3 ro, f3 rd, uint32_t* stk_parent, float* stk_tmax,
                                         uint64_t& w) {
    const uint32_t* nodes = p.nodes;
    const uint32_t cast_stack_depth = 23;
    const uint32_t slots = p.stack_slots;
    f3 t_coeff = divv(splat(1.0f), neg(mk(fabs_(rd.x), fabs_(rd.y), fabs_(rd.z))));
    f3 t_bias = mul(t_coeff, ro);
    bool gx = rd.x > 0.0f, gy = rd.y > 0.0f, gz = rd.z > 0.0f;
    f3 alt = sub(smul(3.0f, t_coeff), t_bias);
    t_bias = mk(gx ? alt.x : t_bias.x, gy ? alt.y : t_bias.y, gz ? alt.z : t_bias.z);
    const uint32_t octant_mask = cxor(0u, gx, gy, gz);
    float t_min = max_elem(sub(smul(2.0f, t_coeff), t_bias));
    float t_max = min_elem(sub(t_coeff, t_bias));
    float h = t_max;
    t_min = gmax(t_min, 0.0f);
    t_max = gmin(t_max, __builtin_sqrtf(3.0f));
    uint32_t parent = 0, idx = 0;
    f3 pos = splat(1.0f);
    uint32_t scale = cast_stack_depth - 1;
    float scale_exp2 = 0.5f;
    {
        f3 c = sub(smul(1.5f, t_coeff), t_bias);
        bool ax = c.x > t_min, ay = c.y > t_min, az = c.z > t_min;
        pos = mk(ax ? 1.5f : pos.x, ay ? 1.5f : pos.y, az ? 1.5f : pos.z);
        idx = cxor(idx, ax, ay, az);
    }
    f3 total = splat(0.0f);
    uint64_t it = 0;
    while (scale < cast_stack_depth) {
        if (++it > p.cap) break;
        f3 t_corner = sub(mul(pos, t_coeff), t_bias);
        float tc_max = min_elem(t_corner);
        if (t_min <= t_max) {
            float tv_max = gmin(t_max, tc_max);
            if (t_min <= tv_max) {
                uint32_t child = word<COUNT>(nodes, parent, idx ^ octant_mask, w);
                if (word<COUNT>(nodes, child, 9, w) >= kLeafMask) {
                    f3 color = unorm_rgb(word<COUNT>(nodes, child, 8, w));
                    total = add(total, muls(color, tv_max - t_min));
                } else {
                    // PUSH
                    if (tc_max < h) {
                        uint32_t slot = 22u - scale;
                        if (slot < slots) {
                            stk_parent[slot * kBlockThreads] = parent;
                            stk_tmax[slot * kBlockThreads] = t_max;
                        }
                    }
                    h = tc_max;
                    parent = child;
                    --scale;
                    scale_exp2 = scale_exp2 * 0.5f;
                    f3 t_center = add(smul(scale_exp2, t_coeff), t_corner);
                    bool ax = t_center.x > t_min, ay = t_center.y > t_min, az = t_center.z > t_min;
                    idx = cxor(0u, ax, ay, az);
                    pos = add(pos, mk(ax ? scale_exp2 : 0.0f, ay ? scale_exp2 : 0.0f, az ? scale_exp2 : 0.0f));
                    t_max = tv_max;
                    continue;
                }
            }
        }
        // ADVANCE
        bool ax = t_corner.x <= tc_max, ay = t_corner.y <= tc_max, az = t_corner.z <= tc_max;
        uint32_t step_mask = cxor(0u, ax, ay, az);
        pos = sub(pos, mk(ax ? scale_exp2 : 0.0f, ay ? scale_exp2 : 0.0f, az ? scale_exp2 : 0.0f));
        t_min = tc_max;
        idx ^= step_mask;
        if ((idx & step_mask) != 0) {
            // POP: highest differing bit of the position in the stepped axes
            uint32_t bx = f2u(pos.x), by = f2u(pos.y), bz = f2u(pos.z);
            uint32_t x0 = bx ^ f2u(pos.x + scale_exp2);
            uint32_t x1 = by ^ f2u(pos.y + scale_exp2);
            uint32_t x2 = bz ^ f2u(pos.z + scale_exp2);
            uint32_t dbits = (ax ? x0 : 0u) | (ay ? x1 : 0u) | (az ? x2 : 0u);
            scale = (f2u((float)dbits) >> 23) - 127u;
            if (scale >= cast_stack_depth) break;  // left the root cube
            scale_exp2 = u2f((scale - cast_stack_depth + 127u) << 23);
            uint32_t slot = 22u - scale;
            parent = slot < slots ? stk_parent[slot * kBlockThreads] : 0u;
            t_max = slot < slots ? stk_tmax[slot * kBlockThreads] : 0.0f;
            bx >>= scale;
            by >>= scale;
            bz >>= scale;
            pos = mk(u2f(bx << scale), u2f(by << scale), u2f(bz << scale));
            idx = (bx & 1u) * 4u + (by & 1u) * 2u + (bz & 1u);
            h = 0.0f;
        }
    }
    return total;
}

// ---------------------------------------------------------------------------
// Octree kernels (main() of svo_naive/svo_df/esvo/svo_rope.comp)
// ---------------------------------------------------------------------------
template <int TRAV, int TW, bool COUNT>
__global__ __launch_bounds__(256) void k_svo(RenderParams p) {
    extern __shared__ uint32_t lds[];
    uint32_t ix, iy;
    block_pixel<TW>(p.view.out_ext[0], p.view.out_ext[1], ix, iy);
    uint64_t words = 0;
    if (ix < p.view.out_ext[0] && iy < p.view.out_ext[1]) {
        f3 rd = ray_dir(p.cam, p.view, ix, iy);
        f3 c;
        if (TRAV == 2) {  // esvo
            uint32_t* stk_parent = lds + threadIdx.x;
            float* stk_tmax = reinterpret_cast<float*>(lds + p.stack_slots * kBlockThreads) + threadIdx.x;
            for (uint32_t s = 0; s < p.stack_slots; ++s) {
                stk_parent[s * kBlockThreads] = 0u;
                stk_tmax[s * kBlockThreads] = 0.0f;
            }
            f3 ro = add(p.cam.translation, splat(1.0f));
            float t0, t1;
            aabb_intersect(splat(1.0f), splat(2.0f), ro, rd, t0, t1);
            ro = add(ro, smul(gmax(t0, 0.0f), rd));
            c = esvo_trace<COUNT>(p, ro, rd, stk_parent, stk_tmax, words);
        } else if (TRAV == 3) {  // svo-df
            uint32_t* stk_node = lds + threadIdx.x;
            uint32_t* stk_child = lds + p.stack_slots * kBlockThreads + threadIdx.x;
            c = df_trace<COUNT>(p, p.cam.translation, rd, stk_node, stk_child, words);
        } else if (TRAV == 1) {
            c = naive_trace<COUNT>(p, p.cam.translation, rd, words);
        } else {
            c = rope_trace<COUNT>(p, p.cam.translation, rd, words);
        }
        if (!COUNT) p.out[(uint64_t)iy * p.out_pitch + ix] = pack_rgba(muls(c, emission_coeff(p.view, rd)));
    }
    if (COUNT) add_words(p.words, words);
}

size_t traversal_lds_bytes(int traversal, uint32_t stack_slots) {
    if (traversal == 2 || traversal == 3) return (size_t)2 * stack_slots * kBlockThreads * sizeof(uint32_t);
    return 0;
}

template <int TW, bool COUNT>
static hipError_t launch_tw(int traversal, const RenderParams& p, hipStream_t s) {
    constexpr uint32_t TH = 64 / TW;
    const uint32_t W = p.view.out_ext[0], H = p.view.out_ext[1];
    const uint32_t blocks = ((W + 2 * TW - 1) / (2 * TW)) * ((H + 2 * TH - 1) / (2 * TH));
    if (blocks == 0) return hipSuccess;
    const size_t lds = traversal_lds_bytes(traversal, p.stack_slots);
    switch (traversal) {
    case 0:
        hipLaunchKernelGGL((k_dda<TW, COUNT>), dim3(blocks), dim3(256), 0, s, p);
        break;
    case 1:
        hipLaunchKernelGGL((k_svo<1, TW, COUNT>), dim3(blocks), dim3(256), 0, s, p);
        break;
    case 2:
        hipLaunchKernelGGL((k_svo<2, TW, COUNT>), dim3(blocks), dim3(256), lds, s, p);
        break;
    case 3:
        hipLaunchKernelGGL((k_svo<3, TW, COUNT>), dim3(blocks), dim3(256), lds, s, p);
        break;
    case 4:
        hipLaunchKernelGGL((k_svo<4, TW, COUNT>), dim3(blocks), dim3(256), 0, s, p);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_traversal(int traversal, int tile_w, bool count_words, const RenderParams& p, hipStream_t s) {
    if (count_words) {
        if (tile_w == 16) return launch_tw<16, true>(traversal, p, s);
        if (tile_w == 32) return launch_tw<32, true>(traversal, p, s);
        return launch_tw<8, true>(traversal, p, s);
    }
    if (tile_w == 16) return launch_tw<16, false>(traversal, p, s);
    if (tile_w == 32) return launch_tw<32, false>(traversal, p, s);
    return launch_tw<8, false>(traversal, p, s);
}

// ---------------------------------------------------------------------------
// Grid upload: linear z-layers -> 4x4x4 bricks + occupancy
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_brick_layer(const uint32_t* __restrict__ staging, uint32_t layers, uint32_t bz,
                                                     uint32_t* __restrict__ bricks, uint8_t* __restrict__ occ,
                                                     uint32_t X, uint32_t Y, uint32_t nbx, uint32_t nby) {
    // one thread per brick of this brick-layer; bricks of one layer are contiguous
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nbx * nby) return;
    const uint32_t bx = i % nbx, by = i / nbx;
    const uint64_t b = ((uint64_t)bz * nby + by) * nbx + bx;
    uint32_t any = 0;
    for (uint32_t k = 0; k < 64; ++k) {
        const uint32_t x = bx * 4 + (k & 3), y = by * 4 + ((k >> 2) & 3), zl = k >> 4;
        uint32_t v = 0;
        if (x < X && y < Y && zl < layers) v = staging[((uint64_t)zl * Y + y) * X + x];
        bricks[b * 64 + k] = v;
        any |= v & 0xFFFFFFu;
    }
    occ[b] = any ? 1 : 0;
}

hipError_t launch_brick_layer(const uint32_t* staging, uint32_t layers, uint32_t bz, uint32_t* bricks, uint8_t* occ,
                              const uint32_t dims[3], const uint32_t nb[3], hipStream_t s) {
    const uint32_t n = nb[0] * nb[1];
    hipLaunchKernelGGL(k_brick_layer, dim3((n + 255) / 256), dim3(256), 0, s, staging, layers, bz, bricks, occ,
                       dims[0], dims[1], nb[0], nb[1]);
    return hipGetLastError();
}

}  // namespace xd
