// Octree traversals on the derived device layout (tagged child words).
//
// Layout: child_w[n*8 + k] = index of child k | kTagLeaf if that child is a
// leaf | kTagEmpty if it is a leaf whose RGB is zero; meta[n] = {color,
// is_leaf_depth}.  A leaf's child words are its ropes, tagged the same way.
// Trees above 2^30 nodes carry 31-bit indices and no empty tag; their empty
// leaves are the indices from a threshold on (empty_mask and empty_min of
// FastSvo cover both).
// The traversal arithmetic is the shaders' (see xd_svo_ref.hpp); only loads
// change:
//   * "is this node a leaf?" comes with the pointer that led to it, so a
//     descent costs one dependent load per level instead of two;
//   * the colour of an empty leaf is known to be 0 without a load (the sum
//     still adds 0 * step, exactly as the shader does);
//   * svo-df fetches a node's eight child words with two 16-byte loads and
//     then steps through its children from registers.
#pragma once

#include "xd_svo_ref.hpp"

namespace xd {

struct FastSvo {
    const uint32_t* cw;
    const uint32_t* meta;  // colours of nodes [0, n), then their is_leaf_depth words
    uint32_t root_tag;     // kTagLeaf / kTagEmpty of node 0
    const uint32_t* top;   // LDS copy of the child words of nodes [0, n_top)
    uint32_t n_top;
    uint32_t n;            // node count (indices read back from a stack are clamped to it)
    uint32_t depth_exact;  // depth fields of internal nodes equal their tree depth
    const uint2* table;    // start table: kTableCells entries {child word, level}, or null
    uint32_t rope_depth;   // ropes carry their targets' depth fields (kRopeDepthShift)
    uint32_t index_mask;   // index bits of a child word
    // a child word wd names an empty leaf iff (wd & empty_mask) >= empty_min:
    // {kTagEmpty, kTagEmpty} where leaves carry the empty tag, else {index
    // mask, first empty leaf index} (see make_fast_svo)
    uint32_t empty_mask;
    uint32_t empty_min;
    uint32_t leaf_base;    // first leaf: the leaves' ropes are stored as six planes
};

// Rope k of a leaf (its k-th child word): in the plane of rope k for leaves
// from leaf_base on, else (a leaf root) in its own grouped record.
__device__ __forceinline__ uint32_t rope_at(const FastSvo& s, uint32_t leaf, uint32_t k) {
    const uint64_t i = leaf >= s.leaf_base
                           ? (uint64_t)s.leaf_base * 8 + (uint64_t)k * (s.n - s.leaf_base) + (leaf - s.leaf_base)
                           : (uint64_t)leaf * 8 + k;
    return *(const __attribute__((address_space(1))) uint32_t*)(s.cw + i);
}

// The iteration cap as a 32-bit count (a runaway guard that well-formed
// input never reaches; 32-bit counting saves two instructions per step).
__device__ __forceinline__ uint32_t cap32_of(uint64_t cap) { return cap > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)cap; }

// A child word names a leaf with zero RGB (whose colour need not be loaded).
__device__ __forceinline__ bool empty_leaf(const FastSvo& s, uint32_t wd) { return (wd & s.empty_mask) >= s.empty_min; }

// FastSvo of a tagged tree.  empty_tag = kTagEmpty marks empty leaves by
// that bit (wd & kTagEmpty is 0 or kTagEmpty, and those layouts set
// empty_start to 0xFFFFFFFF, which no index reaches); empty_tag = 0 by the
// index range from empty_start on.  One comparison covers both.
__device__ __forceinline__ FastSvo make_fast_svo(const uint32_t* cw, const uint32_t* meta, uint32_t root_tag,
                                                 const uint32_t* top, uint32_t n_top, uint32_t n, uint32_t depth_exact,
                                                 const uint2* table, uint32_t rope_depth, uint32_t index_mask,
                                                 uint32_t empty_tag, uint32_t empty_start, uint32_t leaf_base) {
    return FastSvo{cw, meta, root_tag, top, n_top, n, depth_exact, table, rope_depth, index_mask,
                   empty_tag ? empty_tag : index_mask, empty_tag ? empty_tag : empty_start, leaf_base};
}

// Child word k of node n: from LDS for the breadth-first top levels.
__device__ __forceinline__ uint32_t cw_at(const FastSvo& s, uint32_t n, uint32_t k) {
    if (n < s.n_top) return *(const __attribute__((address_space(3))) uint32_t*)(s.top + n * 8 + k);
    return *(const __attribute__((address_space(1))) uint32_t*)(s.cw + (uint64_t)n * 8 + k);
}

// GLSL mod(x, y) for y = 2^k: x / y == x * 2^-k exactly (no rounding in
// either), so the division is a multiplication by the exact reciprocal.
__device__ __forceinline__ float pmod(float x, float y, float inv_y) { return x - y * __builtin_floorf(x * inv_y); }

__device__ __forceinline__ uint32_t colour_if(const FastSvo& s, uint32_t node, uint32_t tag) {
    return empty_leaf(s, tag) ? 0u : s.meta[node];
}

// The child choices of find() at every level, packed: bit (22 - l) of
// level_bits(p) is the comparison p >= offset + extent made at level l.  At
// level l, offset + extent = (2m + 1) * 2^-(l+1) for the integer m that the
// earlier choices fix, and offset is exactly that dyadic value (sums of
// powers of two below 1 are exact in FP32 over the first 23 levels).  For p
// in [0, 1) the comparison is therefore bit 0 of floor(p * 2^(l+1)), which is
// bit (22 - l) of floor(p * 2^23); p < 0 (or NaN) fails every comparison and
// p >= 1 passes every one.
__device__ __forceinline__ uint32_t level_bits(float p) {
    if (!(p > 0.0f)) return 0u;
    if (!(p < 1.0f)) return (1u << 23) - 1u;
    return (uint32_t)(p * 8388608.0f);
}

// svo_naive.comp:8-27 with tagged pointers.  `tag` returns the found node's
// tags.  With a start table the walk begins at the node the table holds for
// the point's level-kTableLevel cell (found from the point's level bits): the same node
// the first six levels of the walk reach, with the same offset and extent,
// which are exact dyadic values.
__device__ __forceinline__ uint32_t fast_find(const FastSvo& s, f3 pos, f3& base, float& side, uint32_t& tag) {
    float extent = 1.0f;
    uint32_t index = 0;
    uint32_t level = 0;
    f3 offset = splat(0.0f);
    tag = s.root_tag;
    if (s.table) {
        const uint32_t bx = level_bits(pos.x), by = level_bits(pos.y), bz = level_bits(pos.z);
        constexpr uint32_t sh = 23u - kTableLevel;
        const uint2 e = s.table[((bz >> sh) << (2 * kTableLevel)) | ((by >> sh) << kTableLevel) | (bx >> sh)];
        index = e.x & s.index_mask;
        tag = e.x;
        level = e.y;
        extent = __builtin_ldexpf(1.0f, -(int)level);
        const uint32_t c = 23u - level;  // 32-bit shifts: level >= 1 here, or the bits are unused
        offset = level == 0 ? splat(0.0f)
                            : smul(extent, mk((float)(bx >> c), (float)(by >> c), (float)(bz >> c)));
    }
    for (;; ++level) {
        if ((tag & kTagLeaf) || level > kMaxLevels) {
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
        const uint32_t wd = cw_at(s, index, child);
        index = wd & s.index_mask;
        tag = wd;
    }
}

// svo_naive.comp:29-71
__device__ __forceinline__ f3 fast_naive(const FastSvo& s, uint64_t cap, f3 ro, f3 rd) {
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
    const uint32_t cap32 = cap32_of(cap);
    uint32_t it = 0;
    while (t < t_max) {
        if (++it > cap32) break;
        f3 pt = add(smul(t, rd), ro);
        f3 offset;
        float side;
        uint32_t tag;
        uint32_t node = fast_find(s, pt, offset, side, tag);
        f3 node_min = sub(mul(offset, rrd), bias);
        f3 node_max = sub(mul(add(offset, splat(side)), rrd), bias);
        float u_min = max_elem(vmin(node_min, node_max));
        float u_max = min_elem(vmax(node_min, node_max));
        u_min = gmax(u_min, 0.0f);
        float step = gmax(u_max - u_min, min_step);
        t = t + step;
        total = add(total, muls(unorm_rgb(colour_if(s, node, tag)), step));
    }
    return total;
}

__device__ __forceinline__ uint32_t pick8(const uint4& lo, const uint4& hi, uint32_t ci) {
    const uint32_t a = (ci & 1u) ? lo.y : lo.x, b = (ci & 1u) ? lo.w : lo.z;
    const uint32_t c = (ci & 1u) ? hi.y : hi.x, d = (ci & 1u) ? hi.w : hi.z;
    const uint32_t e = (ci & 2u) ? b : a, f = (ci & 2u) ? d : c;
    return (ci & 4u) ? f : e;
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint4 to_uint4(u32x4 v) { return make_uint4(v.x, v.y, v.z, v.w); }

__device__ __forceinline__ void load8(const FastSvo& s, uint32_t node, uint4& lo, uint4& hi) {
    // LDS and global loads kept in their own address spaces (a pointer
    // selected between them would become a slower flat load)
    if (node < s.n_top) {
        const __attribute__((address_space(3))) u32x4* p =
            (const __attribute__((address_space(3))) u32x4*)(s.top + node * 8);
        lo = to_uint4(p[0]);
        hi = to_uint4(p[1]);
    } else {
        const __attribute__((address_space(1))) u32x4* p =
            (const __attribute__((address_space(1))) u32x4*)(s.cw + (uint64_t)node * 8);
        lo = to_uint4(p[0]);
        hi = to_uint4(p[1]);
    }
}

// Slab values of the three planes per axis shared by a node's eight child
// boxes: p0 = corner, p1 = corner + side, p2 = p1 + side; T = p * rrd - bias.
// Child box k has, per axis, box_min = T[b] and box_max = T[b + 1] with b the
// axis bit of k -- the same floats svo_df.comp computes from pos and pos+side.
struct Planes {
    float x0, x1, x2, y0, y1, y2, z0, z1, z2;
};

__device__ __forceinline__ Planes node_planes(f3 corner, float side, f3 rrd, f3 bias) {
    Planes P;
    const float px1 = corner.x + side, py1 = corner.y + side, pz1 = corner.z + side;
    P.x0 = corner.x * rrd.x - bias.x;
    P.x1 = px1 * rrd.x - bias.x;
    P.x2 = (px1 + side) * rrd.x - bias.x;
    P.y0 = corner.y * rrd.y - bias.y;
    P.y1 = py1 * rrd.y - bias.y;
    P.y2 = (py1 + side) * rrd.y - bias.y;
    P.z0 = corner.z * rrd.z - bias.z;
    P.z1 = pz1 * rrd.z - bias.z;
    P.z2 = (pz1 + side) * rrd.z - bias.z;
    return P;
}

// t_min / t_max of child box ci of a node: the slab values of node_planes for
// the child's side of each axis (same operations, same floats).
__device__ __forceinline__ void child_trange(f3 corner, float side, f3 rrd, f3 bias, uint32_t ci, float& t_min,
                                             float& t_max) {
    const float px = (ci & 4u) ? corner.x + side : corner.x;
    const float py = (ci & 2u) ? corner.y + side : corner.y;
    const float pz = (ci & 1u) ? corner.z + side : corner.z;
    const float ax = px * rrd.x - bias.x, bx = (px + side) * rrd.x - bias.x;
    const float ay = py * rrd.y - bias.y, by = (py + side) * rrd.y - bias.y;
    const float az = pz * rrd.z - bias.z, bz = (pz + side) * rrd.z - bias.z;
    t_min = max_elem(mk(gmin(ax, bx), gmin(ay, by), gmin(az, bz)));
    t_max = min_elem(mk(gmax(ax, bx), gmax(ay, by), gmax(az, bz)));
}

// Bit k set when child box k of the node is hit (t_min < t_max && t_max > 0,
// svo_df.comp:25-32).  Per axis the low child spans [T0, T1] and the high
// child [T1, T2]; min/max of finite floats are exact, so these are the same
// comparisons as the shader's per-child test.  The two comparisons are taken
// as one, t_max > max(t_min, 0): equal for non-NaN values, and t_min is NaN
// only when all three of its axis minima are, whose axis pairs then give a
// NaN t_max too, so both forms are false there.  The 0 joins the y/z partial
// maxima that the eight children share.
__device__ __forceinline__ uint32_t child_hits(const Planes& P) {
    const float xa0 = gmin(P.x0, P.x1), xa1 = gmax(P.x0, P.x1), xb0 = gmin(P.x1, P.x2), xb1 = gmax(P.x1, P.x2);
    const float ya0 = gmin(P.y0, P.y1), ya1 = gmax(P.y0, P.y1), yb0 = gmin(P.y1, P.y2), yb1 = gmax(P.y1, P.y2);
    const float za0 = gmin(P.z0, P.z1), za1 = gmax(P.z0, P.z1), zb0 = gmin(P.z1, P.z2), zb1 = gmax(P.z1, P.z2);
    uint32_t m = 0;
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) {
        const float lo_yz = gmax(gmax((k & 2u) ? yb0 : ya0, (k & 1u) ? zb0 : za0), 0.0f);
        const float lo = gmax((k & 4u) ? xb0 : xa0, lo_yz);
        const float hi = gmin((k & 4u) ? xb1 : xa1, gmin((k & 2u) ? yb1 : ya1, (k & 1u) ? zb1 : za1));
        m |= (hi > lo) ? (1u << k) : 0u;
    }
    return m;
}

// svo_df.comp:6-66.  The shader walks the children of a node in index order,
// testing each box, and keeps (node, child index) on a stack.  Here a node's
// eight box tests are done at once into a hit mask, and only hit children are
// visited -- still in index order, so leaves are summed in the shader's order
// with the shader's t values.  A stack entry is (node, children still to
// visit, last visited child, depth); a node with none left is not pushed (the
// shader pops it and finds nothing more to do).  Each iteration visits one hit child and then
// enters at most one node -- the child (descent) or the nearest ancestor with
// children left (pop) -- through one shared path, so lanes that descend and
// lanes that pop wait on the same loads.  The child position the shader
// carries in `pos` is recomputed where it is needed: on descent (corner +
// side on the set axes) and on pop (the shader's pos - mod(pos, 2 * side),
// svo_df.comp:50-61, from a point inside the node being finished, which lies
// inside every ancestor).  The popped node's depth is the depth field the
// shader reads (svo_df.comp:58); when every depth field equals the node's
// depth in the tree, the depth kept on the stack is that same value.  When
// they do not, a popped node can be placed elsewhere than where it was
// entered, so the walk follows the shader literally: every non-last child
// pushes, and a popped node's remaining children are tested again.  EXACT
// selects the first walk (the tree's depth fields are exact, s.depth_exact);
// each walk is built on its own so that neither carries the other's code.
template <bool EXACT>
__device__ __forceinline__ f3 fast_df(const FastSvo& s, uint64_t cap, uint32_t slots, f3 ro, f3 rd,
                                      uint32_t* stk_node, uint32_t* stk_child) {
    f3 rrd = divv(splat(1.0f), rd);
    f3 bias = mul(rrd, ro);
    const int cap_sp = (int)slots;
    int sp = 0;
    uint32_t node = 0, depth = 0;
    f3 corner = splat(0.0f);
    float side = 0.5f;
    f3 total = splat(0.0f);
    uint4 lo, hi;
    load8(s, node, lo, hi);
    uint32_t mask = child_hits(node_planes(corner, side, rrd, bias));
    const uint32_t cap32 = cap32_of(cap);
    uint32_t it = 0;
    for (;;) {
        bool pop = true;
        if (mask != 0) {
            if (++it > cap32) break;
            const uint32_t ci = (uint32_t)__builtin_ctz(mask);
            mask &= mask - 1u;
            bool push;
            const uint32_t wd = pick8(lo, hi, ci);
            const bool bx = (ci & 4u) != 0, by = (ci & 2u) != 0, bz = (ci & 1u) != 0;
            if (wd & kTagLeaf) {
                if (!empty_leaf(s, wd)) {  // an empty leaf adds exactly +0
                    float t_min, t_max;
                    child_trange(corner, side, rrd, bias, ci, t_min, t_max);
                    total = add(total, muls(unorm_rgb(s.meta[wd & s.index_mask]), t_max - gmax(t_min, 0.0f)));
                }
                if (mask != 0) continue;  // next sibling, from registers
            } else if (!(push = EXACT ? mask != 0 : ci != 7u) || sp < cap_sp) {
                if (push) {
                    stk_node[sp * kWaveLanes] = node;
                    stk_child[sp * kWaveLanes] = mask | (ci << 8) | (depth << 11);
                    ++sp;
                }
                corner = mk(bx ? corner.x + side : corner.x, by ? corner.y + side : corner.y,
                            bz ? corner.z + side : corner.z);
                side = side * 0.5f;
                node = wd & s.index_mask;
                depth = depth + 1u;
                pop = false;
            } else if (mask != 0) {
                continue;  // stack full: the shader skips this child
            }
        }
        if (pop) {
            if (sp == 0) break;
            --sp;
            node = min(stk_node[sp * kWaveLanes], s.n - 1u);
            const uint32_t e = stk_child[sp * kWaveLanes];
            mask = e & 0xFFu;
            depth = (e >> 11) & 0xFFu;
            const uint32_t d = EXACT ? depth : (s.meta[(uint64_t)s.n + node] & kDepthMask);
            const f3 pos = mk(corner.x + side, corner.y + side, corner.z + side);
            side = __builtin_ldexpf(1.0f, -(int)d) * 0.5f;
            if (EXACT) {
                // The tree's geometry: pos and every corner are multiples of
                // 2^-23 in [0, 1] (at most 23 levels), and s2 = 2^-d.  Then
                // pos * 2^d, its floor times s2 and both subtractions of
                // pos - mod(pos, s2) are exact, so that corner is the
                // product s2 * floor(pos * 2^d).
                const float s2 = side * 2.0f, inv = __builtin_ldexpf(1.0f, (int)d);
                corner = mk(s2 * __builtin_floorf(pos.x * inv), s2 * __builtin_floorf(pos.y * inv),
                            s2 * __builtin_floorf(pos.z * inv));
            } else {
                // side is a power of two, so mod(pos, 2 side) is exact with the reciprocal
                const float s2 = side * 2.0f;
                const float inv = __builtin_ldexpf(1.0f, -__builtin_amdgcn_frexp_expf(s2) + 1);
                corner = sub(pos, mk(pmod(pos.x, s2, inv), pmod(pos.y, s2, inv), pmod(pos.z, s2, inv)));
            }
        }
        load8(s, node, lo, hi);
        if (!pop) {
            mask = child_hits(node_planes(corner, side, rrd, bias));
        } else if (!EXACT) {
            // the shader tests the remaining children at the geometry the
            // depth field gives the popped node, which may differ from the
            // geometry they were first tested at
            const uint32_t ci = (stk_child[sp * kWaveLanes] >> 8) & 7u;
            mask = child_hits(node_planes(corner, side, rrd, bias)) & (0xFEu << ci) & 0xFFu;
        }
    }
    return total;
}

// svo_rope.comp:65-135 (find_relative inlined).  With s.rope_depth the depth
// field of a rope's target comes with the rope word, so a hop needs no load
// of the target's {color, depth} before its walk; the found leaf's colour is
// loaded together with the next rope.
__device__ __forceinline__ f3 fast_rope(const FastSvo& s, uint64_t cap, f3 ro, f3 rd) {
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
    const uint32_t rope_mask = s.index_mask;  // kRopeIndexMask when ropes carry depths
    f3 pos = add(ro, smul(t_min, rd));
    f3 offset;
    float side;
    uint32_t tag;
    uint32_t node = fast_find(s, pos, offset, side, tag);
    f3 node_min = sub(mul(offset, rrd), bias);
    f3 node_max = sub(mul(add(offset, splat(side)), rrd), bias);
    f3 far = vmax(node_min, node_max);
    float u_min = max_elem(vmin(node_min, node_max));
    float u_max = min_elem(far);
    float step = u_max - gmax(u_min, 0.0f);
    f3 total = muls(unorm_rgb(colour_if(s, node, tag)), step);
    f3 mask;
    uint32_t n = neighbor_index(nbx, nby, nbz, far, mask);
    uint32_t wd = rope_at(s, node, n);
    node = wd & rope_mask;
    tag = wd;
    offset = add(offset, muls(mul(mask, sgn), side));
    const uint32_t cap32 = cap32_of(cap);
    uint32_t it = 0;
    while (node != 0) {
        if (++it > cap32) break;
        pos = add(ro, smul(u_max, rd));
        // find_relative (svo_rope.comp:29-48) from the rope's target
        const uint32_t start = node;
        uint32_t start_col = 0;
        int d;
        if (s.rope_depth) {
            d = (int)((wd >> kRopeDepthShift) & 31u);
        } else {
            d = (int)(s.meta[(uint64_t)s.n + node] & kDepthMask);
            start_col = s.meta[node];
        }
        float extent = __builtin_ldexpf(1.0f, -d);
        const float inv = __builtin_ldexpf(1.0f, d);
        offset = sub(offset, mk(pmod(offset.x, extent, inv), pmod(offset.y, extent, inv), pmod(offset.z, extent, inv)));
        for (uint32_t level = 0;; ++level) {
            if ((tag & kTagLeaf) || level > kMaxLevels) break;
            extent = extent * 0.5f;
            bool mx = pos.x >= offset.x + extent;
            bool my = pos.y >= offset.y + extent;
            bool mz = pos.z >= offset.z + extent;
            uint32_t child = (mx ? 4u : 0u) + (my ? 2u : 0u) + (mz ? 1u : 0u);
            offset = add(offset, mul(mk(mx ? 1.0f : 0.0f, my ? 1.0f : 0.0f, mz ? 1.0f : 0.0f), splat(extent)));
            wd = cw_at(s, node, child);
            node = wd & s.index_mask;
            tag = wd;
        }
        side = extent;
        node_min = sub(mul(offset, rrd), bias);
        node_max = sub(mul(add(offset, splat(side)), rrd), bias);
        far = vmax(node_min, node_max);
        u_min = max_elem(vmin(node_min, node_max));
        u_max = min_elem(far);
        step = u_max - gmax(u_min, 0.0f);
        const uint32_t col =
            empty_leaf(s, tag) ? 0u : (node == start && !s.rope_depth ? start_col : s.meta[node]);
        n = neighbor_index(nbx, nby, nbz, far, mask);
        wd = rope_at(s, node, n);  // in flight together with the colour
        total = add(total, muls(unorm_rgb(col), step));
        node = wd & rope_mask;
        tag = wd;
        offset = add(offset, muls(mul(mask, sgn), side));
    }
    return total;
}

// esvo.comp:30-136
__device__ __forceinline__ f3 fast_esvo(const FastSvo& s, uint64_t cap, uint32_t slots, f3 ro, f3 rd,
                                        uint32_t* stk_parent, float* stk_tmax) {
    EsvoRay r = esvo_setup(ro, rd);
    float t_min = r.t_min, t_max = r.t_max, h = r.h;
    f3 pos = r.pos;
    uint32_t idx = r.idx, parent = 0, scale = 22;
    float scale_exp2 = 0.5f;
    f3 total = splat(0.0f);
    uint32_t pend_col = 0;  // the last leaf's term, added at the next leaf so that
    float pend_w = 0.0f;    // its colour load overlaps the walk in between
    const uint32_t cap32 = cap32_of(cap);
    uint32_t it = 0;
    while (scale < 23u) {
        if (++it > cap32) break;
        f3 t_corner = sub(mul(pos, r.t_coeff), r.t_bias);
        float tc_max = min_elem(t_corner);
        if (t_min <= t_max) {
            float tv_max = gmin(t_max, tc_max);
            if (t_min <= tv_max) {
                const uint32_t wd = cw_at(s, parent, idx ^ r.octant_mask);
                const uint32_t child = wd & s.index_mask;
                if (wd & kTagLeaf) {
                    const uint32_t col = colour_if(s, child, wd);
                    total = add(total, muls(unorm_rgb(pend_col), pend_w));  // +0 the first time
                    pend_col = col;
                    pend_w = tv_max - t_min;
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
            parent = slot < slots ? min(stk_parent[slot * kWaveLanes], s.n - 1u) : 0u;
            t_max = slot < slots ? stk_tmax[slot * kWaveLanes] : 0.0f;
            h = 0.0f;
        }
    }
    return add(total, muls(unorm_rgb(pend_col), pend_w));
}

}  // namespace xd
