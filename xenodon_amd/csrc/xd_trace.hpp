// Per-ray traversal device code shared by the plain kernels (one launch
// per output) and the persistent batch kernel (xd_batch.hip).
#pragma once

#include "xd_internal.hpp"
#include "xd_svo_fast.hpp"

namespace xd {

// Pixel of this lane.  WPB = 4: a block is a 2x2 arrangement of wavefront
// tiles; WPB = 1: a block is one tile.  The 8 XCDs receive blocks
// round-robin (block b runs on XCD b % 8).  chunks = 0 keeps that order, so
// neighbouring blocks run on different XCDs; chunks = N >= 1 cuts the image's
// blocks (row-major) into 8N runs and gives XCD x runs x, x + 8, ..., so each
// XCD's L2 holds the volume behind a few bands while every XCD still gets a
// share of each part of the image (N = 1: one contiguous band per XCD).
template <int TW, int WPB>
__device__ __forceinline__ void block_pixel(uint32_t W, uint32_t H, uint32_t chunks, uint32_t& ix, uint32_t& iy) {
    constexpr uint32_t TH = 64 / TW;
    constexpr uint32_t BW = (WPB == 4 ? 2 : 1) * TW, BH = (WPB == 4 ? 2 : 1) * TH;
    const uint32_t nbx = (W + BW - 1) / BW;
    const uint32_t nblocks = nbx * ((H + BH - 1) / BH);
    const uint32_t b = blockIdx.x;
    uint32_t lb = b;
    if (chunks != 0) {
        const uint32_t run = nblocks / (8u * chunks);  // blocks per run
        if (run != 0 && b < run * 8u * chunks) {
            const uint32_t j = b >> 3;  // j-th block this XCD receives
            lb = ((j / run) * 8u + (b & 7u)) * run + j % run;
        }
    }
    const uint32_t bx = lb % nbx, by = lb / nbx;
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    ix = bx * BW + (WPB == 4 ? (w & 1u) * TW : 0u) + (lane % TW);
    iy = by * BH + (WPB == 4 ? (w >> 1) * TH : 0u) + (lane / TW);
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
//
// The step sequence (side_dist additions, masks, t) is the shader's, one
// rounding per operation.  Steps are produced in batches of four so that the
// four texel gathers of a batch are in flight together; their contributions
// are then added in step order, so the sum is the shader's sum.  Texels are
// fetched only inside macro cells whose bitmask bit (staged in LDS) says they
// hold a non-zero RGB; every skipped texel is zero and would add exactly +0.
// ---------------------------------------------------------------------------
// With PIPE the bitmask is read from device memory (for the persistent batch
// kernels, whose waves serve views of different volumes and stage nothing in
// LDS), and the macro words of the next batch of four steps are loaded while
// the current batch's texels are gathered: a batch waits on one memory round
// trip instead of two in a row.  The steps, the texels fetched and the order
// of the sum are unchanged.  Without PIPE the bitmask may be in LDS or in
// device memory.
// Bit of voxel (ux, uy, uz)'s macro cell in the bitmask.  A bitmask holds at
// most kMaxMacroWords * 32 = 2^17 cells, so every operand and partial result
// is below 2^24 and the 24-bit multiply-add (full rate) is exact.
__device__ __forceinline__ uint32_t macro_bit(uint32_t ux, uint32_t uy, uint32_t uz, uint32_t ms, uint32_t mnx,
                                              uint32_t mny) {
    return __umul24(__umul24(uz >> ms, mny) + (uy >> ms), mnx) + (ux >> ms);
}

// Texel index of voxel (ux, uy, uz) in the bricked grid, 32-bit form.  Grids
// take it only when nbz * nby < 2^24 and nbx < 2^24 (else the 64-bit form;
// see fill_params), so both multiplies' operands are below 2^24.
__device__ __forceinline__ uint32_t brick_texel32(uint32_t ux, uint32_t uy, uint32_t uz, uint32_t nbx, uint32_t nby,
                                                  uint32_t in_brick) {
    return ((__umul24(__umul24(uz >> 2, nby) + (uy >> 2), nbx) + (ux >> 2)) << 6) + in_brick;
}

template <bool COUNT, bool WIDE, bool PIPE = false>
__device__ __forceinline__ f3 dda_trace(const RenderParams& p, const uint32_t* __restrict__ lds_macro, f3 ro, f3 rd,
                                        uint64_t& words) {
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
    const uint32_t cap32 = cap32_of(p.cap);
    uint32_t it = 0;
    if (COUNT) {
        while (t < limit) {
            if (++it > cap32) break;
            bool mx = sd.x <= gmin(sd.y, sd.z);
            bool my = sd.y <= gmin(sd.z, sd.x);
            bool mz = sd.z <= gmin(sd.x, sd.y);
            t = min_elem(sd);
            ++words;
            sd = add(sd, mk(mx ? t_delta.x : 0.0f, my ? t_delta.y : 0.0f, mz ? t_delta.z : 0.0f));
        }
        return total;
    }
    const uint32_t X = p.dims[0], Y = p.dims[1], Z = p.dims[2];
    const uint32_t nbx = p.nb[0], nby = p.nb[1];
    const uint32_t ms = p.macro_shift, mnx = p.mn[0], mny = p.mn[1];
    bool live = true;
    if (PIPE) {
        typedef typename std::conditional<WIDE, uint64_t, uint32_t>::type Addr;
        const __attribute__((address_space(1))) uint32_t* gmacro =
            (const __attribute__((address_space(1))) uint32_t*)lds_macro;
        // the next four steps: texel addresses, lengths, and the macro word and
        // bit of each step's cell (word 0 outside the grid)
        struct Steps {
            Addr addr[4];
            float dts[4];
            uint32_t word[4];
            uint32_t bits;
            uint32_t n;
        };
        auto advance = [&](Steps& S) {
            S.bits = 0;
            S.n = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                S.addr[j] = 0;
                S.dts[j] = 0.0f;
                S.word[j] = 0;
                if (live && t < limit && ++it <= cap32) {
                    bool mx = sd.x <= gmin(sd.y, sd.z);
                    bool my = sd.y <= gmin(sd.z, sd.x);
                    bool mz = sd.z <= gmin(sd.x, sd.y);
                    float t0 = min_elem(sd);
                    if ((uint32_t)px < X && (uint32_t)py < Y && (uint32_t)pz < Z) {
                        const uint32_t ux = (uint32_t)px, uy = (uint32_t)py, uz = (uint32_t)pz;
                        const uint32_t bit = macro_bit(ux, uy, uz, ms, mnx, mny);
                        S.word[j] = gmacro[bit >> 5];
                        S.bits |= (bit & 31u) << (5 * j);
                        const uint32_t in_brick = (ux & 3u) + ((uy & 3u) << 2) + ((uz & 3u) << 4);
                        if (WIDE)
                            S.addr[j] = (((uint64_t)(uz >> 2) * nby + (uy >> 2)) * nbx + (ux >> 2)) * 64u + in_brick;
                        else
                            S.addr[j] = brick_texel32(ux, uy, uz, nbx, nby, in_brick);
                    }
                    S.dts[j] = t0 - t;
                    t = t0;
                    sd = add(sd, mk(mx ? t_delta.x : 0.0f, my ? t_delta.y : 0.0f, mz ? t_delta.z : 0.0f));
                    px += mx ? sx : 0;
                    py += my ? sy : 0;
                    pz += mz ? sz : 0;
                    ++S.n;
                } else {
                    live = false;
                }
            }
        };
        Steps cur;
        advance(cur);
        while (cur.n) {
            uint32_t tex[4];
#pragma unroll
            for (int j = 0; j < 4; ++j)
                tex[j] = ((cur.word[j] >> ((cur.bits >> (5 * j)) & 31u)) & 1u) ? p.bricks[cur.addr[j]] : 0u;
            Steps nxt;
            advance(nxt);  // its macro loads are in flight with this batch's gathers
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (tex[j] & 0xFFFFFFu) total = add(total, muls(unorm_rgb(tex[j]), cur.dts[j]));
            cur = nxt;
        }
        return total;
    }
    while (live) {
        constexpr int B = 4;
        typename std::conditional<WIDE, uint64_t, uint32_t>::type addr[B];
        float dts[B];
        bool need[B];
        // the steps as selects rather than branches: a step that does not run
        // (the ray ended, or the iteration cap) leaves the state as it is
        // (x + 0 = x for every side distance, which are +0 or more); its
        // bitmask read goes to word 0 and its texel is not fetched
#pragma unroll
        for (int j = 0; j < B; ++j) {
            const bool go = live && t < limit;
            it += go ? 1u : 0u;
            const bool act = go && it <= cap32;
            live = act;
            const bool mx = act && sd.x <= gmin(sd.y, sd.z);
            const bool my = act && sd.y <= gmin(sd.z, sd.x);
            const bool mz = act && sd.z <= gmin(sd.x, sd.y);
            const float t0 = min_elem(sd);
            const uint32_t ux = (uint32_t)px, uy = (uint32_t)py, uz = (uint32_t)pz;
            const bool inside = act && ux < X && uy < Y && uz < Z;
            const uint32_t bit = inside ? macro_bit(ux, uy, uz, ms, mnx, mny) : 0u;
            need[j] = inside && ((lds_macro[bit >> 5] >> (bit & 31u)) & 1u);
            const uint32_t in_brick = (ux & 3u) + ((uy & 3u) << 2) + ((uz & 3u) << 4);
            if (WIDE)
                addr[j] = (((uint64_t)(uz >> 2) * nby + (uy >> 2)) * nbx + (ux >> 2)) * 64u + in_brick;
            else
                addr[j] = brick_texel32(ux, uy, uz, nbx, nby, in_brick);
            dts[j] = act ? t0 - t : 0.0f;
            t = act ? t0 : t;
            sd = add(sd, mk(mx ? t_delta.x : 0.0f, my ? t_delta.y : 0.0f, mz ? t_delta.z : 0.0f));
            px += mx ? sx : 0;
            py += my ? sy : 0;
            pz += mz ? sz : 0;
        }
        uint32_t tex[B];
#pragma unroll
        for (int j = 0; j < B; ++j) tex[j] = need[j] ? p.bricks[addr[j]] : 0u;
#pragma unroll
        for (int j = 0; j < B; ++j)
            if (tex[j] & 0xFFFFFFu) total = add(total, muls(unorm_rgb(tex[j]), dts[j]));
    }
    return total;
}

// ---------------------------------------------------------------------------
// Octree kernels (main() of svo_naive/svo_df/esvo/svo_rope.comp)
//   TAGGED && !COUNT  fast traversals on the tagged layout (xd_svo_fast.hpp)
//   otherwise         reference-structured traversals (xd_svo_ref.hpp)
// ---------------------------------------------------------------------------
// LDS_TOP: the kernel may stage top-level child words in LDS (plain
// launches; only svo-rope's views do).  Where no view can, the traversal is
// built with no top levels at all, which drops the LDS-or-device-memory
// choice from every child-word read.
template <int TRAV, bool COUNT, bool TAGGED, bool LDS_TOP = true>
__device__ __forceinline__ f3 svo_pixel(const RenderParams& p, f3 rd, uint32_t* lds, const uint32_t* lds_top,
                                        uint64_t& words) {
    // this wavefront's stack region: two arrays of lds_slots x 64 lane-strided
    // words.  The stride is the launch's (all views of a batch share it), not
    // this tree's, so regions of different waves never overlap.
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    uint32_t* stk_a = lds + wave * 2u * p.lds_slots * kWaveLanes + lane;
    uint32_t* stk_b = stk_a + p.lds_slots * kWaveLanes;
    const uint32_t n_top = LDS_TOP && TRAV == 4 ? p.top_nodes : 0u;  // svo-rope
    const FastSvo fs = make_fast_svo(p.cw, p.meta, p.root_tag, lds_top, n_top, p.node_count, p.depth_exact, p.table, p.rope_depth,
                      p.index_mask, p.empty_tag, p.empty_start, p.leaf_base);
    if (TRAV == 2) {  // esvo
        float* stk_tmax = reinterpret_cast<float*>(stk_b);
        for (uint32_t s = 0; s < p.stack_slots; ++s) {
            stk_a[s * kWaveLanes] = 0u;
            stk_tmax[s * kWaveLanes] = 0.0f;
        }
        f3 ro = add(p.cam.translation, splat(1.0f));
        float t0, t1;
        aabb_intersect(splat(1.0f), splat(2.0f), ro, rd, t0, t1);
        ro = add(ro, smul(gmax(t0, 0.0f), rd));
        if (TAGGED && !COUNT) return fast_esvo(fs, p.cap, p.stack_slots, ro, rd, stk_a,
                                                stk_tmax);
        if (TAGGED) return ref_esvo<COUNT>(TaggedNodes{p.cw, p.meta, p.node_count, p.index_mask, p.leaf_base}, p.cap, p.stack_slots, ro, rd, stk_a, stk_tmax,
                                           words);
        return ref_esvo<COUNT>(RawNodes{p.nodes, p.node_count}, p.cap, p.stack_slots, ro, rd, stk_a, stk_tmax, words);
    }
    const f3 ro = p.cam.translation;
    if (TRAV == 3) {  // svo-df
        if (TAGGED && !COUNT)  // depth-field handling fixed per tree (a uniform choice)
            return p.depth_exact ? fast_df<true>(fs, p.cap, p.stack_slots, ro, rd, stk_a, stk_b)
                                 : fast_df<false>(fs, p.cap, p.stack_slots, ro, rd, stk_a, stk_b);
        if (TAGGED) return ref_df<COUNT>(TaggedNodes{p.cw, p.meta, p.node_count, p.index_mask, p.leaf_base}, p.cap, p.stack_slots, ro, rd, stk_a, stk_b, words);
        return ref_df<COUNT>(RawNodes{p.nodes, p.node_count}, p.cap, p.stack_slots, ro, rd, stk_a, stk_b, words);
    }
    if (TRAV == 1) {  // svo-naive
        if (TAGGED && !COUNT) return fast_naive(fs, p.cap, ro, rd);
        if (TAGGED) return ref_naive<COUNT>(TaggedNodes{p.cw, p.meta, p.node_count, p.index_mask, p.leaf_base}, p.cap, ro, rd, words);
        return ref_naive<COUNT>(RawNodes{p.nodes, p.node_count}, p.cap, ro, rd, words);
    }
    // svo-rope
    if (TAGGED && !COUNT) return fast_rope(fs, p.cap, ro, rd);
    if (TAGGED) return ref_rope<COUNT>(TaggedNodes{p.cw, p.meta, p.node_count, p.index_mask, p.leaf_base}, p.cap, ro, rd, words);
    return ref_rope<COUNT>(RawNodes{p.nodes, p.node_count}, p.cap, ro, rd, words);
}

}  // namespace xd
