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
#include "xd_trace.hpp"

namespace xd {

template <int TW, bool COUNT, bool WIDE>
__global__ __launch_bounds__(256) void k_dda(RenderParams p) {
    extern __shared__ uint32_t lds_macro[];  // macro_words, staged per block
    if (!COUNT) {
        for (uint32_t i = threadIdx.x; i < p.macro_words; i += blockDim.x) lds_macro[i] = p.macro[i];
        __syncthreads();
    }
    uint32_t ix, iy;
    block_pixel<TW, 4>(p.view.out_ext[0], p.view.out_ext[1], p.xcd_chunks, ix, iy);
    uint64_t words = 0;
    if (ix < p.view.out_ext[0] && iy < p.view.out_ext[1]) {
        f3 rd = ray_dir(p.cam, p.view, ix, iy);
        float side = max_elem(mk((float)p.dims[0], (float)p.dims[1], (float)p.dims[2]));
        f3 ro = muls(p.cam.translation, side);
        float ec = emission_coeff(p.view, rd) / side;
        f3 total = dda_trace<COUNT, WIDE>(p, lds_macro, ro, rd, words);
        if (!COUNT) p.out[(uint64_t)iy * p.out_pitch + ix] = pack_rgba(muls(total, ec));
    }
    if (COUNT) add_words(p.words, words);
}

// Four wavefronts (a 2x2 arrangement of tiles) per block: the block stages
// the top-level child words in LDS once for four tiles.
template <int TRAV, int TW, bool COUNT, bool TAGGED>
__global__ __launch_bounds__(256) void k_svo(RenderParams p) {
    extern __shared__ uint32_t lds[];
    uint32_t* lds_top = lds + ((TRAV == 2 || TRAV == 3) ? 2 * p.lds_slots * blockDim.x : 0);
    if (TAGGED && !COUNT) {
        for (uint32_t i = threadIdx.x; i < p.top_nodes * 8; i += blockDim.x) lds_top[i] = p.cw[i];
        __syncthreads();
    }
    uint32_t ix, iy;
    block_pixel<TW, 4>(p.view.out_ext[0], p.view.out_ext[1], p.xcd_chunks, ix, iy);
    uint64_t words = 0;
    if (ix < p.view.out_ext[0] && iy < p.view.out_ext[1]) {
        f3 rd = ray_dir(p.cam, p.view, ix, iy);
        f3 c = svo_pixel<TRAV, COUNT, TAGGED>(p, rd, lds, lds_top, words);
        if (!COUNT) p.out[(uint64_t)iy * p.out_pitch + ix] = pack_rgba(muls(c, emission_coeff(p.view, rd)));
    }
    if (COUNT) add_words(p.words, words);
}

size_t traversal_lds_bytes(int traversal, uint32_t stack_slots, uint32_t top_nodes, uint32_t threads) {
    size_t b = (size_t)top_nodes * 8 * sizeof(uint32_t);
    if (traversal == 2 || traversal == 3) b += (size_t)2 * stack_slots * threads * sizeof(uint32_t);
    return b;
}

template <int TW, bool COUNT, bool TAGGED>
static hipError_t launch_tw(int traversal, const RenderParams& p, hipStream_t s) {
    constexpr uint32_t TH = 64 / TW;
    const uint32_t W = p.view.out_ext[0], H = p.view.out_ext[1];
    const uint32_t blocks = ((W + 2 * TW - 1) / (2 * TW)) * ((H + 2 * TH - 1) / (2 * TH));
    if (blocks == 0) return hipSuccess;
    const size_t lds = traversal_lds_bytes(traversal, p.lds_slots, p.top_nodes, 256);
    switch (traversal) {
    case 0: {
        const size_t mlds = COUNT ? 0 : (size_t)p.macro_words * sizeof(uint32_t);
        if (p.wide)
            hipLaunchKernelGGL((k_dda<TW, COUNT, true>), dim3(blocks), dim3(256), mlds, s, p);
        else
            hipLaunchKernelGGL((k_dda<TW, COUNT, false>), dim3(blocks), dim3(256), mlds, s, p);
        break;
    }
    case 1:
        hipLaunchKernelGGL((k_svo<1, TW, COUNT, TAGGED>), dim3(blocks), dim3(256), lds, s, p);
        break;
    case 2:
        hipLaunchKernelGGL((k_svo<2, TW, COUNT, TAGGED>), dim3(blocks), dim3(256), lds, s, p);
        break;
    case 3:
        hipLaunchKernelGGL((k_svo<3, TW, COUNT, TAGGED>), dim3(blocks), dim3(256), lds, s, p);
        break;
    case 4:
        hipLaunchKernelGGL((k_svo<4, TW, COUNT, TAGGED>), dim3(blocks), dim3(256), lds, s, p);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <bool COUNT, bool TAGGED>
static hipError_t launch_tiles(int traversal, int tile_w, const RenderParams& p, hipStream_t s) {
    if (tile_w == 16) return launch_tw<16, COUNT, TAGGED>(traversal, p, s);
    if (tile_w == 32) return launch_tw<32, COUNT, TAGGED>(traversal, p, s);
    return launch_tw<8, COUNT, TAGGED>(traversal, p, s);
}

hipError_t launch_traversal(int traversal, int tile_w, bool count_words, const RenderParams& p, hipStream_t s) {
    const bool tagged = p.cw != nullptr;
    if (count_words)
        return tagged ? launch_tiles<true, true>(traversal, tile_w, p, s) : launch_tiles<true, false>(traversal, tile_w, p, s);
    return tagged ? launch_tiles<false, true>(traversal, tile_w, p, s) : launch_tiles<false, false>(traversal, tile_w, p, s);
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

__global__ __launch_bounds__(256) void k_macro_mask(const uint8_t* __restrict__ occ, uint32_t nbx, uint32_t nby,
                                                    uint32_t nbz, uint32_t shift, uint32_t mnx, uint32_t mny,
                                                    uint32_t mnz, uint32_t* __restrict__ macro, uint32_t words) {
    // one thread per bitmask word (32 macro cells)
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= words) return;
    const uint32_t bshift = shift - 2;  // bricks per macro cell side = 2^(shift-2)
    uint32_t bits = 0;
    for (uint32_t k = 0; k < 32; ++k) {
        const uint32_t cell = w * 32 + k;
        if (cell >= mnx * mny * mnz) break;
        const uint32_t cx = cell % mnx, cy = (cell / mnx) % mny, cz = cell / (mnx * mny);
        uint32_t any = 0;
        for (uint32_t bz = cz << bshift; bz < min(nbz, (cz + 1) << bshift) && !any; ++bz)
            for (uint32_t by = cy << bshift; by < min(nby, (cy + 1) << bshift) && !any; ++by)
                for (uint32_t bx = cx << bshift; bx < min(nbx, (cx + 1) << bshift); ++bx)
                    if (occ[((uint64_t)bz * nby + by) * nbx + bx]) {
                        any = 1;
                        break;
                    }
        bits |= any << k;
    }
    macro[w] = bits;
}

hipError_t launch_macro_mask(const uint8_t* occ, const uint32_t nb[3], uint32_t macro_shift, const uint32_t mn[3],
                             uint32_t* macro, uint32_t words, hipStream_t s) {
    hipLaunchKernelGGL(k_macro_mask, dim3((words + 255) / 256), dim3(256), 0, s, occ, nb[0], nb[1], nb[2],
                       macro_shift, mn[0], mn[1], mn[2], macro, words);
    return hipGetLastError();
}

hipError_t launch_brick_layer(const uint32_t* staging, uint32_t layers, uint32_t bz, uint32_t* bricks, uint8_t* occ,
                              const uint32_t dims[3], const uint32_t nb[3], hipStream_t s) {
    const uint32_t n = nb[0] * nb[1];
    hipLaunchKernelGGL(k_brick_layer, dim3((n + 255) / 256), dim3(256), 0, s, staging, layers, bz, bricks, occ,
                       dims[0], dims[1], nb[0], nb[1]);
    return hipGetLastError();
}

}  // namespace xd
