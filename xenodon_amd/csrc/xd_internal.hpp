// Internal interfaces between the C-ABI runtime and the traversal kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "xd_device.hpp"

namespace xd {

constexpr uint32_t kNodeWords = 10;  // children[8], color, is_leaf_depth
constexpr uint32_t kLeafMask = 0x80000000u;
constexpr uint32_t kDepthMask = 0x7FFFFFFFu;
constexpr uint32_t kMaxLevels = 64;  // find() descent cap (cyclic-tree guard)
constexpr uint32_t kBrick = 4;       // grid bricks are 4x4x4 texels (256 B)
constexpr uint32_t kBlockThreads = 256;

// Everything one traversal launch needs; passed by value as the kernel
// argument (kernarg segment), so a launch needs no device-side setup.
struct RenderParams {
    Camera cam;
    View view;
    uint32_t* out;
    uint32_t out_pitch;
    uint32_t stack_slots;  // LDS stack entries per ray (svo-df, esvo)
    uint64_t cap;          // iteration cap (never reached on well-formed input)
    // grid volume (dda): 4x4x4 bricks of packed RGBA8 + one occupancy byte per brick
    const uint32_t* bricks;
    const uint8_t* occ;
    uint32_t dims[3];
    uint32_t nb[3];
    // octree volume: node_count * 10 u32 words
    const uint32_t* nodes;
    unsigned long long* words;  // word counter (instrumented variant only)
};

// Launch one traversal over the output region.  tile_w in {8, 16, 32}
// selects the per-wavefront pixel tile (8x8, 16x4, 32x2).
hipError_t launch_traversal(int traversal, int tile_w, bool count_words, const RenderParams& p, hipStream_t s);

// Bytes of dynamic LDS a traversal needs per 256-thread block.
size_t traversal_lds_bytes(int traversal, uint32_t stack_slots);

// Grid upload: rewrite `layers` z-layers (X*Y texels each, x fastest) held in
// `staging` into bricks of brick-layer `bz`, and set the occupancy bytes.
hipError_t launch_brick_layer(const uint32_t* staging, uint32_t layers, uint32_t bz, uint32_t* bricks, uint8_t* occ,
                              const uint32_t dims[3], const uint32_t nb[3], hipStream_t s);

}  // namespace xd
