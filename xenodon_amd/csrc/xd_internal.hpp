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
// LDS stacks: every wavefront owns slots x 64 entries per array, lane-strided
constexpr uint32_t kWaveLanes = 64;
// tagged child words of the derived octree layout
constexpr uint32_t kTagLeaf = 0x80000000u;   // the child is a leaf
constexpr uint32_t kTagEmpty = 0x40000000u;  // ... whose RGB is zero
constexpr uint32_t kIndexMask = 0x3FFFFFFFu;
// Tagged child words hold a 30-bit index with both tags.  Trees of 2^30 to
// 2^31 nodes use wide tags: a 31-bit index and the leaf tag only (no empty
// tag; such leaves load their colour, which is 0).
constexpr uint64_t kMaxNarrowTagNodes = 1ull << 30;
constexpr uint64_t kMaxTaggedNodes = 1ull << 31;
constexpr uint32_t kWideIndexMask = 0x7FFFFFFFu;
// Trees of fewer than 2^25 nodes whose depth fields are below 32 keep, in
// bits 25..29 of each leaf's child words (its ropes), the depth field of the
// node the word points to, so svo-rope learns the neighbour's depth without
// loading it.  Such a rope word's index is its low 25 bits.
constexpr uint32_t kRopeDepthShift = 25;
constexpr uint32_t kRopeIndexMask = (1u << kRopeDepthShift) - 1u;
// top tree levels numbered breadth-first and staged in LDS: levels 0..3 hold
// at most 1 + 8 + 64 + 512 = 585 nodes (18.7 KB of child words)
constexpr uint32_t kTopLevels = 4;
// Start table of a tagged octree: for each cell of the level-kTableLevel grid
// (256^3 cells) the node that find() reaches at that level, or the leaf that
// ends the walk above it, with its level.  128 MiB per octree volume.  On the
// 512^3 octree (depth 9) svo-naive ran at 743 / 807 / 880 / 906 Mrays/s plain
// and 1120 / 1286 / 1433 / 1393 batched with tables of levels 6 / 7 / 8 / 9
// (2 MiB / 16 MiB / 128 MiB / 1 GiB).
constexpr uint32_t kTableLevel = 8;
constexpr uint64_t kTableCells = 1ull << (3 * kTableLevel);
constexpr uint64_t kTableBytes = kTableCells * 8;

// Everything one traversal launch needs; passed by value as the kernel
// argument (kernarg segment), so a launch needs no device-side setup.
struct RenderParams {
    Camera cam;
    View view;
    uint32_t* out;
    uint32_t out_pitch;
    uint32_t stack_slots;  // LDS stack entries this tree needs per ray (svo-df, esvo)
    uint32_t lds_slots;    // stack entries per ray in the launch's LDS layout (>= stack_slots)
    uint64_t cap;          // iteration cap (never reached on well-formed input)
    // grid volume (dda): 4x4x4 bricks of packed RGBA8, one occupancy byte per
    // brick, and a bitmask of non-empty macro cells (2^macro_shift voxels a
    // side) that the kernel keeps in LDS
    const uint32_t* bricks;
    const uint8_t* occ;
    const uint32_t* macro;
    uint32_t dims[3];
    uint32_t nb[3];
    uint32_t mn[3];
    uint32_t macro_shift;
    uint32_t macro_words;
    uint32_t wide;  // 64-bit brick addressing: padded texel count >= 2^32, or nbz*nby or nbx >= 2^24
    // octree volume: the .svo words (node_count * 10), used only for trees
    // too large to tag; or the tagged layout: child words + {color, depth}
    const uint32_t* nodes;
    const uint32_t* cw;
    const uint32_t* meta;  // tagged layout: node_count colours, then node_count is_leaf_depth words
    uint32_t root_tag;
    uint32_t node_count;
    uint32_t top_nodes;  // child words of nodes [0, top_nodes) are staged in LDS
    uint32_t depth_exact;  // every internal node's depth field equals its depth in the tree
    const uint2* table;    // start table (tagged layout), kTableCells entries {child word, level}
    uint32_t rope_depth;   // leaves' child words carry their targets' depth fields (kRopeDepthShift)
    uint32_t index_mask;   // index bits of a tagged child word (kIndexMask, kWideIndexMask, kRopeIndexMask)
    uint32_t empty_tag;    // kTagEmpty, or 0 for wide tags
    uint32_t empty_start;  // wide tags: child indices from here on are empty leaves (else 0xFFFFFFFF)
    uint32_t leaf_base;    // tagged layout: first leaf; rope k of leaf l >= leaf_base is word
                           // leaf_base*8 + k*(node_count - leaf_base) + (l - leaf_base) of cw
    uint32_t xcd_chunks;  // plain launches: block runs per XCD (0: hardware order)
    unsigned long long* words;  // word counter (instrumented variant only)
};

// Launch one traversal over the output region.  tile_w in {8, 16, 32}
// selects the per-wavefront pixel tile (8x8, 16x4, 32x2).
hipError_t launch_traversal(int traversal, int tile_w, bool count_words, const RenderParams& p, hipStream_t s);

// Bytes of dynamic LDS a traversal needs per block of `threads` threads.
size_t traversal_lds_bytes(int traversal, uint32_t stack_slots, uint32_t top_nodes, uint32_t threads);

// Largest macro-cell bitmask the DDA kernel stages in LDS (16 KiB).
constexpr uint32_t kMaxMacroWords = 4096;

// Build the macro-cell occupancy bitmask from the brick occupancy bytes.
hipError_t launch_macro_mask(const uint8_t* occ, const uint32_t nb[3], uint32_t macro_shift, const uint32_t mn[3],
                             uint32_t* macro, uint32_t words, hipStream_t s);

// Grid upload: rewrite `layers` z-layers (X*Y texels each, x fastest) held in
// `staging` into bricks of brick-layer `bz`, and set the occupancy bytes.
hipError_t launch_brick_layer(const uint32_t* staging, uint32_t layers, uint32_t bz, uint32_t* bricks, uint8_t* occ,
                              const uint32_t dims[3], const uint32_t nb[3], hipStream_t s);

}  // namespace xd
