// Runtime object layouts shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <unordered_set>

#include "xenodon.h"
#include "xd_internal.hpp"

struct xd_context {
    int device = 0;
    int tile_w = 8;  // per-wavefront tile width: 8 (8x8), 16 (16x4), 32 (32x2)
    int octree_layout = 0;  // 0 tagged (wide tags above 2^30 nodes), 1 .svo words, 2 wide tags always
    uint32_t xcd_chunks = 0;  // plain launches: block runs per XCD (see block_pixel)
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    unsigned long long* d_words = nullptr;
    // live objects created on this context; xd_close frees what is left
    std::unordered_set<struct xd_volume*> volumes;
    std::unordered_set<struct xd_frame*> frames;
    std::unordered_set<struct xd_batch*> batches;
    std::unordered_set<struct xd_server*> servers;
};

struct xd_volume {
    xd_context* ctx = nullptr;
    int kind = XD_VOLUME_GRID;
    uint64_t dims[3] = {0, 0, 0};
    uint64_t side = 0;
    uint64_t node_count = 0;
    uint32_t max_depth = 0;
    uint32_t is_rope = 0;
    uint64_t bytes = 0;
    // grid: 4x4x4 bricks + occupancy byte per brick
    uint32_t* d_bricks = nullptr;
    uint8_t* d_occ = nullptr;
    uint32_t nb[3] = {0, 0, 0};
    uint32_t* d_macro = nullptr;  // non-empty macro cells, 1 bit each
    uint32_t mn[3] = {0, 0, 0};
    uint32_t macro_shift = 0;
    uint32_t macro_words = 0;
    // octree: the tagged layout (child words + {color, depth}) when the tree
    // has at most 2^31 nodes, else the .svo words as they are
    uint32_t* d_nodes = nullptr;
    uint32_t* d_cw = nullptr;
    uint32_t* d_meta = nullptr;  // colours of all nodes, then their is_leaf_depth words
    uint2* d_table = nullptr;  // start table of the tagged layout (see build_start_table)
    uint32_t rope_depth = 0;   // ropes carry their targets' depth fields (see pack_rope_depths)
    uint32_t index_mask = 0;   // index bits of the tagged child words
    uint32_t empty_tag = 0;    // kTagEmpty, or 0 for wide tags
    uint32_t empty_start = 0xFFFFFFFFu;  // wide tags: empty leaves are the indices from here on
    uint32_t root_tag = 0;
    uint32_t depth_exact = 0;  // internal nodes' depth fields match their depth (see octree_depths_exact)
    uint32_t top[xd::kTopLevels] = {0, 0, 0, 0};  // nodes in breadth-first levels 0..L
    uint32_t leaf_base = 0;    // tagged layout: first leaf; the leaves' ropes are stored as planes
};

struct xd_frame {
    xd_context* ctx = nullptr;
    uint32_t w = 0, h = 0;
    uint32_t* d_pixels = nullptr;
};

namespace xd {

int fail(int code, const char* fmt, ...);
int fail_hip(hipError_t e, const char* what);
void clear_error();
const char* traversal_name(int t);
int check_traversal(int traversal, const xd_volume* vol);
int check_view(const xd_uniforms* u);
// Macro-cell side (log2) and per-axis counts of a grid; returns bitmask words.
uint32_t macro_geometry(const uint64_t dims[3], uint32_t* shift, uint32_t mn[3]);
// HBM bytes xd_volume_from_grid / xd_volume_from_svo keep allocated.
uint64_t grid_footprint(const uint64_t dims[3]);
uint64_t svo_footprint(uint64_t node_count, bool raw);
uint64_t iteration_cap(int traversal, const xd_volume* vol, const uint32_t model_dim[3]);
void fill_params(RenderParams& p, const xd_volume* vol, int traversal, const xd_camera* cam, const xd_uniforms* u);
// Host-side derivation of the tagged, sibling-grouped device layout
// (child_w: n*8 words, meta: n*2 words) from .svo node words.
// top[L] = number of nodes in levels 0..L, which hold indices [0, top[L]).
// wide: 31-bit indices and the leaf tag only (trees above 2^30 nodes).  The
// empty leaves take the last indices; with wide tags *empty_start is the
// first of them, which is how a child word marks an empty leaf (else
// *empty_start = 0xFFFFFFFF and the empty tag does).  tree: octree_is_tree(w, n) holds, so the
// numbering runs subtree-parallel.
// *leaf_base: the first leaf's index (internal nodes and the root come first).
void build_tagged_layout(const uint32_t* w, uint64_t n, uint32_t* child_w, uint32_t* meta, uint32_t top[kTopLevels],
                         bool wide, bool tree, uint32_t* empty_start, uint32_t* leaf_base);
// The device form of child_w: the words of nodes [0, leaf_base) as they are,
// then the six ropes of the leaves as planes, rope k of leaf leaf_base + j at
// leaf_base*8 + k*(n - leaf_base) + j (rope_plane_words words in all).
uint64_t rope_plane_words(uint64_t n, uint32_t leaf_base);
void to_rope_planes(const uint32_t* child_w, uint64_t n, uint32_t leaf_base, uint32_t* out);
// True when no internal node's child word names the root and no node is
// named by two child words of internal nodes (every node has one parent).
bool octree_is_tree(const uint32_t* w, uint64_t n);
// True when every internal node reachable from the root is reached only at
// one depth and its is_leaf_depth depth field equals that depth (root = 0),
// as in every tree the converter writes.  Traversals may then track depth
// themselves instead of reading the field back.
bool octree_depths_exact(const uint32_t* w, uint64_t n, bool tree);
// Start table of a tagged layout (child_w as build_tagged_layout writes it,
// root_tag the root's own tags): entry (z << 2L | y << L | x) of the level-L
// grid (L = kTableLevel) = {tagged word of the node find() reaches at level L
// for any point of cell (x, y, z), or of the leaf that ends the walk above it;
// that node's level}.  The root's word is 0 | root_tag.  The walk follows
// child words for at most L levels, so any tree, cyclic ones included, gives
// a table.
void build_start_table(const uint32_t* child_w, uint32_t root_tag, uint32_t index_mask, uint2* table);
// When the tree has fewer than 2^25 nodes and every depth field is below 32,
// ORs into every child word of every leaf (its ropes) the depth field of the
// node it points to, shifted by kRopeDepthShift, and returns true; else
// changes nothing and returns false.  Internal nodes' child words keep their
// 30-bit indices.
bool pack_rope_depths(uint32_t* child_w, const uint32_t* meta, uint64_t n);

}  // namespace xd
