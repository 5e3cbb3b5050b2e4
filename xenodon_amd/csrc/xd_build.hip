// Octree construction on the GPU (`xenodon convert` for grids already in HBM).
//
// Produces the node buffer of build_octree (src/model/OctreeConstruction.h:
// 13-237) bit for bit -- the same buffer as the host converter
// (xd_convert.cpp) -- for the channel-difference heuristic and sparse or rope
// trees.  The reference recurses depth-first and scans every node's region;
// here the tree is built level by level:
//
//   1. statistics pyramid: per-channel min / max / sum and the in-grid voxel
//      count of every 4^3 cell (one per device brick), then of every coarser
//      cell by merging its eight children (integer arithmetic, so the result
//      does not depend on the order of the reduction);
//   2. top-down, one level at a time: the existing nodes of a level (a list
//      in Morton order, child k = x*4 + y*2 + z as in the reference) are
//      classified -- empty leaf if the node's offset is outside the grid
//      (:141-154), leaf if (!split && fully inside) || extent == 1
//      (:156-171), colour = per-channel integer mean (Grid.cpp:86-131) -- and
//      the eight children of every internal node form the next level's list
//      (cells of extent 2 and 1 are classified straight from the grid);
//   3. numbering: subtree sizes bottom-up, subtree starts top-down, so a
//      node's post-order index is start + size - 1 and its index in the
//      reversed buffer N - 1 - that (:92-112);
//   4. emission of the 40-byte nodes, then, for rope trees, six find()s per
//      leaf on the finished buffer (Octree.cpp:116-153, 181-201).
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "xd_runtime.hpp"

namespace xd {
namespace {

constexpr uint32_t LEAF = 0x80000000u;

struct CellStats {
    unsigned long long n;       // in-grid voxels
    unsigned long long sum[4];  // per channel
    uint32_t mn, mx;            // per-channel min / max, one byte each
};

__device__ __forceinline__ uint32_t byte_min(uint32_t a, uint32_t b) {
    uint32_t r = 0;
    for (int c = 0; c < 4; ++c) r |= min((a >> (8 * c)) & 0xFFu, (b >> (8 * c)) & 0xFFu) << (8 * c);
    return r;
}
__device__ __forceinline__ uint32_t byte_max(uint32_t a, uint32_t b) {
    uint32_t r = 0;
    for (int c = 0; c < 4; ++c) r |= max((a >> (8 * c)) & 0xFFu, (b >> (8 * c)) & 0xFFu) << (8 * c);
    return r;
}

struct GridDev {
    const uint32_t* bricks;
    uint32_t X, Y, Z, nbx, nby, nbz;
    __device__ uint32_t texel(uint32_t x, uint32_t y, uint32_t z) const {
        const uint64_t b = ((uint64_t)(z >> 2) * nby + (y >> 2)) * nbx + (x >> 2);
        return bricks[b * 64 + (x & 3u) + ((y & 3u) << 2) + ((z & 3u) << 4)];
    }
};

// Cell stats over the in-grid voxels of the box [o, o + e) (e <= 4).
__device__ CellStats box_stats(const GridDev& g, uint32_t ox, uint32_t oy, uint32_t oz, uint32_t e) {
    CellStats s{};
    s.mn = 0xFFFFFFFFu;
    s.mx = 0;
    for (uint32_t z = oz; z < min(oz + e, g.Z); ++z)
        for (uint32_t y = oy; y < min(oy + e, g.Y); ++y)
            for (uint32_t x = ox; x < min(ox + e, g.X); ++x) {
                const uint32_t v = g.texel(x, y, z);
                s.n += 1;
                for (int c = 0; c < 4; ++c) s.sum[c] += (v >> (8 * c)) & 0xFFu;
                s.mn = byte_min(s.mn, v);
                s.mx = byte_max(s.mx, v);
            }
    return s;
}

// Level of 4^3 cells, one thread per cell (nc cells per axis).
__global__ void k_stats_base(GridDev g, uint32_t nc, CellStats* out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t total = (uint64_t)nc * nc * nc;
    if (i >= total) return;
    const uint32_t x = (uint32_t)(i % nc), y = (uint32_t)((i / nc) % nc), z = (uint32_t)(i / ((uint64_t)nc * nc));
    out[i] = box_stats(g, x * 4, y * 4, z * 4, 4);
}

// Parent level from a child level of 2*nc cells per axis.
__global__ void k_stats_merge(const CellStats* child, uint32_t nc, CellStats* out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t total = (uint64_t)nc * nc * nc;
    if (i >= total) return;
    const uint32_t x = (uint32_t)(i % nc), y = (uint32_t)((i / nc) % nc), z = (uint32_t)(i / ((uint64_t)nc * nc));
    const uint64_t cn = 2ull * nc;
    CellStats s{};
    s.mn = 0xFFFFFFFFu;
    s.mx = 0;
    for (int k = 0; k < 8; ++k) {
        const uint64_t cx = 2ull * x + ((k >> 2) & 1), cy = 2ull * y + ((k >> 1) & 1), cz = 2ull * z + (k & 1);
        const CellStats& c = child[(cz * cn + cy) * cn + cx];
        if (!c.n) continue;
        s.n += c.n;
        for (int q = 0; q < 4; ++q) s.sum[q] += c.sum[q];
        s.mn = byte_min(s.mn, c.mn);
        s.mx = byte_max(s.mx, c.mx);
    }
    out[i] = s;
}

__device__ __forceinline__ void decode(uint64_t code, uint32_t level, uint32_t& x, uint32_t& y, uint32_t& z) {
    x = y = z = 0;
    for (uint32_t l = 0; l < level; ++l) {
        const uint32_t k = (uint32_t)(code >> (3 * (level - 1 - l))) & 7u;
        x = (x << 1) | (k >> 2);
        y = (y << 1) | ((k >> 1) & 1u);
        z = (z << 1) | (k & 1u);
    }
}

// Classifies the nodes of one level.  stats = that level's pyramid (cells of
// extent >= 4), or nullptr to scan the grid (extent 2 and 1).
__global__ void k_classify(GridDev g, const uint64_t* codes, uint64_t count, uint32_t level, uint32_t extent,
                           const CellStats* stats, uint32_t nc, uint32_t chan_diff, uint32_t* colour,
                           uint32_t* internal) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t cx, cy, cz;
    decode(codes[i], level, cx, cy, cz);
    const uint64_t ox = (uint64_t)cx * extent, oy = (uint64_t)cy * extent, oz = (uint64_t)cz * extent;
    uint32_t col = 0, inner = 0;
    if (ox < g.X && oy < g.Y && oz < g.Z) {
        const CellStats s = stats ? stats[((uint64_t)cz * nc + cy) * nc + cx]
                                  : box_stats(g, (uint32_t)ox, (uint32_t)oy, (uint32_t)oz, extent);
        uint32_t diff = 0;
        for (int c = 0; c < 4; ++c) {
            col |= (uint32_t)(uint8_t)(s.sum[c] / s.n) << (8 * c);
            diff = max(diff, ((s.mx >> (8 * c)) & 0xFFu) - ((s.mn >> (8 * c)) & 0xFFu));
        }
        const bool fully = ox + extent <= g.X && oy + extent <= g.Y && oz + extent <= g.Z;
        const bool leaf = (diff <= chan_diff && fully) || extent == 1;
        inner = leaf ? 0u : 1u;
    }
    colour[i] = col;
    internal[i] = inner;
}

__global__ void k_children(const uint64_t* codes, uint64_t count, const uint32_t* internal, const uint32_t* rank,
                           uint64_t* next) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count || !internal[i]) return;
    const uint64_t base = 8ull * rank[i];
    for (uint32_t k = 0; k < 8; ++k) next[base + k] = codes[i] * 8 + k;
}

__global__ void k_sizes(uint64_t count, const uint32_t* internal, const uint32_t* rank, const uint32_t* child_size,
                        uint32_t* size) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t s = 1;
    if (internal[i])
        for (uint32_t k = 0; k < 8; ++k) s += child_size[8ull * rank[i] + k];
    size[i] = s;
}

// Starts of the children of every internal node of the parent level.
__global__ void k_starts(uint64_t parent_count, const uint32_t* internal, const uint32_t* rank,
                         const uint32_t* parent_start, const uint32_t* child_size, uint32_t* child_start) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= parent_count || !internal[i]) return;
    uint32_t run = parent_start[i];
    for (uint32_t k = 0; k < 8; ++k) {
        const uint64_t c = 8ull * rank[i] + k;
        child_start[c] = run;
        run += child_size[c];
    }
}

__global__ void k_emit(uint64_t count, uint32_t level, uint32_t n_total, const uint32_t* internal, const uint32_t* rank,
                       const uint32_t* start, const uint32_t* size, const uint32_t* colour,
                       const uint32_t* child_start, const uint32_t* child_size, uint32_t* nodes) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t f = n_total - 1u - (start[i] + size[i] - 1u);
    uint32_t* w = nodes + (uint64_t)f * 10;
    if (internal[i]) {
        for (uint32_t k = 0; k < 8; ++k) {
            const uint64_t c = 8ull * rank[i] + k;
            w[k] = n_total - 1u - (child_start[c] + child_size[c] - 1u);
        }
        w[9] = level;
    } else {
        for (uint32_t k = 0; k < 8; ++k) w[k] = 0;
        w[9] = LEAF | level;
    }
    w[8] = colour[i];
}

// Octree::find on the finished buffer; ~0 for the reference's nullptr.
__device__ uint64_t find_node(const uint32_t* nodes, uint64_t side, uint64_t px, uint64_t py, uint64_t pz,
                              uint32_t max_depth) {
    uint64_t extent = side;
    if (px >= extent || py >= extent || pz >= extent) return ~0ull;
    uint64_t index = 0, ox = 0, oy = 0, oz = 0;
    for (;;) {
        extent >>= 1;
        if ((nodes[index * 10 + 9] & LEAF) || extent == 0 || max_depth == 0) return index;
        uint32_t ci = 0;
        if (px >= ox + extent) { ci |= 4; ox += extent; }
        if (py >= oy + extent) { ci |= 2; oy += extent; }
        if (pz >= oz + extent) { ci |= 1; oz += extent; }
        index = nodes[index * 10 + ci];
        --max_depth;
    }
}

// Ropes of the leaves of one level (Octree::generate_ropes): children[2a] and
// [2a+1] = the node holding the point one extent beyond the leaf along +a / -a.
__global__ void k_ropes(const uint64_t* codes, uint64_t count, uint32_t level, uint64_t side, uint32_t n_total,
                        const uint32_t* internal, const uint32_t* start, const uint32_t* size, uint32_t* nodes) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count || internal[i]) return;
    uint32_t cx, cy, cz;
    decode(codes[i], level, cx, cy, cz);
    const uint64_t extent = side >> level;
    const uint64_t p[3] = {(uint64_t)cx * extent, (uint64_t)cy * extent, (uint64_t)cz * extent};
    uint32_t r[6];
    for (int a = 0; a < 3; ++a) {
        uint64_t q[3] = {p[0], p[1], p[2]};
        q[a] = p[a] + extent;
        uint64_t f = find_node(nodes, side, q[0], q[1], q[2], level);
        r[2 * a] = f == ~0ull ? 0u : (uint32_t)f;
        q[a] = p[a] - extent;  // wraps below zero like size_t
        f = find_node(nodes, side, q[0], q[1], q[2], level);
        r[2 * a + 1] = f == ~0ull ? 0u : (uint32_t)f;
    }
    const uint32_t fi = n_total - 1u - (start[i] + size[i] - 1u);
    for (int k = 0; k < 6; ++k) nodes[(uint64_t)fi * 10 + k] = r[k];
}

inline unsigned blocks_for(uint64_t n) { return (unsigned)((n + 255) / 256); }

struct Level {
    uint64_t count = 0;
    uint64_t* codes = nullptr;
    uint32_t *colour = nullptr, *internal = nullptr, *rank = nullptr, *size = nullptr, *start = nullptr;
};

}  // namespace
}  // namespace xd

using namespace xd;

extern "C" {

int xd_build_octree_device(xd_context* ctx, const xd_volume* grid, int chan_diff, int tree_type, uint32_t** out_nodes,
                           xd_build_stats* stats) {
    if (!ctx || !grid || !out_nodes || !stats) return fail(XD_ERR_INVALID_ARGUMENT, "null argument");
    *out_nodes = nullptr;
    if (grid->kind != XD_VOLUME_GRID) return fail(XD_ERR_INVALID_ARGUMENT, "the device builder needs a grid volume");
    if (grid->ctx != ctx) return fail(XD_ERR_INVALID_ARGUMENT, "volume belongs to another context");
    if (chan_diff < 0 || chan_diff > 255) return fail(XD_ERR_INVALID_ARGUMENT, "chan_diff must be 0..255");
    if (tree_type != XD_TREE_SPARSE && tree_type != XD_TREE_ROPE)
        return fail(XD_ERR_INVALID_ARGUMENT,
                    "the device builder makes sparse and rope trees; use xd_build_octree for DAGs");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return fail_hip(e, "hipSetDevice");
    hipStream_t s = ctx->stream;
    uint64_t side = 1;
    while (side < std::max({grid->dims[0], grid->dims[1], grid->dims[2]})) side <<= 1;
    uint32_t D = 0;
    while ((1ull << D) < side) ++D;
    const GridDev g{grid->d_bricks, (uint32_t)grid->dims[0], (uint32_t)grid->dims[1], (uint32_t)grid->dims[2],
                    grid->nb[0], grid->nb[1], grid->nb[2]};

    std::vector<void*> owned;
    auto alloc = [&](void** p, size_t bytes) -> hipError_t {
        hipError_t r = hipMalloc(p, std::max<size_t>(bytes, 16));
        if (r == hipSuccess) owned.push_back(*p);
        return r;
    };
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(s);
        for (void* p : owned) (void)hipFree(p);
        owned.clear();
    };
#define XD_TRY(expr, what)                      \
    do {                                        \
        hipError_t _e = (expr);                 \
        if (_e != hipSuccess) {                 \
            cleanup();                          \
            return fail_hip(_e, what);          \
        }                                       \
    } while (0)

    // 1. statistics pyramid: levels B = D-2 (4^3 cells) up to 0
    const int B = (int)D - 2;
    std::vector<CellStats*> pyr(std::max(B + 1, 0), nullptr);
    if (B >= 0) {
        const uint32_t ncb = 1u << B;
        const uint64_t cells = (uint64_t)ncb * ncb * ncb;
        XD_TRY(alloc((void**)&pyr[B], cells * sizeof(CellStats)), "octree build: statistics");
        hipLaunchKernelGGL(k_stats_base, dim3(blocks_for(cells)), dim3(256), 0, s, g, ncb, pyr[B]);
        for (int L = B - 1; L >= 0; --L) {
            const uint32_t nc = 1u << L;
            const uint64_t n = (uint64_t)nc * nc * nc;
            XD_TRY(alloc((void**)&pyr[L], n * sizeof(CellStats)), "octree build: statistics");
            hipLaunchKernelGGL(k_stats_merge, dim3(blocks_for(n)), dim3(256), 0, s, pyr[L + 1], nc, pyr[L]);
        }
        XD_TRY(hipGetLastError(), "octree build: statistics");
    }

    // 2. top-down lists
    std::vector<Level> lv;
    lv.emplace_back();
    lv[0].count = 1;
    XD_TRY(alloc((void**)&lv[0].codes, 8), "octree build: lists");
    XD_TRY(hipMemsetAsync(lv[0].codes, 0, 8, s), "octree build: lists");
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    uint64_t total = 0, leaves = 0;
    for (uint32_t L = 0;; ++L) {
        Level& cur = lv[L];
        const uint64_t n = cur.count;
        total += n;
        if (total > 0xFFFFFFFFull) {
            cleanup();
            return fail(XD_ERR_FORMAT, "octree exceeds 2^32 nodes");
        }
        XD_TRY(alloc((void**)&cur.colour, n * 4), "octree build: lists");
        XD_TRY(alloc((void**)&cur.internal, n * 4), "octree build: lists");
        XD_TRY(alloc((void**)&cur.rank, n * 4), "octree build: lists");
        const uint32_t extent = (uint32_t)std::min<uint64_t>(side >> L, 0xFFFFFFFFull);
        const bool from_pyr = (int)L <= B;
        hipLaunchKernelGGL(k_classify, dim3(blocks_for(n)), dim3(256), 0, s, g, cur.codes, n, L, extent,
                           from_pyr ? pyr[L] : nullptr, 1u << std::min<uint32_t>(L, 31), (uint32_t)chan_diff,
                           cur.colour, cur.internal);
        size_t need = 0;
        XD_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, need, cur.internal, cur.rank, (int)n, s),
               "octree build: scan");
        if (need > tmp_bytes) {
            XD_TRY(alloc(&tmp, need), "octree build: scan");
            tmp_bytes = need;
        }
        XD_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, cur.internal, cur.rank, (int)n, s),
               "octree build: scan");
        uint32_t last_rank = 0, last_internal = 0;
        XD_TRY(hipMemcpyAsync(&last_rank, cur.rank + n - 1, 4, hipMemcpyDeviceToHost, s), "octree build: scan");
        XD_TRY(hipMemcpyAsync(&last_internal, cur.internal + n - 1, 4, hipMemcpyDeviceToHost, s),
               "octree build: scan");
        XD_TRY(hipStreamSynchronize(s), "octree build: classify");
        const uint64_t internals = (uint64_t)last_rank + last_internal;
        leaves += n - internals;
        if (internals == 0) break;
        Level next;
        next.count = internals * 8;
        XD_TRY(alloc((void**)&next.codes, next.count * 8), "octree build: lists");
        hipLaunchKernelGGL(k_children, dim3(blocks_for(n)), dim3(256), 0, s, cur.codes, n, cur.internal, cur.rank,
                           next.codes);
        lv.push_back(next);
    }
    const uint32_t depth = (uint32_t)lv.size() - 1;
    const uint32_t N = (uint32_t)total;

    // 3. numbering
    for (int L = (int)depth; L >= 0; --L) {
        Level& cur = lv[L];
        XD_TRY(alloc((void**)&cur.size, cur.count * 4), "octree build: numbering");
        XD_TRY(alloc((void**)&cur.start, cur.count * 4), "octree build: numbering");
        hipLaunchKernelGGL(k_sizes, dim3(blocks_for(cur.count)), dim3(256), 0, s, cur.count, cur.internal, cur.rank,
                           L < (int)depth ? lv[L + 1].size : nullptr, cur.size);
    }
    XD_TRY(hipMemsetAsync(lv[0].start, 0, 4, s), "octree build: numbering");
    for (uint32_t L = 0; L < depth; ++L)
        hipLaunchKernelGGL(k_starts, dim3(blocks_for(lv[L].count)), dim3(256), 0, s, lv[L].count, lv[L].internal,
                           lv[L].rank, lv[L].start, lv[L + 1].size, lv[L + 1].start);

    // 4. nodes, then ropes
    uint32_t* d_nodes = nullptr;
    XD_TRY(alloc((void**)&d_nodes, (size_t)N * 40), "octree build: nodes");
    for (uint32_t L = 0; L <= depth; ++L) {
        const Level& cur = lv[L];
        hipLaunchKernelGGL(k_emit, dim3(blocks_for(cur.count)), dim3(256), 0, s, cur.count, L, N, cur.internal,
                           cur.rank, cur.start, cur.size, cur.colour, L < depth ? lv[L + 1].start : nullptr,
                           L < depth ? lv[L + 1].size : nullptr, d_nodes);
    }
    XD_TRY(hipGetLastError(), "octree build: emit");
    if (tree_type == XD_TREE_ROPE) {
        for (uint32_t L = 0; L <= depth; ++L) {
            const Level& cur = lv[L];
            hipLaunchKernelGGL(k_ropes, dim3(blocks_for(cur.count)), dim3(256), 0, s, cur.codes, cur.count, L, side,
                               N, cur.internal, cur.start, cur.size, d_nodes);
        }
        XD_TRY(hipGetLastError(), "octree build: ropes");
    }
    uint32_t* host = static_cast<uint32_t*>(malloc((size_t)N * 40));
    if (!host) {
        cleanup();
        return fail(XD_ERR_OUT_OF_MEMORY, "out of host memory for %u nodes", N);
    }
    XD_TRY(hipMemcpyAsync(host, d_nodes, (size_t)N * 40, hipMemcpyDeviceToHost, s), "octree build: download");
    XD_TRY(hipStreamSynchronize(s), "octree build: download");
    cleanup();
#undef XD_TRY
    stats->total_leaves = leaves;
    stats->unique_leaves = leaves;
    stats->total_nodes = N;
    stats->depth = depth;
    stats->side = side;
    stats->node_count = N;
    *out_nodes = host;
    clear_error();
    return XD_OK;
}

}  // extern "C"
