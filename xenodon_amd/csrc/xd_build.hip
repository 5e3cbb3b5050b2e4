// Octree construction on the GPU (`xenodon convert` for grids already in HBM).
//
// Produces the node buffer of build_octree (src/model/OctreeConstruction.h:
// 13-237) bit for bit -- the same buffer as the host converter
// (xd_convert.cpp) -- for both heuristics and sparse, DAG and rope trees.
// The reference recurses depth-first and scans every node's region; here the
// tree is built level by level:
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
//
// The std-dev heuristic (Grid.cpp:139-214) splits when sqrt(S / n) > limit,
// where S is an FP64 sum over the node's voxels taken in z, y, x order.  A
// node of at most 16^3 voxels sums them in that order in one thread, which is
// the reference's S exactly.  A larger node sums them in parallel in
// double-double, which fixes S only to within the rounding error of a
// sequential sum (|S - exact| <= gamma_{n-1} * exact, the terms being >= 0);
// the split is decided on the device when the whole interval falls on one side
// of the limit by a margin, and otherwise that node's S is taken sequentially
// and the split decided on the host with the converter's own expression.
//
// DAGs (HashCache, OctreeConstruction.h:19-30) keep the first node inserted
// of every set of equal 40-byte nodes, children already replaced by their
// kept indices.  Nodes of different depths never compare equal (the depth
// is part of is_leaf_depth), so the sets are found level by level, bottom-up:
// a node's key is its colour, leaf flag and its children's sets; equal keys
// are found by sorting 64-bit key hashes (stably, so within a run of equal
// hashes the nodes stay in post-order, which on one level is Morton order)
// and comparing whole keys within a run.  A set is named by the post-order
// index of its first member; a node is kept when it is its set's first
// member, and its index in the DAG counts the kept nodes before it in
// post-order.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
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

// One voxel's term of the std-dev sum, Grid.cpp:187-206: the four squared
// channel deviations added left to right (the library is built without FP
// contraction, so no product is fused into an add).
__device__ __forceinline__ double sd_term(uint32_t v, const double avg[4]) {
    const double dr = (double)(v & 0xFFu) - avg[0], dg = (double)((v >> 8) & 0xFFu) - avg[1];
    const double db = (double)((v >> 16) & 0xFFu) - avg[2], da = (double)(v >> 24) - avg[3];
    return dr * dr + dg * dg + db * db + da * da;
}

// Channel means as Grid.cpp:178-184 forms them: integer sums over the
// in-grid voxel count, divided in FP64.
__device__ __forceinline__ void sd_avg(const CellStats& s, double avg[4]) {
    const double nd = (double)s.n;
    for (int c = 0; c < 4; ++c) avg[c] = (double)s.sum[c] / nd;
}

// The reference's sequential sum over the in-grid part of [o, o + e).
__device__ double sd_sequential(const GridDev& g, uint64_t ox, uint64_t oy, uint64_t oz, uint64_t e,
                                const double avg[4]) {
    double s = 0;
    const uint64_t x1 = min(ox + e, (uint64_t)g.X), y1 = min(oy + e, (uint64_t)g.Y), z1 = min(oz + e, (uint64_t)g.Z);
    for (uint64_t z = oz; z < z1; ++z)
        for (uint64_t y = oy; y < y1; ++y)
            for (uint64_t x = ox; x < x1; ++x) s += sd_term(g.texel((uint32_t)x, (uint32_t)y, (uint32_t)z), avg);
    return s;
}

// Split decision from sqrt(S / n) against the limit with a relative margin
// that no few-ulp difference in the division or square root can cross:
// 1 split, 0 not, 2 too close to call here.
constexpr double kSdMargin = 1e-12;
__device__ __forceinline__ uint32_t sd_side(double s_lo, double s_hi, double nd, double limit) {
    if (s_hi == 0.0) return 0u;  // every term is zero: sqrt(0) = 0 is not above a limit >= 0
    if (__builtin_sqrt(s_lo / nd) > limit * (1.0 + kSdMargin)) return 1u;
    if (__builtin_sqrt(s_hi / nd) < limit * (1.0 - kSdMargin)) return 0u;
    return 2u;
}

// Classifies the nodes of one level.  stats = that level's pyramid (cells of
// extent >= 4), or nullptr to scan the grid (extent 2 and 1).  heuristic 1
// (std-dev): big = this level's double-double sums (extent >= 32), and nodes
// whose split cannot be decided here are listed in amb (count in amb_n) and
// left internal until the host decides.
__global__ void k_classify(GridDev g, const uint64_t* codes, uint64_t count, uint32_t level, uint32_t extent,
                           const CellStats* stats, uint32_t nc, uint32_t heuristic, uint32_t chan_diff, double sd_limit,
                           const double2* big, uint32_t* amb, uint32_t* amb_n, uint32_t* colour, uint32_t* internal) {
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
        bool split;
        if (extent == 1) {
            split = false;  // a single voxel is a leaf whatever the heuristic says
        } else if (heuristic == 0) {
            split = diff > chan_diff;
        } else {
            const double nd = (double)s.n;
            uint32_t side;
            if (big) {
                // |sequential - exact| <= gamma_{n-1} * exact for terms >= 0;
                // doubled, plus the double-double sum's own error
                const double e = big[i].x + big[i].y;
                const double u = 1.1102230246251565e-16;  // 2^-53
                const double gam = nd * u / (1.0 - nd * u);
                const double bound = 2.0 * gam * e + 0x1p-90 * e;
                side = sd_side(e - bound, e + bound, nd, sd_limit);
            } else {
                double avg[4];
                sd_avg(s, avg);
                const double sum = sd_sequential(g, ox, oy, oz, extent, avg);
                side = sd_side(sum, sum, nd, sd_limit);
            }
            if (side == 2u) amb[atomicAdd(amb_n, 1u)] = (uint32_t)i;
            split = side != 0u;
        }
        inner = (split || !fully) ? 1u : 0u;
    }
    colour[i] = col;
    internal[i] = inner;
}

// Double-double sum of the std-dev terms of node i's voxels (one block per
// node; a level of nodes of extent >= 32).
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

__global__ __launch_bounds__(256) void k_stddev_big(GridDev g, const uint64_t* codes, uint32_t level, uint32_t extent,
                                                    const CellStats* stats, uint32_t nc, double2* out) {
    const uint64_t i = blockIdx.x;
    uint32_t cx, cy, cz;
    decode(codes[i], level, cx, cy, cz);
    const uint64_t ox = (uint64_t)cx * extent, oy = (uint64_t)cy * extent, oz = (uint64_t)cz * extent;
    __shared__ double hs[256], ls[256];
    double hi = 0, lo = 0;
    if (ox < g.X && oy < g.Y && oz < g.Z) {
        const CellStats st = stats[((uint64_t)cz * nc + cy) * nc + cx];
        double avg[4];
        sd_avg(st, avg);
        const uint64_t wx = min(ox + extent, (uint64_t)g.X) - ox, wy = min(oy + extent, (uint64_t)g.Y) - oy;
        const uint64_t wz = min(oz + extent, (uint64_t)g.Z) - oz;
        const uint64_t n = wx * wy * wz;
        for (uint64_t k = threadIdx.x; k < n; k += blockDim.x) {
            const uint64_t x = ox + k % wx, y = oy + (k / wx) % wy, z = oz + k / (wx * wy);
            double s, e;
            two_sum(hi, sd_term(g.texel((uint32_t)x, (uint32_t)y, (uint32_t)z), avg), s, e);
            hi = s;
            lo += e;
        }
    }
    hs[threadIdx.x] = hi;
    ls[threadIdx.x] = lo;
    __syncthreads();
    for (uint32_t w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            double s, e;
            two_sum(hs[threadIdx.x], hs[threadIdx.x + w], s, e);
            hs[threadIdx.x] = s;
            ls[threadIdx.x] += ls[threadIdx.x + w] + e;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[i] = make_double2(hs[0], ls[0]);
}

// The reference's sequential S for the nodes k_classify could not decide.
__global__ void k_stddev_exact(GridDev g, const uint64_t* codes, uint32_t level, uint32_t extent,
                               const CellStats* stats, uint32_t nc, const uint32_t* amb, uint32_t n_amb, double* sums,
                               double* counts) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_amb) return;
    const uint64_t i = amb[j];
    uint32_t cx, cy, cz;
    decode(codes[i], level, cx, cy, cz);
    const uint64_t ox = (uint64_t)cx * extent, oy = (uint64_t)cy * extent, oz = (uint64_t)cz * extent;
    const CellStats st = stats ? stats[((uint64_t)cz * nc + cy) * nc + cx]
                               : box_stats(g, (uint32_t)ox, (uint32_t)oy, (uint32_t)oz, extent);
    double avg[4];
    sd_avg(st, avg);
    sums[j] = sd_sequential(g, ox, oy, oz, extent, avg);
    counts[j] = (double)st.n;
}

// Host decisions for the listed nodes: internal = split || !fully inside.
__global__ void k_stddev_patch(const uint64_t* codes, uint32_t level, uint32_t extent, uint32_t X, uint32_t Y,
                               uint32_t Z, const uint32_t* amb, const uint8_t* split, uint32_t n_amb,
                               uint32_t* internal) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_amb) return;
    const uint64_t i = amb[j];
    uint32_t cx, cy, cz;
    decode(codes[i], level, cx, cy, cz);
    const bool fully = (uint64_t)cx * extent + extent <= X && (uint64_t)cy * extent + extent <= Y &&
                       (uint64_t)cz * extent + extent <= Z;
    internal[i] = (split[j] || !fully) ? 1u : 0u;
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

// ---- DAG: equal nodes of one level, bottom-up ------------------------------

__device__ __forceinline__ uint32_t post_index(const uint32_t* start, const uint32_t* size, uint64_t i) {
    return start[i] + size[i] - 1u;
}

__device__ __forceinline__ uint64_t mix64(uint64_t h, uint64_t v) {
    h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    h ^= h >> 31;
    h *= 0xBF58476D1CE4E5B9ull;
    return h ^ (h >> 29);
}

// Hash of node i's key: leaf flag, colour and, for an internal node, the sets
// of its eight children (cls_child, by post-order index of their first node).
__global__ void k_dag_hash(uint64_t count, const uint32_t* internal, const uint32_t* rank, const uint32_t* colour,
                           const uint32_t* cls_child, uint64_t* keys, uint32_t* idx) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint64_t h = mix64(0x1234567ull + internal[i], colour[i]);
    if (internal[i])
        for (uint32_t k = 0; k < 8; ++k) h = mix64(h, cls_child[8ull * rank[i] + k]);
    keys[i] = h;
    idx[i] = (uint32_t)i;
}

__global__ void k_run_heads(uint64_t count, const uint64_t* keys, uint32_t* head) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count) return;
    head[j] = (j == 0 || keys[j] != keys[j - 1]) ? (uint32_t)j : 0u;
}

__device__ __forceinline__ bool same_key(uint32_t a, uint32_t b, const uint32_t* internal, const uint32_t* rank,
                                         const uint32_t* colour, const uint32_t* cls_child) {
    if (internal[a] != internal[b] || colour[a] != colour[b]) return false;
    if (internal[a])
        for (uint32_t k = 0; k < 8; ++k)
            if (cls_child[8ull * rank[a] + k] != cls_child[8ull * rank[b] + k]) return false;
    return true;
}

// Set of every node: the first node of its sorted run with an equal key.  A
// run is sorted stably from Morton order, which on one level is post-order,
// so that node is the set's first in post-order.
__global__ void k_dag_sets(uint64_t count, const uint32_t* idx, const uint32_t* run_start, const uint32_t* internal,
                           const uint32_t* rank, const uint32_t* colour, const uint32_t* cls_child,
                           const uint32_t* start, const uint32_t* size, uint32_t* cls,
                           unsigned long long* unique_leaves) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count) return;
    const uint32_t i = idx[j];
    uint32_t lead = i;
    for (uint64_t k = run_start[j]; k < j; ++k)
        if (same_key(idx[k], i, internal, rank, colour, cls_child)) {
            lead = idx[k];
            break;
        }
    cls[i] = post_index(start, size, lead);
    if (lead == i && !internal[i]) atomicAdd(unique_leaves, 1ull);
}

__global__ void k_dag_mark(uint64_t count, const uint32_t* start, const uint32_t* size, const uint32_t* cls,
                           uint32_t* kept) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t p = post_index(start, size, i);
    kept[p] = cls[i] == p ? 1u : 0u;
}

// Kept nodes into the reversed buffer of U nodes: a node kept as the k-th in
// post-order lands at U - 1 - k, and so do references to its set.
__global__ void k_emit_dag(uint64_t count, uint32_t level, uint32_t n_kept, const uint32_t* internal,
                           const uint32_t* rank, const uint32_t* start, const uint32_t* size, const uint32_t* colour,
                           const uint32_t* cls, const uint32_t* cls_child, const uint32_t* ins, uint32_t* nodes) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t p = post_index(start, size, i);
    if (cls[i] != p) return;
    uint32_t* w = nodes + (uint64_t)(n_kept - 1u - ins[p]) * 10;
    if (internal[i]) {
        for (uint32_t k = 0; k < 8; ++k) w[k] = n_kept - 1u - ins[cls_child[8ull * rank[i] + k]];
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
    uint32_t* cls = nullptr;  // DAG: post-order index of the first node equal to this one
};

}  // namespace
}  // namespace xd

using namespace xd;

extern "C" {

int xd_build_octree_device(xd_context* ctx, const xd_volume* grid, int heuristic, int chan_diff, double std_dev,
                           int tree_type, uint32_t** out_nodes, xd_build_stats* stats) {
    if (!ctx || !grid || !out_nodes || !stats) return fail(XD_ERR_INVALID_ARGUMENT, "null argument");
    *out_nodes = nullptr;
    if (grid->kind != XD_VOLUME_GRID) return fail(XD_ERR_INVALID_ARGUMENT, "the device builder needs a grid volume");
    if (grid->ctx != ctx) return fail(XD_ERR_INVALID_ARGUMENT, "volume belongs to another context");
    if (heuristic != XD_HEURISTIC_CHAN_DIFF && heuristic != XD_HEURISTIC_STD_DEV)
        return fail(XD_ERR_INVALID_ARGUMENT, "unknown heuristic %d", heuristic);
    if (chan_diff < 0 || chan_diff > 255) return fail(XD_ERR_INVALID_ARGUMENT, "chan_diff must be 0..255");
    if (!(std_dev >= 0.0)) return fail(XD_ERR_INVALID_ARGUMENT, "std_dev must be >= 0");
    if (tree_type != XD_TREE_SPARSE && tree_type != XD_TREE_ROPE && tree_type != XD_TREE_DAG)
        return fail(XD_ERR_INVALID_ARGUMENT, "unknown tree type %d", tree_type);
    const bool sd = heuristic == XD_HEURISTIC_STD_DEV, dag = tree_type == XD_TREE_DAG;
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
        const uint32_t nc = 1u << std::min<uint32_t>(L, 31);
        double2* big = nullptr;
        uint32_t *amb = nullptr, *amb_n = nullptr;
        if (sd) {
            XD_TRY(alloc((void**)&amb, n * 4), "octree build: std-dev");
            XD_TRY(alloc((void**)&amb_n, 4), "octree build: std-dev");
            XD_TRY(hipMemsetAsync(amb_n, 0, 4, s), "octree build: std-dev");
            if (from_pyr && extent >= 32) {
                XD_TRY(alloc((void**)&big, n * sizeof(double2)), "octree build: std-dev");
                hipLaunchKernelGGL(k_stddev_big, dim3((unsigned)n), dim3(256), 0, s, g, cur.codes, L, extent, pyr[L],
                                   nc, big);
            }
        }
        hipLaunchKernelGGL(k_classify, dim3(blocks_for(n)), dim3(256), 0, s, g, cur.codes, n, L, extent,
                           from_pyr ? pyr[L] : nullptr, nc, sd ? 1u : 0u, (uint32_t)chan_diff, std_dev, big, amb,
                           amb_n, cur.colour, cur.internal);
        if (sd) {
            // nodes too close to the limit to decide on the device: the
            // reference's sequential S, and the converter's own comparison
            uint32_t n_amb = 0;
            XD_TRY(hipMemcpyAsync(&n_amb, amb_n, 4, hipMemcpyDeviceToHost, s), "octree build: std-dev");
            XD_TRY(hipStreamSynchronize(s), "octree build: std-dev");
            if (n_amb) {
                double *d_sums = nullptr, *d_counts = nullptr;
                uint8_t* d_split = nullptr;
                XD_TRY(alloc((void**)&d_sums, n_amb * 8ull), "octree build: std-dev");
                XD_TRY(alloc((void**)&d_counts, n_amb * 8ull), "octree build: std-dev");
                XD_TRY(alloc((void**)&d_split, n_amb), "octree build: std-dev");
                hipLaunchKernelGGL(k_stddev_exact, dim3(blocks_for(n_amb)), dim3(256), 0, s, g, cur.codes, L, extent,
                                   from_pyr ? pyr[L] : nullptr, nc, amb, n_amb, d_sums, d_counts);
                std::vector<double> sums(n_amb), counts(n_amb);
                std::vector<uint8_t> split(n_amb);
                XD_TRY(hipMemcpyAsync(sums.data(), d_sums, n_amb * 8ull, hipMemcpyDeviceToHost, s), "octree build: std-dev");
                XD_TRY(hipMemcpyAsync(counts.data(), d_counts, n_amb * 8ull, hipMemcpyDeviceToHost, s),
                       "octree build: std-dev");
                XD_TRY(hipStreamSynchronize(s), "octree build: std-dev");
                for (uint32_t j = 0; j < n_amb; ++j) split[j] = std::sqrt(sums[j] / counts[j]) > std_dev ? 1 : 0;
                XD_TRY(hipMemcpyAsync(d_split, split.data(), n_amb, hipMemcpyHostToDevice, s), "octree build: std-dev");
                hipLaunchKernelGGL(k_stddev_patch, dim3(blocks_for(n_amb)), dim3(256), 0, s, cur.codes, L, extent,
                                   g.X, g.Y, g.Z, amb, d_split, n_amb, cur.internal);
                XD_TRY(hipStreamSynchronize(s), "octree build: std-dev");
            }
        }
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

    // 3b. DAG: the set of every node, bottom-up; kept nodes and their indices
    uint32_t U = N;
    unsigned long long unique_leaves = leaves;
    uint32_t* ins = nullptr;
    if (dag) {
        unsigned long long* d_unique = nullptr;
        XD_TRY(alloc((void**)&d_unique, 8), "octree build: dag");
        XD_TRY(hipMemsetAsync(d_unique, 0, 8, s), "octree build: dag");
        for (int L = (int)depth; L >= 0; --L) {
            Level& cur = lv[L];
            const uint64_t n = cur.count;
            if (n > 0x7FFFFFFFull) {
                cleanup();
                return fail(XD_ERR_FORMAT, "octree level of %llu nodes is too large for the device DAG build",
                            (unsigned long long)n);
            }
            const uint32_t* cls_child = L < (int)depth ? lv[L + 1].cls : nullptr;
            uint64_t *k0 = nullptr, *k1 = nullptr;
            uint32_t *i0 = nullptr, *i1 = nullptr, *head = nullptr, *run = nullptr;
            XD_TRY(alloc((void**)&cur.cls, n * 4), "octree build: dag");
            XD_TRY(alloc((void**)&k0, n * 8), "octree build: dag");
            XD_TRY(alloc((void**)&k1, n * 8), "octree build: dag");
            XD_TRY(alloc((void**)&i0, n * 4), "octree build: dag");
            XD_TRY(alloc((void**)&i1, n * 4), "octree build: dag");
            XD_TRY(alloc((void**)&head, n * 4), "octree build: dag");
            XD_TRY(alloc((void**)&run, n * 4), "octree build: dag");
            hipLaunchKernelGGL(k_dag_hash, dim3(blocks_for(n)), dim3(256), 0, s, n, cur.internal, cur.rank, cur.colour,
                               cls_child, k0, i0);
            size_t need = 0;
            XD_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, need, k0, k1, i0, i1, (int)n, 0, 64, s),
                   "octree build: dag sort");
            if (need > tmp_bytes) {
                XD_TRY(alloc(&tmp, need), "octree build: dag sort");
                tmp_bytes = need;
            }
            XD_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, k0, k1, i0, i1, (int)n, 0, 64, s),
                   "octree build: dag sort");
            hipLaunchKernelGGL(k_run_heads, dim3(blocks_for(n)), dim3(256), 0, s, n, k1, head);
            need = 0;
            XD_TRY(hipcub::DeviceScan::InclusiveScan(nullptr, need, head, run, hipcub::Max(), (int)n, s),
                   "octree build: dag runs");
            if (need > tmp_bytes) {
                XD_TRY(alloc(&tmp, need), "octree build: dag runs");
                tmp_bytes = need;
            }
            XD_TRY(hipcub::DeviceScan::InclusiveScan(tmp, tmp_bytes, head, run, hipcub::Max(), (int)n, s),
                   "octree build: dag runs");
            hipLaunchKernelGGL(k_dag_sets, dim3(blocks_for(n)), dim3(256), 0, s, n, i1, run, cur.internal, cur.rank,
                               cur.colour, cls_child, cur.start, cur.size, cur.cls, d_unique);
            XD_TRY(hipGetLastError(), "octree build: dag");
        }
        uint32_t* kept = nullptr;
        XD_TRY(alloc((void**)&kept, (size_t)N * 4), "octree build: dag");
        XD_TRY(alloc((void**)&ins, (size_t)N * 4), "octree build: dag");
        for (uint32_t L = 0; L <= depth; ++L)
            hipLaunchKernelGGL(k_dag_mark, dim3(blocks_for(lv[L].count)), dim3(256), 0, s, lv[L].count, lv[L].start,
                               lv[L].size, lv[L].cls, kept);
        size_t need = 0;
        XD_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, need, kept, ins, (int)N, s), "octree build: dag scan");
        if (need > tmp_bytes) {
            XD_TRY(alloc(&tmp, need), "octree build: dag scan");
            tmp_bytes = need;
        }
        XD_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, kept, ins, (int)N, s), "octree build: dag scan");
        uint32_t last_ins = 0, last_kept = 0;
        XD_TRY(hipMemcpyAsync(&last_ins, ins + N - 1, 4, hipMemcpyDeviceToHost, s), "octree build: dag scan");
        XD_TRY(hipMemcpyAsync(&last_kept, kept + N - 1, 4, hipMemcpyDeviceToHost, s), "octree build: dag scan");
        XD_TRY(hipMemcpyAsync(&unique_leaves, d_unique, 8, hipMemcpyDeviceToHost, s), "octree build: dag scan");
        XD_TRY(hipStreamSynchronize(s), "octree build: dag scan");
        U = last_ins + last_kept;
    }

    // 4. nodes, then ropes
    uint32_t* d_nodes = nullptr;
    XD_TRY(alloc((void**)&d_nodes, (size_t)U * 40), "octree build: nodes");
    for (uint32_t L = 0; L <= depth; ++L) {
        const Level& cur = lv[L];
        if (dag)
            hipLaunchKernelGGL(k_emit_dag, dim3(blocks_for(cur.count)), dim3(256), 0, s, cur.count, L, U, cur.internal,
                               cur.rank, cur.start, cur.size, cur.colour, cur.cls, L < depth ? lv[L + 1].cls : nullptr,
                               ins, d_nodes);
        else
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
    uint32_t* host = static_cast<uint32_t*>(malloc((size_t)U * 40));
    if (!host) {
        cleanup();
        return fail(XD_ERR_OUT_OF_MEMORY, "out of host memory for %u nodes", U);
    }
    XD_TRY(hipMemcpyAsync(host, d_nodes, (size_t)U * 40, hipMemcpyDeviceToHost, s), "octree build: download");
    XD_TRY(hipStreamSynchronize(s), "octree build: download");
    cleanup();
#undef XD_TRY
    stats->total_leaves = leaves;
    stats->unique_leaves = unique_leaves;
    stats->total_nodes = N;
    stats->depth = depth;
    stats->side = side;
    stats->node_count = U;
    *out_nodes = host;
    clear_error();
    return XD_OK;
}

}  // extern "C"
