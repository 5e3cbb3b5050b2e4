// Octree construction (`xenodon convert`) on the host, multi-threaded.
//
// Produces the node buffer of build_octree (src/model/OctreeConstruction.h:
// 13-237) bit for bit:
//   * side = max over axes of ceil_pow2(dim)                         (:197-211)
//   * a node whose offset is outside the grid on any axis is the empty leaf
//     {children 0, colour 0, LEAF|depth}                             (:141-154)
//   * otherwise its region (clipped to the grid) is scanned; it becomes a leaf
//     if (!split && the region lies fully inside) || extent == 1      (:156-171)
//   * the colour is the per-channel integer mean (Grid.cpp:86-131); chan-diff
//     splits when max over channels of (max - min) > N (Grid.cpp:102-135);
//     std-dev splits when sqrt(sum_voxels sum_channels (x-avg)^2 / n) > S,
//     accumulated in FP64 in z, y, x order (Grid.cpp:139-214)
//   * nodes are emitted in post-order, children x*4+y*2+z, then reversed with
//     internal children remapped to end - child and leaf children zeroed
//                                                                    (:92-112)
//   * --dag keeps the first node (in post-order) of every set of equal nodes,
//     equality over all 40 bytes                                     (:19-30)
//   * --rope: per leaf, the node found at depth <= the leaf's that holds the
//     face-neighbour point, or 0 (Octree.cpp:116-153, 181-201)
//
// Parallelism: the eight subtrees under each of the top two levels are built
// concurrently into private post-order lists that are concatenated in child
// order, which is the reference's sequential order.  DAG deduplication runs
// as one sequential pass over the final post-order list, which assigns the
// same first-occurrence indices as the reference's insert-time cache.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <future>
#include <thread>
#include <unordered_map>
#include <vector>

#include "xenodon.h"
#include "xd_runtime.hpp"

namespace {

constexpr uint32_t LEAF = 0x80000000u;

struct Node {
    uint32_t w[10];
};

struct NodeHash {
    size_t operator()(const Node& n) const {
        uint64_t h = 1469598103934665603ull;
        for (int i = 0; i < 10; ++i) {
            h ^= n.w[i];
            h *= 1099511628211ull;
            h ^= h >> 29;
        }
        return (size_t)h;
    }
};
struct NodeEq {
    bool operator()(const Node& a, const Node& b) const { return memcmp(a.w, b.w, sizeof(a.w)) == 0; }
};

struct Stats {
    uint64_t total_leaves = 0, total_nodes = 0, depth = 0;
    void add(const Stats& o) {
        total_leaves += o.total_leaves;
        total_nodes += o.total_nodes;
        depth = std::max(depth, o.depth);
    }
};

struct Grid {
    const uint8_t* data;
    uint64_t X, Y, Z;
};

struct Scan {
    uint32_t avg;   // packed RGBA8 mean
    bool split;
};

// Grid::vol_scan over [lo, hi) clipped to the grid; rows are scanned with
// byte-wise min/max and 64-bit sums (order-independent integer arithmetic).
void chan_scan_rows(const Grid& g, const uint64_t lo[3], const uint64_t hi[3], uint64_t z0, uint64_t z1,
                    uint64_t sum[4], uint8_t mn[4], uint8_t mx[4]) {
    for (uint64_t z = z0; z < z1; ++z)
        for (uint64_t y = lo[1]; y < hi[1]; ++y) {
            const uint8_t* p = g.data + ((z * g.Y + y) * g.X + lo[0]) * 4;
            const uint64_t n = hi[0] - lo[0];
            uint32_t s[4] = {0, 0, 0, 0};
            uint8_t a[4] = {mn[0], mn[1], mn[2], mn[3]}, b[4] = {mx[0], mx[1], mx[2], mx[3]};
            for (uint64_t i = 0; i < n; ++i) {
                for (int c = 0; c < 4; ++c) {
                    uint8_t v = p[i * 4 + c];
                    s[c] += v;
                    a[c] = v < a[c] ? v : a[c];
                    b[c] = v > b[c] ? v : b[c];
                }
            }
            for (int c = 0; c < 4; ++c) {
                sum[c] += s[c];
                mn[c] = a[c];
                mx[c] = b[c];
            }
        }
}

Scan chan_diff_scan(const Grid& g, const uint64_t off[3], uint64_t extent, int chan_diff, bool parallel) {
    uint64_t lo[3] = {std::min(off[0], g.X), std::min(off[1], g.Y), std::min(off[2], g.Z)};
    uint64_t hi[3] = {std::min(off[0] + extent, g.X), std::min(off[1] + extent, g.Y), std::min(off[2] + extent, g.Z)};
    const uint64_t n = (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
    if (n == 0) return {0u, false};
    uint64_t sum[4] = {0, 0, 0, 0};
    uint8_t mn[4] = {255, 255, 255, 255}, mx[4] = {0, 0, 0, 0};
    const uint64_t nz = hi[2] - lo[2];
    if (parallel && n > (1ull << 22) && nz >= 8) {
        const unsigned T = std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
        std::vector<std::future<void>> fs;
        std::vector<std::array<uint64_t, 4>> S(T);
        std::vector<std::array<uint8_t, 8>> M(T);
        for (unsigned t = 0; t < T; ++t) {
            fs.push_back(std::async(std::launch::async, [&, t] {
                uint64_t s[4] = {0, 0, 0, 0};
                uint8_t a[4] = {255, 255, 255, 255}, b[4] = {0, 0, 0, 0};
                uint64_t z0 = lo[2] + nz * t / T, z1 = lo[2] + nz * (t + 1) / T;
                chan_scan_rows(g, lo, hi, z0, z1, s, a, b);
                for (int c = 0; c < 4; ++c) {
                    S[t][c] = s[c];
                    M[t][c] = a[c];
                    M[t][4 + c] = b[c];
                }
            }));
        }
        for (auto& f : fs) f.get();
        for (unsigned t = 0; t < T; ++t)
            for (int c = 0; c < 4; ++c) {
                sum[c] += S[t][c];
                mn[c] = std::min(mn[c], M[t][c]);
                mx[c] = std::max(mx[c], M[t][4 + c]);
            }
    } else {
        chan_scan_rows(g, lo, hi, lo[2], hi[2], sum, mn, mx);
    }
    uint32_t avg = 0, md = 0;
    for (int c = 0; c < 4; ++c) {
        avg |= (uint32_t)(uint8_t)(sum[c] / n) << (8 * c);
        md = std::max<uint32_t>(md, (uint32_t)(mx[c] - mn[c]));
    }
    return {avg, md > (uint32_t)chan_diff};
}

// Grid::stddev_scan: FP64 accumulation in the reference's z, y, x order.
Scan std_dev_scan(const Grid& g, const uint64_t off[3], uint64_t extent, double threshold) {
    uint64_t lo[3] = {std::min(off[0], g.X), std::min(off[1], g.Y), std::min(off[2], g.Z)};
    uint64_t hi[3] = {std::min(off[0] + extent, g.X), std::min(off[1] + extent, g.Y), std::min(off[2] + extent, g.Z)};
    const uint64_t n = (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
    if (n == 0) return {0u, 0.0 > threshold};
    uint64_t acc[4] = {0, 0, 0, 0};
    for (uint64_t z = lo[2]; z < hi[2]; ++z)
        for (uint64_t y = lo[1]; y < hi[1]; ++y)
            for (uint64_t x = lo[0]; x < hi[0]; ++x) {
                const uint8_t* p = g.data + ((z * g.Y + y) * g.X + x) * 4;
                for (int c = 0; c < 4; ++c) acc[c] += p[c];
            }
    const double nd = (double)n;
    double avg[4];
    for (int c = 0; c < 4; ++c) avg[c] = (double)acc[c] / nd;
    double s = 0;
    for (uint64_t z = lo[2]; z < hi[2]; ++z)
        for (uint64_t y = lo[1]; y < hi[1]; ++y)
            for (uint64_t x = lo[0]; x < hi[0]; ++x) {
                const uint8_t* p = g.data + ((z * g.Y + y) * g.X + x) * 4;
                double dr = (double)p[0] - avg[0], dg = (double)p[1] - avg[1];
                double db = (double)p[2] - avg[2], da = (double)p[3] - avg[3];
                s += dr * dr + dg * dg + db * db + da * da;
            }
    uint32_t a = 0;
    for (int c = 0; c < 4; ++c) a |= (uint32_t)(uint8_t)avg[c] << (8 * c);
    return {a, std::sqrt(s / nd) > threshold};
}

struct Builder {
    Grid g;
    int heuristic;
    int chan_diff;
    double std_dev;

    Scan scan(const uint64_t off[3], uint64_t extent, bool top) const {
        return heuristic == 0 ? chan_diff_scan(g, off, extent, chan_diff, top) : std_dev_scan(g, off, extent, std_dev);
    }

    // Appends the subtree's nodes to `out` in post-order; children hold
    // indices into `out`.  Returns the subtree root's index.
    uint32_t construct(std::vector<Node>& out, Stats& st, const uint64_t off[3], uint64_t extent, uint64_t depth,
                       bool parallel) const {
        st.depth = std::max(st.depth, depth);
        Node node;
        memset(&node, 0, sizeof(node));
        const bool inside_any = off[0] < g.X && off[1] < g.Y && off[2] < g.Z;
        if (!inside_any) {
            node.w[9] = LEAF | (uint32_t)depth;
            ++st.total_nodes;
            ++st.total_leaves;
            out.push_back(node);
            return (uint32_t)(out.size() - 1);
        }
        const bool fully_inside = off[0] + extent <= g.X && off[1] + extent <= g.Y && off[2] + extent <= g.Z;
        const Scan sc = scan(off, extent, parallel);
        if ((!sc.split && fully_inside) || extent == 1) {
            node.w[8] = sc.avg;
            node.w[9] = LEAF | (uint32_t)depth;
            ++st.total_nodes;
            ++st.total_leaves;
            out.push_back(node);
            return (uint32_t)(out.size() - 1);
        }
        const uint64_t h = extent / 2;
        node.w[8] = sc.avg;
        node.w[9] = (uint32_t)depth;
        uint64_t coff[8][3];
        for (int c = 0; c < 8; ++c) {
            coff[c][0] = off[0] + ((c & 4) ? h : 0);
            coff[c][1] = off[1] + ((c & 2) ? h : 0);
            coff[c][2] = off[2] + ((c & 1) ? h : 0);
        }
        if (parallel && depth < 2 && h >= 16) {
            std::vector<Node> sub[8];
            Stats sst[8];
            uint32_t root[8];
            std::vector<std::future<void>> fs;
            for (int c = 0; c < 8; ++c)
                fs.push_back(std::async(std::launch::async, [&, c] {
                    root[c] = construct(sub[c], sst[c], coff[c], h, depth + 1, depth + 1 < 2);
                }));
            for (auto& f : fs) f.get();
            for (int c = 0; c < 8; ++c) {
                const uint32_t base = (uint32_t)out.size();
                for (auto nd : sub[c]) {
                    if (!(nd.w[9] & LEAF))
                        for (int k = 0; k < 8; ++k) nd.w[k] += base;
                    out.push_back(nd);
                }
                node.w[c] = base + root[c];
                st.add(sst[c]);
                std::vector<Node>().swap(sub[c]);
            }
        } else {
            for (int c = 0; c < 8; ++c) node.w[c] = construct(out, st, coff[c], h, depth + 1, false);
        }
        ++st.total_nodes;
        out.push_back(node);
        return (uint32_t)(out.size() - 1);
    }
};

uint64_t ceil_2pow(uint64_t x) {
    --x;
    x |= x >> 1;
    x |= x >> 2;
    x |= x >> 4;
    x |= x >> 8;
    x |= x >> 16;
    x |= x >> 32;
    return ++x;
}

// Octree::find (Octree.cpp:116-153); UINT64_MAX for the reference's nullptr
uint64_t find(const uint32_t* nodes, uint64_t side, const uint64_t pos[3], uint64_t max_depth) {
    uint64_t extent = side;
    if (pos[0] >= extent || pos[1] >= extent || pos[2] >= extent) return UINT64_MAX;
    uint64_t index = 0, off[3] = {0, 0, 0};
    for (;;) {
        extent /= 2;
        if ((nodes[index * 10 + 9] & LEAF) || extent == 0 || max_depth == 0) return index;
        uint32_t ci = 0;
        for (int a = 0; a < 3; ++a)
            if (pos[a] >= off[a] + extent) {
                ci |= 4u >> a;
                off[a] += extent;
            }
        index = nodes[index * 10 + ci];
        --max_depth;
    }
}

struct LeafRef {
    uint64_t index, pos[3], extent, depth;
};

void collect_leaves(const uint32_t* nodes, uint64_t index, const uint64_t pos[3], uint64_t extent, uint64_t depth,
                    std::vector<LeafRef>& out) {
    const uint32_t* nd = nodes + index * 10;
    if (nd[9] & LEAF) {
        out.push_back({index, {pos[0], pos[1], pos[2]}, extent, depth});
        return;
    }
    const uint64_t h = extent / 2;
    for (int c = 0; c < 8; ++c) {
        uint64_t cp[3] = {pos[0] + ((c & 4) ? h : 0), pos[1] + ((c & 2) ? h : 0), pos[2] + ((c & 1) ? h : 0)};
        collect_leaves(nodes, nd[c], cp, h, depth + 1, out);
    }
}

void generate_ropes(uint32_t* nodes, uint64_t side) {
    std::vector<LeafRef> leaves;
    const uint64_t zero[3] = {0, 0, 0};
    collect_leaves(nodes, 0, zero, side, 0, leaves);
    const unsigned T = std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    // Ropes only rewrite leaf children, which find() never reads, so leaves
    // can be processed in any order.
    for (unsigned t = 0; t < T; ++t)
        th.emplace_back([&, t] {
            for (size_t i = t; i < leaves.size(); i += T) {
                const LeafRef& L = leaves[i];
                uint32_t r[6];
                for (int a = 0; a < 3; ++a) {
                    uint64_t p[3] = {L.pos[0], L.pos[1], L.pos[2]};
                    p[a] = L.pos[a] + L.extent;
                    uint64_t f = find(nodes, side, p, L.depth);
                    r[2 * a] = f == UINT64_MAX ? 0u : (uint32_t)f;
                    p[a] = L.pos[a] - L.extent;  // wraps below zero like size_t
                    f = find(nodes, side, p, L.depth);
                    r[2 * a + 1] = f == UINT64_MAX ? 0u : (uint32_t)f;
                }
                memcpy(nodes + L.index * 10, r, sizeof(r));
            }
        });
    for (auto& x : th) x.join();
}

}  // namespace

extern "C" {

int xd_build_octree(const void* rgba8, uint64_t dim_x, uint64_t dim_y, uint64_t dim_z, int heuristic, int chan_diff,
                    double std_dev, int tree_type, uint32_t** out_nodes, xd_build_stats* stats) {
    using xd::fail;
    if (!rgba8 || !out_nodes || !stats) return fail(XD_ERR_INVALID_ARGUMENT, "null argument");
    *out_nodes = nullptr;
    if (dim_x == 0 || dim_y == 0 || dim_z == 0) return fail(XD_ERR_INVALID_ARGUMENT, "grid has a zero dimension");
    if (heuristic != XD_HEURISTIC_CHAN_DIFF && heuristic != XD_HEURISTIC_STD_DEV)
        return fail(XD_ERR_INVALID_ARGUMENT, "unknown heuristic %d", heuristic);
    if (tree_type < XD_TREE_SPARSE || tree_type > XD_TREE_ROPE)
        return fail(XD_ERR_INVALID_ARGUMENT, "unknown tree type %d", tree_type);
    Builder b{{static_cast<const uint8_t*>(rgba8), dim_x, dim_y, dim_z}, heuristic, chan_diff, std_dev};
    const uint64_t side = std::max({ceil_2pow(dim_x), ceil_2pow(dim_y), ceil_2pow(dim_z)});
    std::vector<Node> post;
    Stats st;
    const uint64_t zero[3] = {0, 0, 0};
    try {
        b.construct(post, st, zero, side, 0, true);
    } catch (const std::bad_alloc&) {
        return fail(XD_ERR_OUT_OF_MEMORY, "out of host memory while building the octree");
    }
    if (post.size() > 0xFFFFFFFFull) return fail(XD_ERR_FORMAT, "octree exceeds 2^32 nodes");
    uint64_t unique_leaves = st.total_leaves;
    if (tree_type == XD_TREE_DAG) {
        // first-occurrence dedup in post-order (HashCache, :19-30)
        std::unordered_map<Node, uint32_t, NodeHash, NodeEq> cache;
        cache.reserve(post.size());
        std::vector<uint32_t> canon(post.size());
        std::vector<Node> uniq;
        uniq.reserve(post.size() / 2 + 1);
        unique_leaves = 0;
        for (size_t i = 0; i < post.size(); ++i) {
            Node nd = post[i];
            if (!(nd.w[9] & LEAF))
                for (int k = 0; k < 8; ++k) nd.w[k] = canon[nd.w[k]];
            auto it = cache.find(nd);
            if (it != cache.end()) {
                canon[i] = it->second;
            } else {
                canon[i] = (uint32_t)uniq.size();
                cache.emplace(nd, canon[i]);
                uniq.push_back(nd);
                if (nd.w[9] & LEAF) ++unique_leaves;
            }
        }
        post.swap(uniq);
    }
    const uint64_t n = post.size();
    uint32_t* nodes = static_cast<uint32_t*>(malloc(n * 40));
    if (!nodes) return fail(XD_ERR_OUT_OF_MEMORY, "out of host memory for %llu nodes", (unsigned long long)n);
    // OctreeBuilder::build: reverse, internal children -> end - child, leaf children -> 0
    const uint32_t end = (uint32_t)(n - 1);
    for (uint64_t i = 0; i < n; ++i) {
        const Node& src = post[n - 1 - i];
        uint32_t* dst = nodes + i * 10;
        const bool leaf = (src.w[9] & LEAF) != 0;
        for (int k = 0; k < 8; ++k) dst[k] = leaf ? 0u : end - src.w[k];
        dst[8] = src.w[8];
        dst[9] = src.w[9];
    }
    std::vector<Node>().swap(post);
    if (tree_type == XD_TREE_ROPE) generate_ropes(nodes, side);
    stats->total_leaves = st.total_leaves;
    stats->unique_leaves = unique_leaves;
    stats->total_nodes = st.total_nodes;
    stats->depth = st.depth;
    stats->side = side;
    stats->node_count = n;
    *out_nodes = nodes;
    xd::clear_error();
    return XD_OK;
}

void xd_free(void* p) { free(p); }

}  // extern "C"
