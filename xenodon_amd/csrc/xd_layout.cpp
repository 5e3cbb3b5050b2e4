// Device layout of an octree: tagged child words in sibling-grouped order.
//
// The .svo numbering (reverse post-order, OctreeConstruction.h:92-112)
// scatters the eight children of a node across its subtrees.  On upload the
// nodes are renumbered so that the children of every internal node take eight
// consecutive indices, sibling groups being allocated depth-first from the
// root (node 0 stays node 0, which the rope traversal uses as "no
// neighbour").  Rays that pass through one parent then read child words
// from the same two cache lines, and a subtree occupies one contiguous range.
// In a DAG a shared node keeps the index of its first visit.  The first
// kTopLevels levels are numbered breadth-first ahead of everything else, so
// their child words form one small block that the kernels stage in LDS.
// Finally the numbering is split, keeping order within each part: internal
// nodes, then leaves with colour, then leaves with zero RGB.
//
// child_w[i*8 + k] = new index of child k | kTagLeaf (child is a leaf)
//                    | kTagEmpty (child is a leaf with zero RGB);
// leaves' child words are their ropes, renumbered and tagged the same way.
// Trees above 2^30 nodes use 31-bit indices and no empty tag: an index at or
// past the first empty leaf's marks one.
// meta[i] = {color, is_leaf_depth} of node i.
// On the device the leaves' child words -- their ropes, which only svo-rope
// reads, one word per leaf it crosses -- are stored as six planes
// (to_rope_planes): rope k of leaf leaf_base + j at leaf_base*8 + k*n_leaves + j,
// so the ropes of consecutive leaves share cache lines (32 per line instead
// of the 4 leaves' 32-byte records of the grouped layout).
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "xd_runtime.hpp"

namespace xd {
namespace {

unsigned worker_count() { return std::max(1u, std::min(16u, std::thread::hardware_concurrency())); }

}  // namespace

bool octree_is_tree(const uint32_t* w, uint64_t n) {
    std::vector<std::atomic<uint64_t>> seen((n + 63) / 64);
    for (auto& x : seen) x.store(0, std::memory_order_relaxed);
    std::atomic<bool> ok{true};
    const unsigned T = worker_count();
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; ++t)
        th.emplace_back([&, t] {
            for (uint64_t i = t; i < n && ok.load(std::memory_order_relaxed); i += T) {
                const uint32_t* nd = w + i * kNodeWords;
                if (nd[9] & kLeafMask) continue;
                for (int k = 0; k < 8; ++k) {
                    const uint32_t c = nd[k];
                    const uint64_t bit = 1ull << (c & 63u);
                    if (c == 0 || (seen[c >> 6].fetch_or(bit, std::memory_order_relaxed) & bit)) {
                        ok.store(false, std::memory_order_relaxed);
                        break;
                    }
                }
            }
        });
    for (auto& x : th) x.join();
    return ok.load();
}

namespace {

// The sequential sibling-group numbering below the top levels, for a tree:
// the subtree of the i-th node of `roots` takes the contiguous range of ids
// after those of roots 0..i-1 (8 ids per internal node in it), so the
// subtrees are counted, then numbered, in parallel.  Returns the next id.
uint32_t number_subtrees(const uint32_t* w, const std::vector<uint32_t>& roots, uint32_t next, uint32_t* new_id) {
    const size_t m = roots.size();
    std::vector<uint64_t> size(m, 0);
    const unsigned T = worker_count();
    auto run = [&](auto&& body) {
        std::atomic<size_t> cursor{0};
        std::vector<std::thread> th;
        for (unsigned t = 0; t < T; ++t)
            th.emplace_back([&] {
                std::vector<uint32_t> stack;
                for (size_t i; (i = cursor.fetch_add(1)) < m;) body(i, stack);
            });
        for (auto& x : th) x.join();
    };
    run([&](size_t i, std::vector<uint32_t>& stack) {
        uint64_t count = 0;
        stack.assign(1, roots[i]);
        while (!stack.empty()) {
            const uint32_t* nd = w + (uint64_t)stack.back() * kNodeWords;
            stack.pop_back();
            if (nd[9] & kLeafMask) continue;
            count += 8;
            for (int k = 0; k < 8; ++k) stack.push_back(nd[k]);
        }
        size[i] = count;
    });
    std::vector<uint64_t> start(m + 1, next);
    for (size_t i = 0; i < m; ++i) start[i + 1] = start[i] + size[i];
    run([&](size_t i, std::vector<uint32_t>& stack) {
        uint64_t id = start[i];
        stack.assign(1, roots[i]);
        while (!stack.empty()) {
            const uint32_t* nd = w + (uint64_t)stack.back() * kNodeWords;
            stack.pop_back();
            if (nd[9] & kLeafMask) continue;
            for (int k = 0; k < 8; ++k) new_id[nd[k]] = (uint32_t)(id++);
            for (int k = 7; k >= 0; --k) stack.push_back(nd[k]);  // child 0's group first
        }
    });
    return (uint32_t)start[m];
}

// Stable three-way partition of the numbering: internal nodes (and the root)
// first, then leaves with colour, then leaves with zero RGB.  The child words
// that descents read (those of internal nodes) become one dense block, and a
// child word names an empty leaf exactly when its index is >= *empty_start.
// top[] becomes the count of internal nodes of the top levels (whose child
// words the rope kernel stages in LDS).
void partition_leaves(const uint32_t* w, uint64_t n, uint32_t* new_id, uint32_t top[kTopLevels],
                            uint32_t* empty_start, uint32_t* leaf_base) {
    // class 0: internal nodes (and the root), 1: leaves with colour, 2: empty leaves
    auto cls = [w](uint64_t i) -> int {
        const uint32_t* nd = w + i * kNodeWords;
        if (i == 0 || !(nd[9] & kLeafMask)) return 0;
        return (nd[8] & 0xFFFFFFu) ? 1 : 2;
    };
    std::vector<uint32_t> order(n);  // old index by current new index
    const unsigned T = worker_count();
    auto parallel = [T](uint64_t count, auto&& body) {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < T; ++t) th.emplace_back([&, t] { body(count * t / T, count * (t + 1) / T, t); });
        for (auto& x : th) x.join();
    };
    parallel(n, [&](uint64_t lo, uint64_t hi, unsigned) {
        for (uint64_t i = lo; i < hi; ++i) order[new_id[i]] = (uint32_t)i;
    });
    std::vector<uint64_t> cnt(3 * (T + 1), 0);  // per class, nodes before each slice
    parallel(n, [&](uint64_t lo, uint64_t hi, unsigned t) {
        uint64_t c[3] = {0, 0, 0};
        for (uint64_t k = lo; k < hi; ++k) ++c[cls(order[k])];
        for (int q = 0; q < 3; ++q) cnt[q * (T + 1) + t + 1] = c[q];
    });
    for (int q = 0; q < 3; ++q)
        for (unsigned t = 0; t < T; ++t) cnt[q * (T + 1) + t + 1] += cnt[q * (T + 1) + t];
    const uint64_t base[3] = {0, cnt[T], cnt[T] + cnt[(T + 1) + T]};
    parallel(n, [&](uint64_t lo, uint64_t hi, unsigned t) {
        uint64_t at[3];
        for (int q = 0; q < 3; ++q) at[q] = base[q] + cnt[q * (T + 1) + t];
        for (uint64_t k = lo; k < hi; ++k) {
            const uint32_t i = order[k];
            new_id[i] = (uint32_t)(at[cls(i)]++);
        }
    });
    for (uint32_t L = 0; L < kTopLevels; ++L) {
        uint32_t c = 0;
        for (uint32_t k = 0; k < top[L]; ++k) c += cls(order[k]) == 0 ? 1 : 0;
        top[L] = c;
    }
    *empty_start = (uint32_t)base[2];
    *leaf_base = (uint32_t)base[1];
}

}  // namespace

void build_tagged_layout(const uint32_t* w, uint64_t n, uint32_t* child_w, uint32_t* meta, uint32_t top[kTopLevels],
                         bool wide, bool tree, uint32_t* empty_start, uint32_t* leaf_base) {
    const uint32_t empty_tag = wide ? 0u : kTagEmpty;
    std::vector<uint32_t> new_id(n, 0xFFFFFFFFu);
    uint32_t next = 1;
    new_id[0] = 0;
    // levels 0..kTopLevels-1 in breadth-first order: the lowest indices
    std::vector<uint32_t> level{0};
    top[0] = 1;
    for (uint32_t L = 1; L < kTopLevels; ++L) {
        std::vector<uint32_t> nxt;
        for (uint32_t old : level) {
            const uint32_t* nd = w + (uint64_t)old * kNodeWords;
            if (nd[9] & kLeafMask) continue;
            for (int k = 0; k < 8; ++k) {
                const uint32_t c = nd[k];
                if (new_id[c] == 0xFFFFFFFFu) {
                    new_id[c] = next++;
                    nxt.push_back(c);
                }
            }
        }
        level.swap(nxt);
        top[L] = next;  // nodes of levels 0..L
    }
    // below: sibling groups, depth-first from each node of the last top level
    if (tree) {
        next = number_subtrees(w, level, next, new_id.data());
    } else {
        std::vector<uint32_t> stack;
        uint32_t fresh[8];
        for (auto it = level.rbegin(); it != level.rend(); ++it) stack.push_back(*it);
        while (!stack.empty()) {
            const uint32_t old = stack.back();
            stack.pop_back();
            const uint32_t* nd = w + (uint64_t)old * kNodeWords;
            if (nd[9] & kLeafMask) continue;
            int nf = 0;
            for (int k = 0; k < 8; ++k) {
                const uint32_t c = nd[k];
                if (new_id[c] == 0xFFFFFFFFu) {
                    new_id[c] = next++;
                    fresh[nf++] = c;
                }
            }
            // visit child 0's group first
            for (int k = nf - 1; k >= 0; --k) stack.push_back(fresh[k]);
        }
    }
    // nodes no internal node points to (never traversed) go last
    for (uint64_t i = 0; i < n; ++i)
        if (new_id[i] == 0xFFFFFFFFu) new_id[i] = next++;
    // internal nodes, then leaves with colour, then empty leaves: with wide
    // tags the empty leaves' indices mark them; with narrow tags the tag does,
    // and the order still packs the descents' child words together
    // (svo-naive +4 %, esvo +2 % on the 512x361x512 tree)
    partition_leaves(w, n, new_id.data(), top, empty_start, leaf_base);
    if (!wide) *empty_start = 0xFFFFFFFFu;

    const unsigned T = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; ++t)
        th.emplace_back([&, t] {
            for (uint64_t i = t; i < n; i += T) {
                const uint32_t* nd = w + i * kNodeWords;
                const uint32_t ni = new_id[i];
                for (int k = 0; k < 8; ++k) {
                    const uint32_t c = nd[k];
                    const uint32_t* cn = w + (uint64_t)c * kNodeWords;
                    uint32_t tag = 0;
                    if (cn[9] & kLeafMask) tag = kTagLeaf | ((cn[8] & 0xFFFFFFu) ? 0u : empty_tag);
                    child_w[(uint64_t)ni * 8 + k] = new_id[c] | tag;
                }
                meta[(uint64_t)ni * 2] = nd[8];
                meta[(uint64_t)ni * 2 + 1] = nd[9];
            }
        });
    for (auto& x : th) x.join();
}

namespace {
void fill_table(const uint32_t* child_w, uint32_t mask, uint32_t word, uint32_t level, uint32_t x0, uint32_t y0,
                uint32_t z0, uint2* table) {
    const uint32_t n = 1u << (kTableLevel - level);  // cells per side below this node
    if ((word & kTagLeaf) || level == kTableLevel) {
        for (uint32_t z = z0; z < z0 + n; ++z)
            for (uint32_t y = y0; y < y0 + n; ++y)
                for (uint32_t x = x0; x < x0 + n; ++x)
                    table[((uint64_t)z << (2 * kTableLevel)) | (y << kTableLevel) | x] = make_uint2(word, level);
        return;
    }
    const uint64_t node = word & mask;
    const uint32_t h = n / 2;
    for (uint32_t k = 0; k < 8; ++k)  // child k = 4x + 2y + z
        fill_table(child_w, mask, child_w[node * 8 + k], level + 1, x0 + ((k >> 2) & 1u) * h, y0 + ((k >> 1) & 1u) * h,
                   z0 + (k & 1u) * h, table);
}
}  // namespace

bool pack_rope_depths(uint32_t* child_w, const uint32_t* meta, uint64_t n) {
    if (n > (uint64_t)kRopeIndexMask + 1) return false;
    for (uint64_t i = 0; i < n; ++i)
        if ((meta[i * 2 + 1] & kDepthMask) > 31) return false;
    for (uint64_t i = 0; i < n; ++i) {
        if (!(meta[i * 2 + 1] & kLeafMask)) continue;
        for (int k = 0; k < 8; ++k) {
            const uint64_t c = child_w[i * 8 + k] & kIndexMask;
            child_w[i * 8 + k] |= (meta[c * 2 + 1] & kDepthMask) << kRopeDepthShift;
        }
    }
    return true;
}

uint64_t rope_plane_words(uint64_t n, uint32_t leaf_base) { return (uint64_t)leaf_base * 8 + 6 * (n - leaf_base); }

void to_rope_planes(const uint32_t* child_w, uint64_t n, uint32_t leaf_base, uint32_t* out) {
    const uint64_t nl = n - leaf_base, base = (uint64_t)leaf_base * 8;
    std::memcpy(out, child_w, base * sizeof(uint32_t));
    const unsigned T = worker_count();
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; ++t)
        th.emplace_back([&, t] {
            for (uint64_t j = nl * t / T; j < nl * (t + 1) / T; ++j)
                for (uint64_t k = 0; k < 6; ++k) out[base + k * nl + j] = child_w[(leaf_base + j) * 8 + k];
        });
    for (auto& x : th) x.join();
}

void build_start_table(const uint32_t* child_w, uint32_t root_tag, uint32_t index_mask, uint2* table) {
    fill_table(child_w, index_mask, root_tag, 0, 0, 0, 0, table);
}

bool octree_depths_exact(const uint32_t* w, uint64_t n, bool tree) {
    if (tree) {
        // one parent per node: a reachable node's depth is its parent's + 1,
        // so checking every internal node against its internal children (and
        // the root against 0) is the walk's check, done in parallel; an
        // unreachable internal node can only make the answer false
        if (!(w[9] & kLeafMask) && (w[9] & kDepthMask) != 0) return false;
        std::atomic<bool> ok{true};
        const unsigned T = worker_count();
        std::vector<std::thread> th;
        for (unsigned t = 0; t < T; ++t)
            th.emplace_back([&, t] {
                for (uint64_t i = t; i < n && ok.load(std::memory_order_relaxed); i += T) {
                    const uint32_t* nd = w + i * kNodeWords;
                    if (nd[9] & kLeafMask) continue;
                    const uint32_t d = nd[9] & kDepthMask;
                    for (int k = 0; k < 8; ++k) {
                        const uint32_t* cn = w + (uint64_t)nd[k] * kNodeWords;
                        if (!(cn[9] & kLeafMask) && (cn[9] & kDepthMask) != d + 1) {
                            ok.store(false, std::memory_order_relaxed);
                            break;
                        }
                    }
                }
            });
        for (auto& x : th) x.join();
        return ok.load();
    }
    std::vector<uint8_t> depth(n, 0xFF);
    std::vector<uint32_t> level{0}, next;
    depth[0] = 0;
    for (uint32_t d = 0; !level.empty(); ++d) {
        next.clear();
        for (uint32_t i : level) {
            const uint32_t* nd = w + (uint64_t)i * kNodeWords;
            if (nd[9] & kLeafMask) continue;
            if ((nd[9] & kDepthMask) != d) return false;
            for (int k = 0; k < 8; ++k) {
                const uint32_t c = nd[k];
                if (depth[c] == 0xFF) {
                    depth[c] = (uint8_t)std::min<uint32_t>(d + 1, 0xFE);
                    next.push_back(c);
                } else if (depth[c] != d + 1) {
                    return false;
                }
            }
        }
        level.swap(next);
    }
    return true;
}

}  // namespace xd
